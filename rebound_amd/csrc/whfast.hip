// WHFast in democratic-heliocentric coordinates on the GPU, for one
// simulation (the "whfast" integrator plugin) and for batched ensembles of
// independent planetary systems (ensemble.hip).
//
// Reference: integrator_whfast.c
//   Stumpff / Stiefel functions      :189-235
//   Kepler solver                    :241-427
//   DH interaction / jump / kepler   :494-501, :537-554, :612-617, com :638-651
//   init / step / synchronize        :876-947, :1113-1259, :1051-1111
// and transformations.c:388-477 (inertial <-> democratic heliocentric).
//
// Execution model: one wavefront (64 lanes) integrates 64/G systems; each
// system owns a group of G lanes (G = 2, 4, ..., 64, the smallest power of two
// >= its planet count) and lane k of the group owns planet k+1.  The whole
// state (heliocentric and inertial copies, the star and the barycentre) stays
// in registers for all steps of a launch.  Every sum of the reference that
// runs over bodies in index order (centre of mass, total momentum, the
// planet-planet forces, ...) is evaluated by every lane of the group in the
// same index order from values published in LDS, so each lane holds the same
// bits and the result equals the reference's sequential loop.  The library is
// built without FP contraction; sqrt and division are IEEE.
#include <cmath>
#include <cstring>

#include "internal.h"
#include "whfast_device.h"

namespace rbamd {

// 1/k! as the reference spells them (integrator_whfast.c:158); 0..13 are used
// by the solver, 0..15 by the variational drift.
__constant__ double kInvFact[16] = {1.,
                                    1.,
                                    1. / 2.,
                                    1. / 6.,
                                    1. / 24.,
                                    1. / 120.,
                                    1. / 720.,
                                    1. / 5040.,
                                    1. / 40320.,
                                    1. / 362880.,
                                    1. / 3628800.,
                                    1. / 39916800.,
                                    1. / 479001600.,
                                    1. / 6227020800.,
                                    1. / 87178291200.,
                                    1. / 1307674368000.};

__device__ __forceinline__ double fastabs(double x) { return (x > 0.) ? x : -x; }

// c_0..c_3 Stumpff functions of z by series + quartering (integrator_whfast.c:189-213).
__device__ __forceinline__ void stumpff_cs3(double* cs, double z) {
    // A finite z needs at most ~520 quarterings.  An infinite z never comes
    // below 0.1 (the reference's loop would not end): the functions are
    // returned as NaN at once.
    if (isinf(z)) {
        cs[0] = cs[1] = cs[2] = cs[3] = __builtin_nan("");
        return;
    }
    unsigned int n = 0;
    while (fabs(z) > 0.1) {
        z = z / 4.;
        n++;
    }
    double c_odd = kInvFact[13];
    double c_even = kInvFact[12];
#pragma unroll
    for (int np = 11; np >= 3; np -= 2) {
        c_odd = kInvFact[np] - z * c_odd;
        c_even = kInvFact[np - 1] - z * c_even;
    }
    cs[3] = c_odd;
    cs[2] = c_even;
    cs[1] = kInvFact[1] - z * c_odd;
    cs[0] = kInvFact[0] - z * c_even;
    for (; n > 0; n--) {
        cs[3] = (cs[2] + cs[0] * cs[3]) * 0.25;
        cs[2] = cs[1] * cs[1] * 0.5;
        cs[1] = cs[0] * cs[1];
        cs[0] = 2. * cs[0] * cs[0] - 1.;
    }
}

__device__ __forceinline__ void stiefel_Gs3(double* Gs, double beta, double X) {
    const double X2 = X * X;
    stumpff_cs3(Gs, beta * X2);
    Gs[1] *= X;
    Gs[2] *= X2;
    Gs[3] *= X2 * X;
}

// c_0..c_5 (integrator_whfast.c:161-187), for the variational drift.
__device__ __forceinline__ void stumpff_cs(double* cs, double z) {
    if (isinf(z)) {
        cs[0] = cs[1] = cs[2] = cs[3] = cs[4] = cs[5] = __builtin_nan("");
        return;
    }
    unsigned int n = 0;
    while (fastabs(z) > 0.1) {
        z = z / 4.;
        n++;
    }
    double c_odd = kInvFact[15];
    double c_even = kInvFact[14];
#pragma unroll
    for (int np = 13; np >= 5; np -= 2) {
        c_odd = kInvFact[np] - z * c_odd;
        c_even = kInvFact[np - 1] - z * c_even;
    }
    cs[5] = c_odd;
    cs[4] = c_even;
    cs[3] = kInvFact[3] - z * cs[5];
    cs[2] = kInvFact[2] - z * cs[4];
    cs[1] = kInvFact[1] - z * cs[3];
    for (; n > 0; n--) {
        z = z * 4.;
        cs[5] = (cs[5] + cs[4] + cs[3] * cs[2]) * 0.0625;
        cs[4] = (1. + cs[1]) * cs[3] * 0.125;
        cs[3] = 1. / 6. - z * cs[5];
        cs[2] = 0.5 - z * cs[4];
        cs[1] = 1. - z * cs[3];
    }
    cs[0] = kInvFact[0] - z * cs[2];
}

__device__ __forceinline__ void stiefel_Gs(double* Gs, double beta, double X) {
    const double X2 = X * X;
    stumpff_cs(Gs, beta * X2);
    Gs[1] *= X;
    Gs[2] *= X2;
    double pw = X2 * X;
    Gs[3] *= pw;
    pw *= X;
    Gs[4] *= pw;
    pw *= X;
    Gs[5] *= pw;
}

// What the variational drift of a body needs from its Kepler solve: the
// state before the drift, the solver's scalars and the drift coefficients.
struct KepOut {
    double x0, y0, z0, vx0, vy0, vz0;
    double r0, r0i, ri, beta, eta0, zeta0, X, mu;
    double f, g, fd, gd;
};

// One quartic (Laguerre-Conway) update of X (integrator_whfast.c:292-300).
// Gs receives the Stiefel functions of the X the update started from.
__device__ __forceinline__ double quartic_next(double X, double beta, double r0, double eta0, double zeta0,
                                               double dt, double* Gs) {
    stiefel_Gs3(Gs, beta, X);
    const double f = r0 * X + eta0 * Gs[2] + zeta0 * Gs[3] - dt;
    const double fp = r0 + eta0 * Gs[1] + zeta0 * Gs[2];
    const double fpp = eta0 * Gs[0] + zeta0 * Gs[1];
    const double denom = fp + sqrt(fabs(16. * fp * fp - 20. * f * fpp));
    return (X * denom - 5. * f) / denom;
}

// Universal-variable Kepler drift of one body (integrator_whfast.c:241-390),
// in place.  Returns a bit mask: 1 = the step exceeded one orbital period
// (the reference warns once), 4 = the quartic iteration ran, 8 = the
// bisection fallback ran.
__device__ __forceinline__ int kepler_solver(double& px, double& py, double& pz, double& pvx, double& pvy, double& pvz, double mu,
                             double dt, KepOut* out = nullptr) {
    const double x0 = px, y0 = py, z0 = pz, vx0 = pvx, vy0 = pvy, vz0 = pvz;
    int flags = 0;
    const double r0 = sqrt(x0 * x0 + y0 * y0 + z0 * z0);
    const double r0i = 1. / r0;
    const double v2 = vx0 * vx0 + vy0 * vy0 + vz0 * vz0;
    const double beta = 2. * mu * r0i - v2;
    const double eta0 = x0 * vx0 + y0 * vy0 + z0 * vz0;
    const double zeta0 = mu - beta * r0;
    double X;
    double Gs[4];
    double invperiod = 0;
    double X_per_period = __builtin_nan("");
    if (beta > 0.) {
        const double sqrt_beta = sqrt(beta);
        invperiod = sqrt_beta * beta / (2. * M_PI * mu);
        X_per_period = 2. * M_PI / sqrt_beta;
        if (fabs(dt) * invperiod > 1.) flags |= 1;
        const double dtr0i = dt * r0i;
        X = dtr0i * (1. - dtr0i * eta0 * 0.5 * r0i);
    } else {
        X = 0.;
    }
    unsigned int converged = 0;
    double oldX = X;
    stiefel_Gs3(Gs, beta, X);
    const double e0 = eta0 * Gs[1] + zeta0 * Gs[2];
    double ri = 1. / (r0 + e0);
    X = ri * (X * e0 - eta0 * Gs[2] - zeta0 * Gs[3] + dt);

    if (fastabs(X - oldX) > 0.01 * X_per_period) {
        // Quartic (Laguerre-Conway) iteration with cycle detection (:289-312):
        // converged at the first iterate equal to any earlier one.  The last
        // eight iterates are compared in registers (NaN: empty, equal to
        // nothing); an iterate leaving that window is stored in `older`,
        // which is read only when the loop runs past nine iterations.  Short
        // loops (most of them) touch no memory.
        flags |= 4;
        X = beta * dt / mu;
        double older[56];      // older[k] = X_k once it has left the window
        double w0 = __builtin_nan(""), w1 = w0, w2 = w0, w3 = w0, w4 = w0, w5 = w0, w6 = w0, w7 = w0;
        for (int n_lag = 1; n_lag < 64; n_lag++) {
            // X_{n_lag} from X_{n_lag-1}; w0..w7 hold X_{n_lag-1}..X_{n_lag-8}
            X = quartic_next(X, beta, r0, eta0, zeta0, dt, Gs);
            if (X == w0 || X == w1 || X == w2 || X == w3 || X == w4 || X == w5 || X == w6 || X == w7) {
                converged = 1;
                break;
            }
            for (int k = 1; k < n_lag - 8; k++) {
                if (X == older[k]) {
                    converged = 1;
                    break;
                }
            }
            if (converged) break;
            if (n_lag > 8) older[n_lag - 8] = w7;
            w7 = w6; w6 = w5; w5 = w4; w4 = w3; w3 = w2; w2 = w1; w1 = w0; w0 = X;
        }
        const double e1 = eta0 * Gs[1] + zeta0 * Gs[2];
        ri = 1. / (r0 + e1);
    } else {
        // Newton's method (:313-330).
        double oldX2 = __builtin_nan("");
        for (int n_hg = 1; n_hg < 32; n_hg++) {
            oldX2 = oldX;
            oldX = X;
            stiefel_Gs3(Gs, beta, X);
            const double e1 = eta0 * Gs[1] + zeta0 * Gs[2];
            ri = 1. / (r0 + e1);
            X = ri * (X * e1 - eta0 * Gs[2] - zeta0 * Gs[3] + dt);
            if (X == oldX || X == oldX2) {
                converged = 1;
                break;
            }
        }
    }
    if (converged == 0) {
        // Bisection fallback (:333-369).
        flags |= 8;
        double X_min, X_max;
        if (beta > 0.) {
            X_min = X_per_period * floor(dt * invperiod);
            X_max = X_min + X_per_period;
        } else {
            const double h2 = r0 * r0 * v2 - eta0 * eta0;
            const double q = h2 / mu / (1. + sqrt(1. - h2 * beta / (mu * mu)));
            const double vq = copysign(sqrt(h2) / q, dt);
            X_min = dt / (fastabs(vq * dt) + r0);
            X_max = dt / q;
            if (dt < 0.) {
                const double temp = X_min;
                X_min = X_max;
                X_max = temp;
            }
        }
        X = (X_max + X_min) / 2.;
        // halving exhausts the doubles of any finite interval in < 2200 steps
        int guard = 0;
        do {
            stiefel_Gs3(Gs, beta, X);
            const double s = r0 * X + eta0 * Gs[2] + zeta0 * Gs[3] - dt;
            if (s >= 0.) {
                X_max = X;
            } else {
                X_min = X;
            }
            X = (X_max + X_min) / 2.;
        } while (fastabs((X_max - X_min)) > fastabs((X_max + X_min) * 1e-15) && ++guard < 4096);
        const double e1 = eta0 * Gs[1] + zeta0 * Gs[2];
        ri = 1. / (r0 + e1);
    }
    if (isnan(ri)) {
        ri = 0.;
        Gs[1] = 0.;
        Gs[2] = 0.;
        Gs[3] = 0.;
    }
    const double f = -mu * Gs[2] * r0i;
    const double g = dt - mu * Gs[3];
    const double fd = -mu * Gs[1] * r0i * ri;
    const double gd = -mu * Gs[2] * ri;
    px += f * x0 + g * vx0;
    py += f * y0 + g * vy0;
    pz += f * z0 + g * vz0;
    pvx += fd * x0 + gd * vx0;
    pvy += fd * y0 + gd * vy0;
    pvz += fd * z0 + gd * vz0;
    if (out) {
        *out = KepOut{x0, y0, z0, vx0, vy0, vz0, r0, r0i, ri, beta, eta0, zeta0, X, mu, f, g, fd, gd};
    }
    return flags;
}

// The drift of one body's variational particle d along its Kepler solve k
// (integrator_whfast.c:392-426): the Stiefel functions at the final X, then
// the derivatives of X, r and the drift coefficients.
__device__ __forceinline__ void kepler_var(const KepOut& k, double& dx, double& dy, double& dz, double& dvx,
                                           double& dvy, double& dvz) {
    const double d_x = dx, d_y = dy, d_z = dz, d_vx = dvx, d_vy = dvy, d_vz = dvz;
    double Gs[6];
    stiefel_Gs(Gs, k.beta, k.X);
    const double mu = k.mu, r0 = k.r0, r0i = k.r0i, ri = k.ri, X = k.X, beta = k.beta;
    const double dr0 = (d_x * k.x0 + d_y * k.y0 + d_z * k.z0) * r0i;
    const double dbeta = -2. * mu * dr0 * r0i * r0i - 2. * (d_vx * k.vx0 + d_vy * k.vy0 + d_vz * k.vz0);
    const double deta0 = d_x * k.vx0 + d_y * k.vy0 + d_z * k.vz0 + k.x0 * d_vx + k.y0 * d_vy + k.z0 * d_vz;
    const double dzeta0 = -beta * dr0 - r0 * dbeta;
    const double G3beta = 0.5 * (3. * Gs[5] - X * Gs[4]);
    const double G2beta = 0.5 * (2. * Gs[4] - X * Gs[3]);
    const double G1beta = 0.5 * (Gs[3] - X * Gs[2]);
    const double tbeta = k.eta0 * G2beta + k.zeta0 * G3beta;
    const double dX = -1. * ri * (X * dr0 + Gs[2] * deta0 + Gs[3] * dzeta0 + tbeta * dbeta);
    const double dG1 = Gs[0] * dX + G1beta * dbeta;
    const double dG2 = Gs[1] * dX + G2beta * dbeta;
    const double dG3 = Gs[2] * dX + G3beta * dbeta;
    const double dr = dr0 + Gs[1] * deta0 + Gs[2] * dzeta0 + k.eta0 * dG1 + k.zeta0 * dG2;
    const double df = mu * Gs[2] * dr0 * r0i * r0i - mu * dG2 * r0i;
    const double dg = -mu * dG3;
    const double dfd = -mu * dG1 * r0i * ri + mu * Gs[1] * (dr0 * r0i + dr * ri) * r0i * ri;
    const double dgd = -mu * dG2 * ri + mu * Gs[2] * dr * ri * ri;
    dx += k.f * d_x + k.g * d_vx + df * k.x0 + dg * k.vx0;
    dy += k.f * d_y + k.g * d_vy + df * k.y0 + dg * k.vy0;
    dz += k.f * d_z + k.g * d_vz + df * k.z0 + dg * k.vz0;
    dvx += k.fd * d_x + k.gd * d_vx + dfd * k.x0 + dgd * k.vx0;
    dvy += k.fd * d_y + k.gd * d_vy + dfd * k.y0 + dgd * k.vy0;
    dvz += k.fd * d_z + k.gd * d_vz + dfd * k.z0 + dgd * k.vz0;
}

// ---------------------------------------------------------------- batch Kepler
__global__ void k_kepler_batch(double* __restrict__ s6, const double* __restrict__ mu,
                               const double* __restrict__ dt, int n, int* __restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double* s = s6 + 6 * (size_t)i;
    double x = s[0], y = s[1], z = s[2], vx = s[3], vy = s[4], vz = s[5];
    const int f = kepler_solver(x, y, z, vx, vy, vz, mu[i], dt[i]);
    s[0] = x;
    s[1] = y;
    s[2] = z;
    s[3] = vx;
    s[4] = vy;
    s[5] = vz;
    if (f && flags) atomicOr(flags, f);
}

// ---------------------------------------------------------------- the group kernel
// LDS exchange slots: one per lane.
struct Slot {
    double x, y, z, vx, vy, vz, m;
};

// Sequential sum over the group's planets 1..n-1 (lanes 0..n-2 of the group) of
// slot values, in index order.  Every lane of the group evaluates the same sum.
#define GROUP_LOOP(var, lo, hi) for (int var = (lo); var < (hi); var++)

// Symplectic corrector coefficients of Wisdom, Holman & Touma (1996), the
// values integrator_whfast.c:125-152 spells: a_k = k * sqrt(7/40) and the
// b coefficients of the 3rd to 17th order correctors; kCorr2B is the second
// corrector's coefficient (Wisdom 2006).
__constant__ double kCorrA[9] = {0.,
                                 0.41833001326703777398908601289259374469640768464934,
                                 0.83666002653407554797817202578518748939281536929867,
                                 1.2549900398011133219672580386777812340892230539480,
                                 1.6733200530681510959563440515703749787856307385973,
                                 2.0916500663351888699454300644629687234820384232467,
                                 2.5099800796022266439345160773555624681784461078960,
                                 2.9283100928692644179236020902481562128748537925454,
                                 3.3466401061363021919126881031407499575712614771947};
static constexpr double kCorr2B = 0.03486083443891981449909050107438281205803;
// One corrector of order n is a chain of Z(a, b) operators.  An entry holds
// the signed index of a (a = sign * a_|k| * dt) and the b coefficient for
// inv = +1 (b = inv * coef * dt); the chains of integrator_whfast.c:696-751.
struct CorrZ {
    signed char ka;
    double kb;
};
__constant__ CorrZ kCorr3[2] = {{1, 0.024900596027799867499350357910273437184309981229127},
                                {-1, -0.024900596027799867499350357910273437184309981229127}};
__constant__ CorrZ kCorr5[4] = {{-2, 0.0083001986759332891664501193034244790614366604097090},
                                {-1, -0.041500993379666445832250596517122395307183302048545},
                                {1, 0.041500993379666445832250596517122395307183302048545},
                                {2, -0.0083001986759332891664501193034244790614366604097090}};
__constant__ CorrZ kCorr7[6] = {{-3, -0.0024926811426922105779030593952776964450539008582219},
                                {-2, 0.018270923246702131478062356884535264841652263842597},
                                {-1, -0.053964399093127498721765893493510877532452806339655},
                                {1, 0.053964399093127498721765893493510877532452806339655},
                                {2, -0.018270923246702131478062356884535264841652263842597},
                                {3, 0.0024926811426922105779030593952776964450539008582219}};
__constant__ CorrZ kCorr11[10] = {{-5, -0.00020361579647854651301632818774633716473696537436847},
                                  {-4, 0.0023487215292295354188307328851055489876255097419754},
                                  {-3, -0.012309078592019946317544564763237909911330686448336},
                                  {-2, 0.038121613681288650508647613260247372125243616270670},
                                  {-1, -0.072593394748842738674253180742744961827622366521517},
                                  {1, 0.072593394748842738674253180742744961827622366521517},
                                  {2, -0.038121613681288650508647613260247372125243616270670},
                                  {3, 0.012309078592019946317544564763237909911330686448336},
                                  {4, -0.0023487215292295354188307328851055489876255097419754},
                                  {5, 0.00020361579647854651301632818774633716473696537436847}};
__constant__ CorrZ kCorr17[16] = {{-8, 0.0000043347415473373580190650223498124944896789841432241},
                                  {-7, -0.000076436355227935738363241846979413475106795392377415},
                                  {-6, 0.00063599983075817658983166881625078545864140848560259},
                                  {-5, -0.0033132577069380655655490196833451994080066801611459},
                                  {-4, 0.012071760822342291062449751726959664253913904872527},
                                  {-3, -0.032422198864713580293681523029577130832258806467604},
                                  {-2, 0.065192863576377893658290760803725762027864651086787},
                                  {-1, -0.093056103771425958591541059067553547100903397724386},
                                  {1, 0.093056103771425958591541059067553547100903397724386},
                                  {2, -0.065192863576377893658290760803725762027864651086787},
                                  {3, 0.032422198864713580293681523029577130832258806467604},
                                  {4, -0.012071760822342291062449751726959664253913904872527},
                                  {5, 0.0033132577069380655655490196833451994080066801611459},
                                  {6, -0.00063599983075817658983166881625078545864140848560259},
                                  {7, 0.000076436355227935738363241846979413475106795392377415},
                                  {8, -0.0000043347415473373580190650223498124944896789841432241}};

__device__ __forceinline__ const CorrZ* corrector_chain(int order, int& n) {
    switch (order) {
        case 3: n = 2; return kCorr3;
        case 5: n = 4; return kCorr5;
        case 7: n = 6; return kCorr7;
        case 11: n = 10; return kCorr11;
        case 17: n = 16; return kCorr17;
        default: n = 0; return nullptr;
    }
}

// One wavefront per workgroup and rarely more than two per SIMD (an ensemble of
// 16384 systems is 1317 wave tasks): the register budget allows two waves per
// SIMD, so a step's whole state (about 40 doubles per lane plus the Kepler
// solver's temporaries) stays in VGPRs without spills.
//
// COORD selects the reference's coordinate system: WH_DH (democratic
// heliocentric) or WH_JACOBI.  In Jacobi coordinates the star takes part in
// the interaction (gravity_ignore_terms = BETWEEN_0_AND_1), the transforms are
// the sequential scans of transformations.c:34-81,147-176,178-258, which every
// lane of a group evaluates over the published bodies in the reference's order,
// and the symplectic correctors run at the synchronisation points.
template <int COORD, bool VAR>
__device__ __forceinline__ void whfast_body(const WHParams& P) {
    constexpr bool JAC = COORD == WH_JACOBI;
    static_assert(!VAR || JAC, "variational equations run in Jacobi coordinates");
    __shared__ Slot slot[64];
    __shared__ double aux[64][3];    // per-planet sums of the serial Jacobi gravity / jerk
    __shared__ double aux0[64][3];   // the star's, at the group's first lane
    __shared__ int task_id;
    const int lane = threadIdx.x;
    for (;;) {
        if (lane == 0) task_id = (int)atomicAdd(P.counter, 1u);
        __syncthreads();
        const int t = task_id;
        __syncthreads();
        if (t >= P.ntasks) break;
        const WHTask task = P.tasks[t];
        const int G = task.G;
        const int gl = lane % G;              // lane within the group
        const int gbase = lane - gl;          // first lane of the group
        const int ls = lane / G;              // system within the task
        const bool sys_ok = ls < task.nsys;
        const int sidx = sys_ok ? P.order[task.sys0 + ls] : 0;
        const WHSys S = P.sys[sidx];
        const int N = sys_ok ? S.N : 0;      // bodies incl. the star
        const int me = gl + 1;               // my body index
        const bool live = me < N;
        const int NA = sys_ok ? S.Nact : 0;        // N_active of the whfast sums
        const int NAG = sys_ok ? S.Nact_grav : 0;  // N_active of the gravity loop
        // Every system of a task has the same N, Nact, Nact_grav and
        // testparticle_type (the callers build tasks from systems of one body
        // count), so the loops over a system's bodies take the task's bounds,
        // read from lane 0 (always a system of the task) into scalar
        // registers.  Lanes without a system run those loops on their own
        // slots and discard the results (they store nothing).
        const int Nw = __builtin_amdgcn_readfirstlane(N);
        const int NAw = __builtin_amdgcn_readfirstlane(NA);
        const int NAGw = __builtin_amdgcn_readfirstlane(NAG);
        const int TPw = __builtin_amdgcn_readfirstlane(sys_ok ? S.tp_type : 0);
        const size_t st = P.stride;
        const size_t b0 = (size_t)S.body0;
        // ---- load
        double ix = 0, iy = 0, iz = 0, ivx = 0, ivy = 0, ivz = 0, im = 0;   // inertial planet
        double hx = 0, hy = 0, hz = 0, hvx = 0, hvy = 0, hvz = 0;           // heliocentric / Jacobi planet
        double sx = 0, sy = 0, sz = 0, svx = 0, svy = 0, svz = 0, sm = 0;   // inertial star
        double cx = 0, cy = 0, cz = 0, cvx = 0, cvy = 0, cvz = 0, cm = 0;   // p_jh[0] (barycentre)
        double ax = 0, ay = 0, az = 0;
        double sax = 0, say = 0, saz = 0;                                   // star (Jacobi only)
        if (sys_ok) {
            sx = P.body[0 * st + b0]; sy = P.body[1 * st + b0]; sz = P.body[2 * st + b0];
            svx = P.body[3 * st + b0]; svy = P.body[4 * st + b0]; svz = P.body[5 * st + b0];
            if (JAC) {
                sax = P.body[6 * st + b0]; say = P.body[7 * st + b0]; saz = P.body[8 * st + b0];
            }
            sm = P.body[9 * st + b0];
            cx = P.ph[0 * st + b0]; cy = P.ph[1 * st + b0]; cz = P.ph[2 * st + b0];
            cvx = P.ph[3 * st + b0]; cvy = P.ph[4 * st + b0]; cvz = P.ph[5 * st + b0];
            cm = P.ph[6 * st + b0];
        }
        if (live) {
            const size_t b = b0 + me;
            ix = P.body[0 * st + b]; iy = P.body[1 * st + b]; iz = P.body[2 * st + b];
            ivx = P.body[3 * st + b]; ivy = P.body[4 * st + b]; ivz = P.body[5 * st + b];
            ax = P.body[6 * st + b]; ay = P.body[7 * st + b]; az = P.body[8 * st + b];
            im = P.body[9 * st + b];
            hx = P.ph[0 * st + b]; hy = P.ph[1 * st + b]; hz = P.ph[2 * st + b];
            hvx = P.ph[3 * st + b]; hvy = P.ph[4 * st + b]; hvz = P.ph[5 * st + b];
        }
        double hm = im;   // p_jh[i].m (DH: equal to the inertial mass on this engine)
        if (JAC && live) hm = P.ph[6 * st + b0 + me];
        const unsigned int flags0 = sys_ok ? P.flags[sidx] : 1u;
        int sync = flags0 & 1u;
        int warn = (flags0 >> 1) & 15u;       // warnings and branch bits stay set until the host clears them
        // Steps of this system in this launch (a system may stop before the
        // launch's last step; its lanes keep taking part in the barriers).
        const long my_steps = !sys_ok ? 0 : (S.steps >= 0 && S.steps < P.nsteps ? (long)S.steps : P.nsteps);
        const double dt = S.dt;
        const double Gc = S.G;
        const bool safe = P.safe_mode || (sys_ok && (S.req & 4));

        // Each lane writes its planet's values into LDS; the group then forms
        // the reference's sequential sums over planets 1..NA-1 in index order.
        // A term such as p_h[i].x*m/mtot is evaluated once by its owner lane,
        // with the reference's expression, so every sum has the reference's bits.
        auto publish = [&](const Slot& v) {
            __syncthreads();
            slot[lane] = v;
            __syncthreads();
        };
        // Jacobi: eta of my Kepler problem and of my interaction term, m0 plus
        // the masses p_jh[1..min(me, NA-1)].m added in index order
        // (integrator_whfast.c:595-605 serial form, :446-451).
        double eta_me = 0.;
        auto interior_mass = [&]() {
            publish(Slot{0., 0., 0., 0., 0., 0., hm});
            double eta = sm;
            GROUP_LOOP(k, 1, NAw) {
                if (k <= me) eta += slot[gbase + k - 1].m;
            }
            eta_me = eta;
        };
        if (JAC) interior_mass();

        // ------------------------------------------------ variational particles
        // First-order sets (VAR: Jacobi, default kernel, one system).  Their
        // state stays in device memory (P.pv inertial, P.jv Jacobi); every
        // lane handles its planet's entries and lane 0 the star's and the
        // barycentre's.  Each operation starts with a barrier, so that what
        // another lane wrote before is visible.  The transforms take the real
        // masses, as the reference passes `particles` as p_mass.
        const size_t vs = P.vstride;
        auto vget = [&](const double* a, int f, size_t i) { return a[(size_t)f * vs + i]; };
        auto vput = [&](double* a, int f, size_t i, double x) { a[(size_t)f * vs + i] = x; };
        // transformations.c:34-81 on particles_var (integrator_whfast.c:958-962)
        auto var_from_inertial = [&](bool en) {
            for (int v = 0; v < P.nvar; v++) {
                const size_t b = b0 + (size_t)P.var_index[v];
                __syncthreads();
                double q[6] = {0., 0., 0., 0., 0., 0.}, d0[6] = {0., 0., 0., 0., 0., 0.}, dm = 0.;
                if (live) {
                    for (int f = 0; f < 6; f++) q[f] = vget(P.pv, f, b + me);
                    dm = vget(P.pv, 9, b + me);
                }
                if (sys_ok)
                    for (int f = 0; f < 6; f++) d0[f] = vget(P.pv, f, b);
                publish(Slot{q[0], q[1], q[2], q[3], q[4], q[5], im});
                double eta = sm;
                double s_x = eta * d0[0], s_y = eta * d0[1], s_z = eta * d0[2];
                double s_vx = eta * d0[3], s_vy = eta * d0[4], s_vz = eta * d0[5];
                double j[6] = {0., 0., 0., 0., 0., 0.};
                GROUP_LOOP(k, 1, NAw) {
                    const Slot o = slot[gbase + k - 1];
                    const double ei = 1. / eta;
                    eta += o.m;
                    const double pme = eta * ei;
                    const double jx = o.x - s_x * ei, jy = o.y - s_y * ei, jz = o.z - s_z * ei;
                    const double jvx = o.vx - s_vx * ei, jvy = o.vy - s_vy * ei, jvz = o.vz - s_vz * ei;
                    if (k == me) {
                        j[0] = jx; j[1] = jy; j[2] = jz;
                        j[3] = jvx; j[4] = jvy; j[5] = jvz;
                    }
                    s_x = s_x * pme + o.m * jx;
                    s_y = s_y * pme + o.m * jy;
                    s_z = s_z * pme + o.m * jz;
                    s_vx = s_vx * pme + o.m * jvx;
                    s_vy = s_vy * pme + o.m * jvy;
                    s_vz = s_vz * pme + o.m * jvz;
                }
                const double ei = 1. / eta;
                if (me >= NA) {
                    j[0] = q[0] - s_x * ei; j[1] = q[1] - s_y * ei; j[2] = q[2] - s_z * ei;
                    j[3] = q[3] - s_vx * ei; j[4] = q[4] - s_vy * ei; j[5] = q[5] - s_vz * ei;
                }
                if (!en) continue;
                if (live) {
                    for (int f = 0; f < 6; f++) vput(P.jv, f, b + me, j[f]);
                    vput(P.jv, 9, b + me, dm);
                }
                if (gl == 0 && sys_ok) {
                    const double Mtotali = 1. / eta;
                    vput(P.jv, 0, b, s_x * Mtotali); vput(P.jv, 1, b, s_y * Mtotali); vput(P.jv, 2, b, s_z * Mtotali);
                    vput(P.jv, 3, b, s_vx * Mtotali); vput(P.jv, 4, b, s_vy * Mtotali); vput(P.jv, 5, b, s_vz * Mtotali);
                    vput(P.jv, 9, b, eta);
                }
            }
        };
        // transformations.c:178-258 on particles_var (integrator_whfast.c:662-672, 1089-1092)
        auto var_to_inertial = [&](bool en, bool with_vel) {
            for (int v = 0; v < P.nvar; v++) {
                const size_t b = b0 + (size_t)P.var_index[v];
                __syncthreads();
                double h[6] = {0., 0., 0., 0., 0., 0.}, c[6] = {0., 0., 0., 0., 0., 0.}, ceta = 0.;
                if (live)
                    for (int f = 0; f < 6; f++) h[f] = vget(P.jv, f, b + me);
                if (sys_ok) {
                    for (int f = 0; f < 6; f++) c[f] = vget(P.jv, f, b);
                    ceta = vget(P.jv, 9, b);
                }
                publish(Slot{h[0], h[1], h[2], h[3], h[4], h[5], im});
                double eta = ceta;
                double s_x = c[0] * eta, s_y = c[1] * eta, s_z = c[2] * eta;
                double s_vx = 0., s_vy = 0., s_vz = 0.;
                if (with_vel) {
                    s_vx = c[3] * eta; s_vy = c[4] * eta; s_vz = c[5] * eta;
                }
                double n[6] = {0., 0., 0., 0., 0., 0.};
                if (me >= NA) {
                    const double ei = 1. / eta;
                    n[0] = h[0] + s_x * ei; n[1] = h[1] + s_y * ei; n[2] = h[2] + s_z * ei;
                    if (with_vel) {
                        n[3] = h[3] + s_vx * ei; n[4] = h[4] + s_vy * ei; n[5] = h[5] + s_vz * ei;
                    }
                }
                for (int k = NAw - 1; k > 0; k--) {
                    const Slot o = slot[gbase + k - 1];
                    const double ei = 1. / eta;
                    s_x = (s_x - o.m * o.x) * ei;
                    s_y = (s_y - o.m * o.y) * ei;
                    s_z = (s_z - o.m * o.z) * ei;
                    if (with_vel) {
                        s_vx = (s_vx - o.m * o.vx) * ei;
                        s_vy = (s_vy - o.m * o.vy) * ei;
                        s_vz = (s_vz - o.m * o.vz) * ei;
                    }
                    if (k == me) {
                        n[0] = o.x + s_x; n[1] = o.y + s_y; n[2] = o.z + s_z;
                        if (with_vel) {
                            n[3] = o.vx + s_vx; n[4] = o.vy + s_vy; n[5] = o.vz + s_vz;
                        }
                    }
                    eta -= o.m;
                    s_x *= eta; s_y *= eta; s_z *= eta;
                    if (with_vel) {
                        s_vx *= eta; s_vy *= eta; s_vz *= eta;
                    }
                }
                if (!en) continue;
                const int nf = with_vel ? 6 : 3;
                if (live)
                    for (int f = 0; f < nf; f++) vput(P.pv, f, b + me, n[f]);
                if (gl == 0 && sys_ok) {
                    const double mi = 1. / eta;
                    vput(P.pv, 0, b, s_x * mi); vput(P.pv, 1, b, s_y * mi); vput(P.pv, 2, b, s_z * mi);
                    if (with_vel) {
                        vput(P.pv, 3, b, s_vx * mi); vput(P.pv, 4, b, s_vy * mi); vput(P.pv, 5, b, s_vz * mi);
                    }
                }
            }
        };
        // gravity.c:533-640, first order, all bodies varied: the pairs j < i
        // from (starti, startj) as gravity_ignore_terms selects, i-major, each
        // term added to i and subtracted from j.  Body k's sum therefore holds
        // its terms as i (ascending j), then as j (ascending i), then those of
        // test particles i >= N_active (type 1).  The star's is formed by lane 0.
        auto var_gravity = [&](bool en, int starti, int startj) {
            for (int v = 0; v < P.nvar; v++) {
                const size_t b = b0 + (size_t)P.var_index[v];
                __syncthreads();
                double d[3] = {0., 0., 0.}, dm = 0., ds[3] = {0., 0., 0.}, dsm = 0.;
                if (live) {
                    d[0] = vget(P.pv, 0, b + me); d[1] = vget(P.pv, 1, b + me); d[2] = vget(P.pv, 2, b + me);
                    dm = vget(P.pv, 9, b + me);
                }
                if (sys_ok) {
                    ds[0] = vget(P.pv, 0, b); ds[1] = vget(P.pv, 1, b); ds[2] = vget(P.pv, 2, b);
                    dsm = vget(P.pv, 9, b);
                }
                aux[lane][0] = dm;
                publish(Slot{ix, iy, iz, d[0], d[1], d[2], im});
                // body k: 0 the star, k >= 1 the slot of planet k
                auto term = [&](int I, int J, bool to_i, double& tx, double& ty, double& tz) {
                    const Slot oi = I == 0 ? Slot{sx, sy, sz, ds[0], ds[1], ds[2], sm} : slot[gbase + I - 1];
                    const Slot oj = J == 0 ? Slot{sx, sy, sz, ds[0], ds[1], ds[2], sm} : slot[gbase + J - 1];
                    const double dmi = I == 0 ? dsm : aux[gbase + I - 1][0];
                    const double dmj = J == 0 ? dsm : aux[gbase + J - 1][0];
                    const double dx = oi.x - oj.x;
                    const double dy = oi.y - oj.y;
                    const double dz = oi.z - oj.z;
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    const double _r = sqrt(r2);
                    const double r3inv = 1. / (r2 * _r);
                    const double r5inv = 3. * r3inv / r2;
                    const double ddx = oi.vx - oj.vx;
                    const double ddy = oi.vy - oj.vy;
                    const double ddz = oi.vz - oj.vz;
                    const double dxdx = dx * dx * r5inv - r3inv;
                    const double dydy = dy * dy * r5inv - r3inv;
                    const double dzdz = dz * dz * r5inv - r3inv;
                    const double dxdy = dx * dy * r5inv;
                    const double dxdz = dx * dz * r5inv;
                    const double dydz = dy * dz * r5inv;
                    const double dax = ddx * dxdx + ddy * dxdy + ddz * dxdz;
                    const double day = ddx * dxdy + ddy * dydy + ddz * dydz;
                    const double daz = ddx * dxdz + ddy * dydz + ddz * dzdz;
                    if (to_i) {
                        const double Gmj = Gc * oj.m;
                        const double dGmj = Gc * dmj;
                        tx = Gmj * dax - dGmj * r3inv * dx;
                        ty = Gmj * day - dGmj * r3inv * dy;
                        tz = Gmj * daz - dGmj * r3inv * dz;
                    } else {
                        const double Gmi = Gc * oi.m;
                        const double dGmi = Gc * dmi;
                        tx = Gmi * dax - dGmi * r3inv * dx;
                        ty = Gmi * day - dGmi * r3inv * dy;
                        tz = Gmi * daz - dGmi * r3inv * dz;
                    }
                };
                // sum of body k (k = 0: the star)
                auto body_sum = [&](int k, double& a_x, double& a_y, double& a_z) {
                    a_x = 0.; a_y = 0.; a_z = 0.;
                    double tx, ty, tz;
                    if (k < NAGw) {
                        if (k >= starti)
                            for (int j = startj; j < k; j++) {
                                term(k, j, true, tx, ty, tz);
                                a_x += tx; a_y += ty; a_z += tz;
                            }
                        if (k >= startj) {
                            for (int i = (k + 1 > starti ? k + 1 : starti); i < NAGw; i++) {
                                term(i, k, false, tx, ty, tz);
                                a_x -= tx; a_y -= ty; a_z -= tz;
                            }
                            if (TPw)
                                for (int i = NAGw; i < Nw; i++) {
                                    term(i, k, false, tx, ty, tz);
                                    a_x -= tx; a_y -= ty; a_z -= tz;
                                }
                        }
                    } else {
                        for (int j = startj; j < NAGw; j++) {
                            term(k, j, true, tx, ty, tz);
                            a_x += tx; a_y += ty; a_z += tz;
                        }
                    }
                };
                if (!en) continue;
                if (live) {
                    double a_x, a_y, a_z;
                    body_sum(me, a_x, a_y, a_z);
                    vput(P.pv, 6, b + me, a_x); vput(P.pv, 7, b + me, a_y); vput(P.pv, 8, b + me, a_z);
                }
                if (gl == 0 && sys_ok) {
                    double a_x, a_y, a_z;
                    body_sum(0, a_x, a_y, a_z);
                    vput(P.pv, 6, b, a_x); vput(P.pv, 7, b, a_y); vput(P.pv, 8, b, a_z);
                }
            }
        };

        // ------------------------------------------------ coordinate transforms
        // DH: transformations.c:388-424.  Jacobi: transformations.c:34-81.
        auto from_inertial = [&]() {
            if constexpr (JAC) {
                publish(Slot{ix, iy, iz, ivx, ivy, ivz, im});
                double eta = sm;
                double s_x = eta * sx, s_y = eta * sy, s_z = eta * sz;
                double s_vx = eta * svx, s_vy = eta * svy, s_vz = eta * svz;
                GROUP_LOOP(k, 1, NAw) {
                    const Slot q = slot[gbase + k - 1];
                    const double ei = 1. / eta;
                    eta += q.m;
                    const double pme = eta * ei;
                    const double jx = q.x - s_x * ei, jy = q.y - s_y * ei, jz = q.z - s_z * ei;
                    const double jvx = q.vx - s_vx * ei, jvy = q.vy - s_vy * ei, jvz = q.vz - s_vz * ei;
                    if (k == me) {
                        hx = jx; hy = jy; hz = jz;
                        hvx = jvx; hvy = jvy; hvz = jvz;
                        eta_me = eta;
                    }
                    s_x = s_x * pme + q.m * jx;
                    s_y = s_y * pme + q.m * jy;
                    s_z = s_z * pme + q.m * jz;
                    s_vx = s_vx * pme + q.m * jvx;
                    s_vy = s_vy * pme + q.m * jvy;
                    s_vz = s_vz * pme + q.m * jvz;
                }
                const double ei = 1. / eta;
                if (me >= NA) {
                    hx = ix - s_x * ei; hy = iy - s_y * ei; hz = iz - s_z * ei;
                    hvx = ivx - s_vx * ei; hvy = ivy - s_vy * ei; hvz = ivz - s_vz * ei;
                    eta_me = eta;
                }
                const double Mtotali = 1. / eta;
                cm = eta;
                cx = s_x * Mtotali; cy = s_y * Mtotali; cz = s_z * Mtotali;
                cvx = s_vx * Mtotali; cvy = s_vy * Mtotali; cvz = s_vz * Mtotali;
                hm = im;
            } else {
                publish(Slot{ix * im, iy * im, iz * im, ivx * im, ivy * im, ivz * im, im});
                double x0 = 0., y0 = 0., z0 = 0., vx0 = 0., vy0 = 0., vz0 = 0., m0 = 0.;
                if (0 < NA) {
                    const double m = sm;
                    x0 += sx * m; y0 += sy * m; z0 += sz * m;
                    vx0 += svx * m; vy0 += svy * m; vz0 += svz * m;
                    m0 += m;
                }
                GROUP_LOOP(k, 1, NAw) {
                    const Slot q = slot[gbase + k - 1];
                    x0 += q.x; y0 += q.y; z0 += q.z;
                    vx0 += q.vx; vy0 += q.vy; vz0 += q.vz;
                    m0 += q.m;
                }
                cx = x0 / m0; cy = y0 / m0; cz = z0 / m0;
                cvx = vx0 / m0; cvy = vy0 / m0; cvz = vz0 / m0;
                cm = m0;
                hx = ix - sx; hy = iy - sy; hz = iz - sz;
                hvx = ivx - cvx; hvy = ivy - cvy; hvz = ivz - cvz;
                hm = im;
            }
        };
        // DH: transformations.c:426-477 (positions then velocities).  Jacobi:
        // transformations.c:178-226 (posvel) and :228-258 (positions only).
        // Barriers are executed by every lane; lanes whose system does not
        // take part (en == false) keep their state.  The velocity pass is
        // skipped (with_vel == false) where nothing reads the inertial
        // velocities before they are recomputed.
        auto to_inertial = [&](bool en, bool with_vel) {
            if constexpr (VAR) var_to_inertial(en, with_vel);
            if constexpr (JAC) {
                Slot t{hx, hy, hz, 0., 0., 0., im};
                if (with_vel) {
                    t.vx = hvx; t.vy = hvy; t.vz = hvz;
                }
                publish(t);
                double eta = cm;
                double s_x = cx * eta, s_y = cy * eta, s_z = cz * eta;
                double s_vx = 0., s_vy = 0., s_vz = 0.;
                if (with_vel) {
                    s_vx = cvx * eta; s_vy = cvy * eta; s_vz = cvz * eta;
                }
                double nx = ix, ny = iy, nz = iz, nvx = ivx, nvy = ivy, nvz = ivz;
                if (me >= NA) {
                    const double ei = 1. / eta;
                    nx = hx + s_x * ei; ny = hy + s_y * ei; nz = hz + s_z * ei;
                    if (with_vel) {
                        nvx = hvx + s_vx * ei; nvy = hvy + s_vy * ei; nvz = hvz + s_vz * ei;
                    }
                }
                for (int k = NAw - 1; k > 0; k--) {
                    const Slot q = slot[gbase + k - 1];
                    const double ei = 1. / eta;
                    s_x = (s_x - q.m * q.x) * ei;
                    s_y = (s_y - q.m * q.y) * ei;
                    s_z = (s_z - q.m * q.z) * ei;
                    if (with_vel) {
                        s_vx = (s_vx - q.m * q.vx) * ei;
                        s_vy = (s_vy - q.m * q.vy) * ei;
                        s_vz = (s_vz - q.m * q.vz) * ei;
                    }
                    if (k == me) {
                        nx = q.x + s_x; ny = q.y + s_y; nz = q.z + s_z;
                        if (with_vel) {
                            nvx = q.vx + s_vx; nvy = q.vy + s_vy; nvz = q.vz + s_vz;
                        }
                    }
                    eta -= q.m;
                    s_x *= eta; s_y *= eta; s_z *= eta;
                    if (with_vel) {
                        s_vx *= eta; s_vy *= eta; s_vz *= eta;
                    }
                }
                if (!en) return;
                const double mi = 1. / eta;
                sx = s_x * mi; sy = s_y * mi; sz = s_z * mi;
                if (live) {
                    ix = nx; iy = ny; iz = nz;
                }
                if (with_vel) {
                    svx = s_vx * mi; svy = s_vy * mi; svz = s_vz * mi;
                    if (live) {
                        ivx = nvx; ivy = nvy; ivz = nvz;
                    }
                }
            } else {
                const double mtot = cm;
                const double m0 = sm;   // particles[0].m; particles[i].m == p_h[i].m after the position pass
                Slot t{hx * hm / mtot, hy * hm / mtot, hz * hm / mtot, 0., 0., 0., 0.};
                if (with_vel) {
                    t.vx = hvx * hm / m0;
                    t.vy = hvy * hm / m0;
                    t.vz = hvz * hm / m0;
                }
                publish(t);
                double x0 = 0., y0 = 0., z0 = 0.;
                double vx0 = 0., vy0 = 0., vz0 = 0.;
                GROUP_LOOP(k, 1, NAw) {
                    const Slot q = slot[gbase + k - 1];
                    x0 += q.x;
                    y0 += q.y;
                    z0 += q.z;
                    if (with_vel) {
                        vx0 += q.vx;
                        vy0 += q.vy;
                        vz0 += q.vz;
                    }
                }
                if (!en) return;
                if (me < NA) im = hm;
                sx = cx - x0; sy = cy - y0; sz = cz - z0;
                ix = hx + sx; iy = hy + sy; iz = hz + sz;
                if (with_vel) {
                    ivx = hvx + cvx; ivy = hvy + cvy; ivz = hvz + cvz;
                    svx = cvx - vx0; svy = cvy - vy0; svz = cvz - vz0;
                }
            }
        };
        // integrator_whfast.c:589-636 (DH mu = G m0; Jacobi mu = G eta_i)
        auto kepler = [&](double h, bool en) {
            if (!en || !live) return;
            const double mu = JAC ? eta_me * Gc : sm * Gc;
            if constexpr (VAR) {
                KepOut ko;
                warn |= kepler_solver(hx, hy, hz, hvx, hvy, hvz, mu, h, &ko);
                for (int v = 0; v < P.nvar; v++) {
                    const size_t b = b0 + (size_t)P.var_index[v] + me;
                    double d[6];
                    for (int f = 0; f < 6; f++) d[f] = vget(P.jv, f, b);
                    kepler_var(ko, d[0], d[1], d[2], d[3], d[4], d[5]);
                    for (int f = 0; f < 6; f++) vput(P.jv, f, b, d[f]);
                }
            } else {
                warn |= kepler_solver(hx, hy, hz, hvx, hvy, hvz, mu, h);
            }
        };
        // integrator_whfast.c:638-651
        auto com = [&](double h, bool en) {
            if (!en) return;
            cx += h * cvx;
            cy += h * cvy;
            cz += h * cvz;
            if constexpr (VAR) {
                if (gl == 0)
                    for (int v = 0; v < P.nvar; v++) {
                        const size_t b = b0 + (size_t)P.var_index[v];
                        vput(P.jv, 0, b, vget(P.jv, 0, b) + h * vget(P.jv, 3, b));
                        vput(P.jv, 1, b, vget(P.jv, 1, b) + h * vget(P.jv, 4, b));
                        vput(P.jv, 2, b, vget(P.jv, 2, b) + h * vget(P.jv, 5, b));
                    }
            }
        };
        // integrator_whfast.c:537-554, DH only (masses are the inertial ones,
        // equal to p_h[i].m here: the engine changes no masses)
        auto jump = [&](double h, bool en) {
            if constexpr (!JAC) {
                publish(Slot{hm * hvx, hm * hvy, hm * hvz, 0., 0., 0., 0.});
                double qx = 0, qy = 0, qz = 0;
                GROUP_LOOP(k, 1, NAw) {
                    const Slot q = slot[gbase + k - 1];
                    qx += q.x;
                    qy += q.y;
                    qz += q.z;
                }
                const double m0 = sm;
                if (live && en) {
                    hx += h * (qx / m0);
                    hy += h * (qy / m0);
                    hz += h * (qz / m0);
                }
            }
        };
        // One pair term of BASIC (gravity.c:215-232) or COMPENSATED (:308-410)
        // in the reference's form, and its addition to (a, k).
        auto pair_term = [&](double dx, double dy, double dz, double mj, double& t_x, double& t_y, double& t_z) {
            if (P.compensated) {
                const double r2 = dx * dx + dy * dy + dz * dz + P.soft2;
                const double rr = sqrt(r2);
                const double prefact = Gc / (r2 * rr);
                const double prefactj = -prefact * mj;
                t_x = prefactj * dx;
                t_y = prefactj * dy;
                t_z = prefactj * dz;
            } else {
                const double _r = sqrt(dx * dx + dy * dy + dz * dz + P.soft2);
                const double prefact = -Gc / (_r * _r * _r) * mj;
                t_x = prefact * dx;
                t_y = prefact * dy;
                t_z = prefact * dz;
            }
        };
        auto pair_add = [&](double t_x, double t_y, double t_z, double& a_x, double& a_y, double& a_z, double& k_x,
                            double& k_y, double& k_z) {
            if (P.compensated) {
                double tv, yv;
                yv = t_x - k_x; tv = a_x + yv; k_x = (tv - a_x) - yv; a_x = tv;
                yv = t_y - k_y; tv = a_y + yv; k_y = (tv - a_y) - yv; a_y = tv;
                yv = t_z - k_z; tv = a_z + yv; k_z = (tv - a_z) - yv; a_z = tv;
            } else {
                a_x += t_x;
                a_y += t_y;
                a_z += t_z;
            }
        };
        // Accelerations of BASIC or COMPENSATED gravity, ascending j.  DH
        // (gravity_ignore_terms = INVOLVING_0): planets only.  Jacobi
        // (BETWEEN_0_AND_1): the star is a source for planets 2.. and a
        // target; its sum is ordered over owner-lane terms.  The terms of
        // four sources at a time are independent and evaluated together;
        // they are added in ascending j.
        auto gravity = [&](bool en) {
            Slot pub{ix, iy, iz, 0., 0., 0., im};
            if constexpr (JAC) {
                // star target, my planet as source (pass 0 for me < NAG, pass 1
                // with test particles of type 1); pair 0-1 ignored
                if (live && me >= 2 && (me < NAG || S.tp_type))
                    pair_term(sx - ix, sy - iy, sz - iz, im, pub.vx, pub.vy, pub.vz);
            }
            publish(pub);
            if constexpr (JAC) {
                if (en) {
                    double a_x = 0., a_y = 0., a_z = 0., k_x = 0., k_y = 0., k_z = 0.;
                    for (int j = 2; j < Nw; j++) {
                        if (!(j < NAGw || TPw)) break;
                        const Slot q = slot[gbase + j - 1];
                        pair_add(q.vx, q.vy, q.vz, a_x, a_y, a_z, k_x, k_y, k_z);
                    }
                    sax = a_x; say = a_y; saz = a_z;
                }
            }
            if (!live || !en) return;
            ax = 0.; ay = 0.; az = 0.;
            double kx = 0., ky = 0., kz = 0.;
            if (JAC && me != 1 && NAG > 0) {
                double tx, ty, tz;
                pair_term(ix - sx, iy - sy, iz - sz, sm, tx, ty, tz);
                pair_add(tx, ty, tz, ax, ay, az, kx, ky, kz);
            }
            for (int pass = 0; pass < 2; pass++) {
                const int j0 = pass == 0 ? 1 : (NAGw > 1 ? NAGw : 1);
                const int j1 = pass == 0 ? NAGw : Nw;
                if (pass == 1 && !(TPw && me < NAGw)) break;
                for (int jb = j0; jb < j1; jb += 4) {
                    double tx[4], ty[4], tz[4];
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int j = jb + q;
                        if (j >= j1 || j == me) continue;
                        const Slot o = slot[gbase + j - 1];
                        pair_term(ix - o.x, iy - o.y, iz - o.z, o.m, tx[q], ty[q], tz[q]);
                    }
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int j = jb + q;
                        if (j >= j1 || j == me) continue;
                        pair_add(tx[q], ty[q], tz[q], ax, ay, az, kx, ky, kz);
                    }
                }
            }
        };
        // The Jacobi gravity (gravity.c:108-163) of the modified-kick and lazy
        // kernels visits every pair once, j = 0..N-1 and i = 0..j, and adds
        // to both bodies; its sums are formed by the group's first lane in
        // that order over the published bodies (the star from registers).
        // Body k of the system: 0 the star, k >= 1 the slot of planet k.
        auto bx = [&](int k) { return k == 0 ? sx : slot[gbase + k - 1].x; };
        auto by = [&](int k) { return k == 0 ? sy : slot[gbase + k - 1].y; };
        auto bz = [&](int k) { return k == 0 ? sz : slot[gbase + k - 1].z; };
        auto bm = [&](int k) { return k == 0 ? sm : slot[gbase + k - 1].m; };
        auto acc_of = [&](int k) -> double* { return k == 0 ? aux0[gbase] : aux[gbase + k - 1]; };
        auto gravity_jacobi = [&](bool en) {
            publish(Slot{ix, iy, iz, 0., 0., 0., im});
            if (gl == 0 && sys_ok) {
                double Rx = 0., Ry = 0., Rz = 0., Mt = 0.;
                for (int j = 0; j < Nw; j++) {
                    double* aj = acc_of(j);
                    aj[0] = 0.; aj[1] = 0.; aj[2] = 0.;
                    const double xj = bx(j), yj = by(j), zj = bz(j), mj = bm(j);
                    for (int i = 0; i <= j; i++) {
                        double* ai = acc_of(i);
                        if (j > 1) {
                            const double qx = xj - Rx / Mt;
                            const double qy = yj - Ry / Mt;
                            const double qz = zj - Rz / Mt;
                            const double dr = sqrt(qx * qx + qy * qy + qz * qz);
                            const double w = i < j ? -mj : Mt;
                            const double pf = Gc * w / (dr * dr * dr);
                            ai[0] += pf * qx;
                            ai[1] += pf * qy;
                            ai[2] += pf * qz;
                        }
                        if (i != j && (i != 0 || j != 1)) {
                            const double dx = bx(i) - xj;
                            const double dy = by(i) - yj;
                            const double dz = bz(i) - zj;
                            const double dr = sqrt(dx * dx + dy * dy + dz * dz);
                            const double pf = Gc / (dr * dr * dr);
                            const double pfi = pf * bm(i);
                            const double pfj = pf * mj;
                            ai[0] -= pfj * dx;
                            ai[1] -= pfj * dy;
                            ai[2] -= pfj * dz;
                            aj[0] += pfi * dx;
                            aj[1] += pfi * dy;
                            aj[2] += pfi * dz;
                        }
                    }
                    Rx += mj * xj;
                    Ry += mj * yj;
                    Rz += mj * zj;
                    Mt += mj;
                }
            }
            __syncthreads();
            if (!en) return;
            sax = aux0[gbase][0]; say = aux0[gbase][1]; saz = aux0[gbase][2];
            if (live) {
                ax = aux[lane][0]; ay = aux[lane][1]; az = aux[lane][2];
            }
        };
        // integrator_whfast.c:786-873 (jerk of the Jacobi gravity, same pair
        // order) and :1178-1186: a += dt^2/12 jerk for every body.
        auto modified_kick = [&](bool en) {
            publish(Slot{ix, iy, iz, ax, ay, az, im});
            if (gl == 0 && sys_ok) {
                auto bax = [&](int k) { return k == 0 ? sax : slot[gbase + k - 1].vx; };
                auto bay = [&](int k) { return k == 0 ? say : slot[gbase + k - 1].vy; };
                auto baz = [&](int k) { return k == 0 ? saz : slot[gbase + k - 1].vz; };
                double Rx = 0., Ry = 0., Rz = 0., Mt = 0., Ax = 0., Ay = 0., Az = 0.;
                for (int j = 0; j < Nw; j++) {
                    double* kj = acc_of(j);
                    kj[0] = 0.; kj[1] = 0.; kj[2] = 0.;
                    const double xj = bx(j), yj = by(j), zj = bz(j), mj = bm(j);
                    const double axj = bax(j), ayj = bay(j), azj = baz(j);
                    for (int i = 0; i <= j; i++) {
                        double* ki = acc_of(i);
                        if (j > 1) {
                            const double w = i < j ? -mj : Mt;
                            const double qx = xj - Rx / Mt;
                            const double qy = yj - Ry / Mt;
                            const double qz = zj - Rz / Mt;
                            const double dax = axj - Ax / Mt;
                            const double day = ayj - Ay / Mt;
                            const double daz = azj - Az / Mt;
                            const double dr = sqrt(qx * qx + qy * qy + qz * qz);
                            const double p2 = Gc * w / (dr * dr * dr);
                            ki[0] += p2 * dax;
                            ki[1] += p2 * day;
                            ki[2] += p2 * daz;
                            const double al = dax * qx + day * qy + daz * qz;
                            const double p1 = 3. * al * p2 / (dr * dr);
                            ki[0] -= p1 * qx;
                            ki[1] -= p1 * qy;
                            ki[2] -= p1 * qz;
                        }
                        if (j != i && (i != 0 || j != 1)) {
                            const double dx = xj - bx(i);
                            const double dy = yj - by(i);
                            const double dz = zj - bz(i);
                            const double dax = axj - bax(i);
                            const double day = ayj - bay(i);
                            const double daz = azj - baz(i);
                            const double dr = sqrt(dx * dx + dy * dy + dz * dz);
                            const double al = dax * dx + day * dy + daz * dz;
                            const double p2 = Gc / (dr * dr * dr);
                            const double p2i = p2 * bm(i);
                            const double p2j = p2 * mj;
                            kj[0] -= dax * p2i;
                            kj[1] -= day * p2i;
                            kj[2] -= daz * p2i;
                            ki[0] += dax * p2j;
                            ki[1] += day * p2j;
                            ki[2] += daz * p2j;
                            const double p1 = 3. * al * p2 / (dr * dr);
                            const double p1i = p1 * bm(i);
                            const double p1j = p1 * mj;
                            kj[0] += dx * p1i;
                            kj[1] += dy * p1i;
                            kj[2] += dz * p1i;
                            ki[0] -= dx * p1j;
                            ki[1] -= dy * p1j;
                            ki[2] -= dz * p1j;
                        }
                    }
                    Ax += axj * mj;
                    Ay += ayj * mj;
                    Az += azj * mj;
                    Rx += xj * mj;
                    Ry += yj * mj;
                    Rz += zj * mj;
                    Mt += mj;
                }
            }
            __syncthreads();
            if (!en) return;
            const double pf = dt * dt / 12.;
            sax += pf * aux0[gbase][0];
            say += pf * aux0[gbase][1];
            saz += pf * aux0[gbase][2];
            if (live) {
                ax += pf * aux[lane][0];
                ay += pf * aux[lane][1];
                az += pf * aux[lane][2];
            }
        };
        // reb_simulation_update_acceleration for the kernel in use
        auto grav = [&](bool en) {
            // gravity_ignore_terms BETWEEN_0_AND_1 (integrator_whfast.c:933)
            if constexpr (VAR) var_gravity(en, 2, 0);
            if (JAC && P.jacgrav)
                gravity_jacobi(en);
            else
                gravity(en);
        };
        // DH: integrator_whfast.c:494-501.  Jacobi: :439-483 (accelerations to
        // Jacobi form, transformations.c:147-176, then Eq. 132 of the WHFast
        // paper with the Kepler terms of the interior bodies).
        // integrator_whfast.c:441-447 and :467-490 for the variational sets:
        // accelerations to Jacobi form, the linearised Kepler term of the
        // interior mass, then the kick.  The real body's factors are those of
        // the real kick below (its positions do not change in a kick).
        auto var_interaction = [&](double h, bool en) {
            for (int v = 0; v < P.nvar; v++) {
                const size_t b = b0 + (size_t)P.var_index[v];
                __syncthreads();
                double a[3] = {0., 0., 0.}, sa[3] = {0., 0., 0.};
                if (live) {
                    a[0] = vget(P.pv, 6, b + me); a[1] = vget(P.pv, 7, b + me); a[2] = vget(P.pv, 8, b + me);
                }
                if (sys_ok) {
                    sa[0] = vget(P.pv, 6, b); sa[1] = vget(P.pv, 7, b); sa[2] = vget(P.pv, 8, b);
                }
                publish(Slot{a[0], a[1], a[2], 0., 0., 0., im});
                double eta = sm;
                double s_x = eta * sa[0], s_y = eta * sa[1], s_z = eta * sa[2];
                double jax = 0., jay = 0., jaz = 0.;
                GROUP_LOOP(k, 1, NAw) {
                    const Slot o = slot[gbase + k - 1];
                    const double ei = 1. / eta;
                    eta += o.m;
                    const double pme = eta * ei;
                    const double jx = o.x - s_x * ei, jy = o.y - s_y * ei, jz = o.z - s_z * ei;
                    if (k == me) {
                        jax = jx; jay = jy; jaz = jz;
                    }
                    s_x = s_x * pme + o.m * jx;
                    s_y = s_y * pme + o.m * jy;
                    s_z = s_z * pme + o.m * jz;
                }
                if (me >= NA) {
                    const double ei = 1. / eta;
                    jax = a[0] - s_x * ei; jay = a[1] - s_y * ei; jaz = a[2] - s_z * ei;
                }
                if (!en) continue;
                if (gl == 0 && sys_ok) {
                    const double Mtotali = 1. / eta;
                    vput(P.jv, 6, b, s_x * Mtotali); vput(P.jv, 7, b, s_y * Mtotali); vput(P.jv, 8, b, s_z * Mtotali);
                }
                if (!live) continue;
                const size_t bi = b + me;
                vput(P.jv, 6, bi, jax); vput(P.jv, 7, bi, jay); vput(P.jv, 8, bi, jaz);
                const double dx = vget(P.jv, 0, bi), dy = vget(P.jv, 1, bi), dz = vget(P.jv, 2, bi);
                double dvx = vget(P.jv, 3, bi), dvy = vget(P.jv, 4, bi), dvz = vget(P.jv, 5, bi);
                if (me > 1) {
                    const double px = hx, py = hy, pz = hz;
                    const double rj2i = 1. / (px * px + py * py + pz * pz);
                    const double rji = sqrt(rj2i);
                    const double rj3iM = rji * rj2i * Gc * eta_me;
                    const double prefac1 = h * rj3iM;
                    const double rj5M = rj3iM * rj2i;
                    const double rdr = dx * px + dy * py + dz * pz;
                    const double prefac2 = -h * 3. * rdr * rj5M;
                    dvx += prefac1 * dx + prefac2 * px;
                    dvy += prefac1 * dy + prefac2 * py;
                    dvz += prefac1 * dz + prefac2 * pz;
                }
                dvx += h * jax;
                dvy += h * jay;
                dvz += h * jaz;
                vput(P.jv, 3, bi, dvx); vput(P.jv, 4, bi, dvy); vput(P.jv, 5, bi, dvz);
            }
        };
        auto interaction = [&](double h, bool en) {
            if constexpr (VAR) var_interaction(h, en);
            if constexpr (JAC) {
                publish(Slot{ax, ay, az, 0., 0., 0., im});
                double eta = sm;
                double s_x = eta * sax, s_y = eta * say, s_z = eta * saz;
                double jax = 0., jay = 0., jaz = 0.;
                GROUP_LOOP(k, 1, NAw) {
                    const Slot q = slot[gbase + k - 1];
                    const double ei = 1. / eta;
                    eta += q.m;
                    const double pme = eta * ei;
                    const double jx = q.x - s_x * ei, jy = q.y - s_y * ei, jz = q.z - s_z * ei;
                    if (k == me) {
                        jax = jx; jay = jy; jaz = jz;
                    }
                    s_x = s_x * pme + q.m * jx;
                    s_y = s_y * pme + q.m * jy;
                    s_z = s_z * pme + q.m * jz;
                }
                if (me >= NA) {
                    const double ei = 1. / eta;
                    jax = ax - s_x * ei; jay = ay - s_y * ei; jaz = az - s_z * ei;
                }
                if (!live || !en) return;
                const double px = hx, py = hy, pz = hz;
                hvx += h * jax;
                hvy += h * jay;
                hvz += h * jaz;
                if (me > 1 && !P.jacgrav) {   // the Jacobi gravity holds these terms already
                    const double rj2i = 1. / (px * px + py * py + pz * pz);
                    const double rji = sqrt(rj2i);
                    const double rj3iM = rji * rj2i * Gc * eta_me;
                    const double prefac1 = h * rj3iM;
                    hvx += prefac1 * px;
                    hvy += prefac1 * py;
                    hvz += prefac1 * pz;
                }
            } else {
                if (!live || !en) return;
                hvx += h * ax;
                hvy += h * ay;
                hvz += h * az;
            }
        };
        // Symplectic correctors (Jacobi): integrator_whfast.c:653-751 and
        // :753-784.  Every lane runs the chain; en selects the systems.
        auto corrector_Z = [&](double a, double b, bool en) {
            kepler(a, en);
            to_inertial(en, false);
            grav(en);
            interaction(-b, en);
            kepler(-2. * a, en);
            to_inertial(en, false);
            grav(en);
            interaction(b, en);
            kepler(a, en);
        };
        auto apply_corrector = [&](double inv, bool en) {
            int n = 0;
            const CorrZ* c = corrector_chain(P.corrector, n);
            for (int z = 0; z < n; z++) {
                const int k = c[z].ka;
                const double a = (k < 0 ? -kCorrA[-k] : kCorrA[k]) * dt;
                const double b = (c[z].kb * inv) * dt;
                corrector_Z(a, b, en);
            }
        };
        auto operator_C = [&](double a, double b, bool en) {
            kepler(a, en);
            to_inertial(en, false);
            grav(en);
            interaction(b, en);
            kepler(-a, en);
        };
        auto operator_U = [&](double a, double b, bool en) {
            kepler(a, en);
            operator_C(a, b, en);        // Y(a, b)
            operator_C(-a, -b, en);
            operator_C(a, -b, en);       // Y(a, -b)
            operator_C(-a, b, en);
            kepler(-a, en);
        };
        auto apply_corrector2 = [&](double inv, bool en) {
            const double a = 0.5 * inv * dt;
            const double b = kCorr2B * inv * dt;
            operator_U(a, b, en);
            operator_U(-a, b, en);
        };
        // integrator_whfast.c:1051-1111.  Every lane executes the barriers;
        // `want` selects the systems that take part.  The composition kernel
        // drifts 3/8 of a step, the others half a step.
        auto synchronize = [&](bool want) {
            const bool en = want && !sync;
            const double hs = P.kernel == 2 ? 3. * dt / 8. : dt / 2.;
            kepler(hs, en);
            com(hs, en);
            if (JAC && P.corrector2 && __any(en)) apply_corrector2(-1., en);
            if (JAC && P.corrector && __any(en)) apply_corrector(-1., en);
            to_inertial(en, true);
            if (want) sync = 1;
        };
        auto from_inertial_if = [&](bool en) {
            // the inertial -> heliocentric transform writes only when en
            const double sx0 = cx, sy0 = cy, sz0 = cz, svx0 = cvx, svy0 = cvy, svz0 = cvz, sm0 = cm;
            const double hx0 = hx, hy0 = hy, hz0 = hz, hvx0 = hvx, hvy0 = hvy, hvz0 = hvz, hm0 = hm;
            const double eta0 = eta_me;
            from_inertial();
            if (!en) {
                cx = sx0; cy = sy0; cz = sz0; cvx = svx0; cvy = svy0; cvz = svz0; cm = sm0;
                hx = hx0; hy = hy0; hz = hz0; hvx = hvx0; hvy = hvy0; hvz = hvz0; hm = hm0;
                eta_me = eta0;
            }
            if constexpr (VAR) var_from_inertial(en);
        };
        // tools.c:1313-1380 after a batched step (synchronized: safe mode):
        // a set whose largest coordinate exceeds 1e100 is divided by it (the
        // same IEEE divisions), which changes no MEGNO value; the logarithms
        // of the scales are not kept.
        auto var_rescale = [&](bool en) {
            for (int v = 0; v < P.nvar; v++) {
                const size_t b = b0 + (size_t)P.var_index[v];
                __syncthreads();
                if (gl == 0 && sys_ok) {
                    double scale = 0.;
                    for (int i = 0; i < Nw; i++)
                        for (int f = 0; f < 6; f++) {
                            const double a = fabs(vget(P.pv, f, b + (size_t)i));
                            scale = a > scale ? a : scale;
                        }
                    aux0[gbase][0] = scale;
                }
                __syncthreads();
                const double scale = sys_ok ? aux0[gbase][0] : 0.;
                if (!en || !(scale > 1e100)) continue;
                if (live)
                    for (int f = 0; f < 6; f++) vput(P.pv, f, b + me, vget(P.pv, f, b + me) / scale);
                if (gl == 0)
                    for (int f = 0; f < 6; f++) vput(P.pv, f, b, vget(P.pv, f, b) / scale);
            }
        };
        // integrator_whfast.c:1250-1258: the variational accelerations with
        // every term (gravity_ignore_terms NONE), then tools.c:1483-1502 over
        // particles_var[0 .. N-1] by lane 0; the host adds it to the MEGNO sums.
        auto megno_step = [&](bool en) {
            if constexpr (VAR) {
                var_gravity(en, 1, 0);
                __syncthreads();
                if (en && gl == 0 && sys_ok) {
                    double deltad = 0., delta2 = 0.;
                    for (int i = 0; i < Nw; i++) {
                        double q[9];
                        for (int f = 0; f < 9; f++) q[f] = vget(P.pv, f, b0 + (size_t)i);
                        deltad += q[3] * q[0];
                        deltad += q[4] * q[1];
                        deltad += q[5] * q[2];
                        deltad += q[6] * q[3];
                        deltad += q[7] * q[4];
                        deltad += q[8] * q[5];
                        delta2 += q[0] * q[0];
                        delta2 += q[1] * q[1];
                        delta2 += q[2] * q[2];
                        delta2 += q[3] * q[3];
                        delta2 += q[4] * q[4];
                        delta2 += q[5] * q[5];
                    }
                    const double dd = deltad / delta2;
                    if (P.megno == 1) {
                        *P.megno_out = dd;
                    } else {
                        // the system's time and MEGNO sums (tools.c:1504-1522),
                        // t advanced as the reference's two half steps
                        double* M = P.megno_sys + (size_t)sidx * kMegnoFields;
                        double t = M[0];
                        t += dt / 2.;
                        t += dt / 2.;
                        const double t0 = M[1];
                        double Ys = M[2], Yss = M[3], cov = M[4], var = M[5], mean_t = M[6], mean_Y = M[7];
                        const double n = M[8] + 1.;
                        Ys += dt * 2. * (t - t0) * dd;
                        const double Y = Ys / (t - t0);
                        Yss += Y * dt;
                        const double d_t = t - t0 - mean_t;
                        mean_t += d_t / n;
                        const double meg = t == t0 ? 0. : Yss / (t - t0);
                        const double d_Y = meg - mean_Y;
                        mean_Y += d_Y / n;
                        cov += (n - 1.) / n * (t - t0 - mean_t) * (meg - mean_Y);
                        var += (n - 1.) / n * (t - t0 - mean_t) * (t - t0 - mean_t);
                        M[0] = t; M[2] = Ys; M[3] = Yss; M[4] = cov; M[5] = var; M[6] = mean_t; M[7] = mean_Y;
                        M[8] = n;
                    }
                }
            }
        };

        // integrator_whfast.c:1113-1259
        const double h_first = P.kernel == 2 ? 5. * dt / 8. : dt / 2.;
        bool modified = P.did_modify || (sys_ok && (S.req & 2));
        for (long s = 0; s < P.nsteps; s++) {
            const bool act = s < my_steps;
            const bool recalc = act && (safe || modified);
            if (__any(recalc)) {
                if (recalc && sync == 0) warn |= 2;
                synchronize(recalc);
                from_inertial_if(recalc);
            }
            if (JAC && (P.corrector || P.corrector2)) {
                const bool first = act && sync;
                if (P.corrector && __any(first)) apply_corrector(1., first);
                if (P.corrector2 && __any(first)) apply_corrector2(1., first);
            }
            kepler(sync ? h_first : dt, act);
            com(sync ? h_first : dt, act);
            jump(dt / 2., act);
            // Inertial velocities of the middle of a step are visible only
            // after the system's last step of the launch (or are replaced by a
            // synchronization before anything reads them).
            to_inertial(act, act && s == my_steps - 1);
            grav(act);
            if (!JAC || P.kernel == 0) {
                interaction(dt, act);
                jump(dt / 2., act);
            } else if (P.kernel == 1) {
                // integrator_whfast.c:1178-1187
                modified_kick(act);
                interaction(dt, act);
            } else if (P.kernel == 2) {
                // integrator_whfast.c:1189-1218: kicks of -dt/6, dt/6, dt,
                // -dt/6, dt/6 between drifts of -dt/4, dt/8, -dt/8, dt/4,
                // positions and accelerations refreshed before each kick
                const double kick[5] = {-dt / 6., dt / 6., dt, -dt / 6., dt / 6.};
                const double drift[4] = {-dt / 4., dt / 8., -dt / 8., dt / 4.};
#pragma unroll
                for (int q = 0; q < 5; q++) {
                    if (q > 0) {
                        to_inertial(act, false);
                        grav(act);
                    }
                    interaction(kick[q], act);
                    if (q < 4) {
                        kepler(drift[q], act);
                        com(drift[q], act);
                    }
                }
            } else {
                // integrator_whfast.c:1220-1235 (WHT Eq. 10.6): positions
                // moved by dt^2/12 a, accelerations there, then the kick
                if (act && live) {
                    const double pf = dt * dt / 12.;
                    ix += pf * ax;
                    iy += pf * ay;
                    iz += pf * az;
                }
                grav(act);
                interaction(dt, act);
            }
            if (act) sync = 0;
            if (__any(act && safe)) synchronize(act && safe);
            if (VAR && P.megno) megno_step(act);
            if (VAR && P.megno == 2) var_rescale(act);
            if (act) modified = false;
        }
        synchronize(sys_ok && (P.sync_at_end || (S.req & 1)));
        // the system's flags: any of its lanes' warnings and solver branches
        for (int o = 1; o < G; o <<= 1) warn |= __shfl_xor(warn, o);

        // ---- store
        if (live) {
            const size_t b = b0 + me;
            P.body[0 * st + b] = ix; P.body[1 * st + b] = iy; P.body[2 * st + b] = iz;
            P.body[3 * st + b] = ivx; P.body[4 * st + b] = ivy; P.body[5 * st + b] = ivz;
            P.body[6 * st + b] = ax; P.body[7 * st + b] = ay; P.body[8 * st + b] = az;
            P.body[9 * st + b] = im;
            P.ph[0 * st + b] = hx; P.ph[1 * st + b] = hy; P.ph[2 * st + b] = hz;
            P.ph[3 * st + b] = hvx; P.ph[4 * st + b] = hvy; P.ph[5 * st + b] = hvz;
            P.ph[6 * st + b] = hm;
        }
        if (sys_ok && gl == 0) {
            P.body[0 * st + b0] = sx; P.body[1 * st + b0] = sy; P.body[2 * st + b0] = sz;
            P.body[3 * st + b0] = svx; P.body[4 * st + b0] = svy; P.body[5 * st + b0] = svz;
            P.body[6 * st + b0] = sax; P.body[7 * st + b0] = say; P.body[8 * st + b0] = saz;
            P.ph[0 * st + b0] = cx; P.ph[1 * st + b0] = cy; P.ph[2 * st + b0] = cz;
            P.ph[3 * st + b0] = cvx; P.ph[4 * st + b0] = cvy; P.ph[5 * st + b0] = cvz;
            P.ph[6 * st + b0] = cm;
            P.flags[sidx] = (unsigned)(sync & 1) | (unsigned)((warn & 15) << 1);
        }
        __syncthreads();
    }
}

template <int COORD>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_whfast(WHParams P) {
    whfast_body<COORD, false>(P);
}

// Jacobi coordinates without variational sets, two wavefronts per SIMD: the
// registers capped at 256 (spilling to scratch) make every task of a
// 16384-system ensemble resident at once (54 us per step against 73 at one
// wavefront, safe_mode 0; 81 against 116 with safe_mode 1).
#ifndef RB_WHJAC_WAVES
#define RB_WHJAC_WAVES 2
#endif
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RB_WHJAC_WAVES, RB_WHJAC_WAVES))) void
k_whfast_jac(WHParams P) {
    whfast_body<WH_JACOBI, false>(P);
}

// The variational instance (Jacobi), two wavefronts per SIMD: its registers
// are capped at 256 (spilling to scratch) so that every task of a 16384-system
// map is resident at once (150 us per step against 205 at one wavefront).
#ifndef RB_WHVAR_WAVES
#define RB_WHVAR_WAVES 2
#endif
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RB_WHVAR_WAVES, RB_WHVAR_WAVES))) void
k_whfast_var(WHParams P) {
    whfast_body<WH_JACOBI, true>(P);
}

// Energy of each system in the reference's order (tools.c:108-162):
// kinetic over bodies, then potential over pairs (i<j), one thread per system.
__global__ void k_system_energy(WHParams P, double* __restrict__ out, int nsys) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    const WHSys S = P.sys[s];
    const size_t st = P.stride;
    const double* B = P.body;
    const int N = S.N;
    const int NA = S.Nact_grav;
    const int NI = S.tp_type == 0 ? NA : N;
    double e_kin = 0., e_pot = 0.;
    for (int i = 0; i < NI; i++) {
        const size_t b = (size_t)S.body0 + i;
        const double vx = B[3 * st + b], vy = B[4 * st + b], vz = B[5 * st + b], m = B[9 * st + b];
        e_kin += 0.5 * m * (vx * vx + vy * vy + vz * vz);
    }
    for (int i = 0; i < NA; i++) {
        const size_t bi = (size_t)S.body0 + i;
        for (int j = i + 1; j < NI; j++) {
            const size_t bj = (size_t)S.body0 + j;
            const double dx = B[bi] - B[bj];
            const double dy = B[st + bi] - B[st + bj];
            const double dz = B[2 * st + bi] - B[2 * st + bj];
            e_pot -= S.G * B[9 * st + bj] * B[9 * st + bi] / sqrt(dx * dx + dy * dy + dz * dz);
        }
    }
    out[s] = e_kin + e_pot;
}

__global__ void k_reset_counter(unsigned int* c) { *c = 0u; }

int whfast_launch(const WHParams& P, hipStream_t st) {
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, st, P.counter);
    if (P.coords == WH_JACOBI && P.nvar > 0)
        hipLaunchKernelGGL(k_whfast_var, dim3(P.grid), dim3(64), 0, st, P);
    else if (P.coords == WH_JACOBI)
        hipLaunchKernelGGL(k_whfast_jac, dim3(P.grid), dim3(64), 0, st, P);
    else
        hipLaunchKernelGGL(k_whfast<WH_DH>, dim3(P.grid), dim3(64), 0, st, P);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int whfast_energy_launch(const WHParams& P, double* out, int nsys, hipStream_t st) {
    hipLaunchKernelGGL(k_system_energy, dim3((nsys + 63) / 64), dim3(64), 0, st, P, out, nsys);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// Lanes per system with K planets: one per planet, so a wavefront packs
// floor(64 / K) systems of one planet count.
int whfast_group_size(int K) { return K < 1 ? 1 : (K > 64 ? 64 : K); }

// Persistent workgroups (one wavefront each) that fit on the device at once.
int whfast_max_resident(int device) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_whfast<WH_DH>, 64, 0) != hipSuccess || per_cu < 1)
        per_cu = 8;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1)
        cus = 256;
    return per_cu * cus;
}

// ---------------------------------------------------------------- per-simulation plugin
struct WHFastDevice {
    DevBuf ph;       // 7 * cap doubles (p_h SoA, same stride as the particle block)
    DevBuf meta;     // WHSys, WHTask, order, flags, counter
    DevBuf var;      // variational sets: pv and jv (10 * N_var doubles each), set offsets, MEGNO value
    size_t cap = 0;
};

void whfast_release(struct reb_simulation* r) {
    DeviceState* d = dev_of(r);
    if (d->whfast) {
        d->whfast->ph.release();
        d->whfast->meta.release();
        d->whfast->var.release();
        delete d->whfast;
        d->whfast = nullptr;
    }
}

// integrator_whfast.c:876-947, restricted to what this engine implements.
static bool whfast_check(struct reb_simulation* r, struct reb_integrator_whfast_state* s) {
    if (r->N_var) {
        // integrator_whfast.c:877-910
        if (s->coordinates != REB_INTEGRATOR_WHFAST_COORDINATES_JACOBI) {
            reb_simulation_error(r, "Variational particles are only compatible with Jacobi coordinates.");
            return false;
        }
        if (s->kernel != REB_INTEGRATOR_WHFAST_KERNEL_DEFAULT) {
            reb_simulation_error(r, "Variational particles are only compatible with the standard kernel.");
            return false;
        }
        if (s->corrector2) {
            reb_simulation_error(r, "Variational particles not compatible with 2nd corrector.");
            return false;
        }
        if (s->safe_mode == 0 && r->calculate_megno) {
            reb_simulation_error(r, "MEGNO is not compatible with WHFast's safe_mode=0.");
            return false;
        }
        if (s->N_allocated_var != r->N_var) {
            s->p_jh_var = (struct reb_particle*)realloc(s->p_jh_var, sizeof(struct reb_particle) * r->N_var);
            for (size_t i = s->N_allocated_var; i < r->N_var; i++) memset(&s->p_jh_var[i], 0, sizeof(struct reb_particle));
            s->N_allocated_var = r->N_var;
        }
        for (size_t v = 0; v < r->N_var_config; v++) {
            const struct reb_variational_configuration& vc = r->var_config[v];
            if (vc.order != 1) {
                reb_simulation_error(r, "WHFast only supports first order variational equations.");
                return false;
            }
            if (vc.testparticle >= 0) {
                reb_simulation_error(r, "Test particle variations not supported with WHFast. Use IAS15.");
                return false;
            }
            // the kernel reads entries index .. index + N - 1 of the set
            if (vc.index < 0 || (size_t)vc.index + r->N > r->N_var) {
                reb_simulation_error(r, "A set of variational particles does not match the number of bodies.");
                return false;
            }
        }
        // simulation.c:673-682: variational accelerations of BASIC gravity only
        if (r->gravity != REB_GRAVITY_BASIC) {
            reb_simulation_error(r, "Variational gravity calculation not implemented in selected gravity module. Please use REB_GRAVITY_BASIC.");
            return false;
        }
    }
    if (s->coordinates != REB_INTEGRATOR_WHFAST_COORDINATES_DEMOCRATICHELIOCENTRIC &&
        s->coordinates != REB_INTEGRATOR_WHFAST_COORDINATES_JACOBI) {
        reb_simulation_error(r, "This engine implements WHFast in Jacobi and democratic heliocentric coordinates.");
        return false;
    }
    if (s->kernel != REB_INTEGRATOR_WHFAST_KERNEL_DEFAULT && s->coordinates != REB_INTEGRATOR_WHFAST_COORDINATES_JACOBI) {
        reb_simulation_error(r, "Non-standard kernel requires Jacobi coordinates.");
        return false;
    }
    if (s->kernel > 3) {
        reb_simulation_error(r, "Kernel method must be 0 (default), 1 (exact modified kick), 2 (composition kernel), or 3 (lazy implementer's modified kick). ");
        return false;
    }
    if (s->corrector != 0 && s->coordinates != REB_INTEGRATOR_WHFAST_COORDINATES_JACOBI) {
        reb_simulation_error(r, "Symplectic correctors are only compatible with Jacobi and Barycentric coordinates.");
        return false;
    }
    if (s->corrector != 0 && s->corrector != 3 && s->corrector != 5 && s->corrector != 7 && s->corrector != 11 &&
        s->corrector != 17) {
        reb_simulation_error(r, "First symplectic correctors are only available in the following orders: 0, 3, 5, 7, 11, 17.");
        return false;
    }
    if (s->corrector2 != 0 && s->coordinates != REB_INTEGRATOR_WHFAST_COORDINATES_JACOBI) {
        // the reference's second corrector runs the Jacobi operators whatever the coordinates
        reb_simulation_error(r, "The second symplectic corrector is supported with Jacobi coordinates only.");
        return false;
    }
    if (s->keep_unsynchronized == 1 && s->safe_mode == 1)
        reb_simulation_error(r, "whfast->keep_unsynchronized == 1 is not compatible with safe_mode. Must set whfast->safe_mode = 0.");
    if (s->keep_unsynchronized) {
        reb_simulation_error(r, "keep_unsynchronized is not supported by this engine.");
        return false;
    }
    const bool jac = s->coordinates == REB_INTEGRATOR_WHFAST_COORDINATES_JACOBI;
    // integrator_whfast.c:929-939: the modified-kick and lazy kernels switch
    // the simulation to the Jacobi gravity
    if (s->kernel == 1 || s->kernel == 3) r->gravity = REB_GRAVITY_JACOBI;
    if (r->gravity != REB_GRAVITY_BASIC && r->gravity != REB_GRAVITY_COMPENSATED &&
        !(jac && r->gravity == REB_GRAVITY_JACOBI)) {
        reb_simulation_error(r, "WHFast on this engine evaluates the interaction with BASIC or COMPENSATED gravity "
                                "(or the Jacobi gravity in Jacobi coordinates).");
        return false;
    }
    if (r->additional_forces || dev_of(r)->dev_forces) {
        // the interaction runs inside the group kernel, which has no host callbacks
        reb_simulation_error(r, "additional_forces and device forces are not supported with WHFast on this engine.");
        return false;
    }
    if (r->N < 1) return false;
    if (r->N - 1 > 64) {
        reb_simulation_error(r, "WHFast on this engine handles up to 64 bodies besides the central one per simulation.");
        return false;
    }
    if (r->N_ghost_x || r->N_ghost_y || r->N_ghost_z) {
        reb_simulation_error(r, "Ghost boxes are not supported with WHFast.");
        return false;
    }
    if (s->kernel != 1 && s->kernel != 3)
        r->gravity_ignore_terms = jac ? REB_GRAVITY_IGNORE_TERMS_BETWEEN_0_AND_1 : REB_GRAVITY_IGNORE_TERMS_INVOLVING_0;
    if (s->N_allocated != r->N) {
        s->N_allocated = r->N;
        s->p_jh = (struct reb_particle*)realloc(s->p_jh, sizeof(struct reb_particle) * r->N);
        memset(s->p_jh, 0, sizeof(struct reb_particle) * r->N);
        r->did_modify_particles = 1;
    }
    return true;
}

static bool whfast_prepare(struct reb_simulation* r, WHParams& P) {
    DeviceState* d = dev_of(r);
    if (!ensure_device(r)) return false;
    if (!d->whfast) d->whfast = new WHFastDevice();
    WHFastDevice* w = d->whfast;
    if (w->cap != d->cap) {
        const bool fresh = w->cap == 0;
        if (!w->ph.ensure(sizeof(double) * 7 * d->cap)) return false;
        if (!RB_HIP(r, hipMemsetAsync(w->ph.p, 0, sizeof(double) * 7 * d->cap, d->stream))) return false;
        w->cap = d->cap;
        auto* s = (struct reb_integrator_whfast_state*)r->integrator.state;
        if (fresh && s && s->p_jh && s->N_allocated == r->N && r->N > 0 && !r->did_modify_particles) {
            // The integrator state holds p_jh of its own (a simulation read
            // from a binary file): continue from it, as the reference does.
            std::vector<double> h(7 * (size_t)r->N);
            for (size_t i = 0; i < r->N; i++) {
                const struct reb_particle& q = s->p_jh[i];
                const double f[7] = {q.x, q.y, q.z, q.vx, q.vy, q.vz, q.m};
                for (int k = 0; k < 7; k++) h[(size_t)k * r->N + i] = f[k];
            }
            for (int k = 0; k < 7; k++)
                if (!RB_HIP(r, hipMemcpyAsync(w->ph.as<double>() + (size_t)k * d->cap, h.data() + (size_t)k * r->N,
                                              sizeof(double) * r->N, hipMemcpyHostToDevice, d->stream)))
                    return false;
            if (!stream_sync(r)) return false;
        } else {
            r->did_modify_particles = 1;
        }
    }
    const size_t meta_bytes = sizeof(WHSys) + sizeof(WHTask) + 2 * sizeof(int) + 2 * sizeof(unsigned int) + 64;
    if (!w->meta.ensure(meta_bytes)) return false;
    char* base = w->meta.as<char>();
    WHSys sys;
    sys.body0 = 0;
    sys.N = (int)r->N;
    const size_t NA = (r->N_active == SIZE_MAX || r->testparticle_type == 1) ? r->N : r->N_active;
    sys.Nact = (int)NA;
    sys.Nact_grav = (int)((r->N_active == SIZE_MAX) ? r->N : r->N_active);
    sys.tp_type = r->testparticle_type;
    sys.steps = -1;
    sys.req = 0;
    sys.pad = 0;
    sys.G = r->G;
    sys.dt = r->dt;
    WHTask task{whfast_group_size((int)r->N - 1), 0, 1};
    struct Meta {
        WHSys sys;
        WHTask task;
        int order;
        unsigned int flags;
        unsigned int counter;
    } h;
    h.sys = sys;
    h.task = task;
    h.order = 0;
    h.flags = r->is_synchronized ? 1u : 0u;
    h.counter = 0;
    // Small, per-call metadata; the copy is ordered on the stream before the launch.
    static thread_local Meta staging;
    if (!stream_sync(r)) return false;
    staging = h;
    if (!RB_HIP(r, hipMemcpyAsync(base, &staging, sizeof(Meta), hipMemcpyHostToDevice, d->stream))) return false;
    Meta* dm = reinterpret_cast<Meta*>(base);
    P.sys = &dm->sys;
    P.tasks = &dm->task;
    P.order = &dm->order;
    P.flags = &dm->flags;
    P.counter = &dm->counter;
    P.ntasks = 1;
    P.grid = 1;
    P.body = d->soa.as<double>();
    P.ph = w->ph.as<double>();
    P.stride = d->cap;
    P.G = r->G;
    P.soft2 = r->softening * r->softening;
    P.compensated = r->gravity == REB_GRAVITY_COMPENSATED ? 1 : 0;
    return true;
}

static bool whfast_finish(struct reb_simulation* r, struct reb_integrator_whfast_state* s, const WHParams& P) {
    DeviceState* d = dev_of(r);
    unsigned int flags = 0;
    if (!RB_HIP(r, hipMemcpyAsync(&flags, P.flags, sizeof(flags), hipMemcpyDeviceToHost, d->stream))) return false;
    if (!stream_sync(r)) return false;
    r->is_synchronized = flags & 1u;
    if ((flags & 2u) && !(r->messages_timestep_warning & 1)) {
        r->messages_timestep_warning |= 1;
        reb_simulation_warning(r, "Possible convergence issue. Timestep in Kepler solver is larger than one orbital period.");
    }
    if ((flags & 4u) && s->recalculate_coordinates_but_not_synchronized_warning == 0) {
        reb_simulation_warning(r, "Particles were modified while simulation was not synchronized.");
        s->recalculate_coordinates_but_not_synchronized_warning++;
    }
    mark_device_newer(r);
    return true;
}

// Refresh the public p_jh copy (heliocentric state) from the device.
static void whfast_mirror_pjh(struct reb_simulation* r, struct reb_integrator_whfast_state* s) {
    DeviceState* d = dev_of(r);
    if (!d->whfast || !s->p_jh || r->N == 0) return;
    std::vector<double> h(7 * r->N);
    for (int f = 0; f < 7; f++)
        (void)hipMemcpyAsync(h.data() + f * r->N, d->whfast->ph.as<double>() + (size_t)f * d->cap,
                             sizeof(double) * r->N, hipMemcpyDeviceToHost, d->stream);
    (void)hipStreamSynchronize(d->stream);
    for (size_t i = 0; i < r->N; i++) {
        s->p_jh[i].x = h[0 * r->N + i];
        s->p_jh[i].y = h[1 * r->N + i];
        s->p_jh[i].z = h[2 * r->N + i];
        s->p_jh[i].vx = h[3 * r->N + i];
        s->p_jh[i].vy = h[4 * r->N + i];
        s->p_jh[i].vz = h[5 * r->N + i];
        s->p_jh[i].m = h[6 * r->N + i];
    }
}

// The host copy of p_jh made current (before a binary snapshot is written).
void whfast_refresh_host(struct reb_simulation* r) {
    auto* s = (struct reb_integrator_whfast_state*)r->integrator.state;
    DeviceState* d = dev_of(r);
    if (!s || !d->whfast || !d->whfast->cap) return;
    whfast_mirror_pjh(r, s);
}

// The variational sets go to the device before each launch and come back
// after it (r->particles_var and the integrator's p_jh_var stay the host's:
// the sets are small and the plugin steps once per launch).  Layout: pv then
// jv, 10 fields of N_var doubles each (x y z vx vy vz ax ay az m), then the
// sets' offsets and the MEGNO value.
static bool var_upload(struct reb_simulation* r, struct reb_integrator_whfast_state* s, WHParams& P) {
    DeviceState* d = dev_of(r);
    WHFastDevice* w = d->whfast;
    const size_t nv = r->N_var;
    const size_t nd = 20 * nv + 1 + (r->N_var_config + 1) / 2;
    if (!w->var.ensure(sizeof(double) * nd)) {
        reb_simulation_error(r, "Out of device memory for the variational particles.");
        return false;
    }
    std::vector<double> h(nd, 0.);
    auto pack = [&](const struct reb_particle* q, double* o) {
        for (size_t i = 0; i < nv; i++) {
            const double f[10] = {q[i].x, q[i].y, q[i].z, q[i].vx, q[i].vy, q[i].vz, q[i].ax, q[i].ay, q[i].az, q[i].m};
            for (int k = 0; k < 10; k++) o[(size_t)k * nv + i] = f[k];
        }
    };
    pack(r->particles_var, h.data());
    pack(s->p_jh_var, h.data() + 10 * nv);
    int* vi = reinterpret_cast<int*>(h.data() + 20 * nv + 1);
    for (size_t v = 0; v < r->N_var_config; v++) vi[v] = r->var_config[v].index;
    if (!RB_HIP(r, hipMemcpyAsync(w->var.p, h.data(), sizeof(double) * nd, hipMemcpyHostToDevice, d->stream)))
        return false;
    double* base = w->var.as<double>();
    P.nvar = (int)r->N_var_config;
    P.var_index = reinterpret_cast<const int*>(base + 20 * nv + 1);
    P.pv = base;
    P.jv = base + 10 * nv;
    P.vstride = nv;
    P.megno = r->calculate_megno ? 1 : 0;
    P.megno_out = base + 20 * nv;
    return stream_sync(r);
}

static bool var_download(struct reb_simulation* r, struct reb_integrator_whfast_state* s, double* megno_dd) {
    DeviceState* d = dev_of(r);
    const size_t nv = r->N_var;
    std::vector<double> h(20 * nv + 1);
    if (!RB_HIP(r, hipMemcpyAsync(h.data(), d->whfast->var.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                                  d->stream)) ||
        !stream_sync(r))
        return false;
    auto unpack = [&](const double* o, struct reb_particle* q) {
        for (size_t i = 0; i < nv; i++) {
            q[i].x = o[0 * nv + i]; q[i].y = o[1 * nv + i]; q[i].z = o[2 * nv + i];
            q[i].vx = o[3 * nv + i]; q[i].vy = o[4 * nv + i]; q[i].vz = o[5 * nv + i];
            q[i].ax = o[6 * nv + i]; q[i].ay = o[7 * nv + i]; q[i].az = o[8 * nv + i];
            q[i].m = o[9 * nv + i];
        }
    };
    unpack(h.data(), r->particles_var);
    unpack(h.data() + 10 * nv, s->p_jh_var);
    *megno_dd = h[20 * nv];
    return true;
}

// n steps in one launch (n == 1 for the per-step plugin call).
bool whfast_run(struct reb_simulation* r, struct reb_integrator_whfast_state* s, long nsteps, int sync_at_end,
                double* megno_dd = nullptr) {
    if (!whfast_check(r, s)) return false;
    WHParams P;
    if (!whfast_prepare(r, P)) return false;
    if (r->N_var && !var_upload(r, s, P)) return false;
    P.nsteps = nsteps;
    P.safe_mode = s->safe_mode ? 1 : 0;
    P.did_modify = r->did_modify_particles ? 1 : 0;
    P.sync_at_end = sync_at_end;
    P.coords = s->coordinates == REB_INTEGRATOR_WHFAST_COORDINATES_JACOBI ? WH_JACOBI : WH_DH;
    P.corrector = (int)s->corrector;
    P.corrector2 = s->corrector2 ? 1 : 0;
    P.kernel = (int)s->kernel;
    P.jacgrav = r->gravity == REB_GRAVITY_JACOBI ? 1 : 0;
    if (whfast_launch(P, dev_of(r)->stream)) return RB_HIP(r, hipGetLastError());
    if (!whfast_finish(r, s, P)) return false;
    double dd = 0.;
    if (r->N_var && !var_download(r, s, &dd)) return false;
    if (megno_dd) *megno_dd = dd;
    return true;
}

void whfast_step_device(struct reb_simulation* r, struct reb_integrator_whfast_state* s) {
    double dd = 0.;
    if (!whfast_run(r, s, 1, 0, &dd)) return;
    r->t += r->dt / 2.;
    r->t += r->dt / 2.;
    r->dt_last_done = r->dt;
    if (r->calculate_megno && r->N_var) {
        // integrator_whfast.c:1250-1258 (the device formed deltad/delta2);
        // the reference leaves gravity_ignore_terms at INVOLVING_0 afterwards
        r->gravity_ignore_terms = REB_GRAVITY_IGNORE_TERMS_INVOLVING_0;
        const double dY = r->dt * 2. * (r->t - r->megno_initial_t) * dd;
        megno_update(r, dY, r->dt);
    }
}

void whfast_synchronize_device(struct reb_simulation* r, struct reb_integrator_whfast_state* s) {
    if (r->is_synchronized) return;
    if (!whfast_run(r, s, 0, 1)) return;
    whfast_mirror_pjh(r, s);
}

}  // namespace rbamd

using namespace rbamd;

extern "C" {

void* rb_whfast_create(void) {
    auto* s = (struct reb_integrator_whfast_state*)calloc(1, sizeof(struct reb_integrator_whfast_state));
    s->coordinates = REB_INTEGRATOR_WHFAST_COORDINATES_JACOBI;
    s->safe_mode = 1;
    return s;
}

void rb_whfast_free(void* p) {
    auto* s = (struct reb_integrator_whfast_state*)p;
    free(s->p_jh);
    free(s->p_jh_var);
    free(s);
}

void rb_whfast_step(struct reb_simulation* r, void* state) {
    whfast_step_device(r, (struct reb_integrator_whfast_state*)state);
}

void rb_whfast_synchronize(struct reb_simulation* r, void* state) {
    auto* s = (struct reb_integrator_whfast_state*)state;
    if (!whfast_check(r, s)) return;
    whfast_synchronize_device(r, s);
}

REB_API int reb_amd_kepler_solver_batch(double* state6, const double* mu, const double* dt, size_t n) {
    if (n == 0) return 0;
    if (reb_amd_device_count() <= 0) {
        set_last_error("reb_amd_kepler_solver_batch: no HIP device");
        return 1;
    }
    double *d6 = nullptr, *dmu = nullptr, *ddt = nullptr;
    int* dflag = nullptr;
    int rc = 0;
    if (hipMalloc(&d6, sizeof(double) * 6 * n) != hipSuccess || hipMalloc(&dmu, sizeof(double) * n) != hipSuccess ||
        hipMalloc(&ddt, sizeof(double) * n) != hipSuccess || hipMalloc(&dflag, sizeof(int)) != hipSuccess) {
        rc = 2;
    } else {
        rc |= hipMemcpy(d6, state6, sizeof(double) * 6 * n, hipMemcpyHostToDevice) != hipSuccess;
        rc |= hipMemcpy(dmu, mu, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess;
        rc |= hipMemcpy(ddt, dt, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess;
        rc |= hipMemset(dflag, 0, sizeof(int)) != hipSuccess;
        hipLaunchKernelGGL(k_kepler_batch, dim3((n + 63) / 64), dim3(64), 0, 0, d6, dmu, ddt, (int)n, dflag);
        rc |= hipGetLastError() != hipSuccess;
        rc |= hipMemcpy(state6, d6, sizeof(double) * 6 * n, hipMemcpyDeviceToHost) != hipSuccess;
    }
    if (d6) (void)hipFree(d6);
    if (dmu) (void)hipFree(dmu);
    if (ddt) (void)hipFree(ddt);
    if (dflag) (void)hipFree(dflag);
    if (rc) set_last_error("reb_amd_kepler_solver_batch: HIP failure");
    return rc;
}

REB_API void reb_integrator_whfast_kepler_solver(struct reb_particle* const p, double mu, double dt,
                                                 const struct reb_simulation* const r) {
    double s[6] = {p->x, p->y, p->z, p->vx, p->vy, p->vz};
    if (reb_amd_kepler_solver_batch(s, &mu, &dt, 1)) {
        if (r) reb_simulation_error((struct reb_simulation*)r, "Kepler solver: no HIP device.");
        return;
    }
    p->x = s[0];
    p->y = s[1];
    p->z = s[2];
    p->vx = s[3];
    p->vy = s[4];
    p->vz = s[5];
}

}  // extern "C"
