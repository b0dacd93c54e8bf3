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

// 1/k! as the reference spells them (integrator_whfast.c:158); only 0..13 are used.
__constant__ double kInvFact[14] = {1.,
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
                                    1. / 6227020800.};

__device__ __forceinline__ double fastabs(double x) { return (x > 0.) ? x : -x; }

// c_0..c_3 Stumpff functions of z by series + quartering (integrator_whfast.c:189-213).
__device__ __forceinline__ void stumpff_cs3(double* cs, double z) {
    // A finite z needs at most ~520 quarterings; the cap only stops a
    // non-finite z from spinning forever (the result is then NaN anyway).
    unsigned int n = 0;
    while (fabs(z) > 0.1 && n < 1100) {
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

// Universal-variable Kepler drift of one body (integrator_whfast.c:241-390),
// in place.  Returns a bit mask: 1 = the step exceeded one orbital period
// (the reference warns once).
__device__ int kepler_solver(double& px, double& py, double& pz, double& pvx, double& pvy, double& pvz, double mu,
                             double dt) {
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
        // Quartic (Laguerre-Conway) iteration with cycle detection (:289-312).
        X = beta * dt / mu;
        double prevX[64 + 1];
        for (int n_lag = 1; n_lag < 64; n_lag++) {
            stiefel_Gs3(Gs, beta, X);
            const double f = r0 * X + eta0 * Gs[2] + zeta0 * Gs[3] - dt;
            const double fp = r0 + eta0 * Gs[1] + zeta0 * Gs[2];
            const double fpp = eta0 * Gs[0] + zeta0 * Gs[1];
            const double denom = fp + sqrt(fabs(16. * fp * fp - 20. * f * fpp));
            X = (X * denom - 5. * f) / denom;
            for (int i = 1; i < n_lag; i++) {
                if (X == prevX[i]) {
                    n_lag = 64;
                    converged = 1;
                    break;
                }
            }
            prevX[n_lag] = X;
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
    return flags;
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

// One wavefront per workgroup and rarely more than two per SIMD (an ensemble of
// 16384 systems is 1317 wave tasks): the register budget allows two waves per
// SIMD, so a step's whole state (about 40 doubles per lane plus the Kepler
// solver's temporaries) stays in VGPRs without spills.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_whfast_dh(WHParams P) {
    __shared__ Slot slot[64];
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
        const size_t st = P.stride;
        const size_t b0 = (size_t)S.body0;
        // ---- load
        double ix = 0, iy = 0, iz = 0, ivx = 0, ivy = 0, ivz = 0, im = 0;   // inertial planet
        double hx = 0, hy = 0, hz = 0, hvx = 0, hvy = 0, hvz = 0;           // heliocentric planet
        double sx = 0, sy = 0, sz = 0, svx = 0, svy = 0, svz = 0, sm = 0;   // inertial star
        double cx = 0, cy = 0, cz = 0, cvx = 0, cvy = 0, cvz = 0, cm = 0;   // p_h[0] (barycentre)
        double ax = 0, ay = 0, az = 0;
        if (sys_ok) {
            sx = P.body[0 * st + b0]; sy = P.body[1 * st + b0]; sz = P.body[2 * st + b0];
            svx = P.body[3 * st + b0]; svy = P.body[4 * st + b0]; svz = P.body[5 * st + b0];
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
        double hm = im;   // p_h[i].m
        const unsigned int flags0 = sys_ok ? P.flags[sidx] : 1u;
        int sync = flags0 & 1u;
        int warn = (flags0 >> 1) & 3u;        // warnings stay set until the host clears them
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
        auto publish_inertial = [&]() { publish(Slot{ix, iy, iz, ivx, ivy, ivz, im}); };
        // transformations.c:388-424 (inertial -> democratic heliocentric)
        auto from_inertial = [&]() {
            publish(Slot{ix * im, iy * im, iz * im, ivx * im, ivy * im, ivz * im, im});
            double x0 = 0., y0 = 0., z0 = 0., vx0 = 0., vy0 = 0., vz0 = 0., m0 = 0.;
            if (0 < NA) {
                const double m = sm;
                x0 += sx * m; y0 += sy * m; z0 += sz * m;
                vx0 += svx * m; vy0 += svy * m; vz0 += svz * m;
                m0 += m;
            }
            GROUP_LOOP(k, 1, NA) {
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
        };
        // transformations.c:426-477 (democratic heliocentric -> inertial,
        // positions then velocities).  Barriers are executed by every lane;
        // lanes whose system does not take part (en == false) keep their state.
        // The velocity pass is skipped (with_vel == false) where nothing reads
        // the inertial velocities before they are recomputed.
        auto to_inertial = [&](bool en, bool with_vel) {
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
            GROUP_LOOP(k, 1, NA) {
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
        };
        // integrator_whfast.c:612-617 and :638-651
        auto kepler_com = [&](double h, bool en) {
            if (!en) return;
            if (live) warn |= kepler_solver(hx, hy, hz, hvx, hvy, hvz, sm * Gc, h);
            cx += h * cvx;
            cy += h * cvy;
            cz += h * cvz;
        };
        // integrator_whfast.c:537-554 (masses are the inertial ones, equal to
        // p_h[i].m here: the engine changes no masses)
        auto jump = [&](double h, bool en) {
            publish(Slot{hm * hvx, hm * hvy, hm * hvz, 0., 0., 0., 0.});
            double qx = 0, qy = 0, qz = 0;
            GROUP_LOOP(k, 1, NA) {
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
        };
        // BASIC (gravity.c:215-232) or COMPENSATED (:308-410) with
        // gravity_ignore_terms = INVOLVING_0: planets only, ascending j.
        auto gravity = [&](bool en) {
            publish_inertial();
            if (!live || !en) return;
            ax = 0.; ay = 0.; az = 0.;
            double kx = 0., ky = 0., kz = 0.;
            for (int pass = 0; pass < 2; pass++) {
                const int j0 = pass == 0 ? 1 : (NAG > 1 ? NAG : 1);
                const int j1 = pass == 0 ? NAG : N;
                if (pass == 1 && !(S.tp_type && me < NAG)) break;
                for (int j = j0; j < j1; j++) {
                    if (j == me) continue;
                    const Slot q = slot[gbase + j - 1];
                    const double dx = ix - q.x;
                    const double dy = iy - q.y;
                    const double dz = iz - q.z;
                    if (P.compensated) {
                        const double r2 = dx * dx + dy * dy + dz * dz + P.soft2;
                        const double rr = sqrt(r2);
                        const double prefact = Gc / (r2 * rr);
                        const double prefactj = -prefact * q.m;
                        double t, yv;
                        yv = prefactj * dx - kx; t = ax + yv; kx = (t - ax) - yv; ax = t;
                        yv = prefactj * dy - ky; t = ay + yv; ky = (t - ay) - yv; ay = t;
                        yv = prefactj * dz - kz; t = az + yv; kz = (t - az) - yv; az = t;
                    } else {
                        const double _r = sqrt(dx * dx + dy * dy + dz * dz + P.soft2);
                        const double prefact = -Gc / (_r * _r * _r) * q.m;
                        ax += prefact * dx;
                        ay += prefact * dy;
                        az += prefact * dz;
                    }
                }
            }
        };
        // integrator_whfast.c:1051-1111 (default kernel, no correctors).  Every
        // lane executes the barriers; `want` selects the systems that take part.
        auto synchronize = [&](bool want) {
            const bool en = want && !sync;
            kepler_com(dt / 2., en);
            to_inertial(en, true);
            if (want) sync = 1;
        };
        auto from_inertial_if = [&](bool en) {
            // the inertial -> heliocentric transform writes only when en
            const double sx0 = cx, sy0 = cy, sz0 = cz, svx0 = cvx, svy0 = cvy, svz0 = cvz, sm0 = cm;
            const double hx0 = hx, hy0 = hy, hz0 = hz, hvx0 = hvx, hvy0 = hvy, hvz0 = hvz, hm0 = hm;
            from_inertial();
            if (!en) {
                cx = sx0; cy = sy0; cz = sz0; cvx = svx0; cvy = svy0; cvz = svz0; cm = sm0;
                hx = hx0; hy = hy0; hz = hz0; hvx = hvx0; hvy = hvy0; hvz = hvz0; hm = hm0;
            }
        };

        // integrator_whfast.c:1113-1259 (DH, default kernel, no correctors)
        bool modified = P.did_modify || (sys_ok && (S.req & 2));
        for (long s = 0; s < P.nsteps; s++) {
            const bool act = s < my_steps;
            const bool recalc = act && (safe || modified);
            if (__any(recalc)) {
                if (recalc && sync == 0) warn |= 2;
                synchronize(recalc);
                from_inertial_if(recalc);
            }
            kepler_com(sync ? dt / 2. : dt, act);
            jump(dt / 2., act);
            // Inertial velocities of the middle of a step are visible only
            // after the system's last step of the launch (or are replaced by a
            // synchronization before anything reads them).
            to_inertial(act, act && s == my_steps - 1);
            gravity(act);
            if (live && act) {
                hvx += dt * ax;
                hvy += dt * ay;
                hvz += dt * az;
            }
            jump(dt / 2., act);
            if (act) sync = 0;
            if (__any(act && safe)) synchronize(act && safe);
            if (act) modified = false;
        }
        synchronize(sys_ok && (P.sync_at_end || (S.req & 1)));

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
            P.body[6 * st + b0] = 0.; P.body[7 * st + b0] = 0.; P.body[8 * st + b0] = 0.;
            P.ph[0 * st + b0] = cx; P.ph[1 * st + b0] = cy; P.ph[2 * st + b0] = cz;
            P.ph[3 * st + b0] = cvx; P.ph[4 * st + b0] = cvy; P.ph[5 * st + b0] = cvz;
            P.ph[6 * st + b0] = cm;
            P.flags[sidx] = (unsigned)(sync & 1) | (unsigned)((warn & 3) << 1);
        }
        __syncthreads();
    }
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
    hipLaunchKernelGGL(k_whfast_dh, dim3(P.grid), dim3(64), 0, st, P);
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
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_whfast_dh, 64, 0) != hipSuccess || per_cu < 1)
        per_cu = 8;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1)
        cus = 256;
    return per_cu * cus;
}

// ---------------------------------------------------------------- per-simulation plugin
struct WHFastDevice {
    DevBuf ph;       // 7 * cap doubles (p_h SoA, same stride as the particle block)
    DevBuf meta;     // WHSys, WHTask, order, flags, counter
    size_t cap = 0;
};

void whfast_release(struct reb_simulation* r) {
    DeviceState* d = dev_of(r);
    if (d->whfast) {
        d->whfast->ph.release();
        d->whfast->meta.release();
        delete d->whfast;
        d->whfast = nullptr;
    }
}

// integrator_whfast.c:876-947, restricted to what this engine implements.
static bool whfast_check(struct reb_simulation* r, struct reb_integrator_whfast_state* s) {
    if (r->N_var) {
        reb_simulation_error(r, "Variational particles are not supported by this engine.");
        return false;
    }
    if (s->coordinates != REB_INTEGRATOR_WHFAST_COORDINATES_DEMOCRATICHELIOCENTRIC) {
        reb_simulation_error(r, "This engine implements WHFast in democratic heliocentric coordinates "
                                "(set coordinates = REB_INTEGRATOR_WHFAST_COORDINATES_DEMOCRATICHELIOCENTRIC).");
        return false;
    }
    if (s->kernel != REB_INTEGRATOR_WHFAST_KERNEL_DEFAULT || s->corrector != 0 || s->corrector2 != 0) {
        reb_simulation_error(r, "WHFast kernels and symplectic correctors are not supported by this engine.");
        return false;
    }
    if (s->keep_unsynchronized == 1 && s->safe_mode == 1)
        reb_simulation_error(r, "whfast->keep_unsynchronized == 1 is not compatible with safe_mode. Must set whfast->safe_mode = 0.");
    if (s->keep_unsynchronized) {
        reb_simulation_error(r, "keep_unsynchronized is not supported by this engine.");
        return false;
    }
    if (r->gravity != REB_GRAVITY_BASIC && r->gravity != REB_GRAVITY_COMPENSATED) {
        reb_simulation_error(r, "WHFast on this engine evaluates the interaction with BASIC or COMPENSATED gravity.");
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
    r->gravity_ignore_terms = REB_GRAVITY_IGNORE_TERMS_INVOLVING_0;
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
        if (!w->ph.ensure(sizeof(double) * 7 * d->cap)) return false;
        if (!RB_HIP(r, hipMemsetAsync(w->ph.p, 0, sizeof(double) * 7 * d->cap, d->stream))) return false;
        w->cap = d->cap;
        r->did_modify_particles = 1;
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

// n steps in one launch (n == 1 for the per-step plugin call).
bool whfast_run(struct reb_simulation* r, struct reb_integrator_whfast_state* s, long nsteps, int sync_at_end) {
    if (!whfast_check(r, s)) return false;
    WHParams P;
    if (!whfast_prepare(r, P)) return false;
    P.nsteps = nsteps;
    P.safe_mode = s->safe_mode ? 1 : 0;
    P.did_modify = r->did_modify_particles ? 1 : 0;
    P.sync_at_end = sync_at_end;
    if (whfast_launch(P, dev_of(r)->stream)) return RB_HIP(r, hipGetLastError());
    return whfast_finish(r, s, P);
}

void whfast_step_device(struct reb_simulation* r, struct reb_integrator_whfast_state* s) {
    if (!whfast_run(r, s, 1, 0)) return;
    r->t += r->dt / 2.;
    r->t += r->dt / 2.;
    r->dt_last_done = r->dt;
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
