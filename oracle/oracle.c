/*
 * CPU oracle -- test infrastructure only (see oracle.h).  Restates the
 * reference's arithmetic expression by expression; compiled with
 * -ffp-contract=off like the reference's ISO C build.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define COL(s, i, c) ((s)[(size_t)(i) * ORC_NCOL + (c)])
enum { X = 0, Y, Z, VX, VY, VZ, AX, AY, AZ, M, RR };

/* ------------------------------------------------------------------ RNG
 * glibc rand_r (restated in rebound.c:238-259) and tools.c:64-105 */
int orc_rand_r(unsigned int* seed) {
    unsigned int next = *seed;
    int result;
    next *= 1103515245;
    next += 12345;
    result = (unsigned int)(next / 65536) % 2048;
    next *= 1103515245;
    next += 12345;
    result <<= 10;
    result ^= (unsigned int)(next / 65536) % 1024;
    next *= 1103515245;
    next += 12345;
    result <<= 10;
    result ^= (unsigned int)(next / 65536) % 1024;
    *seed = next;
    return result;
}

double orc_random_uniform(unsigned int* seed, double min, double max) {
    return ((double)orc_rand_r(seed)) / ((double)(2147483647)) * (max - min) + min;
}

double orc_random_powerlaw(unsigned int* seed, double min, double max, double slope) {
    double y = orc_random_uniform(seed, 0., 1.);
    if (slope == -1) return exp(y * log(max / min) + log(min));
    return pow((pow(max, slope + 1.) - pow(min, slope + 1.)) * y + pow(min, slope + 1.), 1. / (slope + 1.));
}

double orc_random_normal(unsigned int* seed, double variance) {
    double v1 = 0., v2 = 0., rsq = 1.;
    while (rsq >= 1. || rsq < 1.0e-12) {
        v1 = 2. * ((double)orc_rand_r(seed)) / ((double)(2147483647)) - 1.0;
        v2 = 2. * ((double)orc_rand_r(seed)) / ((double)(2147483647)) - 1.0;
        rsq = v1 * v1 + v2 * v2;
    }
    return v1 * sqrt(-2. * log(rsq) / rsq * variance);
}

double orc_random_rayleigh(unsigned int* seed, double sigma) {
    double y = orc_random_uniform(seed, 0., 1.);
    return sigma * sqrt(-2 * log(y));
}

/* ------------------------------------------------------------------ inputs */
void orc_plummer(unsigned int* seed, size_t N, double M_, double R, double* out) {
    double E = 3. / 64. * M_PI * M_ * M_ / R;
    for (size_t i = 0; i < N; i++) {
        double* p = out + i * ORC_NCOL;
        memset(p, 0, sizeof(double) * ORC_NCOL);
        double _r = pow(pow(orc_random_uniform(seed, 0, 1), -2. / 3.) - 1., -1. / 2.);
        double x2 = orc_random_uniform(seed, 0, 1);
        double x3 = orc_random_uniform(seed, 0, 2. * M_PI);
        p[Z] = (1. - 2. * x2) * _r;
        p[X] = sqrt(_r * _r - p[Z] * p[Z]) * cos(x3);
        p[Y] = sqrt(_r * _r - p[Z] * p[Z]) * sin(x3);
        double x5, g, q;
        do {
            x5 = orc_random_uniform(seed, 0., 1.);
            q = orc_random_uniform(seed, 0., 1.);
            g = q * q * pow(1. - q * q, 7. / 2.);
        } while (0.1 * x5 > g);
        double ve = pow(2., 1. / 2.) * pow(1. + _r * _r, -1. / 4.);
        double v = q * ve;
        double x6 = orc_random_uniform(seed, 0., 1.);
        double x7 = orc_random_uniform(seed, 0., 2. * M_PI);
        p[VZ] = (1. - 2. * x6) * v;
        p[VX] = sqrt(v * v - p[VZ] * p[VZ]) * cos(x7);
        p[VY] = sqrt(v * v - p[VZ] * p[VZ]) * sin(x7);
        p[X] *= 3. * M_PI / 64. * M_ * M_ / E;
        p[Y] *= 3. * M_PI / 64. * M_ * M_ / E;
        p[Z] *= 3. * M_PI / 64. * M_ * M_ / E;
        p[VX] *= sqrt(E * 64. / 3. / M_PI / M_);
        p[VY] *= sqrt(E * 64. / 3. / M_PI / M_);
        p[VZ] *= sqrt(E * 64. / 3. / M_PI / M_);
        p[M] = M_ / (double)N;
    }
}

void orc_move_to_com(double* s, size_t N) {
    /* reb_simulation_com: pairwise mass-weighted merge in index order (tools.c:370-408) */
    double c[10] = {0};  /* x y z vx vy vz ax ay az m */
    for (size_t i = 0; i < N; i++) {
        const double* p = s + i * ORC_NCOL;
        for (int k = 0; k < 9; k++) c[k] = c[k] * c[9] + p[k] * p[M];
        c[9] += p[M];
        if (c[9] > 0.)
            for (int k = 0; k < 9; k++) c[k] /= c[9];
    }
    for (size_t i = 0; i < N; i++)
        for (int k = 0; k < 6; k++) COL(s, i, k) -= c[k];
}

int orc_particle_from_orbit(double G, const double* pr, double m, double a, double e, double inc, double Omega,
                            double omega, double f, double* o) {
    if (e == 1.) return 1;
    if (e < 0.) return 2;
    if (e > 1.) {
        if (a > 0.) return 3;
    } else {
        if (a < 0.) return 4;
    }
    if (e * cos(f) < -1.) return 5;
    if (pr[6] < 1.E-308) return 6;
    double r = a * (1 - e * e) / (1 + e * cos(f));
    double v0 = sqrt(G * (m + pr[6]) / a / (1. - e * e));
    double cO = cos(Omega), sO = sin(Omega), co = cos(omega), so = sin(omega);
    double cf = cos(f), sf = sin(f), ci = cos(inc), si = sin(inc);
    o[0] = pr[0] + r * (cO * (co * cf - so * sf) - sO * (so * cf + co * sf) * ci);
    o[1] = pr[1] + r * (sO * (co * cf - so * sf) + cO * (so * cf + co * sf) * ci);
    o[2] = pr[2] + r * (so * cf + co * sf) * si;
    o[3] = pr[3] + v0 * ((e + cf) * (-ci * co * sO - cO * so) - sf * (co * cO - ci * so * sO));
    o[4] = pr[4] + v0 * ((e + cf) * (ci * co * cO - sO * so) - sf * (co * sO + ci * so * cO));
    o[5] = pr[5] + v0 * ((e + cf) * co * si - sf * si * so);
    return 0;
}

/* ------------------------------------------------------------------ gravity */
static int skip_pair(size_t i, size_t j, int ignore) {
    if (ignore == 1 && ((j == 1 && i == 0) || (i == 1 && j == 0))) return 1;
    if (ignore == 2 && (j == 0 || i == 0)) return 1;
    return i == j;
}

void orc_gravity_basic(const double* s, size_t N, size_t N_active, int tp_type, int ignore, double G, double soft,
                       const double* gbs, int ngb, double* acc) {
    const double softening2 = soft * soft;
    if (N_active > N) N_active = N;
    for (size_t i = 0; i < N; i++) acc[3 * i] = acc[3 * i + 1] = acc[3 * i + 2] = 0;
    for (int g = 0; g < ngb; g++) {
        const double gx = gbs[3 * g], gy = gbs[3 * g + 1], gz = gbs[3 * g + 2];
        for (size_t i = 0; i < N; i++) {
            for (size_t j = 0; j < N_active; j++) {
                if (skip_pair(i, j, ignore)) continue;
                const double dx = (gx + COL(s, i, X)) - COL(s, j, X);
                const double dy = (gy + COL(s, i, Y)) - COL(s, j, Y);
                const double dz = (gz + COL(s, i, Z)) - COL(s, j, Z);
                const double _r = sqrt(dx * dx + dy * dy + dz * dz + softening2);
                const double prefact = -G / (_r * _r * _r) * COL(s, j, M);
                acc[3 * i] += prefact * dx;
                acc[3 * i + 1] += prefact * dy;
                acc[3 * i + 2] += prefact * dz;
            }
        }
        if (tp_type) {
            for (size_t i = 0; i < N_active; i++) {
                for (size_t j = N_active; j < N; j++) {
                    if (ignore == 1 && (j == 1 && i == 0)) continue;
                    if (ignore == 2 && (j == 0 || i == 0)) continue;
                    const double dx = (gx + COL(s, i, X)) - COL(s, j, X);
                    const double dy = (gy + COL(s, i, Y)) - COL(s, j, Y);
                    const double dz = (gz + COL(s, i, Z)) - COL(s, j, Z);
                    const double _r = sqrt(dx * dx + dy * dy + dz * dz + softening2);
                    const double prefact = -G / (_r * _r * _r) * COL(s, j, M);
                    acc[3 * i] += prefact * dx;
                    acc[3 * i + 1] += prefact * dy;
                    acc[3 * i + 2] += prefact * dz;
                }
            }
        }
    }
}

static void kahan(double term, double* a, double* c) {
    double yv = term - *c;
    double tv = *a + yv;
    *c = (tv - *a) - yv;
    *a = tv;
}

/* The accelerations of the target rows rows[0..nrows) alone, with the
 * arithmetic and order of orc_gravity_basic (compensated = 0; ghost boxes
 * gbs) or orc_gravity_compensated (compensated = 1; no ghost boxes): for
 * checking large N against a sample of targets. */
void orc_gravity_rows(const double* s, size_t N, size_t N_active, int tp_type, int ignore, double G, double soft,
                      const double* gbs, int ngb, int compensated, const int64_t* rows, size_t nrows, double* acc) {
    const double softening2 = soft * soft;
    if (N_active > N) N_active = N;
    if (compensated) ngb = 1;
    for (size_t q = 0; q < nrows; q++) {
        const size_t i = (size_t)rows[q];
        double a[3] = {0., 0., 0.}, c[3] = {0., 0., 0.};
        for (int g = 0; g < ngb; g++) {
            const double gx = compensated ? 0. : gbs[3 * g], gy = compensated ? 0. : gbs[3 * g + 1],
                         gz = compensated ? 0. : gbs[3 * g + 2];
            for (int pass = 0; pass < 2; pass++) {
                const size_t j0 = pass == 0 ? 0 : N_active, j1 = pass == 0 ? N_active : N;
                if (pass == 1 && (!tp_type || i >= N_active)) break;
                for (size_t j = j0; j < j1; j++) {
                    if (pass == 0 || compensated) {
                        if (skip_pair(i, j, ignore)) continue;
                    } else {
                        if (ignore == 1 && (j == 1 && i == 0)) continue;
                        if (ignore == 2 && (j == 0 || i == 0)) continue;
                    }
                    if (compensated) {
                        const double dx = COL(s, i, X) - COL(s, j, X);
                        const double dy = COL(s, i, Y) - COL(s, j, Y);
                        const double dz = COL(s, i, Z) - COL(s, j, Z);
                        const double r2 = dx * dx + dy * dy + dz * dz + softening2;
                        const double r = sqrt(r2);
                        const double prefact = G / (r2 * r);
                        const double prefactj = -prefact * COL(s, j, M);
                        kahan(prefactj * dx, &a[0], &c[0]);
                        kahan(prefactj * dy, &a[1], &c[1]);
                        kahan(prefactj * dz, &a[2], &c[2]);
                    } else {
                        const double dx = (gx + COL(s, i, X)) - COL(s, j, X);
                        const double dy = (gy + COL(s, i, Y)) - COL(s, j, Y);
                        const double dz = (gz + COL(s, i, Z)) - COL(s, j, Z);
                        const double _r = sqrt(dx * dx + dy * dy + dz * dz + softening2);
                        const double prefact = -G / (_r * _r * _r) * COL(s, j, M);
                        a[0] += prefact * dx;
                        a[1] += prefact * dy;
                        a[2] += prefact * dz;
                    }
                }
            }
        }
        acc[3 * q] = a[0];
        acc[3 * q + 1] = a[1];
        acc[3 * q + 2] = a[2];
    }
}

void orc_gravity_compensated(const double* s, size_t N, size_t N_active, int tp_type, int ignore, double G,
                             double soft, double* acc) {
    const double softening2 = soft * soft;
    if (N_active > N) N_active = N;
    double* cs = calloc(3 * N + 3, sizeof(double));
    for (size_t i = 0; i < 3 * N; i++) acc[i] = 0.;
    /* targets i < N_active: sources j < N_active (gravity.c:308-343);
     * targets i >= N_active: sources j < N_active (:346-378) */
    for (size_t i = 0; i < N; i++) {
        for (size_t j = 0; j < N_active; j++) {
            if (skip_pair(i, j, ignore)) continue;
            const double dx = COL(s, i, X) - COL(s, j, X);
            const double dy = COL(s, i, Y) - COL(s, j, Y);
            const double dz = COL(s, i, Z) - COL(s, j, Z);
            const double r2 = dx * dx + dy * dy + dz * dz + softening2;
            const double r = sqrt(r2);
            const double prefact = G / (r2 * r);
            const double prefactj = -prefact * COL(s, j, M);
            kahan(prefactj * dx, &acc[3 * i], &cs[3 * i]);
            kahan(prefactj * dy, &acc[3 * i + 1], &cs[3 * i + 1]);
            kahan(prefactj * dz, &acc[3 * i + 2], &cs[3 * i + 2]);
        }
    }
    if (tp_type) { /* :379-410 active targets feel test particles, in test-particle order */
        for (size_t j = 0; j < N_active; j++) {
            for (size_t i = N_active; i < N; i++) {
                if (skip_pair(i, j, ignore)) continue;
                const double dx = COL(s, i, X) - COL(s, j, X);
                const double dy = COL(s, i, Y) - COL(s, j, Y);
                const double dz = COL(s, i, Z) - COL(s, j, Z);
                const double r2 = dx * dx + dy * dy + dz * dz + softening2;
                const double r = sqrt(r2);
                const double prefact = G / (r2 * r);
                const double prefacti = prefact * COL(s, i, M);
                kahan(prefacti * dx, &acc[3 * j], &cs[3 * j]);
                kahan(prefacti * dy, &acc[3 * j + 1], &cs[3 * j + 1]);
                kahan(prefacti * dz, &acc[3 * j + 2], &cs[3 * j + 2]);
            }
        }
    }
    free(cs);
}

double orc_energy(const double* s, size_t N, size_t N_active, int tp_type, double G) {
    if (N_active > N) N_active = N;
    double e_kin = 0., e_pot = 0.;
    size_t N_interact = (tp_type == 0) ? N_active : N;
    for (size_t i = 0; i < N_interact; i++) {
        const double* p = s + i * ORC_NCOL;
        e_kin += 0.5 * p[M] * (p[VX] * p[VX] + p[VY] * p[VY] + p[VZ] * p[VZ]);
    }
    for (size_t i = 0; i < N_active; i++) {
        for (size_t j = i + 1; j < N_interact; j++) {
            double dx = COL(s, i, X) - COL(s, j, X);
            double dy = COL(s, i, Y) - COL(s, j, Y);
            double dz = COL(s, i, Z) - COL(s, j, Z);
            e_pot -= G * COL(s, j, M) * COL(s, i, M) / sqrt(dx * dx + dy * dy + dz * dz);
        }
    }
    return e_kin + e_pot;
}

/* ------------------------------------------------------------------ leapfrog */
static void drift(double* s, size_t N, double dt, double* t) {
    for (size_t i = 0; i < N; i++) {
        COL(s, i, X) += dt * COL(s, i, VX);
        COL(s, i, Y) += dt * COL(s, i, VY);
        COL(s, i, Z) += dt * COL(s, i, VZ);
    }
    *t += dt;
}
static void kick(double* s, size_t N, double dt) {
    for (size_t i = 0; i < N; i++) {
        COL(s, i, VX) += dt * COL(s, i, AX);
        COL(s, i, VY) += dt * COL(s, i, AY);
        COL(s, i, VZ) += dt * COL(s, i, AZ);
    }
}

void orc_leapfrog(double* s, size_t N, double G, double soft, double dt, long steps, double* t, long every,
                  double* e_out) {
    double* acc = malloc(sizeof(double) * 3 * N + 8);
    const double gb[3] = {0, 0, 0};
    long k = 0;
    if (e_out && every > 0) e_out[k++] = orc_energy(s, N, N, 0, G);
    for (long st = 0; st < steps; st++) {
        drift(s, N, dt * 0.5, t);
        orc_gravity_basic(s, N, N, 0, 0, G, soft, gb, 1, acc);
        for (size_t i = 0; i < N; i++) {
            COL(s, i, AX) = acc[3 * i];
            COL(s, i, AY) = acc[3 * i + 1];
            COL(s, i, AZ) = acc[3 * i + 2];
        }
        kick(s, N, dt);
        drift(s, N, dt * 0.5, t);
        if (e_out && every > 0 && (st + 1) % every == 0) e_out[k++] = orc_energy(s, N, N, 0, G);
    }
    free(acc);
}

/* Algorithmic floating-point operations of the WHFast path, counted as the
 * reference's source writes them: + - * / and sqrt are one each (constant
 * products such as 2.*M_PI folded), comparisons, fabs and floor are free.
 * Gravity pairs count 20 (gravity.c:222-230). */
static unsigned long long g_flops = 0;
#define FL(n) (g_flops += (unsigned long long)(n))
unsigned long long orc_flops(int reset) {
    unsigned long long v = g_flops;
    if (reset) g_flops = 0;
    return v;
}

/* ------------------------------------------------------------------ Kepler (integrator_whfast.c:161-427) */
static const double invfactorial[14] = {1., 1., 1. / 2., 1. / 6., 1. / 24., 1. / 120., 1. / 720., 1. / 5040.,
                                        1. / 40320., 1. / 362880., 1. / 3628800., 1. / 39916800.,
                                        1. / 479001600., 1. / 6227020800.};

static double fastabs(double x) { return (x > 0.) ? x : -x; }

static void stumpff_cs3(double* cs, double z) {
    unsigned int n = 0;
    while (fabs(z) > 0.1) {
        z = z / 4.;
        n++;
        FL(1);
    }
    FL(24 + 9 * n);
    double c_odd = invfactorial[13];
    double c_even = invfactorial[12];
    for (int np = 11; np >= 3; np -= 2) {
        c_odd = invfactorial[np] - z * c_odd;
        c_even = invfactorial[np - 1] - z * c_even;
    }
    cs[3] = c_odd;
    cs[2] = c_even;
    cs[1] = invfactorial[1] - z * c_odd;
    cs[0] = invfactorial[0] - z * c_even;
    for (; n > 0; n--) {
        cs[3] = (cs[2] + cs[0] * cs[3]) * 0.25;
        cs[2] = cs[1] * cs[1] * 0.5;
        cs[1] = cs[0] * cs[1];
        cs[0] = 2. * cs[0] * cs[0] - 1.;
    }
}

static void stiefel_Gs3(double* Gs, double beta, double X_) {
    double X2 = X_ * X_;
    FL(6);
    stumpff_cs3(Gs, beta * X2);
    Gs[1] *= X_;
    Gs[2] *= X2;
    Gs[3] *= X2 * X_;
}

int orc_kepler(double* p, double mu, double dt) {
    const double p1[6] = {p[0], p[1], p[2], p[3], p[4], p[5]};
    int warn = 0;
    const double r0 = sqrt(p1[0] * p1[0] + p1[1] * p1[1] + p1[2] * p1[2]);
    const double r0i = 1. / r0;
    const double v2 = p1[3] * p1[3] + p1[4] * p1[4] + p1[5] * p1[5];
    const double beta = 2. * mu * r0i - v2;
    const double eta0 = p1[0] * p1[3] + p1[1] * p1[4] + p1[2] * p1[5];
    const double zeta0 = mu - beta * r0;
    FL(22);
    double Xv, Gs[6];
    double invperiod = 0;
    double X_per_period = nan("");
    if (beta > 0.) {
        const double sqrt_beta = sqrt(beta);
        invperiod = sqrt_beta * beta / (2. * M_PI * mu);
        X_per_period = 2. * M_PI / sqrt_beta;
        if (fabs(dt) * invperiod > 1.) warn = 1;
        const double dtr0i = dt * r0i;
        Xv = dtr0i * (1. - dtr0i * eta0 * 0.5 * r0i);
        FL(13);
    } else {
        Xv = 0.;
    }
    unsigned int converged = 0;
    double oldX = Xv;
    stiefel_Gs3(Gs, beta, Xv);
    const double eg = eta0 * Gs[1] + zeta0 * Gs[2];
    double ri = 1. / (r0 + eg);
    Xv = ri * (Xv * eg - eta0 * Gs[2] - zeta0 * Gs[3] + dt);
    FL(14);
    if (fastabs(Xv - oldX) > 0.01 * X_per_period) {
        Xv = beta * dt / mu;
        FL(2);
        double prevX[65];
        for (int n_lag = 1; n_lag < 64; n_lag++) {
            stiefel_Gs3(Gs, beta, Xv);
            const double f = r0 * Xv + eta0 * Gs[2] + zeta0 * Gs[3] - dt;
            const double fp = r0 + eta0 * Gs[1] + zeta0 * Gs[2];
            const double fpp = eta0 * Gs[0] + zeta0 * Gs[1];
            const double denom = fp + sqrt(fabs(16. * fp * fp - 20. * f * fpp));
            Xv = (Xv * denom - 5. * f) / denom;
            FL(24);
            for (int i = 1; i < n_lag; i++) {
                if (Xv == prevX[i]) {
                    n_lag = 64;
                    converged = 1;
                    break;
                }
            }
            prevX[n_lag] = Xv;
        }
        const double e1 = eta0 * Gs[1] + zeta0 * Gs[2];
        ri = 1. / (r0 + e1);
        FL(5);
    } else {
        double oldX2 = nan("");
        for (int n_hg = 1; n_hg < 32; n_hg++) {
            oldX2 = oldX;
            oldX = Xv;
            stiefel_Gs3(Gs, beta, Xv);
            const double e1 = eta0 * Gs[1] + zeta0 * Gs[2];
            ri = 1. / (r0 + e1);
            Xv = ri * (Xv * e1 - eta0 * Gs[2] - zeta0 * Gs[3] + dt);
            FL(12);
            if (Xv == oldX || Xv == oldX2) {
                converged = 1;
                break;
            }
        }
    }
    if (converged == 0) {
        double X_min, X_max;
        if (beta > 0.) {
            X_min = X_per_period * floor(dt * invperiod);
            X_max = X_min + X_per_period;
            FL(3);
        } else {
            double h2 = r0 * r0 * v2 - eta0 * eta0;
            double q = h2 / mu / (1. + sqrt(1. - h2 * beta / (mu * mu)));
            double vq = copysign(sqrt(h2) / q, dt);
            X_min = dt / (fastabs(vq * dt) + r0);
            X_max = dt / q;
            FL(18);
            if (dt < 0.) {
                double temp = X_min;
                X_min = X_max;
                X_max = temp;
            }
        }
        Xv = (X_max + X_min) / 2.;
        FL(2);
        do {
            stiefel_Gs3(Gs, beta, Xv);
            double sv = r0 * Xv + eta0 * Gs[2] + zeta0 * Gs[3] - dt;
            if (sv >= 0.) {
                X_max = Xv;
            } else {
                X_min = Xv;
            }
            Xv = (X_max + X_min) / 2.;
            FL(11);
        } while (fastabs((X_max - X_min)) > fastabs((X_max + X_min) * 1e-15));
        const double e1 = eta0 * Gs[1] + zeta0 * Gs[2];
        ri = 1. / (r0 + e1);
        FL(5);
    }
    if (isnan(ri)) {
        ri = 0.;
        Gs[1] = 0.;
        Gs[2] = 0.;
        Gs[3] = 0.;
    }
    FL(33);
    double f = -mu * Gs[2] * r0i;
    double g = dt - mu * Gs[3];
    double fd = -mu * Gs[1] * r0i * ri;
    double gd = -mu * Gs[2] * ri;
    p[0] += f * p1[0] + g * p1[3];
    p[1] += f * p1[1] + g * p1[4];
    p[2] += f * p1[2] + g * p1[5];
    p[3] += fd * p1[0] + gd * p1[3];
    p[4] += fd * p1[1] + gd * p1[4];
    p[5] += fd * p1[2] + gd * p1[5];
    return warn;
}

/* ------------------------------------------------------------------ WHFast DH */
#define PH(ph, i, c) ((ph)[(size_t)(i) * 7 + (c)])

static void dh_from_inertial(const double* s, double* ph, size_t N, size_t NA) {
    /* transformations.c:388-424 */
    FL(13 * NA + 6 + 6 * (N - 1));
    double x0 = 0., y0 = 0., z0 = 0., vx0 = 0., vy0 = 0., vz0 = 0., m0 = 0.;
    for (size_t i = 0; i < NA; i++) {
        double m = COL(s, i, M);
        x0 += COL(s, i, X) * m;
        y0 += COL(s, i, Y) * m;
        z0 += COL(s, i, Z) * m;
        vx0 += COL(s, i, VX) * m;
        vy0 += COL(s, i, VY) * m;
        vz0 += COL(s, i, VZ) * m;
        m0 += m;
    }
    PH(ph, 0, 0) = x0 / m0;
    PH(ph, 0, 1) = y0 / m0;
    PH(ph, 0, 2) = z0 / m0;
    PH(ph, 0, 3) = vx0 / m0;
    PH(ph, 0, 4) = vy0 / m0;
    PH(ph, 0, 5) = vz0 / m0;
    PH(ph, 0, 6) = m0;
    for (size_t i = 1; i < N; i++) {
        PH(ph, i, 0) = COL(s, i, X) - COL(s, 0, X);
        PH(ph, i, 1) = COL(s, i, Y) - COL(s, 0, Y);
        PH(ph, i, 2) = COL(s, i, Z) - COL(s, 0, Z);
        PH(ph, i, 3) = COL(s, i, VX) - PH(ph, 0, 3);
        PH(ph, i, 4) = COL(s, i, VY) - PH(ph, 0, 4);
        PH(ph, i, 5) = COL(s, i, VZ) - PH(ph, 0, 5);
        PH(ph, i, 6) = COL(s, i, M);
    }
}

static void dh_to_inertial(double* s, const double* ph, size_t N, size_t NA) {
    /* transformations.c:426-477 */
    FL(18 * (NA - 1) + 6 + 6 * (N - 1));
    const double mtot = PH(ph, 0, 6);
    double x0 = 0., y0 = 0., z0 = 0.;
    for (size_t i = 1; i < NA; i++) {
        double m = PH(ph, i, 6);
        x0 += PH(ph, i, 0) * m / mtot;
        y0 += PH(ph, i, 1) * m / mtot;
        z0 += PH(ph, i, 2) * m / mtot;
        COL(s, i, M) = m;
    }
    COL(s, 0, X) = PH(ph, 0, 0) - x0;
    COL(s, 0, Y) = PH(ph, 0, 1) - y0;
    COL(s, 0, Z) = PH(ph, 0, 2) - z0;
    for (size_t i = 1; i < N; i++) {
        COL(s, i, X) = PH(ph, i, 0) + COL(s, 0, X);
        COL(s, i, Y) = PH(ph, i, 1) + COL(s, 0, Y);
        COL(s, i, Z) = PH(ph, i, 2) + COL(s, 0, Z);
    }
    const double m0 = COL(s, 0, M);
    for (size_t i = 1; i < N; i++) {
        COL(s, i, VX) = PH(ph, i, 3) + PH(ph, 0, 3);
        COL(s, i, VY) = PH(ph, i, 4) + PH(ph, 0, 4);
        COL(s, i, VZ) = PH(ph, i, 5) + PH(ph, 0, 5);
    }
    double vx0 = 0., vy0 = 0., vz0 = 0.;
    for (size_t i = 1; i < NA; i++) {
        double m = COL(s, i, M);
        vx0 += PH(ph, i, 3) * m / m0;
        vy0 += PH(ph, i, 4) * m / m0;
        vz0 += PH(ph, i, 5) * m / m0;
    }
    COL(s, 0, VX) = PH(ph, 0, 3) - vx0;
    COL(s, 0, VY) = PH(ph, 0, 4) - vy0;
    COL(s, 0, VZ) = PH(ph, 0, 5) - vz0;
}

static void dh_kepler_com(const double* s, double* ph, size_t N, double G, double h) {
    const double m0 = COL(s, 0, M);
    FL(6 + (N - 1));
    for (size_t i = 1; i < N; i++) {
        double st[6];
        for (int k = 0; k < 6; k++) st[k] = PH(ph, i, k);
        orc_kepler(st, m0 * G, h);
        for (int k = 0; k < 6; k++) PH(ph, i, k) = st[k];
    }
    PH(ph, 0, 0) += h * PH(ph, 0, 3);
    PH(ph, 0, 1) += h * PH(ph, 0, 4);
    PH(ph, 0, 2) += h * PH(ph, 0, 5);
}

static void dh_jump(const double* s, double* ph, size_t N, size_t NA, double h) {
    /* integrator_whfast.c:537-554 */
    const double m0 = COL(s, 0, M);
    FL(6 * (NA - 1) + 6 * (N - 1));
    double px = 0, py = 0, pz = 0;
    for (size_t i = 1; i < NA; i++) {
        const double m = COL(s, i, M);
        px += m * PH(ph, i, 3);
        py += m * PH(ph, i, 4);
        pz += m * PH(ph, i, 5);
    }
    for (size_t i = 1; i < N; i++) {
        PH(ph, i, 0) += h * (px / m0);
        PH(ph, i, 1) += h * (py / m0);
        PH(ph, i, 2) += h * (pz / m0);
    }
}

void orc_whfast_dh(double* s, double* ph, size_t N, double G, double soft, double dt, long steps, int safe_mode,
                   int* is_sync, int did_modify) {
    const size_t NA = N;
    double* acc = malloc(sizeof(double) * 3 * N + 8);
    const double gb[3] = {0, 0, 0};
    for (long st = 0; st < steps; st++) {
        if (safe_mode || did_modify) {
            if (*is_sync == 0) {
                dh_kepler_com(s, ph, N, G, dt / 2.);
                dh_to_inertial(s, ph, N, NA);
                *is_sync = 1;
            }
            dh_from_inertial(s, ph, N, NA);
        }
        if (*is_sync) {
            dh_kepler_com(s, ph, N, G, dt / 2.);
        } else {
            dh_kepler_com(s, ph, N, G, dt);
        }
        dh_jump(s, ph, N, NA, dt / 2.);
        dh_to_inertial(s, ph, N, NA);
        orc_gravity_basic(s, N, N, 0, 2, G, soft, gb, 1, acc);
        FL(20 * (N - 1) * (N > 1 ? N - 2 : 0) + 6 * (N - 1));
        for (size_t i = 0; i < N; i++) {
            COL(s, i, AX) = acc[3 * i];
            COL(s, i, AY) = acc[3 * i + 1];
            COL(s, i, AZ) = acc[3 * i + 2];
        }
        for (size_t i = 1; i < N; i++) {
            PH(ph, i, 3) += dt * COL(s, i, AX);
            PH(ph, i, 4) += dt * COL(s, i, AY);
            PH(ph, i, 5) += dt * COL(s, i, AZ);
        }
        dh_jump(s, ph, N, NA, dt / 2.);
        *is_sync = 0;
        if (safe_mode) {
            dh_kepler_com(s, ph, N, G, dt / 2.);
            dh_to_inertial(s, ph, N, NA);
            *is_sync = 1;
        }
        did_modify = 0;
    }
    /* reb_simulation_steps ends with a synchronize (simulation.c:510) */
    if (*is_sync == 0) {
        dh_kepler_com(s, ph, N, G, dt / 2.);
        dh_to_inertial(s, ph, N, NA);
        *is_sync = 1;
    }
    free(acc);
}

/* ------------------------------------------------------------------ WHFast Jacobi */
typedef struct {
    double* s;        /* inertial bodies, ORC_NCOL columns */
    double* pj;       /* Jacobi bodies, 7 columns */
    double* acc;      /* 3N scratch */
    double* ja;       /* 3N Jacobi accelerations */
    double* jerk;     /* 3N jerk (modified kick kernel) */
    size_t N, NA, NAG;
    int tp, comp;
    int kernel;       /* 0 default, 1 modified kick, 2 composition, 3 lazy */
    int jacgrav;      /* r->gravity == REB_GRAVITY_JACOBI (kernels 1 and 3) */
    double G, soft, dt;
} jacctx;

static void jac_from_inertial(jacctx* c) {
    /* transformations.c:34-81, p_mass = the inertial bodies */
    const double* s = c->s;
    double* pj = c->pj;
    double eta = COL(s, 0, M);
    double sx = eta * COL(s, 0, X), sy = eta * COL(s, 0, Y), sz = eta * COL(s, 0, Z);
    double svx = eta * COL(s, 0, VX), svy = eta * COL(s, 0, VY), svz = eta * COL(s, 0, VZ);
    for (size_t i = 1; i < c->NA; i++) {
        const double ei = 1. / eta;
        const double mi = COL(s, i, M);
        eta += mi;
        const double pme = eta * ei;
        PH(pj, i, 6) = mi;
        PH(pj, i, 0) = COL(s, i, X) - sx * ei;
        PH(pj, i, 1) = COL(s, i, Y) - sy * ei;
        PH(pj, i, 2) = COL(s, i, Z) - sz * ei;
        PH(pj, i, 3) = COL(s, i, VX) - svx * ei;
        PH(pj, i, 4) = COL(s, i, VY) - svy * ei;
        PH(pj, i, 5) = COL(s, i, VZ) - svz * ei;
        sx = sx * pme + mi * PH(pj, i, 0);
        sy = sy * pme + mi * PH(pj, i, 1);
        sz = sz * pme + mi * PH(pj, i, 2);
        svx = svx * pme + mi * PH(pj, i, 3);
        svy = svy * pme + mi * PH(pj, i, 4);
        svz = svz * pme + mi * PH(pj, i, 5);
    }
    const double ei = 1. / eta;
    for (size_t i = c->NA; i < c->N; i++) {
        PH(pj, i, 6) = COL(s, i, M);
        PH(pj, i, 0) = COL(s, i, X) - sx * ei;
        PH(pj, i, 1) = COL(s, i, Y) - sy * ei;
        PH(pj, i, 2) = COL(s, i, Z) - sz * ei;
        PH(pj, i, 3) = COL(s, i, VX) - svx * ei;
        PH(pj, i, 4) = COL(s, i, VY) - svy * ei;
        PH(pj, i, 5) = COL(s, i, VZ) - svz * ei;
    }
    const double Mi = 1. / eta;
    PH(pj, 0, 6) = eta;
    PH(pj, 0, 0) = sx * Mi;
    PH(pj, 0, 1) = sy * Mi;
    PH(pj, 0, 2) = sz * Mi;
    PH(pj, 0, 3) = svx * Mi;
    PH(pj, 0, 4) = svy * Mi;
    PH(pj, 0, 5) = svz * Mi;
}

static void jac_to_inertial(jacctx* c, int with_vel) {
    /* transformations.c:178-226 (with_vel) and :228-258 (positions) */
    double* s = c->s;
    const double* pj = c->pj;
    double eta = PH(pj, 0, 6);
    double sx = PH(pj, 0, 0) * eta, sy = PH(pj, 0, 1) * eta, sz = PH(pj, 0, 2) * eta;
    double svx = PH(pj, 0, 3) * eta, svy = PH(pj, 0, 4) * eta, svz = PH(pj, 0, 5) * eta;
    for (size_t i = c->N - 1; i >= c->NA; i--) {
        const double ei = 1. / eta;
        COL(s, i, X) = PH(pj, i, 0) + sx * ei;
        COL(s, i, Y) = PH(pj, i, 1) + sy * ei;
        COL(s, i, Z) = PH(pj, i, 2) + sz * ei;
        if (with_vel) {
            COL(s, i, VX) = PH(pj, i, 3) + svx * ei;
            COL(s, i, VY) = PH(pj, i, 4) + svy * ei;
            COL(s, i, VZ) = PH(pj, i, 5) + svz * ei;
        }
    }
    for (size_t i = c->NA - 1; i > 0; i--) {
        const double ei = 1. / eta;
        const double mi = COL(s, i, M);
        sx = (sx - mi * PH(pj, i, 0)) * ei;
        sy = (sy - mi * PH(pj, i, 1)) * ei;
        sz = (sz - mi * PH(pj, i, 2)) * ei;
        COL(s, i, X) = PH(pj, i, 0) + sx;
        COL(s, i, Y) = PH(pj, i, 1) + sy;
        COL(s, i, Z) = PH(pj, i, 2) + sz;
        if (with_vel) {
            svx = (svx - mi * PH(pj, i, 3)) * ei;
            svy = (svy - mi * PH(pj, i, 4)) * ei;
            svz = (svz - mi * PH(pj, i, 5)) * ei;
            COL(s, i, VX) = PH(pj, i, 3) + svx;
            COL(s, i, VY) = PH(pj, i, 4) + svy;
            COL(s, i, VZ) = PH(pj, i, 5) + svz;
        }
        eta -= mi;
        sx *= eta;
        sy *= eta;
        sz *= eta;
        svx *= eta;
        svy *= eta;
        svz *= eta;
    }
    const double mi = 1. / eta;
    COL(s, 0, X) = sx * mi;
    COL(s, 0, Y) = sy * mi;
    COL(s, 0, Z) = sz * mi;
    if (with_vel) {
        COL(s, 0, VX) = svx * mi;
        COL(s, 0, VY) = svy * mi;
        COL(s, 0, VZ) = svz * mi;
    }
}

static void jac_kepler(jacctx* c, double h) {
    /* integrator_whfast.c:589-606, the serial form of eta */
    double eta = COL(c->s, 0, M);
    for (size_t i = 1; i < c->N; i++) {
        if (i < c->NA) eta += PH(c->pj, i, 6);
        double st[6];
        for (int k = 0; k < 6; k++) st[k] = PH(c->pj, i, k);
        orc_kepler(st, eta * c->G, h);
        for (int k = 0; k < 6; k++) PH(c->pj, i, k) = st[k];
    }
}

static void jac_com(jacctx* c, double h) {
    /* integrator_whfast.c:638-641 */
    PH(c->pj, 0, 0) += h * PH(c->pj, 0, 3);
    PH(c->pj, 0, 1) += h * PH(c->pj, 0, 4);
    PH(c->pj, 0, 2) += h * PH(c->pj, 0, 5);
}

static void jac_gravity_jacobi(jacctx* c) {
    /* gravity.c:108-163: Jacobi and direct terms, each pair visited once in
     * the order j = 0..N-1, i = 0..j, both bodies updated; no softening */
    double* s = c->s;
    const size_t N = c->N;
    const double G = c->G;
    double Rx = 0., Ry = 0., Rz = 0., Mt = 0.;
    for (size_t j = 0; j < N; j++) {
        COL(s, j, AX) = 0.;
        COL(s, j, AY) = 0.;
        COL(s, j, AZ) = 0.;
        for (size_t i = 0; i < j + 1; i++) {
            if (j > 1) {
                const double qx = COL(s, j, X) - Rx / Mt;
                const double qy = COL(s, j, Y) - Ry / Mt;
                const double qz = COL(s, j, Z) - Rz / Mt;
                const double dr = sqrt(qx * qx + qy * qy + qz * qz);
                const double w = i < j ? -COL(s, j, M) : Mt;
                const double pf = G * w / (dr * dr * dr);
                COL(s, i, AX) += pf * qx;
                COL(s, i, AY) += pf * qy;
                COL(s, i, AZ) += pf * qz;
            }
            if (i != j && (i != 0 || j != 1)) {
                const double dx = COL(s, i, X) - COL(s, j, X);
                const double dy = COL(s, i, Y) - COL(s, j, Y);
                const double dz = COL(s, i, Z) - COL(s, j, Z);
                const double dr = sqrt(dx * dx + dy * dy + dz * dz);
                const double pf = G / (dr * dr * dr);
                const double pfi = pf * COL(s, i, M);
                const double pfj = pf * COL(s, j, M);
                COL(s, i, AX) -= pfj * dx;
                COL(s, i, AY) -= pfj * dy;
                COL(s, i, AZ) -= pfj * dz;
                COL(s, j, AX) += pfi * dx;
                COL(s, j, AY) += pfi * dy;
                COL(s, j, AZ) += pfi * dz;
            }
        }
        Rx += COL(s, j, M) * COL(s, j, X);
        Ry += COL(s, j, M) * COL(s, j, Y);
        Rz += COL(s, j, M) * COL(s, j, Z);
        Mt += COL(s, j, M);
    }
}

void orc_gravity_jacobi(double* s, size_t N, double G) {
    jacctx c;
    memset(&c, 0, sizeof(c));
    c.s = s;
    c.N = N;
    c.G = G;
    jac_gravity_jacobi(&c);
}

static void jac_jerk(jacctx* c) {
    /* integrator_whfast.c:786-873: time derivative of the Jacobi-gravity
     * accelerations, same visiting order as the gravity above */
    const double* s = c->s;
    double* jk = c->jerk;
    const size_t N = c->N;
    const double G = c->G;
    double Rx = 0., Ry = 0., Rz = 0., Mt = 0., Ax = 0., Ay = 0., Az = 0.;
    for (size_t j = 0; j < N; j++) {
        jk[3 * j] = jk[3 * j + 1] = jk[3 * j + 2] = 0.;
        for (size_t i = 0; i < j + 1; i++) {
            if (j > 1) {
                const double w = i < j ? -COL(s, j, M) : Mt;
                const double qx = COL(s, j, X) - Rx / Mt;
                const double qy = COL(s, j, Y) - Ry / Mt;
                const double qz = COL(s, j, Z) - Rz / Mt;
                const double dax = COL(s, j, AX) - Ax / Mt;
                const double day = COL(s, j, AY) - Ay / Mt;
                const double daz = COL(s, j, AZ) - Az / Mt;
                const double dr = sqrt(qx * qx + qy * qy + qz * qz);
                const double p2 = G * w / (dr * dr * dr);
                jk[3 * i] += p2 * dax;
                jk[3 * i + 1] += p2 * day;
                jk[3 * i + 2] += p2 * daz;
                const double al = dax * qx + day * qy + daz * qz;
                const double p1 = 3. * al * p2 / (dr * dr);
                jk[3 * i] -= p1 * qx;
                jk[3 * i + 1] -= p1 * qy;
                jk[3 * i + 2] -= p1 * qz;
            }
            if (j != i && (i != 0 || j != 1)) {
                const double dx = COL(s, j, X) - COL(s, i, X);
                const double dy = COL(s, j, Y) - COL(s, i, Y);
                const double dz = COL(s, j, Z) - COL(s, i, Z);
                const double dax = COL(s, j, AX) - COL(s, i, AX);
                const double day = COL(s, j, AY) - COL(s, i, AY);
                const double daz = COL(s, j, AZ) - COL(s, i, AZ);
                const double dr = sqrt(dx * dx + dy * dy + dz * dz);
                const double al = dax * dx + day * dy + daz * dz;
                const double p2 = G / (dr * dr * dr);
                const double p2i = p2 * COL(s, i, M);
                const double p2j = p2 * COL(s, j, M);
                jk[3 * j] -= dax * p2i;
                jk[3 * j + 1] -= day * p2i;
                jk[3 * j + 2] -= daz * p2i;
                jk[3 * i] += dax * p2j;
                jk[3 * i + 1] += day * p2j;
                jk[3 * i + 2] += daz * p2j;
                const double p1 = 3. * al * p2 / (dr * dr);
                const double p1i = p1 * COL(s, i, M);
                const double p1j = p1 * COL(s, j, M);
                jk[3 * j] += dx * p1i;
                jk[3 * j + 1] += dy * p1i;
                jk[3 * j + 2] += dz * p1i;
                jk[3 * i] -= dx * p1j;
                jk[3 * i + 1] -= dy * p1j;
                jk[3 * i + 2] -= dz * p1j;
            }
        }
        Ax += COL(s, j, AX) * COL(s, j, M);
        Ay += COL(s, j, AY) * COL(s, j, M);
        Az += COL(s, j, AZ) * COL(s, j, M);
        Rx += COL(s, j, X) * COL(s, j, M);
        Ry += COL(s, j, Y) * COL(s, j, M);
        Rz += COL(s, j, Z) * COL(s, j, M);
        Mt += COL(s, j, M);
    }
}

static void jac_gravity(jacctx* c) {
    /* reb_simulation_update_acceleration: the Jacobi gravity for kernels 1
     * and 3, else gravity_ignore_terms = BETWEEN_0_AND_1 */
    if (c->jacgrav) {
        jac_gravity_jacobi(c);
        return;
    }
    const double gb[3] = {0, 0, 0};
    if (c->comp)
        orc_gravity_compensated(c->s, c->N, c->NAG, c->tp, 1, c->G, c->soft, c->acc);
    else
        orc_gravity_basic(c->s, c->N, c->NAG, c->tp, 1, c->G, c->soft, gb, 1, c->acc);
    for (size_t i = 0; i < c->N; i++) {
        COL(c->s, i, AX) = c->acc[3 * i];
        COL(c->s, i, AY) = c->acc[3 * i + 1];
        COL(c->s, i, AZ) = c->acc[3 * i + 2];
    }
}

static void jac_interaction(jacctx* c, double h) {
    /* transformations.c:147-176, then integrator_whfast.c:439-483 */
    const double* s = c->s;
    double* ja = c->ja;
    double eta = COL(s, 0, M);
    double sx = eta * COL(s, 0, AX), sy = eta * COL(s, 0, AY), sz = eta * COL(s, 0, AZ);
    for (size_t i = 1; i < c->NA; i++) {
        const double ei = 1. / eta;
        const double mi = COL(s, i, M);
        eta += mi;
        const double pme = eta * ei;
        ja[3 * i] = COL(s, i, AX) - sx * ei;
        ja[3 * i + 1] = COL(s, i, AY) - sy * ei;
        ja[3 * i + 2] = COL(s, i, AZ) - sz * ei;
        sx = sx * pme + mi * ja[3 * i];
        sy = sy * pme + mi * ja[3 * i + 1];
        sz = sz * pme + mi * ja[3 * i + 2];
    }
    const double ei = 1. / eta;
    for (size_t i = c->NA; i < c->N; i++) {
        ja[3 * i] = COL(s, i, AX) - sx * ei;
        ja[3 * i + 1] = COL(s, i, AY) - sy * ei;
        ja[3 * i + 2] = COL(s, i, AZ) - sz * ei;
    }
    double et = COL(s, 0, M);
    for (size_t i = 1; i < c->N; i++) {
        const double px = PH(c->pj, i, 0), py = PH(c->pj, i, 1), pz = PH(c->pj, i, 2);
        if (i < c->NA) et += PH(c->pj, i, 6);
        PH(c->pj, i, 3) += h * ja[3 * i];
        PH(c->pj, i, 4) += h * ja[3 * i + 1];
        PH(c->pj, i, 5) += h * ja[3 * i + 2];
        if (i > 1 && !c->jacgrav) {   /* Jacobi gravity already holds these terms */
            const double rj2i = 1. / (px * px + py * py + pz * pz);
            const double rji = sqrt(rj2i);
            const double rj3iM = rji * rj2i * c->G * et;
            const double prefac1 = h * rj3iM;
            PH(c->pj, i, 3) += prefac1 * px;
            PH(c->pj, i, 4) += prefac1 * py;
            PH(c->pj, i, 5) += prefac1 * pz;
        }
    }
}

/* Symplectic corrector coefficients: Wisdom, Holman & Touma (1996), as
 * integrator_whfast.c:125-152 spells them. */
static const double ca1 = 0.41833001326703777398908601289259374469640768464934;
static const double ca2 = 0.83666002653407554797817202578518748939281536929867;
static const double ca3 = 1.2549900398011133219672580386777812340892230539480;
static const double ca4 = 1.6733200530681510959563440515703749787856307385973;
static const double ca5 = 2.0916500663351888699454300644629687234820384232467;
static const double ca6 = 2.5099800796022266439345160773555624681784461078960;
static const double ca7 = 2.9283100928692644179236020902481562128748537925454;
static const double ca8 = 3.3466401061363021919126881031407499575712614771947;
static const double cb31 = -0.024900596027799867499350357910273437184309981229127;
static const double cb51 = -0.0083001986759332891664501193034244790614366604097090;
static const double cb52 = 0.041500993379666445832250596517122395307183302048545;
static const double cb71 = 0.0024926811426922105779030593952776964450539008582219;
static const double cb72 = -0.018270923246702131478062356884535264841652263842597;
static const double cb73 = 0.053964399093127498721765893493510877532452806339655;
static const double cb111 = 0.00020361579647854651301632818774633716473696537436847;
static const double cb112 = -0.0023487215292295354188307328851055489876255097419754;
static const double cb113 = 0.012309078592019946317544564763237909911330686448336;
static const double cb114 = -0.038121613681288650508647613260247372125243616270670;
static const double cb115 = 0.072593394748842738674253180742744961827622366521517;
static const double cb171 = -0.0000043347415473373580190650223498124944896789841432241;
static const double cb172 = 0.000076436355227935738363241846979413475106795392377415;
static const double cb173 = -0.00063599983075817658983166881625078545864140848560259;
static const double cb174 = 0.0033132577069380655655490196833451994080066801611459;
static const double cb175 = -0.012071760822342291062449751726959664253913904872527;
static const double cb176 = 0.032422198864713580293681523029577130832258806467604;
static const double cb177 = -0.065192863576377893658290760803725762027864651086787;
static const double cb178 = 0.093056103771425958591541059067553547100903397724386;
static const double c2b = 0.03486083443891981449909050107438281205803;

static void jac_Z(jacctx* c, double a, double b) {
    /* integrator_whfast.c:653-678 */
    jac_kepler(c, a);
    jac_to_inertial(c, 0);
    jac_gravity(c);
    jac_interaction(c, -b);
    jac_kepler(c, -2. * a);
    jac_to_inertial(c, 0);
    jac_gravity(c);
    jac_interaction(c, b);
    jac_kepler(c, a);
}

static void jac_corrector(jacctx* c, double inv, int order) {
    /* integrator_whfast.c:696-751 */
    const double dt = c->dt;
    if (order == 3) {
        jac_Z(c, ca1 * dt, -inv * cb31 * dt);
        jac_Z(c, -ca1 * dt, inv * cb31 * dt);
    }
    if (order == 5) {
        jac_Z(c, -ca2 * dt, -inv * cb51 * dt);
        jac_Z(c, -ca1 * dt, -inv * cb52 * dt);
        jac_Z(c, ca1 * dt, inv * cb52 * dt);
        jac_Z(c, ca2 * dt, inv * cb51 * dt);
    }
    if (order == 7) {
        jac_Z(c, -ca3 * dt, -inv * cb71 * dt);
        jac_Z(c, -ca2 * dt, -inv * cb72 * dt);
        jac_Z(c, -ca1 * dt, -inv * cb73 * dt);
        jac_Z(c, ca1 * dt, inv * cb73 * dt);
        jac_Z(c, ca2 * dt, inv * cb72 * dt);
        jac_Z(c, ca3 * dt, inv * cb71 * dt);
    }
    if (order == 11) {
        const double A[5] = {ca5, ca4, ca3, ca2, ca1};
        const double B[5] = {cb111, cb112, cb113, cb114, cb115};
        for (int k = 0; k < 5; k++) jac_Z(c, -A[k] * dt, -inv * B[k] * dt);
        for (int k = 4; k >= 0; k--) jac_Z(c, A[k] * dt, inv * B[k] * dt);
    }
    if (order == 17) {
        const double A[8] = {ca8, ca7, ca6, ca5, ca4, ca3, ca2, ca1};
        const double B[8] = {cb171, cb172, cb173, cb174, cb175, cb176, cb177, cb178};
        for (int k = 0; k < 8; k++) jac_Z(c, -A[k] * dt, -inv * B[k] * dt);
        for (int k = 7; k >= 0; k--) jac_Z(c, A[k] * dt, inv * B[k] * dt);
    }
}

static void jac_C(jacctx* c, double a, double b) {
    /* integrator_whfast.c:753-766 */
    jac_kepler(c, a);
    jac_to_inertial(c, 0);
    jac_gravity(c);
    jac_interaction(c, b);
    jac_kepler(c, -a);
}

static void jac_U(jacctx* c, double a, double b) {
    /* integrator_whfast.c:768-778: U = K(a) Y(a,b) Y(a,-b) K(-a), Y(a,b) = C(a,b) C(-a,-b) */
    jac_kepler(c, a);
    jac_C(c, a, b);
    jac_C(c, -a, -b);
    jac_C(c, a, -b);
    jac_C(c, -a, b);
    jac_kepler(c, -a);
}

static void jac_corrector2(jacctx* c, double inv) {
    /* integrator_whfast.c:779-784 */
    const double a = 0.5 * inv * c->dt;
    const double b = c2b * inv * c->dt;
    jac_U(c, a, b);
    jac_U(c, -a, b);
}

static void jac_synchronize(jacctx* c, int corrector, int corrector2) {
    /* integrator_whfast.c:1051-1111: a half drift, 3/8 of one for the
     * composition kernel */
    const double h = c->kernel == 2 ? 3. * c->dt / 8. : c->dt / 2.;
    jac_kepler(c, h);
    jac_com(c, h);
    if (corrector2) jac_corrector2(c, -1.);
    if (corrector) jac_corrector(c, -1., corrector);
    jac_to_inertial(c, 1);
}

void orc_whfast_jacobi(double* s, double* pj, size_t N, size_t N_active, int tp_type, double G, double soft,
                       int compensated, double dt, long steps, int safe_mode, int corrector, int corrector2,
                       int kernel, int* is_sync, int did_modify) {
    jacctx c;
    c.s = s;
    c.pj = pj;
    c.acc = malloc(sizeof(double) * 3 * N + 8);
    c.ja = malloc(sizeof(double) * 3 * N + 8);
    c.jerk = malloc(sizeof(double) * 3 * N + 8);
    c.kernel = kernel;
    c.jacgrav = kernel == 1 || kernel == 3;
    c.N = N;
    c.NAG = N_active > N ? N : N_active;
    c.NA = tp_type == 1 ? N : c.NAG;
    c.tp = tp_type;
    c.comp = compensated;
    c.G = G;
    c.soft = soft;
    c.dt = dt;
    for (long st = 0; st < steps; st++) {
        /* integrator_whfast.c:1113-1259 (default kernel) */
        if (safe_mode || did_modify) {
            if (*is_sync == 0) {
                jac_synchronize(&c, corrector, corrector2);
                *is_sync = 1;
            }
            jac_from_inertial(&c);
        }
        if (*is_sync) {
            if (corrector) jac_corrector(&c, 1., corrector);
            if (corrector2) jac_corrector2(&c, 1.);
            const double h = kernel == 2 ? 5. * dt / 8. : dt / 2.;
            jac_kepler(&c, h);
            jac_com(&c, h);
        } else {
            jac_kepler(&c, dt);
            jac_com(&c, dt);
        }
        jac_to_inertial(&c, 1);
        jac_gravity(&c);
        if (kernel == 1) {
            /* integrator_whfast.c:1178-1187 */
            jac_jerk(&c);
            for (size_t i = 0; i < N; i++) {
                const double pf = dt * dt / 12.;
                COL(s, i, AX) += pf * c.jerk[3 * i];
                COL(s, i, AY) += pf * c.jerk[3 * i + 1];
                COL(s, i, AZ) += pf * c.jerk[3 * i + 2];
            }
            jac_interaction(&c, dt);
        } else if (kernel == 2) {
            /* integrator_whfast.c:1189-1218: interaction and drift stages of
             * the composition kernel, a gravity evaluation before each kick */
            const double kick[5] = {-dt / 6., dt / 6., dt, -dt / 6., dt / 6.};
            const double drift[4] = {-dt / 4., dt / 8., -dt / 8., dt / 4.};
            for (int q = 0; q < 5; q++) {
                if (q > 0) {
                    jac_to_inertial(&c, 0);
                    jac_gravity(&c);
                }
                jac_interaction(&c, kick[q]);
                if (q < 4) {
                    jac_kepler(&c, drift[q]);
                    jac_com(&c, drift[q]);
                }
            }
        } else if (kernel == 3) {
            /* integrator_whfast.c:1220-1235 (WHT Eq. 10.6) */
            for (size_t i = 1; i < N; i++) {
                const double pf = dt * dt / 12.;
                COL(s, i, X) += pf * COL(s, i, AX);
                COL(s, i, Y) += pf * COL(s, i, AY);
                COL(s, i, Z) += pf * COL(s, i, AZ);
            }
            jac_gravity(&c);
            jac_interaction(&c, dt);
        } else {
            jac_interaction(&c, dt);
        }
        *is_sync = 0;
        if (safe_mode) {
            jac_synchronize(&c, corrector, corrector2);
            *is_sync = 1;
        }
        did_modify = 0;
    }
    /* reb_simulation_steps ends with a synchronize (simulation.c:510) */
    if (*is_sync == 0) {
        jac_synchronize(&c, corrector, corrector2);
        *is_sync = 1;
    }
    free(c.acc);
    free(c.ja);
    free(c.jerk);
}

/* ------------------------------------------------------------------ SEI + shear boundary */
void orc_sei_operator(double* s, size_t N, double OMEGA, double OMEGAZ, double dt, int with_kick) {
    const double sindt = sin(OMEGA * (-dt / 2.));
    const double tandt = tan(OMEGA * (-dt / 4.));
    const double sindtz = sin(OMEGAZ * (-dt / 2.));
    const double tandtz = tan(OMEGAZ * (-dt / 4.));
    for (size_t i = 0; i < N; i++) {
        double* p = s + i * ORC_NCOL;
        if (with_kick) {
            p[VX] += p[AX] * dt;
            p[VY] += p[AY] * dt;
            p[VZ] += p[AZ] * dt;
        }
        const double zx = p[Z] * OMEGAZ;
        const double zy = p[VZ];
        const double zt1 = zx - tandtz * zy;
        const double zyt = sindtz * zt1 + zy;
        const double zxt = zt1 - tandtz * zyt;
        p[Z] = zxt / OMEGAZ;
        p[VZ] = zyt;
        const double aO = 2. * p[VY] + 4. * p[X] * OMEGA;
        const double bO = p[Y] * OMEGA - 2. * p[VX];
        const double ys = (p[Y] * OMEGA - bO) / 2.;
        const double xs = (p[X] * OMEGA - aO);
        const double xst1 = xs - tandt * ys;
        const double yst = sindt * xst1 + ys;
        const double xst = xst1 - tandt * yst;
        p[X] = (xst + aO) / OMEGA;
        p[Y] = (yst * 2. + bO) / OMEGA - 3. / 4. * aO * dt;
        p[VX] = yst;
        p[VY] = -xst * 2. - 3. / 2. * aO;
    }
}

void orc_shear_wrap(double* s, size_t N, double OMEGA, double bx, double by, double bz, double t) {
    const double offsetp1 = -fmod(-1.5 * OMEGA * bx * t + by / 2., by) - by / 2.;
    const double offsetm1 = -fmod(1.5 * OMEGA * bx * t - by / 2., by) + by / 2.;
    for (size_t i = 0; i < N; i++) {
        double* p = s + i * ORC_NCOL;
        while (p[X] > bx / 2.) {
            p[X] -= bx;
            p[Y] += offsetp1;
            p[VY] += 3. / 2. * OMEGA * bx;
        }
        while (p[X] < -bx / 2.) {
            p[X] += bx;
            p[Y] += offsetm1;
            p[VY] -= 3. / 2. * OMEGA * bx;
        }
        while (p[Y] > by / 2.) p[Y] -= by;
        while (p[Y] < -by / 2.) p[Y] += by;
        while (p[Z] > bz / 2.) p[Z] -= bz;
        while (p[Z] < -bz / 2.) p[Z] += bz;
    }
}

void orc_ghostbox_shear(double OMEGA, double bx, double by, double bz, double t, int i, int j, int k, double* gb) {
    gb[3] = 0.;
    gb[4] = -1.5 * (double)i * OMEGA * bx;
    gb[5] = 0.;
    double shift;
    if (i == 0) {
        shift = -fmod(gb[4] * t, by);
    } else if (i > 0) {
        shift = -fmod(gb[4] * t - by / 2., by) - by / 2.;
    } else {
        shift = -fmod(gb[4] * t + by / 2., by) + by / 2.;
    }
    gb[0] = bx * (double)i;
    gb[1] = by * (double)j - shift;
    gb[2] = bz * (double)k;
}

/* ------------------------------------------------------------------ tree (tree.c) */
typedef struct cell {
    double x, y, z, w, m, mx, my, mz;
    struct cell* oct[8];
    int pt;
} cell;

typedef struct {
    const double* s;
    size_t N;
    double root_size;
    int nrx, nry, nrz;
    int err;
} treectx;

static int octant(const double* p, const cell* c) {
    int o = 0;
    if (p[X] < c->x) o += 1;
    if (p[Y] < c->y) o += 2;
    if (p[Z] < c->z) o += 4;
    return o;
}

static cell* add_to_cell(treectx* T, cell* node, int pt, cell* parent, int o) {
    const double* s = T->s;
    if (node == NULL) {
        node = calloc(1, sizeof(cell));
        const double* p = s + (size_t)pt * ORC_NCOL;
        if (parent == NULL) {
            node->w = T->root_size;
            double bx = T->root_size * (double)T->nrx, by = T->root_size * (double)T->nry,
                   bz = T->root_size * (double)T->nrz;
            int i = ((int)floor((p[X] + bx / 2.) / T->root_size)) % T->nrx;
            int j = ((int)floor((p[Y] + by / 2.) / T->root_size)) % T->nry;
            int k = ((int)floor((p[Z] + bz / 2.) / T->root_size)) % T->nrz;
            node->x = -bx / 2. + T->root_size * (0.5 + (double)i);
            node->y = -by / 2. + T->root_size * (0.5 + (double)j);
            node->z = -bz / 2. + T->root_size * (0.5 + (double)k);
        } else {
            node->w = parent->w / 2.;
            node->x = parent->x + node->w / 2. * ((o >> 0) % 2 == 0 ? 1. : -1);
            node->y = parent->y + node->w / 2. * ((o >> 1) % 2 == 0 ? 1. : -1);
            node->z = parent->z + node->w / 2. * ((o >> 2) % 2 == 0 ? 1. : -1);
        }
        if (node->w <= 0.0) {
            T->err = 1;
            free(node);
            return NULL;
        }
        node->pt = pt;
        return node;
    }
    if (node->pt >= 0) {
        const double* a = s + (size_t)node->pt * ORC_NCOL;
        const double* b = s + (size_t)pt * ORC_NCOL;
        int o1 = octant(a, node), o2 = octant(b, node);
        if (o1 == o2 && a[X] == b[X] && a[Y] == b[Y] && a[Z] == b[Z]) {
            T->err = 2;
            return node;
        }
        node->oct[o1] = add_to_cell(T, node->oct[o1], node->pt, node, o1);
        node->oct[o2] = add_to_cell(T, node->oct[o2], pt, node, o2);
        node->pt = -2;
    } else {
        node->pt--;
        int o = octant(s + (size_t)pt * ORC_NCOL, node);
        node->oct[o] = add_to_cell(T, node->oct[o], pt, node, o);
    }
    return node;
}

static void moments(const double* s, cell* c) {
    if (c->pt < 0) {
        c->m = 0;
        c->mx = 0;
        c->my = 0;
        c->mz = 0;
        for (int o = 0; o < 8; o++) {
            cell* d = c->oct[o];
            if (d) {
                moments(s, d);
                double d_m = d->m;
                c->mx += d->mx * d_m;
                c->my += d->my * d_m;
                c->mz += d->mz * d_m;
                c->m += d_m;
            }
        }
        double m_tot = c->m;
        if (m_tot > 0) {
            c->mx /= m_tot;
            c->my /= m_tot;
            c->mz /= m_tot;
        }
    } else {
        const double* p = s + (size_t)c->pt * ORC_NCOL;
        c->m = p[M];
        c->mx = p[X];
        c->my = p[Y];
        c->mz = p[Z];
    }
}

static void delete_cell(cell* c) {
    if (!c) return;
    for (int o = 0; o < 8; o++) delete_cell(c->oct[o]);
    free(c);
}

static int rootbox_index(const treectx* T, const double* p) {
    int i = ((int)floor((p[X] + T->root_size * (double)T->nrx / 2.) / T->root_size) + T->nrx) % T->nrx;
    int j = ((int)floor((p[Y] + T->root_size * (double)T->nry / 2.) / T->root_size) + T->nry) % T->nry;
    int k = ((int)floor((p[Z] + T->root_size * (double)T->nrz / 2.) / T->root_size) + T->nrz) % T->nrz;
    return (k * T->nry + j) * T->nrx + i;
}

static cell** build(treectx* T) {
    int nroot = T->nrx * T->nry * T->nrz;
    cell** roots = calloc(nroot, sizeof(cell*));
    for (size_t i = 0; i < T->N; i++) {
        const double* p = T->s + i * ORC_NCOL;
        if (fabs(p[X]) > T->root_size * (double)T->nrx / 2. || fabs(p[Y]) > T->root_size * (double)T->nry / 2. ||
            fabs(p[Z]) > T->root_size * (double)T->nrz / 2.) {
            T->err = 3;
            return roots;
        }
        if (!isfinite(p[X]) || !isfinite(p[Y]) || !isfinite(p[Z])) {
            T->err = 4;
            return roots;
        }
        int rb = rootbox_index(T, p);
        roots[rb] = add_to_cell(T, roots[rb], (int)i, NULL, 0);
    }
    for (int r = 0; r < nroot; r++)
        if (roots[r]) moments(T->s, roots[r]);
    return roots;
}

static void dump(const cell* c, double* out, long cap, long* n) {
    if (!c) return;
    if (out && *n < cap) {
        double row[9] = {c->x, c->y, c->z, c->w, c->m, c->mx, c->my, c->mz, (double)c->pt};
        memcpy(out + 9 * (*n), row, sizeof(row));
    }
    (*n)++;
    for (int o = 0; o < 8; o++) dump(c->oct[o], out, cap, n);
}

long orc_tree_dump(const double* s, size_t N, double root_size, int nrx, int nry, int nrz, double* cells, long cap,
                   long* rootstart) {
    treectx T = {s, N, root_size, nrx, nry, nrz, 0};
    cell** roots = build(&T);
    long n = 0;
    int nroot = nrx * nry * nrz;
    for (int r = 0; r < nroot; r++) {
        if (rootstart) rootstart[r] = n;
        dump(roots[r], cells, cap, &n);
    }
    if (rootstart) rootstart[nroot] = n;
    for (int r = 0; r < nroot; r++) delete_cell(roots[r]);
    free(roots);
    return T.err ? -1 : n;
}

typedef struct {
    const double* s;
    double G, soft2, theta2;
    long inter;
} walkctx;

static void walk(walkctx* W, int pt, const cell* node, const double* gb, double* a) {
    const double dx = gb[0] - node->mx;
    const double dy = gb[1] - node->my;
    const double dz = gb[2] - node->mz;
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (node->pt < 0) {
        if (node->w * node->w > W->theta2 * r2) {
            for (int o = 0; o < 8; o++)
                if (node->oct[o]) walk(W, pt, node->oct[o], gb, a);
        } else {
            double _r = sqrt(r2 + W->soft2);
            double prefact = -W->G / (_r * _r * _r) * node->m;
            a[0] += prefact * dx;
            a[1] += prefact * dy;
            a[2] += prefact * dz;
            W->inter++;
        }
    } else {
        if (node->pt == pt) return;
        double _r = sqrt(r2 + W->soft2);
        double prefact = -W->G / (_r * _r * _r) * node->m;
        a[0] += prefact * dx;
        a[1] += prefact * dy;
        a[2] += prefact * dz;
        W->inter++;
    }
}

int orc_tree_gravity(const double* s, size_t N, double root_size, int nrx, int nry, int nrz, double G, double soft,
                     double opening_angle2, const double* gbs, int ngb, double* acc, long* interactions) {
    treectx T = {s, N, root_size, nrx, nry, nrz, 0};
    cell** roots = build(&T);
    int nroot = nrx * nry * nrz;
    walkctx W = {s, G, soft * soft, opening_angle2, 0};
    for (size_t i = 0; i < 3 * N; i++) acc[i] = 0.;
    if (!T.err) {
        for (int g = 0; g < ngb; g++) {
            for (size_t i = 0; i < N; i++) {
                double gb[3] = {gbs[3 * g] + COL(s, i, X), gbs[3 * g + 1] + COL(s, i, Y), gbs[3 * g + 2] + COL(s, i, Z)};
                for (int r = 0; r < nroot; r++)
                    if (roots[r]) walk(&W, (int)i, roots[r], gb, acc + 3 * i);
            }
        }
    }
    for (int r = 0; r < nroot; r++) delete_cell(roots[r]);
    free(roots);
    if (interactions) *interactions = W.inter;
    return T.err;
}

/* ------------------------------------------------------------------ collisions (collision.c) */
typedef struct {
    const double* s;
    double second_r;
    int64_t* pairs;
    double* gbo;
    long cap, n;
} collctx;

static void nn_in_cell(collctx* C, const double* gb, const double* gbunmod, size_t p1, double p1_r, const cell* c) {
    const double* s = C->s;
    if (c->pt >= 0) {
        if ((size_t)c->pt == p1) return;
        const double* p2 = s + (size_t)c->pt * ORC_NCOL;
        double dx = gb[0] - p2[X];
        double dy = gb[1] - p2[Y];
        double dz = gb[2] - p2[Z];
        double r2 = dx * dx + dy * dy + dz * dz;
        double rp = p1_r + p2[RR];
        if (r2 > rp * rp) return;
        double dvx = gb[3] - p2[VX];
        double dvy = gb[4] - p2[VY];
        double dvz = gb[5] - p2[VZ];
        if (dvx * dx + dvy * dy + dvz * dz > 0) return;
        if (C->n < C->cap) {
            C->pairs[2 * C->n] = (int64_t)p1;
            C->pairs[2 * C->n + 1] = c->pt;
            memcpy(C->gbo + 6 * C->n, gbunmod, 6 * sizeof(double));
        }
        C->n++;
    } else {
        double dx = gb[0] - c->x;
        double dy = gb[1] - c->y;
        double dz = gb[2] - c->z;
        double r2 = dx * dx + dy * dy + dz * dz;
        double rp = p1_r + C->second_r + 0.86602540378443 * c->w;
        if (r2 < rp * rp) {
            for (int o = 0; o < 8; o++)
                if (c->oct[o]) nn_in_cell(C, gb, gbunmod, p1, p1_r, c->oct[o]);
        }
    }
}

static double second_largest_radius(const double* s, size_t N) {
    /* simulation.c:766-779 (serial) */
    double largest1 = -1.0, largest2 = -1.0;
    long p1 = -1, p2 = -1;
    for (size_t i = 0; i < N; i++) {
        if (COL(s, i, RR) > largest1) {
            largest2 = largest1;
            p2 = p1;
            largest1 = COL(s, i, RR);
            p1 = (long)i;
        } else if (COL(s, i, RR) > largest2) {
            largest2 = COL(s, i, RR);
            p2 = (long)i;
        }
    }
    return p2 >= 0 ? COL(s, p2, RR) : 0.;
}

long orc_collision_tree(const double* s, size_t N, double root_size, int nrx, int nry, int nrz, const double* gbs,
                        int ngb, int64_t* pairs, double* gbo, long cap) {
    treectx T = {s, N, root_size, nrx, nry, nrz, 0};
    cell** roots = build(&T);
    int nroot = nrx * nry * nrz;
    collctx C = {s, second_largest_radius(s, N), pairs, gbo, cap, 0};
    if (!T.err) {
        for (size_t i = 0; i < N; i++) {
            const double* p1 = s + i * ORC_NCOL;
            for (int g = 0; g < ngb; g++) {
                const double* gu = gbs + 6 * g;
                double gb[6] = {gu[0] + p1[X], gu[1] + p1[Y], gu[2] + p1[Z],
                                gu[3] + p1[VX], gu[4] + p1[VY], gu[5] + p1[VZ]};
                for (int r = 0; r < nroot; r++)
                    if (roots[r]) nn_in_cell(&C, gb, gu, i, p1[RR], roots[r]);
            }
        }
    }
    for (int r = 0; r < nroot; r++) delete_cell(roots[r]);
    free(roots);
    return T.err ? -1 : C.n;
}

long orc_collision_direct(const double* s, size_t N, const double* gbs, int ngb, int64_t* pairs, double* gbo,
                          long cap) {
    long n = 0;
    for (int g = 0; g < ngb; g++) {
        const double* gu = gbs + 6 * g;
        for (size_t i = 0; i < N; i++) {
            const double* p1 = s + i * ORC_NCOL;
            double gb[6] = {gu[0] + p1[X], gu[1] + p1[Y], gu[2] + p1[Z], gu[3] + p1[VX], gu[4] + p1[VY], gu[5] + p1[VZ]};
            for (size_t j = 0; j < N; j++) {
                if (i == j) continue;
                const double* p2 = s + j * ORC_NCOL;
                double dx = gb[0] - p2[X];
                double dy = gb[1] - p2[Y];
                double dz = gb[2] - p2[Z];
                double sr = p1[RR] + p2[RR];
                double r2 = dx * dx + dy * dy + dz * dz;
                if (r2 > sr * sr) continue;
                double dvx = gb[3] - p2[VX];
                double dvy = gb[4] - p2[VY];
                double dvz = gb[5] - p2[VZ];
                if (dvx * dx + dvy * dy + dvz * dz > 0) continue;
                if (n < cap) {
                    pairs[2 * n] = (int64_t)i;
                    pairs[2 * n + 1] = (int64_t)j;
                    memcpy(gbo + 6 * n, gu, 6 * sizeof(double));
                }
                n++;
            }
        }
    }
    return n;
}

/* Swept-sphere overlap (collision.c:150-171, :510-534): closest distance of
 * two bodies on straight lines over the last step, MIN as collision.c:43. */
#define ORC_MIN(a, b) ((a) > (b) ? (b) : (a))
static int line_overlap(double dx1, double dy1, double dz1, double dvx1, double dvy1, double dvz1, double dt,
                        double rsum) {
    const double r1 = (dx1 * dx1 + dy1 * dy1 + dz1 * dz1);
    const double dx2 = dx1 - dt * dvx1;
    const double dy2 = dy1 - dt * dvy1;
    const double dz2 = dz1 - dt * dvz1;
    const double r2 = (dx2 * dx2 + dy2 * dy2 + dz2 * dz2);
    const double t_closest = (dx1 * dvx1 + dy1 * dvy1 + dz1 * dvz1) / (dvx1 * dvx1 + dvy1 * dvy1 + dvz1 * dvz1);
    double rmin2_ab = ORC_MIN(r1, r2);
    if (t_closest / dt >= 0. && t_closest / dt <= 1.) {
        const double dx3 = dx1 - t_closest * dvx1;
        const double dy3 = dy1 - t_closest * dvy1;
        const double dz3 = dz1 - t_closest * dvz1;
        const double r3 = (dx3 * dx3 + dy3 * dy3 + dz3 * dz3);
        rmin2_ab = ORC_MIN(rmin2_ab, r3);
    }
    return !(rmin2_ab > rsum * rsum);
}

long orc_collision_line(const double* s, size_t N, const double* gbs, int ngb, double dt, int64_t* pairs, double* gbo,
                        long cap) {
    /* collision.c:125-196: ghost box, i, j > i */
    long n = 0;
    for (int g = 0; g < ngb; g++) {
        const double* gu = gbs + 6 * g;
        for (size_t i = 0; i < N; i++) {
            const double* p1 = s + i * ORC_NCOL;
            const double gb[6] = {gu[0] + p1[X], gu[1] + p1[Y], gu[2] + p1[Z],
                                  gu[3] + p1[VX], gu[4] + p1[VY], gu[5] + p1[VZ]};
            for (size_t j = i + 1; j < N; j++) {
                const double* p2 = s + j * ORC_NCOL;
                if (!line_overlap(gb[0] - p2[X], gb[1] - p2[Y], gb[2] - p2[Z], gb[3] - p2[VX], gb[4] - p2[VY],
                                  gb[5] - p2[VZ], dt, p1[RR] + p2[RR]))
                    continue;
                if (n < cap) {
                    pairs[2 * n] = (int64_t)i;
                    pairs[2 * n + 1] = (int64_t)j;
                    memcpy(gbo + 6 * n, gu, 6 * sizeof(double));
                }
                n++;
            }
        }
    }
    return n;
}

typedef struct {
    const double* s;
    double dt, maxdrift;
    int64_t* pairs;
    double* gbo;
    long cap, n;
} linectx;

static void line_in_cell(linectx* C, const double* gb, const double* gbunmod, size_t p1, double p1_r,
                         double p1_r_plus_dtv, const cell* c) {
    /* collision.c:506-566 */
    const double* s = C->s;
    if (c->pt >= 0) {
        if ((size_t)c->pt == p1) return;
        const double* p2 = s + (size_t)c->pt * ORC_NCOL;
        if (!line_overlap(gb[0] - p2[X], gb[1] - p2[Y], gb[2] - p2[Z], gb[3] - p2[VX], gb[4] - p2[VY], gb[5] - p2[VZ],
                          C->dt, p1_r + p2[RR]))
            return;
        if (C->n < C->cap) {
            C->pairs[2 * C->n] = (int64_t)p1;
            C->pairs[2 * C->n + 1] = c->pt;
            memcpy(C->gbo + 6 * C->n, gbunmod, 6 * sizeof(double));
        }
        C->n++;
    } else {
        const double dx = gb[0] - c->x;
        const double dy = gb[1] - c->y;
        const double dz = gb[2] - c->z;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double rp = p1_r_plus_dtv + C->maxdrift + 0.86602540378443 * c->w;
        if (r2 < rp * rp) {
            for (int o = 0; o < 8; o++)
                if (c->oct[o]) line_in_cell(C, gb, gbunmod, p1, p1_r, p1_r_plus_dtv, c->oct[o]);
        }
    }
}

long orc_collision_linetree(const double* s, size_t N, double root_size, int nrx, int nry, int nrz, const double* gbs,
                            int ngb, double dt, int64_t* pairs, double* gbo, long cap) {
    /* collision.c:270-331 */
    double vmax2 = 0.;
    for (size_t i = 0; i < N; i++) {
        const double* p = s + i * ORC_NCOL;
        const double v2 = p[VX] * p[VX] + p[VY] * p[VY] + p[VZ] * p[VZ];
        vmax2 = vmax2 > v2 ? vmax2 : v2;
    }
    treectx T = {s, N, root_size, nrx, nry, nrz, 0};
    cell** roots = build(&T);
    int nroot = nrx * nry * nrz;
    linectx C = {s, dt, dt * sqrt(vmax2), pairs, gbo, cap, 0};
    if (!T.err) {
        for (size_t i = 0; i < N; i++) {
            const double* p1 = s + i * ORC_NCOL;
            const double p1_r_plus_dtv = p1[RR] + dt * sqrt(p1[VX] * p1[VX] + p1[VY] * p1[VY] + p1[VZ] * p1[VZ]);
            for (int g = 0; g < ngb; g++) {
                const double* gu = gbs + 6 * g;
                const double gb[6] = {gu[0] + p1[X], gu[1] + p1[Y], gu[2] + p1[Z],
                                      gu[3] + p1[VX], gu[4] + p1[VY], gu[5] + p1[VZ]};
                for (int r = 0; r < nroot; r++)
                    if (roots[r]) line_in_cell(&C, gb, gu, i, p1[RR], p1_r_plus_dtv, roots[r]);
            }
        }
    }
    for (int r = 0; r < nroot; r++) delete_cell(roots[r]);
    free(roots);
    return T.err ? -1 : C.n;
}

void orc_collision_shuffle(int64_t* pairs, double* gbo, long n, unsigned int* seed) {
    for (long i = 0; i < n; i++) {
        long nw = (long)((unsigned int)orc_rand_r(seed) % (unsigned long)n);
        int64_t a0 = pairs[2 * i], a1 = pairs[2 * i + 1];
        double g[6];
        memcpy(g, gbo + 6 * i, sizeof(g));
        pairs[2 * i] = pairs[2 * nw];
        pairs[2 * i + 1] = pairs[2 * nw + 1];
        memcpy(gbo + 6 * i, gbo + 6 * nw, sizeof(g));
        pairs[2 * nw] = a0;
        pairs[2 * nw + 1] = a1;
        memcpy(gbo + 6 * nw, g, sizeof(g));
    }
}

double orc_hardsphere(double* s, int64_t i1, int64_t i2, const double* gb, double eps_c, double min_v, long* log_n) {
    double* P1 = s + (size_t)i1 * ORC_NCOL;
    double* P2 = s + (size_t)i2 * ORC_NCOL;
    double p1[ORC_NCOL], p2[ORC_NCOL];
    memcpy(p1, P1, sizeof(p1));
    memcpy(p2, P2, sizeof(p2));
    double x21 = p1[X] + gb[0] - p2[X];
    double y21 = p1[Y] + gb[1] - p2[Y];
    double z21 = p1[Z] + gb[2] - p2[Z];
    double rp = p1[RR] + p2[RR];
    double oldvyouter = (x21 > 0) ? p1[VY] : p2[VY];
    if (rp * rp < x21 * x21 + y21 * y21 + z21 * z21) return 0.;
    double vx21 = p1[VX] + gb[3] - p2[VX];
    double vy21 = p1[VY] + gb[4] - p2[VY];
    double vz21 = p1[VZ] + gb[5] - p2[VZ];
    if (vx21 * x21 + vy21 * y21 + vz21 * z21 > 0) return 0.;
    double theta = atan2(z21, y21);
    double stheta = sin(theta), ctheta = cos(theta);
    double vy21n = ctheta * vy21 + stheta * vz21;
    double y21n = ctheta * y21 + stheta * z21;
    double phi = atan2(y21n, x21);
    double cphi = cos(phi), sphi = sin(phi);
    double vx21nn = cphi * vx21 + sphi * vy21n;
    double eps = eps_c < 0 ? 1 : eps_c;
    double dvx2 = -(1.0 + eps) * vx21nn;
    double minr = (p1[RR] > p2[RR]) ? p2[RR] : p1[RR];
    double maxr = (p1[RR] < p2[RR]) ? p2[RR] : p1[RR];
    double mindv = minr * min_v;
    double _r = sqrt(x21 * x21 + y21 * y21 + z21 * z21);
    mindv *= 1. - (_r - maxr) / minr;
    if (mindv > maxr * min_v) mindv = maxr * min_v;
    if (dvx2 < mindv) dvx2 = mindv;
    double dvx2n = cphi * dvx2;
    double dvy2n = sphi * dvx2;
    double dvy2nn = ctheta * dvy2n;
    double dvz2nn = stheta * dvy2n;
    const double p2pf = p1[M] / (p1[M] + p2[M]);
    P2[VX] -= p2pf * dvx2n;
    P2[VY] -= p2pf * dvy2nn;
    P2[VZ] -= p2pf * dvz2nn;
    const double p1pf = p2[M] / (p1[M] + p2[M]);
    P1[VX] += p1pf * dvx2n;
    P1[VY] += p1pf * dvy2nn;
    P1[VZ] += p1pf * dvz2nn;
    (*log_n)++;
    if (x21 > 0) return -fabs(x21) * (oldvyouter - P1[VY]) * p1[M];
    return -fabs(x21) * (oldvyouter - P2[VY]) * p2[M];
}
