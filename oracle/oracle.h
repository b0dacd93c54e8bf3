/*
 * CPU oracle: a plain C restatement of REBOUND 5.0.0's force-and-step
 * arithmetic, used ONLY as the checker of the test suite, of smoke() and of
 * the benchmark's CPU-baseline leg.  The product (librebound_amd.so) never
 * links or calls it.
 *
 * State arrays use rows of 11 doubles: x y z vx vy vz ax ay az m r.
 * Every function cites the reference lines it restates; tests pin each one
 * against golden vectors written by the compiled reference
 * (tests/golden/gen_golden.c) or, where noted, against closed forms.
 */
#ifndef REBOUND_ORACLE_H
#define REBOUND_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#define ORC_NCOL 11

/* tools.c:64-105, rebound.c:238-259 */
int orc_rand_r(unsigned int* seed);
double orc_random_uniform(unsigned int* seed, double min, double max);
double orc_random_normal(unsigned int* seed, double variance);
double orc_random_powerlaw(unsigned int* seed, double min, double max, double slope);
double orc_random_rayleigh(unsigned int* seed, double sigma);

/* tools.c:463-502 (N bodies appended to `out`, N x 11) */
void orc_plummer(unsigned int* seed, size_t N, double M, double R, double* out);
/* tools.c:370-420 + :201-372 (no variational particles) */
void orc_move_to_com(double* s, size_t N);
/* tools.c:940-997, returns err code, writes x y z vx vy vz into out6 */
int orc_particle_from_orbit(double G, const double* primary7 /* x y z vx vy vz m */, double m, double a, double e,
                            double inc, double Omega, double omega, double f, double* out6);

/* gravity.c:215-232 + :258-277 (OpenMP per-target order), ghost boxes gbs[ngb][3] */
void orc_gravity_basic(const double* s, size_t N, size_t N_active, int tp_type, int ignore, double G, double soft,
                       const double* gbs, int ngb, double* acc3);
/* gravity.c:308-410 (OpenMP form) */
/* Target rows rows[0..nrows) of orc_gravity_basic (compensated = 0) or
 * orc_gravity_compensated (compensated = 1), same arithmetic and order. */
void orc_gravity_rows(const double* s, size_t N, size_t N_active, int tp_type, int ignore, double G, double soft,
                      const double* gbs, int ngb, int compensated, const int64_t* rows, size_t nrows, double* acc);
void orc_gravity_compensated(const double* s, size_t N, size_t N_active, int tp_type, int ignore, double G,
                             double soft, double* acc3);
/* tools.c:108-162 */
double orc_energy(const double* s, size_t N, size_t N_active, int tp_type, double G);

/* integrator_leapfrog.c:72-107: `steps` order-2 DKD steps with BASIC gravity
 * (no ghost boxes), t advanced as the reference does; energies every
 * `every` steps written to e_out (may be NULL). */
void orc_leapfrog(double* s, size_t N, double G, double soft, double dt, long steps, double* t, long every,
                  double* e_out);

/* integrator_whfast.c:241-427; state6 updated in place; returns 1 if the
 * step exceeded one orbital period */
int orc_kepler(double* state6, double mu, double dt);

/* WHFast, democratic heliocentric, default kernel, no correctors:
 * integrator_whfast.c:1113-1259 + :1051-1111, transformations.c:388-477,
 * interaction with BASIC gravity and IGNORE_TERMS_INVOLVING_0.
 * `steps` calls of step() then synchronize(), as reb_simulation_steps does.
 * ph7 holds the heliocentric state (N x 7: x y z vx vy vz m) across calls;
 * *is_sync the synchronisation flag; did_modify as the simulation field. */
void orc_whfast_dh(double* s, double* ph7, size_t N, double G, double soft, double dt, long steps, int safe_mode,
                   int* is_sync, int did_modify);

/* WHFast in Jacobi coordinates (integrator_whfast.c:876-1259 with the
 * default kernel): interaction with BASIC or COMPENSATED gravity and
 * IGNORE_TERMS_BETWEEN_0_AND_1, first symplectic corrector of order
 * `corrector` (0, 3, 5, 7, 11, 17) and the second corrector when `corrector2`.
 * N_active and tp_type as the simulation fields.  pj7 holds the Jacobi state
 * (N x 7) across calls; otherwise as orc_whfast_dh. */
/* gravity.c:108-163 into the acceleration columns of s */
void orc_gravity_jacobi(double* s, size_t N, double G);
void orc_whfast_jacobi(double* s, double* pj7, size_t N, size_t N_active, int tp_type, double G, double soft,
                       int compensated, double dt, long steps, int safe_mode, int corrector, int corrector2,
                       int kernel, int* is_sync, int did_modify);

/* SEI operators (integrator_sei.c:126-174): H0 over dt for every body,
 * preceded by the phi1 kick v += a dt when with_kick; the sin/tan constants
 * are computed as integrator_sei.c:91-101 does. */
void orc_sei_operator(double* s, size_t N, double OMEGA, double OMEGAZ, double dt, int with_kick);

/* boundary.c:78-114 (SHEAR wrap) and :145-201 (ghost box offsets) */
void orc_shear_wrap(double* s, size_t N, double OMEGA, double bx, double by, double bz, double t);
void orc_ghostbox_shear(double OMEGA, double bx, double by, double bz, double t, int i, int j, int k, double* gb6);

/* ---------------------------------------------------------------- tree (tree.c) */
/* Builds the reference octree (tree.c:64-271), computes monopoles
 * (tree.c:147-229) and returns the number of cells; `cells` receives the
 * pre-order dump (x y z w m mx my mz pt) of each root box in root order
 * (cap rows available, may be NULL to only count). Returns -1 on error
 * (non-finite position, duplicate position, outside box). */
long orc_tree_dump(const double* s, size_t N, double root_size, int nrx, int nry, int nrz, double* cells, long cap,
                   long* rootstart);
/* gravity.c:47-106 with tree.c:275-328 (no boundary wrap here), ghost boxes gbs[ngb][3] */
int orc_tree_gravity(const double* s, size_t N, double root_size, int nrx, int nry, int nrz, double G, double soft,
                     double opening_angle2, const double* gbs, int ngb, double* acc3, long* interactions);

/* ---------------------------------------------------------------- collisions (collision.c) */
/* Tree search (collision.c:197-269, 422-503) in the serial detection order;
 * gbs[ngb][6] is the inner ring of ghost boxes.  pairs[cap][2], gbo[cap][6].
 * Returns the number of collisions (may exceed cap; only cap are stored). */
long orc_collision_tree(const double* s, size_t N, double root_size, int nrx, int nry, int nrz, const double* gbs,
                        int ngb, int64_t* pairs, double* gbo, long cap);
/* Direct search (collision.c:64-124) in the reference's order */
long orc_collision_direct(const double* s, size_t N, const double* gbs, int ngb, int64_t* pairs, double* gbo,
                          long cap);
/* LINE (collision.c:125-196) and LINETREE (:270-331, :506-566) searches with
 * the swept-sphere test over the last step dt (dt_last_done), same output
 * layout as orc_collision_direct. */
long orc_collision_line(const double* s, size_t N, const double* gbs6, int ngb, double dt, int64_t* pairs,
                        double* gbo6, long cap);
long orc_collision_linetree(const double* s, size_t N, double root_size, int nrx, int nry, int nrz,
                            const double* gbs6, int ngb, double dt, int64_t* pairs, double* gbo6, long cap);
/* shuffle (collision.c:336-342) */
void orc_collision_shuffle(int64_t* pairs, double* gbo, long n, unsigned int* seed);
/* hardsphere (collision.c:573-665) with constant coefficient of restitution
 * (eps < 0: the reference default 1); returns the change of collisions_plog */
double orc_hardsphere(double* s, int64_t p1, int64_t p2, const double* gb6, double eps, double min_v,
                      long* log_n);

/* Algorithmic floating-point operations counted by the WHFast path since the
 * last reset (+ - * / sqrt one each, gravity pairs 20). */
unsigned long long orc_flops(int reset);

#endif
