/*
 * CPU-baseline timing driver for the reference's OpenMP build
 * (oracle/_ref/omp/librebound.so, compiled from the reference's sources).
 * Prints one JSON line.  Only the step loop is timed (CLOCK_MONOTONIC).
 *
 *   ref_bench plummer N STEPS SOFTENING DT basic|compensated SEED
 *       Plummer sphere (tools.c:463-502), leapfrog (integrator_leapfrog.c:97-209),
 *       direct gravity in the reference's OpenMP per-i form (gravity.c:215-232).
 *   ref_bench ensemble NSYS STEPS SEED SAFE_MODE KMIN KMAX
 *   ref_bench megno NSYS STEPS SEED KMIN KMAX
 *       WHFast democratic-heliocentric planetary systems, one simulation per
 *       system, `omp parallel for` over systems (each integrated serially).
 *   ref_bench sheet N STEPS ROOT_SIZE N_ROOT SEED
 *       Shearing sheet of hard spheres (the ring recipe, SEI, SHEAR with two
 *       ghost boxes, TREE gravity with OpenMP over particles (gravity.c:47-106),
 *       TREE collisions and the hard-sphere resolver with restitution 0.5).
 *       One warm-up step, then STEPS timed steps.
 *   ref_bench serve JOBFILE SCALE
 *       The serving job list (one job per line: `whfast SEED STEPS` or
 *       `nbody SEED N STEPS SOFTENING DT`), run by a pool of OpenMP threads,
 *       one job per thread at a time, each job integrated serially for
 *       max(1, STEPS/SCALE) steps.  The step loops are timed; the cost of the
 *       full job list is the measured time times SCALE (work is linear in the
 *       step count).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <time.h>
#include <omp.h>
#include "rebound.h"

static double now(void){
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9*ts.tv_nsec;
}

static int bench_plummer(int argc, char** argv){
    if (argc < 8){ fprintf(stderr, "plummer N STEPS SOFT DT MODE SEED\n"); return 1; }
    size_t N = strtoull(argv[2], 0, 10);
    int steps = atoi(argv[3]);
    double soft = atof(argv[4]);
    double dt = atof(argv[5]);
    int comp = strcmp(argv[6], "compensated") == 0;
    unsigned int seed = (unsigned int)strtoul(argv[7], 0, 10);
    struct reb_simulation* r = reb_simulation_create();
    r->rand_seed = seed;
    r->softening = soft;
    r->G = 1.;
    reb_simulation_add_plummer(r, N, 1., 1.);
    reb_simulation_move_to_com(r);
    reb_simulation_set_integrator(r, "leapfrog");
    r->gravity = comp ? REB_GRAVITY_COMPENSATED : REB_GRAVITY_BASIC;
    r->dt = dt;
    reb_simulation_steps(r, 1);                 /* warm-up step (not timed) */
    double t0 = now();
    reb_simulation_steps(r, steps);
    double t1 = now();
    double pairs = (double)N*(double)(N-1)*steps;
    printf("{\"workload\": \"plummer\", \"N\": %zu, \"steps\": %d, \"mode\": \"%s\", "
           "\"seconds\": %.6f, \"pairs_per_s\": %.6e, \"threads\": %d}\n",
           N, steps, comp ? "compensated" : "basic", t1-t0, pairs/(t1-t0), omp_get_max_threads());
    reb_simulation_free(r);
    return 0;
}

/* Planetary-system generator shared (by construction) with the product's
 * ensemble workload: star m=1, K planets, a_1=1, log-uniform period ratios in
 * [1.3,2.5], log-uniform masses in [1e-6,1e-3], Rayleigh e (0.02) and i (0.01),
 * uniform angles, primary = star, then move_to_com; dt = P_1/20. */
static struct reb_simulation* planetary_system(struct reb_simulation* g, int kmin, int kmax, int safe_mode){
    struct reb_simulation* r = reb_simulation_create();
    int K = kmin + (int)floor(reb_random_uniform(g, 0., 1.)*(double)(kmax-kmin+1));
    if (K > kmax) K = kmax;
    struct reb_particle star = {0}; star.m = 1.;
    reb_simulation_add(r, star);
    double a = 1.;
    for (int k=0; k<K; k++){
        if (k > 0){
            double ratio = exp(reb_random_uniform(g, log(1.3), log(2.5)));
            a *= pow(ratio, 2./3.);
        }
        double m  = exp(reb_random_uniform(g, log(1e-6), log(1e-3)));
        double e  = reb_random_rayleigh(g, 0.02);
        double in = reb_random_rayleigh(g, 0.01);
        double Om = reb_random_uniform(g, 0., 2.*M_PI);
        double om = reb_random_uniform(g, 0., 2.*M_PI);
        double f  = reb_random_uniform(g, 0., 2.*M_PI);
        reb_simulation_add(r, reb_particle_from_orbit(r->G, r->particles[0], m, a, e, in, Om, om, f));
    }
    reb_simulation_move_to_com(r);
    struct reb_integrator_whfast_state* wh = reb_simulation_set_integrator(r, "whfast");
    wh->coordinates = REB_INTEGRATOR_WHFAST_COORDINATES_DEMOCRATICHELIOCENTRIC;
    wh->safe_mode = safe_mode;
    r->dt = 2.*M_PI*sqrt(1./(1.+r->particles[1].m))/20.;
    return r;
}

/* MEGNO map: the ensemble's systems in Jacobi coordinates with safe_mode 1,
 * reb_simulation_init_megno_seed(r, s) for system s (tools.c:1431-1470), each
 * integrated for STEPS steps; prints the MEGNO of the first systems. */
static int bench_megno(int argc, char** argv){
    if (argc < 7){ fprintf(stderr, "megno NSYS STEPS SEED KMIN KMAX\n"); return 1; }
    int nsys = atoi(argv[2]);
    long steps = atol(argv[3]);
    unsigned int seed = (unsigned int)strtoul(argv[4], 0, 10);
    int kmin = atoi(argv[5]), kmax = atoi(argv[6]);
    struct reb_simulation* g = reb_simulation_create();
    g->rand_seed = seed;
    struct reb_simulation** sims = malloc(sizeof(*sims)*nsys);
    for (int s=0; s<nsys; s++){
        sims[s] = planetary_system(g, kmin, kmax, 1);
        struct reb_integrator_whfast_state* wh = sims[s]->integrator.state;
        wh->coordinates = REB_INTEGRATOR_WHFAST_COORDINATES_JACOBI;
        reb_simulation_init_megno_seed(sims[s], (unsigned int)s);
    }
    double t0 = now();
#pragma omp parallel for schedule(dynamic, 4)
    for (int s=0; s<nsys; s++){
        reb_simulation_steps(sims[s], (size_t)steps);
    }
    double t1 = now();
    printf("{\"workload\": \"megno\", \"systems\": %d, \"steps\": %ld, \"seconds\": %.6f, "
           "\"system_steps_per_s\": %.6e, \"threads\": %d, \"megno_first\": [",
           nsys, steps, t1-t0, (double)nsys*(double)steps/(t1-t0), omp_get_max_threads());
    for (int s=0; s<nsys && s<8; s++) printf("%s%.17g", s ? ", " : "", reb_simulation_megno(sims[s]));
    printf("]}\n");
    for (int s=0; s<nsys; s++) reb_simulation_free(sims[s]);
    free(sims);
    reb_simulation_free(g);
    return 0;
}

static int bench_ensemble(int argc, char** argv){
    if (argc < 8){ fprintf(stderr, "ensemble NSYS STEPS SEED SAFE KMIN KMAX\n"); return 1; }
    int nsys = atoi(argv[2]);
    long steps = atol(argv[3]);
    unsigned int seed = (unsigned int)strtoul(argv[4], 0, 10);
    int safe = atoi(argv[5]);
    int kmin = atoi(argv[6]), kmax = atoi(argv[7]);
    struct reb_simulation* g = reb_simulation_create();
    g->rand_seed = seed;
    struct reb_simulation** sims = malloc(sizeof(*sims)*nsys);
    for (int s=0; s<nsys; s++) sims[s] = planetary_system(g, kmin, kmax, safe);
    double* e0 = malloc(sizeof(double)*nsys);
    for (int s=0; s<nsys; s++) e0[s] = reb_simulation_energy(sims[s]);
    double t0 = now();
#pragma omp parallel for schedule(dynamic, 4)
    for (int s=0; s<nsys; s++){
        reb_simulation_steps(sims[s], (size_t)steps);
    }
    double t1 = now();
    double maxerr = 0., sumerr = 0.;
    for (int s=0; s<nsys; s++){
        double de = fabs((reb_simulation_energy(sims[s]) - e0[s])/e0[s]);
        if (de > maxerr) maxerr = de;
        sumerr += de;
        reb_simulation_free(sims[s]);
    }
    printf("{\"workload\": \"ensemble\", \"systems\": %d, \"steps\": %ld, \"safe_mode\": %d, "
           "\"seconds\": %.6f, \"system_steps_per_s\": %.6e, \"threads\": %d, "
           "\"max_rel_energy_error\": %.3e, \"mean_rel_energy_error\": %.3e}\n",
           nsys, steps, safe, t1-t0, (double)nsys*(double)steps/(t1-t0), omp_get_max_threads(),
           maxerr, sumerr/nsys);
    free(sims); free(e0);
    reb_simulation_free(g);
    return 0;
}

static double cor_half(const struct reb_simulation* const r, double v){
    (void)r; (void)v;
    return 0.5;
}

static int collisions_seen = 0;
static enum REB_COLLISION_RESOLVE_OUTCOME count_hardsphere(struct reb_simulation* const r, struct reb_collision c){
    collisions_seen++;
    return reb_collision_resolve_hardsphere(r, c);
}

static int bench_sheet(int argc, char** argv){
    if (argc < 7){ fprintf(stderr, "sheet N STEPS ROOT_SIZE N_ROOT SEED\n"); return 1; }
    size_t N = strtoull(argv[2], 0, 10);
    int steps = atoi(argv[3]);
    double root_size = atof(argv[4]);
    int nr = atoi(argv[5]);
    unsigned int seed = (unsigned int)strtoul(argv[6], 0, 10);
    struct reb_simulation* r = reb_simulation_create();
    r->rand_seed = seed;
    r->OMEGA = 1.;
    r->G = 1.546/400.;
    r->softening = 0.1;
    r->dt = 1e-3*2.*M_PI;
    r->root_size = root_size; r->N_root_x = nr; r->N_root_y = nr; r->N_root_z = 1;
    r->N_ghost_x = 2; r->N_ghost_y = 2; r->N_ghost_z = 0;
    r->boundary = REB_BOUNDARY_SHEAR;
    r->gravity = REB_GRAVITY_TREE;
    r->opening_angle2 = 0.25;
    reb_simulation_set_integrator(r, "sei");
    double bx = root_size*nr, by = root_size*nr;
    for (size_t i=0; i<N; i++){
        struct reb_particle p = {0};
        p.x = reb_random_uniform(r, -bx/2., bx/2.);
        p.y = reb_random_uniform(r, -by/2., by/2.);
        p.z = reb_random_normal(r, 1.);
        double R = reb_random_powerlaw(r, 1., 4., -3.);
        p.vx = reb_random_normal(r, 3.2*3.2);
        p.vy = -1.5*p.x*r->OMEGA + reb_random_normal(r, 3.2*3.2);
        p.vz = reb_random_normal(r, 3.2*3.2)*0.5;
        p.r = R;
        p.m = 400.*4./3.*M_PI*R*R*R;
        reb_simulation_add(r, p);
    }
    r->coefficient_of_restitution = cor_half;
    r->collision = REB_COLLISION_TREE;
    r->collision_resolve = count_hardsphere;
    reb_simulation_steps(r, 1);                 /* warm-up step (not timed) */
    collisions_seen = 0;
    int64_t logged0 = r->collisions_log_n;
    double t0 = now();
    reb_simulation_steps(r, steps);
    double t1 = now();
    /* collisions_per_step: the reference's own counter (approaching pairs that
     * were resolved, collision.c:655-663); resolver_calls_per_step: every pair
     * the search reported */
    printf("{\"workload\": \"sheet\", \"N\": %zu, \"steps\": %d, \"seconds\": %.6f, \"steps_per_s\": %.6e, "
           "\"collisions_per_step\": %.1f, \"resolver_calls_per_step\": %.1f, \"threads\": %d}\n",
           N, steps, t1-t0, steps/(t1-t0), (double)(r->collisions_log_n - logged0)/steps,
           (double)collisions_seen/steps, omp_get_max_threads());
    reb_simulation_free(r);
    return 0;
}

static int bench_serve(int argc, char** argv){
    if (argc < 4){ fprintf(stderr, "serve JOBFILE SCALE\n"); return 1; }
    FILE* f = fopen(argv[2], "r");
    if (!f){ fprintf(stderr, "cannot open %s\n", argv[2]); return 1; }
    double scale = atof(argv[3]);
    int cap = 1024, n = 0;
    struct reb_simulation** sims = malloc(sizeof(*sims)*cap);
    long* steps = malloc(sizeof(long)*cap);
    int* kind = malloc(sizeof(int)*cap);
    char word[32];
    while (fscanf(f, "%31s", word) == 1){
        if (n == cap){ cap *= 2; sims = realloc(sims, sizeof(*sims)*cap); steps = realloc(steps, sizeof(long)*cap);
                       kind = realloc(kind, sizeof(int)*cap); }
        unsigned int seed; long st;
        if (strcmp(word, "whfast") == 0){
            if (fscanf(f, "%u %ld", &seed, &st) != 2) break;
            struct reb_simulation* g = reb_simulation_create();
            g->rand_seed = seed;
            sims[n] = planetary_system(g, 2, 10, 0);
            reb_simulation_free(g);
            kind[n] = 0;
        } else {
            size_t N; double soft, dt;
            if (fscanf(f, "%u %zu %ld %lf %lf", &seed, &N, &st, &soft, &dt) != 5) break;
            struct reb_simulation* r = reb_simulation_create();
            r->rand_seed = seed; r->G = 1.; r->softening = soft;
            reb_simulation_add_plummer(r, N, 1., 1.);
            reb_simulation_move_to_com(r);
            reb_simulation_set_integrator(r, "leapfrog");
            r->gravity = REB_GRAVITY_BASIC;
            r->dt = dt;
            sims[n] = r;
            kind[n] = 1;
        }
        long s = (long)floor((double)st/scale);
        steps[n] = s < 1 ? 1 : s;
        n++;
    }
    fclose(f);
    double t0 = now();
#pragma omp parallel for schedule(dynamic, 1)
    for (int k=0; k<n; k++){
        reb_simulation_steps(sims[k], (size_t)steps[k]);
    }
    double t1 = now();
    int nw = 0, nn = 0;
    for (int k=0; k<n; k++){ if (kind[k]) nn++; else nw++; reb_simulation_free(sims[k]); }
    printf("{\"workload\": \"serve\", \"jobs\": %d, \"whfast_jobs\": %d, \"nbody_jobs\": %d, "
           "\"scale\": %g, \"seconds\": %.6f, \"projected_seconds\": %.6f, \"jobs_per_s\": %.6e, \"threads\": %d}\n",
           n, nw, nn, scale, t1-t0, (t1-t0)*scale, (double)n/((t1-t0)*scale), omp_get_max_threads());
    free(sims); free(steps); free(kind);
    return 0;
}

int main(int argc, char** argv){
    if (argc < 2){ fprintf(stderr, "usage: ref_bench plummer|ensemble ...\n"); return 1; }
    if (strcmp(argv[1], "plummer") == 0) return bench_plummer(argc, argv);
    if (strcmp(argv[1], "ensemble") == 0) return bench_ensemble(argc, argv);
    if (strcmp(argv[1], "megno") == 0) return bench_megno(argc, argv);
    if (strcmp(argv[1], "serve") == 0) return bench_serve(argc, argv);
    if (strcmp(argv[1], "sheet") == 0) return bench_sheet(argc, argv);
    fprintf(stderr, "unknown workload %s\n", argv[1]);
    return 1;
}
