/*
 * Golden-vector generator.  Links the reference library compiled from
 * /root/reference/src by oracle/Makefile (target `ref`) and writes small .npy
 * fixtures into tests/golden/.  Runs only where the reference sources exist;
 * the fixtures it writes are committed and are what the GPU box sees.
 *
 * Build + run: python tests/golden/make_golden.py
 *
 * Every fixture here is produced by the reference's own functions:
 *   rand_r / reb_random_*                tools.c:64-105, rebound.c:238-259
 *   reb_simulation_add_plummer           tools.c:463-502
 *   reb_gravity_{basic,compensated}      gravity.c:167-531 (through update_acceleration)
 *   leapfrog + energy                    integrator_leapfrog.c:97-209, tools.c:108-162
 *   whfast kepler solver                 integrator_whfast.c:241-427
 *   whfast DH integration                integrator_whfast.c:1113-1259
 *   whfast Jacobi + symplectic correctors integrator_whfast.c:653-784, 1051-1259
 *   tree build / moments / walk          tree.c:64-328, gravity.c:47-106
 *   collision search + hardsphere        collision.c:49-405, 573-665 (DIRECT, TREE, LINE, LINETREE)
 *   SEI + shear boundary                 integrator_sei.c:86-174, boundary.c:35-201
 *   binary files / simulation archives   binarydata.c:42-932, simulationarchive.c:45-666
 *   variational equations + MEGNO        integrator_whfast.c:392-490, tools.c:1313-1522
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <stdint.h>
#include "rebound.h"
#include "tree.h"
#include "boundary.h"
#ifdef _OPENMP
#include <omp.h>
#endif

static const char* outdir = ".";
/* Raw state dumps (x y z vx vy vz of body 0..N-1, in order) go to this
 * folder; make_golden.py hashes each with SHA-256 and deletes it. */
static const char* rawdir = NULL;

/* ---------------------------------------------------------------- npy out */
static void write_npy(const char* name, const char* descr, size_t elsize,
                      const void* data, int ndim, const size_t* shape){
    char path[1024];
    snprintf(path, sizeof(path), "%s/%s", outdir, name);
    FILE* f = fopen(path, "wb");
    if (!f){ perror(path); exit(1); }
    char shp[256] = {0};
    size_t n = 1;
    if (ndim == 1){
        snprintf(shp, sizeof(shp), "(%zu,)", shape[0]);
    }else{
        int off = snprintf(shp, sizeof(shp), "(");
        for (int d=0; d<ndim; d++){
            off += snprintf(shp+off, sizeof(shp)-off, "%zu%s", shape[d], d==ndim-1?")":", ");
        }
    }
    for (int d=0; d<ndim; d++) n *= shape[d];
    char header[512];
    int hl = snprintf(header, sizeof(header),
        "{'descr': '%s', 'fortran_order': False, 'shape': %s, }", descr, shp);
    int total = 10 + hl + 1;
    int pad = (64 - total % 64) % 64;
    fwrite("\x93NUMPY\x01\x00", 1, 8, f);
    uint16_t hlen = (uint16_t)(hl + pad + 1);
    fwrite(&hlen, 2, 1, f);
    fwrite(header, 1, hl, f);
    for (int i=0; i<pad; i++) fputc(' ', f);
    fputc('\n', f);
    fwrite(data, elsize, n, f);
    fclose(f);
}
static void wd(const char* name, const double* d, size_t n0, size_t n1){
    size_t shape[2] = {n0, n1};
    write_npy(name, "<f8", 8, d, n1 ? 2 : 1, shape);
}
static void wi(const char* name, const int64_t* d, size_t n0, size_t n1){
    size_t shape[2] = {n0, n1};
    write_npy(name, "<i8", 8, d, n1 ? 2 : 1, shape);
}

/* state layout used by every fixture: x y z vx vy vz ax ay az m r */
#define NCOL 11
static double* state_of(struct reb_simulation* r){
    double* s = malloc(sizeof(double)*NCOL*r->N);
    for (size_t i=0; i<r->N; i++){
        struct reb_particle p = r->particles[i];
        double row[NCOL] = {p.x,p.y,p.z,p.vx,p.vy,p.vz,p.ax,p.ay,p.az,p.m,p.r};
        memcpy(s+NCOL*i, row, sizeof(row));
    }
    return s;
}
static void dump_state(const char* name, struct reb_simulation* r){
    double* s = state_of(r);
    wd(name, s, r->N, NCOL);
    free(s);
}
static double* acc_of(struct reb_simulation* r){
    double* a = malloc(sizeof(double)*3*r->N);
    for (size_t i=0; i<r->N; i++){
        a[3*i+0] = r->particles[i].ax;
        a[3*i+1] = r->particles[i].ay;
        a[3*i+2] = r->particles[i].az;
    }
    return a;
}

/* ------------------------------------------------------------ fixtures */
static void gen_rng(void){
    unsigned int seeds[3] = {1u, 42u, 12345u};
    int64_t out[3*64];
    for (int s=0; s<3; s++){
        unsigned int seed = seeds[s];
        for (int i=0; i<64; i++) out[s*64+i] = rand_r(&seed);
    }
    wi("rand_r_streams.npy", out, 3, 64);

    struct reb_simulation* r = reb_simulation_create();
    r->rand_seed = 7;
    double st[4*256];
    for (int i=0; i<256; i++) st[0*256+i] = reb_random_uniform(r, -1., 3.);
    for (int i=0; i<256; i++) st[1*256+i] = reb_random_normal(r, 2.);
    for (int i=0; i<256; i++) st[2*256+i] = reb_random_powerlaw(r, 1., 4., -3.);
    for (int i=0; i<256; i++) st[3*256+i] = reb_random_rayleigh(r, 0.02);
    wd("random_streams_seed7.npy", st, 4, 256);
    reb_simulation_free(r);
}

static struct reb_simulation* plummer(size_t N, unsigned int seed, double soft){
    struct reb_simulation* r = reb_simulation_create();
    r->rand_seed = seed;
    r->softening = soft;
    r->G = 1.;
    reb_simulation_add_plummer(r, N, 1., 1.);
    reb_simulation_move_to_com(r);
    return r;
}

static void gen_plummer(void){
    /* P1 state, BASIC + COMPENSATED accelerations, leapfrog energy history */
    struct reb_simulation* r = plummer(1024, 1u, 0.01);
    dump_state("plummer1024_seed1_state.npy", r);
    r->gravity = REB_GRAVITY_BASIC;
    reb_simulation_update_acceleration(r);
    double* a = acc_of(r); wd("plummer1024_seed1_acc_basic.npy", a, r->N, 3); free(a);
    r->gravity = REB_GRAVITY_COMPENSATED;
    reb_simulation_update_acceleration(r);
    a = acc_of(r); wd("plummer1024_seed1_acc_comp.npy", a, r->N, 3); free(a);

    /* test-particle variant: the last quarter are test particles of type 1 */
    r->N_active = 768; r->testparticle_type = 1;
    r->gravity = REB_GRAVITY_BASIC;
    reb_simulation_update_acceleration(r);
    a = acc_of(r); wd("plummer1024_seed1_acc_basic_tp1.npy", a, r->N, 3); free(a);
    r->gravity = REB_GRAVITY_COMPENSATED;
    reb_simulation_update_acceleration(r);
    a = acc_of(r); wd("plummer1024_seed1_acc_comp_tp1.npy", a, r->N, 3); free(a);
    r->N_active = SIZE_MAX; r->testparticle_type = 0;

    /* 1000 leapfrog steps, energy every 10 steps */
    r->gravity = REB_GRAVITY_BASIC;
    reb_simulation_set_integrator(r, "leapfrog");
    r->dt = 1e-3;
    double e[101];
    e[0] = reb_simulation_energy(r);
    for (int k=1; k<=100; k++){
        reb_simulation_steps(r, 10);
        e[k] = reb_simulation_energy(r);
    }
    wd("plummer1024_seed1_lf_energy.npy", e, 101, 0);
    dump_state("plummer1024_seed1_lf_final.npy", r);
    reb_simulation_free(r);

    /* split-precision error model input: N=4096 */
    r = plummer(4096, 2u, 0.005);
    dump_state("plummer4096_seed2_state.npy", r);
    reb_simulation_update_acceleration(r);
    a = acc_of(r); wd("plummer4096_seed2_acc_basic.npy", a, r->N, 3); free(a);
    reb_simulation_free(r);
}

/* ghost boxes make the serial and OpenMP BASIC orders differ; the OpenMP
 * per-i order is the one emitted by the OMP binary of this generator */
static void gen_ghostbox_basic(const char* tag){
    struct reb_simulation* r = reb_simulation_create();
    r->rand_seed = 3;
    r->boundary = REB_BOUNDARY_PERIODIC;
    r->root_size = 10.; r->N_root_x = 1; r->N_root_y = 1; r->N_root_z = 1;
    r->N_ghost_x = 1; r->N_ghost_y = 1; r->N_ghost_z = 0;
    r->softening = 0.05;
    for (int i=0; i<256; i++){
        struct reb_particle p = {0};
        p.x = reb_random_uniform(r, -5., 5.);
        p.y = reb_random_uniform(r, -5., 5.);
        p.z = reb_random_uniform(r, -5., 5.);
        p.m = reb_random_uniform(r, 0.5, 1.5)/256.;
        reb_simulation_add(r, p);
    }
    char name[128];
    snprintf(name, sizeof(name), "periodic256_state.npy");
    dump_state(name, r);
    reb_simulation_update_acceleration(r);
    double* a = acc_of(r);
    snprintf(name, sizeof(name), "periodic256_acc_basic_%s.npy", tag);
    wd(name, a, r->N, 3); free(a);
    reb_simulation_free(r);
}

static void gen_plummer_omp(void){
    struct reb_simulation* r = plummer(1024, 1u, 0.01);
    r->gravity = REB_GRAVITY_BASIC;
    reb_simulation_update_acceleration(r);
    double* a = acc_of(r); wd("plummer1024_seed1_acc_basic_omp.npy", a, r->N, 3); free(a);
    r->gravity = REB_GRAVITY_COMPENSATED;
    reb_simulation_update_acceleration(r);
    a = acc_of(r); wd("plummer1024_seed1_acc_comp_omp.npy", a, r->N, 3); free(a);
    reb_simulation_free(r);
}

/* Kepler-solver vectors: (x y z vx vy vz mu dt) -> (x y z vx vy vz) */
static void gen_kepler(void){
    struct reb_simulation* r = reb_simulation_create();
    r->rand_seed = 11;
    const int NK = 640;
    double* in = malloc(sizeof(double)*NK*8);
    double* out = malloc(sizeof(double)*NK*6);
    struct reb_particle star = {0}; star.m = 1.;
    for (int k=0; k<NK; k++){
        struct reb_particle p;
        double mu = reb_random_uniform(r, 0.5, 2.);
        double dt;
        int cls = k % 8;
        if (cls < 4){       /* elliptic, small to moderate steps */
            double e = cls==3 ? reb_random_uniform(r, 0.9, 0.99) : reb_random_uniform(r, 0., 0.6);
            double a = reb_random_uniform(r, 0.5, 3.);
            p = reb_particle_from_orbit(mu, star, 0., a, e, reb_random_uniform(r, 0., M_PI),
                    reb_random_uniform(r, 0., 2.*M_PI), reb_random_uniform(r, 0., 2.*M_PI),
                    reb_random_uniform(r, 0., 2.*M_PI));
            double P = 2.*M_PI*sqrt(a*a*a/mu);
            dt = P*reb_random_uniform(r, -0.2, 0.2);
        }else if (cls < 6){ /* elliptic, steps of several orbits (quartic / bisection branches) */
            double e = reb_random_uniform(r, 0., 0.8);
            double a = reb_random_uniform(r, 0.5, 3.);
            p = reb_particle_from_orbit(mu, star, 0., a, e, reb_random_uniform(r, 0., M_PI),
                    reb_random_uniform(r, 0., 2.*M_PI), reb_random_uniform(r, 0., 2.*M_PI),
                    reb_random_uniform(r, 0., 2.*M_PI));
            double P = 2.*M_PI*sqrt(a*a*a/mu);
            dt = P*reb_random_uniform(r, 1.5, 7.);
        }else if (cls == 6){ /* hyperbolic */
            double e = reb_random_uniform(r, 1.1, 5.);
            double a = -reb_random_uniform(r, 0.5, 3.);
            double fmax = acos(-1./e)*0.9;
            p = reb_particle_from_orbit(mu, star, 0., a, e, reb_random_uniform(r, 0., M_PI),
                    reb_random_uniform(r, 0., 2.*M_PI), reb_random_uniform(r, 0., 2.*M_PI),
                    reb_random_uniform(r, -fmax, fmax));
            dt = reb_random_uniform(r, -2., 2.);
        }else{               /* (almost) straight line: huge velocity */
            p = (struct reb_particle){0};
            p.x = reb_random_uniform(r, -1., 1.); p.y = reb_random_uniform(r, -1., 1.); p.z = reb_random_uniform(r, -1., 1.);
            p.vx = 1e8*reb_random_uniform(r, -1., 1.); p.vy = 1e8*reb_random_uniform(r, -1., 1.); p.vz = 1e8*reb_random_uniform(r, -1., 1.);
            dt = reb_random_uniform(r, 1e-3, 1e-2);
        }
        double row[8] = {p.x,p.y,p.z,p.vx,p.vy,p.vz,mu,dt};
        memcpy(in+8*k, row, sizeof(row));
        reb_integrator_whfast_kepler_solver(&p, mu, dt, NULL);
        double o[6] = {p.x,p.y,p.z,p.vx,p.vy,p.vz};
        memcpy(out+6*k, o, sizeof(o));
    }
    wd("kepler_in.npy", in, NK, 8);
    wd("kepler_out.npy", out, NK, 6);
    free(in); free(out);
    reb_simulation_free(r);
}

/* Planetary systems as in the batched-ensemble workload; 32 systems, K=2..8 */
static void gen_whfast(void){
    const int NS = 32;
    struct reb_simulation* g = reb_simulation_create();
    g->rand_seed = 2024;
    size_t maxb = 9;
    double* init = calloc(NS*maxb*NCOL, sizeof(double));
    int64_t* meta = calloc(NS*2, sizeof(int64_t));
    double* dts = calloc(NS, sizeof(double));
    for (int mode=0; mode<2; mode++){
        double* e = calloc(NS*11, sizeof(double));
        double* fin = calloc(NS*maxb*NCOL, sizeof(double));
        unsigned int gseed = 2024;
        g->rand_seed = gseed;
        for (int s=0; s<NS; s++){
            struct reb_simulation* r = reb_simulation_create();
            int K = 2 + s % 7;
            struct reb_particle star = {0}; star.m = 1.;
            reb_simulation_add(r, star);
            double a = 1.;
            for (int k=0; k<K; k++){
                if (k>0){
                    double ratio = exp(reb_random_uniform(g, log(1.3), log(2.5)));
                    a *= pow(ratio, 2./3.);
                }
                double m = exp(reb_random_uniform(g, log(1e-6), log(1e-3)));
                double e_ = reb_random_rayleigh(g, 0.02);
                double inc = reb_random_rayleigh(g, 0.01);
                double Om = reb_random_uniform(g, 0., 2.*M_PI);
                double om = reb_random_uniform(g, 0., 2.*M_PI);
                double f  = reb_random_uniform(g, 0., 2.*M_PI);
                struct reb_particle p = reb_particle_from_orbit(r->G, r->particles[0], m, a, e_, inc, Om, om, f);
                reb_simulation_add(r, p);
            }
            reb_simulation_move_to_com(r);
            struct reb_integrator_whfast_state* wh = reb_simulation_set_integrator(r, "whfast");
            wh->coordinates = REB_INTEGRATOR_WHFAST_COORDINATES_DEMOCRATICHELIOCENTRIC;
            wh->safe_mode = mode;
            r->dt = 2.*M_PI*sqrt(1./(1.+r->particles[1].m))/20.;  /* inner period / 20 */
            if (mode==0){
                meta[2*s] = K; meta[2*s+1] = K+1;
                dts[s] = r->dt;
                double* st = state_of(r);
                memcpy(init+s*maxb*NCOL, st, sizeof(double)*NCOL*r->N);
                free(st);
            }
            e[s*11] = reb_simulation_energy(r);
            for (int c=1; c<=10; c++){
                reb_simulation_steps(r, 100);
                e[s*11+c] = reb_simulation_energy(r);
            }
            double* st = state_of(r);
            memcpy(fin+s*maxb*NCOL, st, sizeof(double)*NCOL*r->N);
            free(st);
            reb_simulation_free(r);
        }
        char name[128];
        snprintf(name, sizeof(name), "whfast_dh32_safe%d_energy.npy", mode);
        wd(name, e, NS, 11);
        snprintf(name, sizeof(name), "whfast_dh32_safe%d_final.npy", mode);
        size_t shape[3] = {NS, maxb, NCOL};
        write_npy(name, "<f8", 8, fin, 3, shape);
        free(e); free(fin);
    }
    size_t shape[3] = {NS, maxb, NCOL};
    write_npy("whfast_dh32_init.npy", "<f8", 8, init, 3, shape);
    wi("whfast_dh32_meta.npy", meta, NS, 2);
    wd("whfast_dh32_dt.npy", dts, NS, 0);
    free(init); free(meta); free(dts);
    reb_simulation_free(g);
}

/* --------------------------------------------------- shearing sheet */
static double cor_half(const struct reb_simulation* const r, double v){ (void)r; (void)v; return 0.5; }

static int64_t* rec_pairs = NULL;
static double*  rec_gb = NULL;
static size_t   rec_n = 0, rec_cap = 0;
static enum REB_COLLISION_RESOLVE_OUTCOME rec_resolve(struct reb_simulation* const r, struct reb_collision c){
    if (rec_n == rec_cap){
        rec_cap = rec_cap ? 2*rec_cap : 1024;
        rec_pairs = realloc(rec_pairs, sizeof(int64_t)*2*rec_cap);
        rec_gb = realloc(rec_gb, sizeof(double)*6*rec_cap);
    }
    rec_pairs[2*rec_n] = (int64_t)c.p1; rec_pairs[2*rec_n+1] = (int64_t)c.p2;
    double g[6] = {c.gb.x,c.gb.y,c.gb.z,c.gb.vx,c.gb.vy,c.gb.vz};
    memcpy(rec_gb+6*rec_n, g, sizeof(g));
    rec_n++;
    return reb_collision_resolve_hardsphere(r, c);
}

/* Small sheet following the shearing-sheet workload recipe (Omega=1, rho=400,
 * G such that G*rho/Omega^2 = 1.546, radii R^-3 in [1,4]). */
static struct reb_simulation* sheet(size_t N, double root_size, int nrx, int nry, unsigned int seed){
    struct reb_simulation* r = reb_simulation_create();
    r->rand_seed = seed;
    r->OMEGA = 1.;
    r->G = 1.546/400.;
    r->softening = 0.1;
    r->dt = 1e-3*2.*M_PI;
    r->root_size = root_size; r->N_root_x = nrx; r->N_root_y = nry; r->N_root_z = 1;
    r->N_ghost_x = 2; r->N_ghost_y = 2; r->N_ghost_z = 0;
    r->boundary = REB_BOUNDARY_SHEAR;
    r->gravity = REB_GRAVITY_TREE;
    r->opening_angle2 = 0.25;
    reb_simulation_set_integrator(r, "sei");
    double bx = root_size*nrx, by = root_size*nry;
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
    return r;
}

/* x y z vx vy vz of body 0..N-1, in order, appended to the raw file `name`
 * (one record per call): make_golden.py hashes each record with SHA-256. */
static void dump_raw6(const char* name, const struct reb_simulation* r){
    if (!rawdir) return;
    char path[1024];
    snprintf(path, sizeof(path), "%s/%s", rawdir, name);
    FILE* f = fopen(path, "ab");
    if (!f){ perror(path); exit(1); }
    for (size_t i=0; i<r->N; i++){
        const double v[6] = {r->particles[i].x, r->particles[i].y, r->particles[i].z,
                             r->particles[i].vx, r->particles[i].vy, r->particles[i].vz};
        fwrite(v, 8, 6, f);
    }
    fclose(f);
}

/* The benchmark's full-size sheet (200000 bodies, 10 x 10 root boxes of 136,
 * seed 99, tree gravity and tree collisions with the hard-sphere resolver):
 * after each of three reb_simulation_steps(r, 1) calls, the bodies in order
 * (hashed), collisions_log_n, collisions_plog and rand_seed. */
static void gen_sheet_full(void){
    struct reb_simulation* r = sheet(200000, 136., 10, 10, 99u);
    r->collision = REB_COLLISION_TREE;
    r->collision_resolve = reb_collision_resolve_hardsphere;
    int64_t meta[3][3];
    double plog[3];
    for (int s=0; s<3; s++){
        reb_simulation_steps(r, 1);
        dump_raw6("sheet200k.raw", r);
        meta[s][0] = 0;
        meta[s][1] = (int64_t)r->collisions_log_n;
        meta[s][2] = (int64_t)r->rand_seed;
        plog[s] = r->collisions_plog;
    }
    wi("sheet200k_meta.npy", &meta[0][0], 3, 3);
    wd("sheet200k_plog.npy", plog, 3, 0);
    reb_simulation_free(r);
}

/* The direct-summation benchmark at full size: Plummer sphere of 65536
 * bodies (seed 42, softening 0.005, G = 1), leapfrog with dt 5e-4 and BASIC
 * gravity.  The bodies in order (hashed) after construction and after each
 * of two steps. */
static void gen_plummer_full(void){
    struct reb_simulation* r = plummer(65536, 42u, 0.005);
    reb_simulation_set_integrator(r, "leapfrog");
    r->gravity = REB_GRAVITY_BASIC;
    r->dt = 5e-4;
    dump_raw6("plummer65536_seed42.raw", r);
    for (int s=1; s<3; s++){
        reb_simulation_steps(r, 1);
        dump_raw6("plummer65536_seed42.raw", r);
    }
    reb_simulation_free(r);
}

/* pre-order dump of the reference tree: center, w, m, com, pt */
static void dump_cell(struct reb_treecell* c, double** buf, size_t* n, size_t* cap){
    if (!c) return;
    if (*n == *cap){ *cap = *cap ? 2**cap : 4096; *buf = realloc(*buf, sizeof(double)*9**cap); }
    double row[9] = {c->x, c->y, c->z, c->w, c->m, c->mx, c->my, c->mz, (double)c->pt};
    memcpy(*buf + 9**n, row, sizeof(row));
    (*n)++;
    for (int o=0; o<8; o++) dump_cell(c->oct[o], buf, n, cap);
}

static void gen_sheet(void){
    struct reb_simulation* r = sheet(2000, 50., 2, 2, 99u);
    dump_state("sheet2000_init.npy", r);
    int64_t seedinit = r->rand_seed;
    wi("sheet2000_seed_init.npy", &seedinit, 1, 0);
    /* tree + tree gravity on the initial state (boundary wrap is a no-op here) */
    reb_tree_construct(r);
    reb_tree_calculate_gravity_data(r);
    double* buf = NULL; size_t n = 0, cap = 0;
    size_t Nroot = r->N_root_x*r->N_root_y*r->N_root_z;
    int64_t* rootstart = calloc(Nroot+1, sizeof(int64_t));
    for (size_t i=0; i<Nroot; i++){ rootstart[i] = (int64_t)n; dump_cell(r->tree_root[i], &buf, &n, &cap); }
    rootstart[Nroot] = (int64_t)n;
    reb_tree_delete(r);
    wd("sheet2000_tree.npy", buf, n, 9);
    wi("sheet2000_tree_rootstart.npy", rootstart, Nroot+1, 0);
    free(buf); free(rootstart);
    reb_simulation_update_acceleration(r);
    double* a = acc_of(r); wd("sheet2000_acc_tree.npy", a, r->N, 3); free(a);
    /* opening_angle2 = 0 opens every cell: tree gravity equals direct summation (order differs) */
    /* five SEI steps with tree collisions + hardsphere, collisions recorded in resolve order */
    r->collision = REB_COLLISION_TREE;
    r->collision_resolve = rec_resolve;
    int64_t counts[5];
    for (int s=0; s<5; s++){
        size_t before = rec_n;
        reb_simulation_steps(r, 1);
        counts[s] = (int64_t)(rec_n - before);
        char name[128];
        snprintf(name, sizeof(name), "sheet2000_step%d_state.npy", s+1);
        dump_state(name, r);
    }
    wi("sheet2000_collision_pairs.npy", rec_pairs, rec_n, 2);
    wd("sheet2000_collision_gb.npy", rec_gb, rec_n, 6);
    wi("sheet2000_collision_counts.npy", counts, 5, 0);
    int64_t seedafter = r->rand_seed;
    wi("sheet2000_seed_after.npy", &seedafter, 1, 0);
    double plog[2] = {r->collisions_plog, (double)r->collisions_log_n};
    wd("sheet2000_plog.npy", plog, 2, 0);
    reb_simulation_free(r);

    /* DIRECT search on the same initial state, one step */
    rec_n = 0;
    r = sheet(2000, 50., 2, 2, 99u);
    r->collision = REB_COLLISION_DIRECT;
    r->collision_resolve = rec_resolve;
    reb_simulation_steps(r, 1);
    wi("sheet2000_direct_pairs.npy", rec_pairs, rec_n, 2);
    wd("sheet2000_direct_gb.npy", rec_gb, rec_n, 6);
    dump_state("sheet2000_direct_step1_state.npy", r);
    reb_simulation_free(r);
}

/* Ensemble workload generator (the draw order of the product's
 * planetary_systems and of oracle/ref_bench.c): K uniform in {kmin..kmax},
 * then per planet period ratio (k>0), mass, e, inc, Omega, omega, f. */
static struct reb_simulation* ens_system(struct reb_simulation* g, int kmin, int kmax){
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
    return r;
}

/* An ensemble of NS systems: initial bodies, dt, energies after 0, n1, n1+n2,
 * n1+n2+n3 steps (reb_simulation_steps calls) and the final bodies. */
static void gen_ensemble(const char* tag, unsigned int seed, int NS, int kmin, int kmax, int safe, const int* nsteps){
    struct reb_simulation* g = reb_simulation_create();
    g->rand_seed = seed;
    size_t cap = (size_t)NS*(kmax+1);
    double* init = calloc(cap*NCOL, sizeof(double));
    double* fin = calloc(cap*NCOL, sizeof(double));
    int64_t* nb = calloc(NS, sizeof(int64_t));
    double* dts = calloc(NS, sizeof(double));
    double* en = calloc(NS*4, sizeof(double));
    size_t row = 0;
    for (int s=0; s<NS; s++){
        struct reb_simulation* r = ens_system(g, kmin, kmax);
        struct reb_integrator_whfast_state* wh = reb_simulation_set_integrator(r, "whfast");
        wh->coordinates = REB_INTEGRATOR_WHFAST_COORDINATES_DEMOCRATICHELIOCENTRIC;
        wh->safe_mode = safe;
        r->dt = r->N > 1 ? 2.*M_PI*sqrt(1./(1.+r->particles[1].m))/20. : 1e-3;
        nb[s] = (int64_t)r->N;
        dts[s] = r->dt;
        double* st = state_of(r);
        memcpy(init+row*NCOL, st, sizeof(double)*NCOL*r->N);
        free(st);
        en[s*4] = reb_simulation_energy(r);
        for (int c=0; c<3; c++){
            reb_simulation_steps(r, (size_t)nsteps[c]);
            en[s*4+1+c] = reb_simulation_energy(r);
        }
        st = state_of(r);
        memcpy(fin+row*NCOL, st, sizeof(double)*NCOL*r->N);
        free(st);
        row += r->N;
        reb_simulation_free(r);
    }
    char name[128];
    snprintf(name, sizeof(name), "%s_init.npy", tag); wd(name, init, row, NCOL);
    snprintf(name, sizeof(name), "%s_final.npy", tag); wd(name, fin, row, NCOL);
    snprintf(name, sizeof(name), "%s_nbodies.npy", tag); wi(name, nb, NS, 0);
    snprintf(name, sizeof(name), "%s_dt.npy", tag); wd(name, dts, NS, 0);
    snprintf(name, sizeof(name), "%s_energy.npy", tag); wd(name, en, NS, 4);
    free(init); free(fin); free(nb); free(dts); free(en);
    reb_simulation_free(g);
}

/* LINE and LINETREE searches (collision.c:125-196, :270-331, :506-566) on
 * the small sheet: collisions in resolve order (hard spheres applied) and the
 * bodies after each step. */
/* The search input of a step: the bodies, time and seed at the first resolve
 * call of the step (no collision has been resolved yet). */
static int pre_pending = 0;
static double* pre_state = NULL;
static double pre_t = 0.;
static enum REB_COLLISION_RESOLVE_OUTCOME rec_resolve_pre(struct reb_simulation* const r, struct reb_collision c){
    if (pre_pending){
        pre_pending = 0;
        pre_state = state_of(r);
        pre_t = r->t;
    }
    return rec_resolve(r, c);
}

static void gen_line(void){
    const int kinds[2] = {REB_COLLISION_LINE, REB_COLLISION_LINETREE};
    const char* tags[2] = {"line", "linetree"};
    const int nsteps[2] = {1, 3};
    for (int k=0; k<2; k++){
        rec_n = 0;
        struct reb_simulation* r = sheet(2000, 50., 2, 2, 99u);
        r->collision = kinds[k];
        r->collision_resolve = rec_resolve_pre;
        int64_t counts[3] = {0, 0, 0};
        char name[128];
        for (int st=0; st<nsteps[k]; st++){
            size_t before = rec_n;
            if (st == 0){
                pre_pending = 1;
                /* the shuffle draws from the seed the search starts with */
                int64_t seed = (int64_t)r->rand_seed;
                snprintf(name, sizeof(name), "sheet2000_%s_seed_before.npy", tags[k]); wi(name, &seed, 1, 0);
            }
            reb_simulation_steps(r, 1);
            if (st == 0 && pre_state){
                snprintf(name, sizeof(name), "sheet2000_%s_search_state.npy", tags[k]);
                wd(name, pre_state, r->N, NCOL);
                double meta[2] = {pre_t, r->dt_last_done};
                snprintf(name, sizeof(name), "sheet2000_%s_search_t_dt.npy", tags[k]); wd(name, meta, 2, 0);
                free(pre_state);
                pre_state = NULL;
            }
            counts[st] = (int64_t)(rec_n - before);
            snprintf(name, sizeof(name), "sheet2000_%s_step%d_state.npy", tags[k], st+1);
            dump_state(name, r);
        }
        snprintf(name, sizeof(name), "sheet2000_%s_pairs.npy", tags[k]); wi(name, rec_pairs, rec_n, 2);
        snprintf(name, sizeof(name), "sheet2000_%s_gb.npy", tags[k]); wd(name, rec_gb, rec_n, 6);
        snprintf(name, sizeof(name), "sheet2000_%s_counts.npy", tags[k]); wi(name, counts, nsteps[k], 0);
        reb_simulation_free(r);
    }
}

/* Binary files and archives (binarydata.c, simulationarchive.c): a leapfrog
 * simulation saved before any step, and a WHFast archive with a snapshot
 * every 7 steps (unsynchronized, safe_mode 0) over 30 steps; the bodies at
 * each snapshot's step and after the 30 steps. */
static void gen_archive(void){
    char path[512];
    struct reb_simulation* r = reb_simulation_create();
    r->rand_seed = 77u;
    r->G = 1.; r->softening = 0.01; r->dt = 0.01;
    reb_simulation_add_plummer(r, 64, 1., 1.);
    reb_simulation_move_to_com(r);
    struct reb_integrator_leapfrog_state* lf = reb_simulation_set_integrator(r, "leapfrog");
    lf->order = 4;
    snprintf(path, sizeof(path), "%s/archive_leapfrog.bin", outdir);
    remove(path);
    reb_simulation_save_to_file(r, path);
    dump_state("archive_leapfrog_state.npy", r);
    reb_simulation_free(r);

    struct reb_simulation* g = reb_simulation_create();
    g->rand_seed = 515u;
    r = ens_system(g, 6, 6);
    reb_simulation_free(g);
    struct reb_integrator_whfast_state* wh = reb_simulation_set_integrator(r, "whfast");
    wh->safe_mode = 0;
    wh->corrector = 5;
    r->dt = 2.*M_PI*sqrt(1./(1.+r->particles[1].m))/20.;
    snprintf(path, sizeof(path), "%s/archive_whfast.bin", outdir);
    remove(path);
    reb_simulation_save_to_file_step(r, path, 7);
    reb_simulation_steps(r, 30);
    dump_state("archive_whfast_final_state.npy", r);
    reb_simulation_free(r);
    /* every snapshot as the reference reads it back */
    for (int k=0; k<5; k++){
        struct reb_simulation* q = reb_simulation_create_from_file(path, k);
        char name[128];
        snprintf(name, sizeof(name), "archive_whfast_snap%d_state.npy", k);
        dump_state(name, q);
        int64_t meta[3] = {(int64_t)q->steps_done, (int64_t)q->is_synchronized, (int64_t)q->N};
        snprintf(name, sizeof(name), "archive_whfast_snap%d_meta.npy", k);
        wi(name, meta, 3, 0);
        reb_simulation_free(q);
    }
}

/* WHFast in Jacobi coordinates (the reference's default): NS systems of
 * 1..8 planets, each integrated under every mode below for three
 * reb_simulation_steps calls of `nsteps` steps.  Per mode: energies after 0,
 * 1, 2, 3 calls and the final bodies; the modes as rows of an int table
 * (safe_mode, corrector, corrector2, compensated, N_active offset from N
 * (0: all active), testparticle_type). */
static void gen_whfast_jacobi(void){
    const int NS = 16;
    const int nsteps = 100;
    enum { NMODE = 8 };
    static const int64_t modes[NMODE][6] = {
        {1, 0, 0, 0, 0, 0},
        {0, 0, 0, 0, 0, 0},
        {0, 11, 0, 0, 0, 0},
        {1, 17, 1, 0, 0, 0},
        {0, 3, 1, 1, 0, 0},
        {0, 5, 0, 0, 2, 0},
        {1, 7, 0, 0, 2, 1},
        {0, 0, 0, 1, 3, 1},
    };
    struct reb_simulation* g = reb_simulation_create();
    g->rand_seed = 4242u;
    struct reb_simulation* sys[NS];
    size_t rows = 0;
    int64_t nb[NS];
    double dts[NS];
    for (int s=0; s<NS; s++){
        sys[s] = ens_system(g, 1 + s % 8, 1 + s % 8);
        nb[s] = (int64_t)sys[s]->N;
        dts[s] = 2.*M_PI*sqrt(1./(1.+sys[s]->particles[1].m))/20.;
        rows += sys[s]->N;
    }
    double* init = calloc(rows*NCOL, sizeof(double));
    size_t row = 0;
    for (int s=0; s<NS; s++){
        double* st = state_of(sys[s]);
        memcpy(init+row*NCOL, st, sizeof(double)*NCOL*sys[s]->N);
        free(st);
        row += sys[s]->N;
    }
    wd("whfast_jac16_init.npy", init, rows, NCOL);
    wi("whfast_jac16_nbodies.npy", nb, NS, 0);
    wd("whfast_jac16_dt.npy", dts, NS, 0);
    wi("whfast_jac16_modes.npy", &modes[0][0], NMODE, 6);
    for (int md=0; md<NMODE; md++){
        double* fin = calloc(rows*NCOL, sizeof(double));
        double* en = calloc(NS*4, sizeof(double));
        row = 0;
        for (int s=0; s<NS; s++){
            struct reb_simulation* r = reb_simulation_copy(sys[s]);
            struct reb_integrator_whfast_state* wh = reb_simulation_set_integrator(r, "whfast");
            wh->coordinates = REB_INTEGRATOR_WHFAST_COORDINATES_JACOBI;
            wh->safe_mode = (unsigned int)modes[md][0];
            wh->corrector = (unsigned int)modes[md][1];
            wh->corrector2 = (unsigned int)modes[md][2];
            r->gravity = modes[md][3] ? REB_GRAVITY_COMPENSATED : REB_GRAVITY_BASIC;
            if (modes[md][4] && (int64_t)r->N > modes[md][4]) r->N_active = r->N - (size_t)modes[md][4];
            r->testparticle_type = (int)modes[md][5];
            r->dt = dts[s];
            en[s*4] = reb_simulation_energy(r);
            for (int c=0; c<3; c++){
                reb_simulation_steps(r, (size_t)nsteps);
                en[s*4+1+c] = reb_simulation_energy(r);
            }
            double* st = state_of(r);
            memcpy(fin+row*NCOL, st, sizeof(double)*NCOL*r->N);
            free(st);
            row += r->N;
            reb_simulation_free(r);
        }
        char name[128];
        snprintf(name, sizeof(name), "whfast_jac16_mode%d_final.npy", md); wd(name, fin, rows, NCOL);
        snprintf(name, sizeof(name), "whfast_jac16_mode%d_energy.npy", md); wd(name, en, NS, 4);
        free(fin); free(en);
    }
    for (int s=0; s<NS; s++) reb_simulation_free(sys[s]);
    free(init);
    reb_simulation_free(g);
}

/* The WHFast kernels (modified kick, composition, lazy implementer's) in
 * Jacobi coordinates on the same 16 systems as the Jacobi fixtures: modes
 * {kernel, safe_mode, corrector, corrector2, compensated}; bodies after three
 * reb_simulation_steps calls of 100 steps, energies, and r->gravity after. */
static void gen_whfast_kernels(void){
    const int NS = 16;
    const int nsteps = 100;
    enum { NMODE = 6 };
    static const int64_t modes[NMODE][5] = {
        {1, 0, 0, 0, 0},
        {1, 1, 5, 1, 0},
        {2, 0, 0, 0, 0},
        {2, 1, 3, 0, 1},
        {3, 0, 0, 0, 0},
        {3, 0, 7, 1, 0},
    };
    struct reb_simulation* g = reb_simulation_create();
    g->rand_seed = 4242u;
    struct reb_simulation* sys[NS];
    size_t rows = 0;
    double dts[NS];
    for (int s=0; s<NS; s++){
        sys[s] = ens_system(g, 1 + s % 8, 1 + s % 8);
        dts[s] = 2.*M_PI*sqrt(1./(1.+sys[s]->particles[1].m))/20.;
        rows += sys[s]->N;
    }
    wi("whfast_kern16_modes.npy", &modes[0][0], NMODE, 5);
    for (int md=0; md<NMODE; md++){
        double* fin = calloc(rows*NCOL, sizeof(double));
        double* en = calloc(NS*4, sizeof(double));
        int64_t grav[NS];
        size_t row = 0;
        for (int s=0; s<NS; s++){
            struct reb_simulation* r = reb_simulation_copy(sys[s]);
            struct reb_integrator_whfast_state* wh = reb_simulation_set_integrator(r, "whfast");
            wh->coordinates = REB_INTEGRATOR_WHFAST_COORDINATES_JACOBI;
            wh->kernel = (unsigned int)modes[md][0];
            wh->safe_mode = (unsigned int)modes[md][1];
            wh->corrector = (unsigned int)modes[md][2];
            wh->corrector2 = (unsigned int)modes[md][3];
            r->gravity = modes[md][4] ? REB_GRAVITY_COMPENSATED : REB_GRAVITY_BASIC;
            r->dt = dts[s];
            en[s*4] = reb_simulation_energy(r);
            for (int c=0; c<3; c++){
                reb_simulation_steps(r, (size_t)nsteps);
                en[s*4+1+c] = reb_simulation_energy(r);
            }
            grav[s] = (int64_t)r->gravity;
            double* st = state_of(r);
            memcpy(fin+row*NCOL, st, sizeof(double)*NCOL*r->N);
            free(st);
            row += r->N;
            reb_simulation_free(r);
        }
        char name[128];
        snprintf(name, sizeof(name), "whfast_kern16_mode%d_final.npy", md); wd(name, fin, rows, NCOL);
        snprintf(name, sizeof(name), "whfast_kern16_mode%d_energy.npy", md); wd(name, en, NS, 4);
        snprintf(name, sizeof(name), "whfast_kern16_mode%d_gravity.npy", md); wi(name, grav, NS, 0);
        free(fin); free(en);
    }
    for (int s=0; s<NS; s++) reb_simulation_free(sys[s]);
    reb_simulation_free(g);
}

/* reb_collision_resolve_merge on random three-body states: inputs, the
 * collision (p1, p2), track_energy_offset and N_active; outputs the bodies,
 * the outcome, energy_offset and G. */
static void gen_merge(void){
    const int NC = 2000;
    struct reb_simulation* g = reb_simulation_create();
    g->rand_seed = 77;
    double* in = calloc((size_t)NC*3*NCOL, sizeof(double));
    double* out = calloc((size_t)NC*3*NCOL, sizeof(double));
    int64_t* meta = calloc((size_t)NC*4, sizeof(int64_t));
    double* res = calloc((size_t)NC*3, sizeof(double));
    for (int c=0; c<NC; c++){
        struct reb_simulation* r = reb_simulation_create();
        r->G = reb_random_uniform(g, 0.5, 2.);
        for (int k=0; k<3; k++){
            struct reb_particle p = {0};
            p.x = reb_random_normal(g, 1.); p.y = reb_random_normal(g, 1.); p.z = reb_random_normal(g, 1.);
            p.vx = reb_random_normal(g, 1.); p.vy = reb_random_normal(g, 1.); p.vz = reb_random_normal(g, 1.);
            p.m = reb_random_uniform(g, 0.1, 3.);
            p.r = reb_random_uniform(g, 0.01, 0.5);
            reb_simulation_add(r, p);
        }
        int p1 = (int)(reb_random_uniform(g, 0., 3.)); if (p1 > 2) p1 = 2;
        int p2 = (p1 + 1 + (int)(reb_random_uniform(g, 0., 2.))%2) % 3;
        r->track_energy_offset = c % 2;
        if (c % 3 == 0) r->N_active = 1;
        r->energy_offset = 0.25;
        double* st = state_of(r);
        memcpy(in + (size_t)c*3*NCOL, st, sizeof(double)*3*NCOL);
        free(st);
        struct reb_collision col = {0};
        col.p1 = p1; col.p2 = p2;
        int outcome = (int)reb_collision_resolve_merge(r, col);
        st = state_of(r);
        memcpy(out + (size_t)c*3*NCOL, st, sizeof(double)*3*NCOL);
        free(st);
        meta[4*c] = p1; meta[4*c+1] = p2; meta[4*c+2] = r->track_energy_offset;
        meta[4*c+3] = (r->N_active == SIZE_MAX) ? -1 : (int64_t)r->N_active;
        res[3*c] = outcome; res[3*c+1] = r->energy_offset; res[3*c+2] = r->G;
        reb_simulation_free(r);
    }
    size_t shape[3] = {NC, 3, NCOL};
    write_npy("merge_in.npy", "<f8", 8, in, 3, shape);
    write_npy("merge_out.npy", "<f8", 8, out, 3, shape);
    wi("merge_meta.npy", meta, NC, 4);
    wd("merge_res.npy", res, NC, 3);
    free(in); free(out); free(meta); free(res);
    reb_simulation_free(g);
}

/* Boundaries during integration (OpenMP build: its per-target sums over
 * ghost boxes are the order the device follows).  OPEN removes bodies that
 * leave the box; PERIODIC wraps them, with ghost boxes in the force. */
static struct reb_simulation* boxed(int boundary, int gravity, size_t N, double root, int nr, unsigned int seed){
    struct reb_simulation* r = reb_simulation_create();
    r->rand_seed = seed;
    r->boundary = boundary;
    r->gravity = gravity;
    r->root_size = root; r->N_root_x = nr; r->N_root_y = nr; r->N_root_z = 1;
    if (boundary == REB_BOUNDARY_PERIODIC){ r->N_ghost_x = 1; r->N_ghost_y = 1; r->N_ghost_z = 0; }
    r->softening = 0.1;
    r->opening_angle2 = 0.25;
    r->dt = 0.01;
    reb_simulation_set_integrator(r, "leapfrog");
    double b = root*nr;
    for (size_t i=0; i<N; i++){
        struct reb_particle p = {0};
        p.x = reb_random_uniform(r, -b/2., b/2.);
        p.y = reb_random_uniform(r, -b/2., b/2.);
        p.z = reb_random_uniform(r, -root/2., root/2.);
        p.vx = reb_random_normal(r, 4.);
        p.vy = reb_random_normal(r, 4.);
        p.vz = reb_random_normal(r, 1.);
        p.m = reb_random_uniform(r, 0.5, 1.5)*1e-2;
        reb_simulation_add(r, p);
    }
    return r;
}

static void gen_boundaries(void){
    /* OPEN: 200 bodies in a 10-box; the N history every 10 steps and the end state */
    struct reb_simulation* r = boxed(REB_BOUNDARY_OPEN, REB_GRAVITY_BASIC, 200, 10., 1, 21u);
    dump_state("open200_init.npy", r);
    int64_t nh[11]; nh[0] = (int64_t)r->N;
    for (int c=1; c<=10; c++){ reb_simulation_steps(r, 10); nh[c] = (int64_t)r->N; }
    wi("open200_nhist.npy", nh, 11, 0);
    dump_state("open200_final.npy", r);
    reb_simulation_free(r);
    /* PERIODIC, direct gravity with one ring of ghost boxes */
    r = boxed(REB_BOUNDARY_PERIODIC, REB_GRAVITY_BASIC, 300, 4., 3, 22u);
    dump_state("periodic300_init.npy", r);
    reb_simulation_steps(r, 50);
    dump_state("periodic300_final_basic.npy", r);
    reb_simulation_free(r);
    /* PERIODIC, tree gravity */
    r = boxed(REB_BOUNDARY_PERIODIC, REB_GRAVITY_TREE, 300, 4., 3, 22u);
    reb_simulation_steps(r, 20);
    dump_state("periodic300_final_tree.npy", r);
    reb_simulation_free(r);
}

/* ---------------------------------------------------------------- integrate loop */
/* Scenarios of reb_simulation_integrate (simulation.c:241-502) and of the
 * collision outcomes it drives (collision.c:336-404).  Each writes the bodies
 * (11 columns) and a meta row
 *   {t, dt, dt_last_done, steps_done, status, N, collisions_log_n, energy_offset}. */
static void dump_meta(const char* name, const struct reb_simulation* r, int status){
    const double m[8] = {r->t, r->dt, r->dt_last_done, (double)r->steps_done, (double)status, (double)r->N,
                         (double)r->collisions_log_n, r->energy_offset};
    wd(name, m, 8, 0);
}

/* heartbeat log: {t, N, particles[k].x, particles[k].vy, energy_offset} per call */
static double hb_log[5*8192];
static size_t hb_n = 0;
static size_t hb_body = 0;
static void hb_record(struct reb_simulation* r){
    if (hb_n >= 8192) return;
    double* o = hb_log + 5*hb_n;
    o[0] = r->t;
    o[1] = (double)r->N;
    o[2] = r->N > hb_body ? r->particles[hb_body].x : 0.;
    o[3] = r->N > hb_body ? r->particles[hb_body].vy : 0.;
    o[4] = r->energy_offset;
    hb_n++;
}
static void pre_kick(struct reb_simulation* r){ r->particles[0].vx += 1e-7; }
static void post_damp(struct reb_simulation* r){
    for (size_t i=0; i<r->N; i++){
        r->particles[i].vx *= 0.9999;
        r->particles[i].vy *= 0.9999;
        r->particles[i].vz *= 0.9999;
    }
}

/* collision callback log: {p1, p2, outcome} per call.  Mode 0 removes both
 * bodies; mode 1 cycles the outcomes 1, 2, 3 (remove p1, p2, both). */
static int64_t cb_log[3*8192];
static size_t cb_n = 0;
static int cb_mode = 0;
static enum REB_COLLISION_RESOLVE_OUTCOME cb_resolve(struct reb_simulation* const r, struct reb_collision c){
    r->collisions_log_n++;
    const int out = cb_mode == 0 ? 3 : 1 + (int)(cb_n % 3);
    if (cb_n < 8192){
        cb_log[3*cb_n] = (int64_t)c.p1;
        cb_log[3*cb_n+1] = (int64_t)c.p2;
        cb_log[3*cb_n+2] = out;
    }
    cb_n++;
    return (enum REB_COLLISION_RESOLVE_OUTCOME)out;
}

static struct reb_simulation* two_planets(double m1, double a1, double e1, double m2, double a2, double e2){
    struct reb_simulation* r = reb_simulation_create();
    struct reb_particle star = {0};
    star.m = 1.;
    reb_simulation_add(r, star);
    reb_simulation_add(r, reb_particle_from_orbit(r->G, r->particles[0], m1, a1, e1, 0.02, 0.3, 0.1, 0.5));
    if (m2 > 0.) reb_simulation_add(r, reb_particle_from_orbit(r->G, r->particles[0], m2, a2, e2, 0.03, 1.3, 2.1, 4.));
    reb_simulation_move_to_com(r);
    return r;
}

static void gen_integrate(void){
    /* leapfrog, exact finish time: tmax not a multiple of dt, then past
     * tmax without exact finish, then backwards with exact finish */
    struct reb_simulation* r = plummer(256, 3u, 0.01);
    reb_simulation_set_integrator(r, "leapfrog");
    r->gravity = REB_GRAVITY_BASIC;
    r->dt = 1e-3;
    dump_state("integ_lf_init.npy", r);
    int st = reb_simulation_integrate(r, 0.0123456);
    dump_state("integ_lf_a_state.npy", r); dump_meta("integ_lf_a_meta.npy", r, st);
    r->exact_finish_time = 0;
    st = reb_simulation_integrate(r, 0.05);
    dump_state("integ_lf_b_state.npy", r); dump_meta("integ_lf_b_meta.npy", r, st);
    r->exact_finish_time = 1;
    st = reb_simulation_integrate(r, 0.0301);
    dump_state("integ_lf_c_state.npy", r); dump_meta("integ_lf_c_meta.npy", r, st);
    reb_simulation_free(r);

    /* WHFast (Jacobi, safe mode off), exact finish time, then without */
    r = two_planets(1e-3, 1., 0.05, 3e-4, 1.6, 0.1);
    reb_simulation_add(r, reb_particle_from_orbit(r->G, r->particles[0], 1e-5, 2.5, 0.02, 0.01, 2.3, 0.7, 1.));
    reb_simulation_move_to_com(r);
    struct reb_integrator_whfast_state* wh = reb_simulation_set_integrator(r, "whfast");
    wh->safe_mode = 0;
    r->dt = 0.0314;
    dump_state("integ_wh_init.npy", r);
    st = reb_simulation_integrate(r, 7.77);
    dump_state("integ_wh_a_state.npy", r); dump_meta("integ_wh_a_meta.npy", r, st);
    r->exact_finish_time = 0;
    st = reb_simulation_integrate(r, 20.);
    dump_state("integ_wh_b_state.npy", r); dump_meta("integ_wh_b_meta.npy", r, st);
    reb_simulation_free(r);

    /* exit_max_distance: an eccentric planet passes 1.4 from the origin */
    r = two_planets(1e-3, 1., 0.5, 0., 0., 0.);
    reb_simulation_set_integrator(r, "leapfrog");
    r->gravity = REB_GRAVITY_BASIC;
    r->dt = 1e-3;
    r->exit_max_distance = 1.4;
    dump_state("integ_escape_init.npy", r);
    st = reb_simulation_integrate(r, 100.);
    dump_state("integ_escape_state.npy", r); dump_meta("integ_escape_meta.npy", r, st);
    reb_simulation_free(r);

    /* exit_min_distance: crossing orbits */
    r = two_planets(1e-3, 1., 0.0, 1e-3, 1.5, 0.4);
    reb_simulation_set_integrator(r, "leapfrog");
    r->gravity = REB_GRAVITY_BASIC;
    r->dt = 1e-3;
    r->exit_min_distance = 0.25;
    dump_state("integ_encounter_init.npy", r);
    st = reb_simulation_integrate(r, 200.);
    dump_state("integ_encounter_state.npy", r); dump_meta("integ_encounter_meta.npy", r, st);
    reb_simulation_free(r);

    /* hooks: heartbeat log, pre-timestep kick of body 0, post-timestep damping */
    r = plummer(128, 4u, 0.01);
    reb_simulation_set_integrator(r, "leapfrog");
    r->gravity = REB_GRAVITY_BASIC;
    r->dt = 1e-3;
    dump_state("integ_hooks_init.npy", r);
    hb_n = 0; hb_body = 5;
    r->heartbeat = hb_record;
    r->pre_timestep_modifications = pre_kick;
    r->post_timestep_modifications = post_damp;
    st = reb_simulation_integrate(r, 0.0205);
    dump_state("integ_hooks_state.npy", r); dump_meta("integ_hooks_meta.npy", r, st);
    wd("integ_hooks_log.npy", hb_log, hb_n, 5);
    reb_simulation_free(r);

    /* halt: the default resolver (none set) with DIRECT search, and the halt
     * resolver set explicitly with TREE search in a periodic box */
    for (int v=0; v<2; v++){
        r = reb_simulation_create();
        r->rand_seed = 31u + (unsigned)v;
        reb_simulation_set_integrator(r, "leapfrog");
        r->gravity = REB_GRAVITY_BASIC;
        r->dt = 0.01;
        r->root_size = 10.;
        r->N_root_x = r->N_root_y = r->N_root_z = 1;
        r->boundary = REB_BOUNDARY_PERIODIC;
        r->collision = v == 0 ? REB_COLLISION_DIRECT : REB_COLLISION_TREE;
        if (v == 1) r->collision_resolve = reb_collision_resolve_halt;
        for (int i=0; i<100; i++){
            struct reb_particle p = {0};
            p.x = reb_random_uniform(r, -5., 5.);
            p.y = reb_random_uniform(r, -5., 5.);
            p.z = reb_random_uniform(r, -5., 5.);
            p.vx = reb_random_normal(r, 1.);
            p.vy = reb_random_normal(r, 1.);
            p.vz = reb_random_normal(r, 1.);
            p.m = 1e-3;
            p.r = 0.2;
            reb_simulation_add(r, p);
        }
        char name[64];
        snprintf(name, sizeof(name), "integ_halt%d_init.npy", v); dump_state(name, r);
        st = reb_simulation_integrate(r, 10.);
        snprintf(name, sizeof(name), "integ_halt%d_state.npy", v); dump_state(name, r);
        snprintf(name, sizeof(name), "integ_halt%d_meta.npy", v); dump_meta(name, r, st);
        reb_simulation_free(r);
    }

    /* collision outcomes that remove bodies (test_collisions.py:104-190):
     * remove both until no body is left, and the cycled outcomes 1, 2, 3,
     * for TREE and DIRECT searches (TREE with tree gravity) */
    for (int v=0; v<4; v++){
        const int tree = v & 1, one = v >> 1;
        r = reb_simulation_create();
        r->rand_seed = 77u + (unsigned)v;
        r->root_size = 50000.;
        reb_simulation_set_integrator(r, "leapfrog");
        r->boundary = REB_BOUNDARY_OPEN;
        r->gravity = tree ? REB_GRAVITY_TREE : REB_GRAVITY_BASIC;
        r->collision = tree ? REB_COLLISION_TREE : REB_COLLISION_DIRECT;
        r->collision_resolve = cb_resolve;
        const int n = one ? 50 : 10;
        const double w = one ? 2. : 20.;
        for (int i=0; i<n; i++){
            struct reb_particle p = {0};
            p.m = 1.;
            p.r = 100.;
            p.x = reb_random_uniform(r, -w, w);
            p.y = reb_random_uniform(r, -w, w);
            p.z = reb_random_uniform(r, -w, w);
            reb_simulation_add(r, p);
        }
        r->dt = 0.001;
        char name[64];
        const char* tag = one ? (tree ? "one_tree" : "one_direct") : (tree ? "both_tree" : "both_direct");
        snprintf(name, sizeof(name), "integ_%s_init.npy", tag); dump_state(name, r);
        cb_mode = one; cb_n = 0;
        if (one){
            st = reb_simulation_integrate(r, r->dt);
            snprintf(name, sizeof(name), "integ_%s_a_meta.npy", tag); dump_meta(name, r, st);
            st = reb_simulation_integrate(r, 2.*r->dt);
        }else{
            st = reb_simulation_integrate(r, 1000.);
        }
        snprintf(name, sizeof(name), "integ_%s_meta.npy", tag); dump_meta(name, r, st);
        if (r->N){ snprintf(name, sizeof(name), "integ_%s_state.npy", tag); dump_state(name, r); }
        snprintf(name, sizeof(name), "integ_%s_log.npy", tag); wi(name, cb_log, cb_n, 3);
        reb_simulation_free(r);
    }

}

/* The merge resolver on a stepping simulation, DIRECT and TREE searches,
 * periodic box with one ring of ghost boxes, energy offset tracked.  Written
 * by the OpenMP build on one thread: with ghost boxes its per-target force
 * sums are the order the device follows (the serial build's half-matrix loop
 * adds the ghost-box terms in another order). */
static void gen_integrate_merge(void){
    struct reb_simulation* r;
    int st;
    for (int v=0; v<2; v++){
        r = reb_simulation_create();
        r->rand_seed = 41u;
        reb_simulation_set_integrator(r, "leapfrog");
        r->gravity = REB_GRAVITY_BASIC;
        r->dt = 0.01;
        r->root_size = 4.;
        r->N_root_x = 2; r->N_root_y = 2; r->N_root_z = 1;
        r->N_ghost_x = 1; r->N_ghost_y = 1; r->N_ghost_z = 0;
        r->boundary = REB_BOUNDARY_PERIODIC;
        r->collision = v == 0 ? REB_COLLISION_DIRECT : REB_COLLISION_TREE;
        r->collision_resolve = reb_collision_resolve_merge;
        r->track_energy_offset = 1;
        for (int i=0; i<300; i++){
            struct reb_particle p = {0};
            p.x = reb_random_uniform(r, -4., 4.);
            p.y = reb_random_uniform(r, -4., 4.);
            p.z = reb_random_uniform(r, -2., 2.);
            p.vx = reb_random_normal(r, 0.5);
            p.vy = reb_random_normal(r, 0.5);
            p.vz = reb_random_normal(r, 0.5);
            p.m = reb_random_uniform(r, 0.5, 1.5)*1e-3;
            p.r = reb_random_uniform(r, 0.05, 0.15);
            reb_simulation_add(r, p);
        }
        char name[64];
        snprintf(name, sizeof(name), "integ_merge%d_init.npy", v); dump_state(name, r);
        hb_n = 0; hb_body = 0;
        r->heartbeat = hb_record;
        st = reb_simulation_integrate(r, 0.3);
        snprintf(name, sizeof(name), "integ_merge%d_state.npy", v); dump_state(name, r);
        snprintf(name, sizeof(name), "integ_merge%d_meta.npy", v); dump_meta(name, r, st);
        snprintf(name, sizeof(name), "integ_merge%d_log.npy", v); wd(name, hb_log, hb_n, 5);
        reb_simulation_free(r);
    }
}

/* A host integrator registered by name (rebound.c:86-113): symplectic Euler
 * through reb_simulation_update_acceleration. */
static void euler_step(struct reb_simulation* r, void* state){
    (void)state;
    reb_simulation_update_acceleration(r);
    const double dt = r->dt;
    for (size_t i=0; i<r->N; i++){
        struct reb_particle* p = &r->particles[i];
        p->vx += dt*p->ax; p->vy += dt*p->ay; p->vz += dt*p->az;
        p->x += dt*p->vx; p->y += dt*p->vy; p->z += dt*p->vz;
    }
    r->t += dt;
    r->dt_last_done = dt;
}

static void gen_custom_integrator(void){
    struct reb_integrator euler = {0};
    euler.step = euler_step;
    reb_integrator_register(euler, "test_symplectic_euler");
    struct reb_simulation* r = plummer(64, 6u, 0.01);
    reb_simulation_set_integrator(r, "test_symplectic_euler");
    r->gravity = REB_GRAVITY_BASIC;
    r->dt = 1e-3;
    r->exact_finish_time = 0;
    dump_state("integ_custom_init.npy", r);
    int st = reb_simulation_integrate(r, 0.01);
    dump_state("integ_custom_state.npy", r); dump_meta("integ_custom_meta.npy", r, st);
    reb_simulation_free(r);
}

/* The shearing-sheet wrap (boundary.c:78-114) on bodies placed up to 2.7 box
 * widths outside the box, at three times. */
static void gen_shear_wrap(void){
    struct reb_simulation* r = sheet(600, 50., 2, 2, 17u);
    const double bx = 100., by = 100., bz = 50.;
    for (size_t i=0; i<r->N; i++){
        r->particles[i].x = reb_random_uniform(r, -2.7*bx, 2.7*bx);
        r->particles[i].y = reb_random_uniform(r, -2.7*by, 2.7*by);
        r->particles[i].z = reb_random_uniform(r, -2.7*bz, 2.7*bz);
    }
    dump_state("shear_wrap_in.npy", r);
    const double ts[3] = {0., 3.7, 1234.56789};
    double* out = malloc(sizeof(double)*3*NCOL*r->N);
    double* in = state_of(r);
    for (int k=0; k<3; k++){
        for (size_t i=0; i<r->N; i++){
            struct reb_particle* p = &r->particles[i];
            const double* row = in + NCOL*i;
            p->x = row[0]; p->y = row[1]; p->z = row[2]; p->vx = row[3]; p->vy = row[4]; p->vz = row[5];
        }
        r->t = ts[k];
        reb_boundary_check(r);
        double* st = state_of(r);
        memcpy(out + (size_t)k*NCOL*r->N, st, sizeof(double)*NCOL*r->N);
        free(st);
    }
    size_t shape[3] = {3, r->N, NCOL};
    write_npy("shear_wrap_out.npy", "<f8", 8, out, 3, shape);
    wd("shear_wrap_t.npy", ts, 3, 0);
    free(out); free(in);
    reb_simulation_free(r);
}

/* Variational equations and MEGNO with WHFast in Jacobi coordinates
 * (integrator_whfast.c:392-490, 1250-1258; tools.c:201-375, 1313-1522;
 * gravity.c:533-640).  Per case: the bodies as added (<tag>_init), the
 * variational particles before the first step (<tag>_var0), and at each
 * checkpoint the bodies, particles_var and p_jh_var (x y z vx vy vz ax ay az m)
 * and a meta row: t, megno, lyapunov, the MEGNO sums (Ys Yss cov_Yt var_t
 * mean_t mean_Y initial_t n), gravity_ignore_terms, is_synchronized,
 * rand_seed, then the sets' lrescale.  A negative checkpoint synchronizes. */
#define VCOL 10
static double* var_rows(const struct reb_particle* q, size_t n){
    double* s = malloc(sizeof(double)*VCOL*(n ? n : 1));
    for (size_t i=0; i<n; i++){
        double row[VCOL] = {q[i].x,q[i].y,q[i].z,q[i].vx,q[i].vy,q[i].vz,q[i].ax,q[i].ay,q[i].az,q[i].m};
        memcpy(s+VCOL*i, row, sizeof(row));
    }
    return s;
}
static void dump_var_case(const char* tag, int k, struct reb_simulation* r){
    char name[160];
    snprintf(name, sizeof(name), "%s_k%d_state.npy", tag, k); dump_state(name, r);
    double* v = var_rows(r->particles_var, r->N_var);
    snprintf(name, sizeof(name), "%s_k%d_var.npy", tag, k); wd(name, v, r->N_var, VCOL);
    free(v);
    struct reb_integrator_whfast_state* wh = r->integrator.state;
    v = var_rows(wh->p_jh_var, wh->N_allocated_var);
    snprintf(name, sizeof(name), "%s_k%d_pjhvar.npy", tag, k); wd(name, v, wh->N_allocated_var, VCOL);
    free(v);
    double meta[32] = {r->t, reb_simulation_megno(r), reb_simulation_lyapunov(r), r->megno_Ys, r->megno_Yss,
                       r->megno_cov_Yt, r->megno_var_t, r->megno_mean_t, r->megno_mean_Y, r->megno_initial_t,
                       (double)r->megno_n, (double)r->gravity_ignore_terms, (double)r->is_synchronized,
                       (double)r->rand_seed};
    size_t nm = 14;
    for (size_t c=0; c<r->N_var_config; c++) meta[nm++] = r->var_config[c].lrescale;
    snprintf(name, sizeof(name), "%s_k%d_meta.npy", tag, k); wd(name, meta, nm, 0);
}
static void run_var_case(const char* tag, struct reb_simulation* r, const int64_t* ck, int nck){
    char name[160];
    double* v = var_rows(r->particles_var, r->N_var);
    snprintf(name, sizeof(name), "%s_var0.npy", tag); wd(name, v, r->N_var, VCOL);
    free(v);
    snprintf(name, sizeof(name), "%s_ck.npy", tag); wi(name, ck, (size_t)nck, 0);
    for (int k=0; k<nck; k++){
        if (ck[k] < 0) reb_simulation_synchronize(r);
        else reb_simulation_steps(r, (size_t)ck[k]);
        dump_var_case(tag, k, r);
    }
}
static struct reb_simulation* megno_bodies(const char* tag, int chaotic, int four){
    struct reb_simulation* r = reb_simulation_create();
    reb_simulation_add_fmt(r, "m", 1.);
    if (chaotic){
        reb_simulation_add_fmt(r, "m P", 1e-4, 1.);
        reb_simulation_add_fmt(r, "m P", 1e-4, 1.17);
    }else if (four){
        reb_simulation_add_fmt(r, "m a e inc", 1e-3, 1.0, 0.05, 0.02);
        reb_simulation_add_fmt(r, "m a e inc Omega", 3e-4, 1.6, 0.1, 0.05, 1.0);
        reb_simulation_add_fmt(r, "m a e inc f", 5e-5, 2.5, 0.02, 0.1, 2.0);
    }else{
        reb_simulation_add_fmt(r, "m a e inc", 1e-3, 1.5, 0.1, 0.1);
        reb_simulation_add_fmt(r, "m a e inc", 1e-3, 15., 0.1, 0.1);
    }
    char name[160];
    snprintf(name, sizeof(name), "%s_init.npy", tag); dump_state(name, r);
    return r;
}
static void gen_megno(void){
    /* A: rebound/tests/test_megno.py:33-42 (regular), dt = 0.07 P_1 */
    {
        struct reb_simulation* r = megno_bodies("megno_reg", 0, 0);
        reb_simulation_set_integrator(r, "whfast");
        reb_simulation_init_megno_seed(r, 0);
        r->dt = reb_orbit_from_particle(r->G, r->particles[1], r->particles[0]).P*0.07;
        const int64_t ck[5] = {1, 9, 90, 900, 2000};
        run_var_case("megno_reg", r, ck, 5);
        reb_simulation_free(r);
    }
    /* B: test_megno.py:58-80 (chaotic), centre-of-mass frame after init_megno */
    {
        struct reb_simulation* r = megno_bodies("megno_chaos", 1, 0);
        reb_simulation_set_integrator(r, "whfast");
        reb_simulation_init_megno_seed(r, 0);
        reb_simulation_move_to_com(r);
        r->dt = 0.05;
        const int64_t ck[5] = {1, 9, 90, 900, 1000};
        run_var_case("megno_chaos", r, ck, 5);
        reb_simulation_free(r);
    }
    /* C: A with the 11th-order symplectic corrector */
    {
        struct reb_simulation* r = megno_bodies("megno_corr", 0, 0);
        struct reb_integrator_whfast_state* wh = reb_simulation_set_integrator(r, "whfast");
        wh->corrector = 11;
        reb_simulation_init_megno_seed(r, 7);
        r->dt = reb_orbit_from_particle(r->G, r->particles[1], r->particles[0]).P*0.07;
        const int64_t ck[4] = {1, 9, 90, 400};
        run_var_case("megno_corr", r, ck, 4);
        reb_simulation_free(r);
    }
    /* D: four bodies, two sets (one with a mass variation), safe_mode 0:
     * unsynchronized checkpoints, then explicit synchronizations */
    {
        struct reb_simulation* r = megno_bodies("var_4b", 0, 1);
        struct reb_integrator_whfast_state* wh = reb_simulation_set_integrator(r, "whfast");
        wh->safe_mode = 0;
        const int a = reb_simulation_add_variation_1st_order(r, -1);
        const int b = reb_simulation_add_variation_1st_order(r, -1);
        r->particles_var[a+1].x = 1.;
        r->particles_var[a+2].vy = 0.5;
        r->particles_var[b+2].vy = 1.;
        r->particles_var[b+3].m = 1e-5;
        r->particles_var[b+0].x = -0.25;
        r->dt = 0.03;
        const int64_t ck[6] = {1, 49, -1, 50, 100, -1};
        run_var_case("var_4b", r, ck, 6);
        reb_simulation_free(r);
    }
    /* E: rescaling (tools.c:1313-1380): a set started near 1e100 */
    {
        struct reb_simulation* r = megno_bodies("var_rescale", 0, 0);
        reb_simulation_set_integrator(r, "whfast");
        const int a = reb_simulation_add_variation_1st_order(r, -1);
        r->rand_seed = 3;
        for (size_t i=0; i<r->N; i++){
            r->particles_var[a+i].x = 4e99*reb_random_normal(r, 1.);
            r->particles_var[a+i].vz = 4e99*reb_random_normal(r, 1.);
        }
        r->dt = 0.2;
        const int64_t ck[4] = {1, 1, 8, 90};
        run_var_case("var_rescale", r, ck, 4);
        reb_simulation_free(r);
    }
}

int main(int argc, char** argv){
    if (argc < 3){ fprintf(stderr, "usage: gen_golden OUTDIR serial|omp\n"); return 1; }
    outdir = argv[1];
    rawdir = getenv("GEN_GOLDEN_RAWDIR");
    if (argc > 3 && strcmp(argv[3], "megno") == 0){
        /* only the variational and MEGNO fixtures */
        if (strcmp(argv[2], "serial") == 0) gen_megno();
        return 0;
    }
    if (strcmp(argv[2], "omp") == 0){
        gen_plummer_omp();
        gen_ghostbox_basic("omp");
        gen_boundaries();
#ifdef _OPENMP
        /* the TREE collision search appends collisions from every thread in
         * arrival order (collision.c:228,475); one thread gives the serial
         * order, which the device search reproduces */
        omp_set_num_threads(1);
#endif
        gen_integrate_merge();
        return 0;
    }
    gen_rng();
    gen_plummer();
    gen_ghostbox_basic("serial");
    gen_kepler();
    gen_whfast();
    {
        const int n_a[3] = {37, 1, 64};
        gen_ensemble("ensemble200_seed99", 99u, 200, 1, 12, 0, n_a);
        const int n_b[3] = {20, 5, 25};
        gen_ensemble("ensemble8_big_seed5", 5u, 8, 30, 64, 1, n_b);
    }
    gen_sheet();
    gen_merge();
    gen_whfast_jacobi();
    gen_whfast_kernels();
    gen_sheet_full();
    gen_plummer_full();
    gen_line();
    gen_archive();
    gen_integrate();
    gen_custom_integrator();
    gen_shear_wrap();
    gen_megno();
    return 0;
}
