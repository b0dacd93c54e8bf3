/*
 * rebound.h -- drop-in C ABI of the MI355X force-and-step engine.
 *
 * This header declares the subset of REBOUND's public C interface that the
 * force-and-step path needs (gravity BASIC / COMPENSATED / TREE, collisions
 * DIRECT / TREE, integrators "leapfrog", "sei", "whfast"), with struct
 * layouts and enum values identical to REBOUND 5.0.0's src/rebound.h on LP64,
 * so host C code written against REBOUND compiles and links unchanged.
 *
 * Layout contracts (each checked field-by-field by tests/test_abi.py):
 *   struct reb_particle            rebound.h:86-104   (112 bytes)
 *   struct reb_vec3d / reb_vec6d   rebound.h:121-134
 *   struct reb_collision           rebound.h:144-149  (72 bytes)
 *   struct reb_integrator          rebound.h:170-179
 *   struct reb_integrator_configuration rebound.h:183-187
 *   enum REB_STATUS                rebound.h:217-233
 *   struct reb_simulation          rebound.h:238-413  (non-MPI build)
 *   struct reb_integrator_whfast_state   integrator_whfast.h:28-56
 *   struct reb_integrator_leapfrog_state integrator_leapfrog.h:32-34
 *
 * MI355X-specific knobs (precision modes, batched ensembles, the serving
 * loop) live in rebound_amd.h and never add fields to these structs.
 */
#ifndef REBOUND_AMD_REBOUND_H
#define REBOUND_AMD_REBOUND_H

#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

#define REB_API __attribute__((visibility("default")))
#define REB_STRING_SIZE_MAX 256

struct reb_simulation;
struct reb_simulationarchive;
struct reb_server_data;
struct reb_treecell;
struct reb_display_data;
struct reb_display_settings;
struct reb_variational_configuration;
struct reb_name_hash_item;
struct reb_ode;
struct reb_binarydata_field_descriptor;

/* One body.  11 doubles + 3 pointers. */
struct reb_particle {
    double x, y, z;          /* position */
    double vx, vy, vz;       /* velocity */
    double ax, ay, az;       /* acceleration (written by the force evaluation) */
    double m;                /* mass */
    double r;                /* physical radius (collisions) */
    const char* name;        /* optional name; not owned by this engine */
    void* ap;                /* extra properties of downstream libraries */
    struct reb_simulation* sim;
};

struct reb_vec3d { double x, y, z; };
struct reb_vec6d { double x, y, z, vx, vy, vz; };

/* A detected overlap: p1, p2 and the ghost-box offset applied to p1. */
struct reb_collision {
    size_t p1;
    size_t p2;
    struct reb_vec6d gb;
    size_t ri;               /* root box of p2 (tree search only) */
};

/* Integrator plugin interface. */
struct reb_integrator {
    char* documentation;
    void (*step)(struct reb_simulation* r, void* state);
    void (*synchronize)(struct reb_simulation* r, void* state);
    void* (*create)(void);
    void (*free)(void* state);
    void (*did_add_particle)(struct reb_simulation* r);
    void (*will_remove_particle)(struct reb_simulation* r, size_t index);
    const struct reb_binarydata_field_descriptor* field_descriptor_list;
};

struct reb_integrator_configuration {
    const char* name;
    struct reb_integrator callbacks;
    void* state;
};

enum REB_COLLISION_RESOLVE_OUTCOME {
    REB_COLLISION_RESOLVE_OUTCOME_REMOVE_NONE = 0,
    REB_COLLISION_RESOLVE_OUTCOME_REMOVE_P1 = 1,
    REB_COLLISION_RESOLVE_OUTCOME_REMOVE_P2 = 2,
    REB_COLLISION_RESOLVE_OUTCOME_REMOVE_BOTH = 3,
};

enum REB_STATUS {
    REB_STATUS_SINGLE_STEP = -10,
    REB_STATUS_SCREENSHOT_READY = -5,
    REB_STATUS_SCREENSHOT = -4,
    REB_STATUS_PAUSED = -3,
    REB_STATUS_LAST_STEP = -2,
    REB_STATUS_RUNNING = -1,
    REB_STATUS_SUCCESS = 0,
    REB_STATUS_GENERIC_ERROR = 1,
    REB_STATUS_NO_PARTICLES = 2,
    REB_STATUS_ENCOUNTER = 3,
    REB_STATUS_ESCAPE = 4,
    REB_STATUS_USER = 5,
    REB_STATUS_SIGINT = 6,
    REB_STATUS_COLLISION = 7,
};

/* Public integrator states (set options through reb_simulation_set_integrator's
 * return value, as with REBOUND). */
struct reb_integrator_leapfrog_state {
    unsigned int order;      /* 2 (default), 4, 6 or 8 */
};

struct reb_integrator_whfast_state {
    unsigned int corrector;
    unsigned int corrector2;
    enum {
        REB_INTEGRATOR_WHFAST_KERNEL_DEFAULT = 0,
        REB_INTEGRATOR_WHFAST_KERNEL_MODIFIEDKICK = 1,
        REB_INTEGRATOR_WHFAST_KERNEL_COMPOSITION = 2,
        REB_INTEGRATOR_WHFAST_KERNEL_LAZY = 3,
    } kernel;
    enum {
        REB_INTEGRATOR_WHFAST_COORDINATES_JACOBI = 0,
        REB_INTEGRATOR_WHFAST_COORDINATES_DEMOCRATICHELIOCENTRIC = 1,
        REB_INTEGRATOR_WHFAST_COORDINATES_WHDS = 2,
        REB_INTEGRATOR_WHFAST_COORDINATES_BARYCENTRIC = 3,
    } coordinates;
    unsigned int safe_mode;            /* 1 (default): synchronize every step */
    unsigned int keep_unsynchronized;
    size_t N_allocated;
    struct reb_particle* __restrict__ p_jh;     /* heliocentric copy, refreshed on return */
    size_t N_allocated_var;
    struct reb_particle* __restrict__ p_jh_var;
    unsigned int recalculate_coordinates_but_not_synchronized_warning;
};

/* The simulation.  Field order, types and padding follow REBOUND exactly;
 * fields outside the force-and-step path are kept (and zero) so that code
 * compiled against REBOUND sees the same offsets. */
struct reb_simulation {
    double  t;
    double  G;
    double  softening;
    double  OMEGA;
    double  OMEGAZ;
    double  dt;
    double  dt_last_done;
    uint64_t steps_done;
    unsigned int is_synchronized;
    unsigned int did_modify_particles;

    size_t  N;
    size_t  N_allocated;
    struct reb_particle* particles;

    size_t N_map;
    size_t* map;

    size_t  N_var;
    struct reb_particle* particles_var;
    size_t  N_var_config;
    struct reb_variational_configuration* var_config;

    size_t  N_active;
    int     testparticle_type;
    int     testparticle_hidewarnings;
    char**  name_list;
    size_t  N_name_list;
    struct reb_name_hash_item* name_hash_table;
    struct reb_vec3d* gravity_cs;
    size_t N_allocated_gravity_cs;
    struct reb_treecell** tree_root;
    double opening_angle2;
    enum REB_STATUS status;
    int     exact_finish_time;

    int force_is_velocity_dependent;
    enum {
        REB_GRAVITY_IGNORE_TERMS_NONE = 0,
        REB_GRAVITY_IGNORE_TERMS_BETWEEN_0_AND_1 = 1,
        REB_GRAVITY_IGNORE_TERMS_INVOLVING_0 = 2,
    } gravity_ignore_terms;
    double output_timing_last;

    int save_messages;
    char** messages;
    int messages_var_rescale_warning;
    int messages_timestep_warning;

    double exit_max_distance;
    double exit_min_distance;
    double usleep;
    struct reb_display_settings* display_settings;
    struct reb_display_data* display_data;
    struct reb_server_data* server_data;
    int track_energy_offset;
    double energy_offset;
    double walltime;
    double walltime_last_step;
    double walltime_last_steps;
    double walltime_last_steps_sum;
    int walltime_last_steps_N;
    uint32_t python_unit_l;
    uint32_t python_unit_m;
    uint32_t python_unit_t;

    double  root_size;
    size_t  N_root_x;
    size_t  N_root_y;
    size_t  N_root_z;
    int  N_ghost_x;
    int  N_ghost_y;
    int  N_ghost_z;

    struct reb_collision* collisions;
    size_t N_allocated_collisions;
    size_t N_collisions;
    size_t N_targets;
    double minimum_collision_velocity;
    double collisions_plog;
    int64_t collisions_log_n;

    int calculate_megno;
    double megno_Ys;
    double megno_Yss;
    double megno_cov_Yt;
    double megno_var_t;
    double megno_mean_t;
    double megno_mean_Y;
    double megno_initial_t;
    int64_t megno_n;

    unsigned int rand_seed;

    int    simulationarchive_version;
    double simulationarchive_auto_interval;
    double simulationarchive_auto_walltime;
    uint64_t simulationarchive_auto_step;
    double simulationarchive_next;
    uint64_t simulationarchive_next_step;
    char*  simulationarchive_filename;

    enum {
        REB_COLLISION_NONE = 0,
        REB_COLLISION_DIRECT = 1,
        REB_COLLISION_TREE = 2,
        REB_COLLISION_LINE = 4,
        REB_COLLISION_LINETREE = 5,
    } collision;
    struct reb_integrator_configuration integrator;
    enum {
        REB_BOUNDARY_NONE = 0,
        REB_BOUNDARY_OPEN = 1,
        REB_BOUNDARY_PERIODIC = 2,
        REB_BOUNDARY_SHEAR = 3,
    } boundary;
    enum {
        REB_GRAVITY_NONE = 0,
        REB_GRAVITY_BASIC = 1,
        REB_GRAVITY_COMPENSATED = 2,
        REB_GRAVITY_TREE = 3,
        REB_GRAVITY_JACOBI = 5,
        REB_GRAVITY_CUSTOM = 7,
    } gravity;

    void (*gravity_custom)(struct reb_simulation* const r);

    struct reb_ode** odes;
    size_t N_odes;
    size_t N_allocated_odes;

    void (*additional_forces)(struct reb_simulation* const r);
    void (*pre_timestep_modifications)(struct reb_simulation* const r);
    void (*post_timestep_modifications)(struct reb_simulation* const r);
    void (*heartbeat)(struct reb_simulation* r);
    int (*key_callback)(struct reb_simulation* r, int key);
    double (*coefficient_of_restitution)(const struct reb_simulation* const r, double v);
    enum REB_COLLISION_RESOLVE_OUTCOME (*collision_resolve)(struct reb_simulation* const r, struct reb_collision);
    void (*free_particle_ap)(struct reb_particle* p);
    void (*extras_cleanup)(struct reb_simulation* r);
    void* extras;
};

/* ---------------------------------------------------------------- lifecycle
 * simulation.c:98-122 (create), :124-197 (free), :709-715 (copy) */
REB_API struct reb_simulation* reb_simulation_create(void);
REB_API void reb_simulation_free(struct reb_simulation* const r);
REB_API struct reb_simulation* reb_simulation_copy(struct reb_simulation* r);
/* Binary files and simulation archives in REBOUND's format (rebound.h:425,
 * :466-472): a new file holds the simulation; an existing archive gets a
 * snapshot appended.  snapshot counts from 0; negative values from the end. */
REB_API struct reb_simulation* reb_simulation_create_from_file(char* filename, int64_t snapshot);
REB_API void reb_simulation_save_to_file(struct reb_simulation* r, const char* filename);
REB_API void reb_simulation_save_to_file_interval(struct reb_simulation* const r, const char* filename, double interval);
REB_API void reb_simulation_save_to_file_walltime(struct reb_simulation* const r, const char* filename, double walltime);
REB_API void reb_simulation_save_to_file_step(struct reb_simulation* const r, const char* filename, uint64_t step);
REB_API size_t reb_simulation_struct_size(void);                     /* simulation.c:65-67 */

/* Archive objects (rebound.h:427,624-626; simulationarchive.c:45-375): an
 * index of the snapshots of an archive file, kept open, from which any
 * snapshot is read.  A missing file gives NULL. */
REB_API struct reb_simulationarchive* reb_simulationarchive_create_from_file(const char* filename);
REB_API void reb_simulationarchive_free(struct reb_simulationarchive* sa);
REB_API struct reb_simulation* reb_simulation_create_from_simulationarchive(struct reb_simulationarchive* sa,
                                                                           int64_t snapshot);
/* The same with the reference's REB_BINARYDATA_* codes (binarydata.h:31-45)
 * OR-ed into *warnings instead of messages (simulationarchive.h:58-60, the
 * calls its Python package makes). */
REB_API struct reb_simulationarchive* reb_simulationarchive_create_from_file_with_messages(const char* filename,
                                                                                         int* warnings);
REB_API struct reb_simulationarchive* reb_simulationarchive_create_from_buffer_with_messages(char* buffer,
                                                                                           size_t size,
                                                                                           int* warnings);
REB_API void reb_simulation_init_from_simulationarchive_with_messages(struct reb_simulation* r,
                                                                     struct reb_simulationarchive* sa,
                                                                     int64_t snapshot, int* warnings);
/* Comparison through the binary format (simulation.c:618-709): 0 when equal
 * (fields named walltime* aside), 1 when not; output_option 1 prints the
 * differing fields.  diff_char returns them as a string to free with
 * reb_free. */
REB_API int reb_simulation_diff(struct reb_simulation* r1, struct reb_simulation* r2, int output_option);
REB_API char* reb_simulation_diff_char(struct reb_simulation* r1, struct reb_simulation* r2);
/* The simulation as one binary snapshot in a malloc'd buffer
 * (binarydata.c:582-606), and a copy made through that format into a freshly
 * created simulation (simulation.c:711-717). */
REB_API void reb_binarydata_simulation_to_stream(struct reb_simulation* r, char** bufp, size_t* sizep);
REB_API void reb_simulation_copy_with_messages(struct reb_simulation* r_copy, struct reb_simulation* r,
                                               int* warnings);
/* Frees memory the library returned (rebound.c:215-217). */
REB_API void reb_free(void* p);

/* Build, version and hash strings (rebound.c:41-43); the reference's Python
 * package reads reb_version_str and reb_build_str with ctypes in_dll. */
REB_API extern const char* reb_build_str;
REB_API extern const char* reb_version_str;
REB_API extern const char* reb_githash_str;

/* ---------------------------------------------------------------- integrators
 * simulation.c:201-238, rebound.c:60-113 */
REB_API void* reb_simulation_set_integrator(struct reb_simulation* r, const char* name);
REB_API void reb_integrator_register(const struct reb_integrator integrator, const char* name);
REB_API char** reb_integrators_registered(void);
REB_API extern const struct reb_integrator reb_integrator_none;

/* ---------------------------------------------------------------- messages
 * rebound.c:118-190 */
REB_API void reb_exit(const char* const msg);
REB_API void reb_simulation_stop(struct reb_simulation* const r);
REB_API void reb_simulation_info(struct reb_simulation* const r, const char* const msg);
REB_API void reb_simulation_warning(struct reb_simulation* const r, const char* const msg);
REB_API void reb_simulation_error(struct reb_simulation* const r, const char* const msg);
REB_API int reb_pop_message(char** messages, char* const buf);        /* rebound_internal.h:90, used by ctypes callers */

/* ---------------------------------------------------------------- stepping
 * simulation.c:465-513 (integrate, steps), :633-637 (synchronize),
 * :640-689 (update_acceleration) */
REB_API void reb_simulation_steps(struct reb_simulation* const r, size_t N_steps);
REB_API enum REB_STATUS reb_simulation_integrate(struct reb_simulation* const r, double tmax);
REB_API void reb_simulation_synchronize(struct reb_simulation* r);
REB_API void reb_simulation_update_acceleration(struct reb_simulation* r);

/* ---------------------------------------------------------------- diagnostics
 * tools.c:108-162 (energy), :380-420 (com), :201-... (move_to_com) */
REB_API double reb_simulation_energy(struct reb_simulation* const r);
REB_API struct reb_vec3d reb_simulation_angular_momentum(const struct reb_simulation* const r);
REB_API struct reb_particle reb_simulation_com(struct reb_simulation* r);
REB_API struct reb_particle reb_particle_com_of_pair(struct reb_particle p1, struct reb_particle p2);
REB_API struct reb_particle reb_simulation_com_range(struct reb_simulation* r, size_t first, size_t last);
REB_API void reb_simulation_move_to_com(struct reb_simulation* const r);
REB_API void reb_simulation_move_to_hel(struct reb_simulation* const r);

/* ---------------------------------------------------------------- particles
 * particle.c:47-76 (add), :333-373 (remove), tools.c:463-502 (plummer),
 * tools.c:940-1002 (from_orbit) */
REB_API void reb_simulation_add(struct reb_simulation* const r, struct reb_particle pt);
REB_API void reb_simulation_add_plummer(struct reb_simulation* r, size_t _N, double M, double R);
REB_API struct reb_particle reb_particle_from_orbit_err(double G, struct reb_particle primary, double m, double a, double e, double i, double Omega, double omega, double f, int* err);
REB_API struct reb_particle reb_particle_from_orbit(double G, struct reb_particle primary, double m, double a, double e, double i, double Omega, double omega, double f);
REB_API struct reb_particle reb_particle_nan(void);
REB_API void reb_simulation_remove_all_particles(struct reb_simulation* const r);
REB_API int reb_simulation_remove_particle(struct reb_simulation* const r, size_t index);
REB_API int reb_particle_cmp(struct reb_particle p1, struct reb_particle p2);
/* particle.c:128-136 (-1 when p is not one of its simulation's particles),
 * :375-410; tools.c:454-461, :1524-1569 (-1 when the counts differ) */
REB_API int reb_simulation_particle_index(struct reb_particle* p);
REB_API int reb_simulation_particle_var_index(struct reb_particle* p);
REB_API struct reb_particle reb_simulation_jacobi_com(struct reb_particle* p);
REB_API void reb_particle_iadd(struct reb_particle* p1, struct reb_particle* p2);
REB_API void reb_particle_isub(struct reb_particle* p1, struct reb_particle* p2);
REB_API void reb_particle_imul(struct reb_particle* p1, double value);
REB_API double reb_particle_distance(struct reb_particle* p1, struct reb_particle* p2);
REB_API void reb_simulation_imul(struct reb_simulation* r, double scalar_pos, double scalar_vel);
REB_API int reb_simulation_iadd(struct reb_simulation* r, struct reb_simulation* r2);
REB_API int reb_simulation_isub(struct reb_simulation* r, struct reb_simulation* r2);
/* simulation.c:801-861: bulk copies; any of the arrays may be NULL */
REB_API void reb_simulation_get_serialized_particle_data(struct reb_simulation* r, double* m, double* radius,
                                                         double (*xyz)[3], double (*vxvyvz)[3],
                                                         double (*xyzvxvyvz)[6]);
REB_API void reb_simulation_set_serialized_particle_data(struct reb_simulation* r, double* m, double* radius,
                                                         double (*xyz)[3], double (*vxvyvz)[3],
                                                         double (*xyzvxvyvz)[6]);

/* ---------------------------------------------------------------- orbits
 * rebound.h:632-660 (layout), tools.c:504-560 (anomalies), :1004-1208
 * (elements of a particle about a primary; err 1: massless primary, 2: the
 * particle sits on the primary), :1571-1583 (spherical coordinates) */
struct reb_orbit {
    double d, v, h, P, n, a, e, inc, Omega, omega, pomega, f, M, l, theta, T, rhill;
    double pal_h, pal_k, pal_ix, pal_iy;
    struct reb_vec3d hvec;
    struct reb_vec3d evec;
};
REB_API struct reb_orbit reb_orbit_from_particle(double G, struct reb_particle p, struct reb_particle primary);
REB_API struct reb_orbit reb_orbit_from_particle_err(double G, struct reb_particle p, struct reb_particle primary,
                                                     int* err);
REB_API double reb_mod2pi(double f);
REB_API double reb_M_to_E(double e, double M);
REB_API double reb_E_to_f(double e, double E);
REB_API double reb_M_to_f(double e, double M);
REB_API struct reb_vec3d reb_tools_spherical_to_xyz(const double magnitude, const double theta, const double phi);
REB_API void reb_tools_xyz_to_spherical(const struct reb_vec3d xyz, double* magnitude, double* theta, double* phi);
/* tools.c:48-56: djb2 hash of a string (unit names) */
REB_API uint32_t reb_hash(const char* str);
/* The reference's OpenMP thread count; the engine's force and step run on the
 * GPU, so this is accepted and has no effect. */
REB_API void reb_omp_set_num_threads(int num_threads);

/* ---------------------------------------------------------------- collisions
 * collision.c:573-737 */
REB_API enum REB_COLLISION_RESOLVE_OUTCOME reb_collision_resolve_halt(struct reb_simulation* const r, struct reb_collision c);
REB_API enum REB_COLLISION_RESOLVE_OUTCOME reb_collision_resolve_hardsphere(struct reb_simulation* const r, struct reb_collision c);
REB_API enum REB_COLLISION_RESOLVE_OUTCOME reb_collision_resolve_merge(struct reb_simulation* const r, struct reb_collision c);

/* ---------------------------------------------------------------- random numbers
 * tools.c:58-105 (glibc rand_r stream on r->rand_seed) */
REB_API double reb_random_uniform(struct reb_simulation* r, double min, double max);
REB_API double reb_random_powerlaw(struct reb_simulation* r, double min, double max, double slope);
REB_API double reb_random_normal(struct reb_simulation* r, double variance);
REB_API double reb_random_rayleigh(struct reb_simulation* r, double sigma);

/* ---------------------------------------------------------------- boundary
 * boundary.c:145-201 */
REB_API struct reb_vec6d reb_boundary_get_ghostbox(struct reb_simulation* const r, int i, int j, int k);

/* ---------------------------------------------------------------- variational equations, MEGNO
 * rebound.h:702-734, tools.c:1313-1522.  First-order sets of all bodies are
 * integrated by WHFast in Jacobi coordinates (default kernel, safe_mode for
 * MEGNO), on the GPU, as the reference's integrator_whfast.c:392-490. */
struct reb_variational_configuration {
    struct reb_simulation* sim;
    int order;              /* 1 or 2 */
    int index;              /* first entry in particles_var */
    int testparticle;       /* -1: every body varied */
    int index_1st_order_a;  /* order 2 only */
    int index_1st_order_b;
    double lrescale;        /* sum of the logarithms of the rescalings */
};
REB_API int reb_simulation_add_variation_1st_order(struct reb_simulation* const r, int testparticle);
REB_API int reb_simulation_add_variation_2nd_order(struct reb_simulation* const r, int testparticle, size_t index_1st_order_a, size_t index_1st_order_b);
REB_API void reb_simulation_rescale_var(struct reb_simulation* const r);
REB_API void reb_simulation_init_megno(struct reb_simulation* const r);
REB_API void reb_simulation_init_megno_seed(struct reb_simulation* const r, unsigned int seed);
REB_API double reb_simulation_megno(struct reb_simulation* r);
REB_API double reb_simulation_lyapunov(struct reb_simulation* r);

/* ---------------------------------------------------------------- WHFast helper
 * integrator_whfast.c:241-427.  Advances one body on a Kepler orbit (mu=G*M)
 * by dt; evaluated on the GPU. */
REB_API void reb_integrator_whfast_kepler_solver(struct reb_particle* const p, double mu, double dt, const struct reb_simulation* const r);

#ifdef __cplusplus
}
#endif

#endif /* REBOUND_AMD_REBOUND_H */
