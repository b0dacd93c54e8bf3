/*
 * rebound_amd.h -- MI355X additions beside the drop-in REBOUND ABI.
 *
 * Nothing here changes a struct of rebound.h.  Per-simulation knobs are set
 * through functions that take the simulation pointer; batched and serving
 * entry points take plain pointers and sizes (no framework types).
 */
#ifndef REBOUND_AMD_H
#define REBOUND_AMD_H

#include <stddef.h>
#include <stdint.h>
#include "rebound.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ precision
 * Arithmetic of the direct-summation force (REB_GRAVITY_BASIC and
 * REB_GRAVITY_COMPENSATED; gravity.c:167-531).
 *
 *  FP64   (default) FP64 throughout.  r^-3 from the hardware reciprocal square
 *         root refined to full FP64 accuracy, fused multiply-adds, source
 *         range split across wavefronts.  Per particle
 *         |a - a_ref| <= 1e-14 * sum_j |a_ij|.
 *  EXACT  IEEE-754 sqrt and division, no contraction, each target summing its
 *         sources in ascending index order exactly as the reference's
 *         OpenMP loop: accelerations are bit-identical to the reference.
 *  SPLIT  Displacement and r^2 in FP64, r^-1 from the FP32 reciprocal square
 *         root on the vector units (about 2^-23 relative), cubed and scaled by
 *         m_j in FP64, accumulation in FP64 with Kahan compensation applied
 *         per tile of sources.  Per particle |a - a_ref| <= 1e-7 *
 *         sum_j |a_ij|.  r^2 + softening^2 must lie in FP32's normal range:
 *         about 1e-19 < r < 1.8e19 in simulation units (SI distances of the
 *         Solar System are well inside it).
 */
enum REB_AMD_PRECISION {
    REB_AMD_PRECISION_FP64 = 0,
    REB_AMD_PRECISION_EXACT = 1,
    REB_AMD_PRECISION_SPLIT = 2,
};
REB_API int reb_amd_set_precision(struct reb_simulation* r, int precision);
REB_API int reb_amd_get_precision(const struct reb_simulation* r);

/* ------------------------------------------------------------ device
 * 0 when no HIP device is usable (every stepping call then records an error
 * on the simulation instead of computing). */
REB_API int reb_amd_device_count(void);
REB_API int reb_amd_device_name(char* buf, size_t len);
REB_API const char* reb_amd_version(void);
REB_API const char* reb_amd_last_error(void);

/* Kernel timing switch for reb_amd_kernel_time (off by default). */
REB_API void reb_amd_set_profiling(struct reb_simulation* r, int on);

/* With profiling on, each launch of the direct-gravity pair kernel is
 * bracketed by HIP events on the simulation's stream; this waits for the
 * stream and returns the summed kernel time (ms) and the number of launches
 * since profiling was switched on or last read. */
REB_API int reb_amd_kernel_time(struct reb_simulation* r, double* total_ms, int* launches);
/* With profiling on: the ordered body pairs the timed direct-gravity kernels
 * evaluated since profiling was switched on or last read (then reset).  On
 * the symmetric path (every body active, no ghost boxes, nothing ignored,
 * N >= 8192) the timed kernel covers the pairs between superblocks of 1024
 * bodies, N^2 minus the pairs inside superblocks; otherwise N(N-1). */
REB_API double reb_amd_kernel_pairs(struct reb_simulation* r);

/* ------------------------------------------------------------ device forces
 * The reference's additional_forces callback (rebound.h:403, called at the end
 * of reb_simulation_update_acceleration, simulation.c:640-689) runs on the
 * host, so every force evaluation would move the bodies to the host and back.
 * A device force is the same hook without that round trip: the callback gets
 * device pointers to the bodies (structure of arrays) and the simulation's HIP
 * stream, and enqueues its own kernels there, adding to ax, ay, az.  It runs
 * after gravity and before additional_forces when both are set.  No
 * counterpart in the reference; for callers whose extra forces are GPU code
 * (REBOUNDx-style physics). */
struct reb_amd_device_bodies {
    double* x;
    double* y;
    double* z;
    double* vx;
    double* vy;
    double* vz;
    double* ax;
    double* ay;
    double* az;
    double* m;
    double* r;
    size_t N;
    void* stream;      /* hipStream_t of the simulation: enqueue work here, do not synchronize */
};
typedef void (*reb_amd_device_forces_fn)(struct reb_simulation* r, const struct reb_amd_device_bodies* b,
                                         void* user);
/* fn == NULL removes the device force.  Not supported with WHFast (its
 * interaction runs inside one kernel). */
REB_API void reb_amd_set_device_forces(struct reb_simulation* r, reb_amd_device_forces_fn fn, void* user);
/* A built-in device force: linear drag a -= gamma v on every body, with
 * `user` pointing to gamma (a double that must outlive the simulation's use). */
REB_API void reb_amd_device_force_drag(struct reb_simulation* r, const struct reb_amd_device_bodies* b, void* user);

/* ------------------------------------------------------------ collisions
 * Sets r->coefficient_of_restitution to a C callback that returns `eps` for
 * every impact velocity (the reference's callback, integrator-agnostic, with
 * no per-collision call into a scripting language). */
REB_API void reb_amd_set_coefficient_of_restitution(struct reb_simulation* r, double eps);
REB_API double reb_amd_coefficient_of_restitution_constant(const struct reb_simulation* const r, double v);

/* ------------------------------------------------------------ residency
 * By default every reb_simulation_steps / integrate / update_acceleration /
 * energy call uploads r->particles first and downloads it on return (host
 * code may write r->particles between calls, as with REBOUND).  A resident
 * simulation keeps its state on the GPU between calls: the host copy is
 * refreshed only by reb_amd_sync_host, and host edits must be announced with
 * reb_amd_mark_host_modified (reb_simulation_add does so itself). */
REB_API void reb_amd_set_resident(struct reb_simulation* r, int on);
REB_API int reb_amd_sync_host(struct reb_simulation* r);
REB_API void reb_amd_mark_host_modified(struct reb_simulation* r);
REB_API int reb_amd_synchronize_device(struct reb_simulation* r);

/* ------------------------------------------------------------ raw kernels
 * Entry points on device-resident arrays (hipMalloc'd pointers), for callers
 * that keep their own state on the GPU.  `stream` is a hipStream_t (NULL:
 * default stream).  Return 0 on success, nonzero on a launch error; they do
 * not synchronize.
 *
 * reb_amd_gravity_direct replaces reb_gravity_basic_calculate_acceleration
 * (gravity.c:167-282) and, with compensated=1,
 * reb_gravity_compensated_calculate_acceleration (gravity.c:284-531), without
 * ghost boxes.  ignore_terms takes REB_GRAVITY_IGNORE_TERMS_* values.
 * `workspace` must hold reb_amd_gravity_workspace_bytes(N, precision) bytes. */
REB_API size_t reb_amd_gravity_workspace_bytes(size_t N, int precision);
REB_API int reb_amd_gravity_direct(const double* x, const double* y, const double* z, const double* m,
                                   double* ax, double* ay, double* az,
                                   size_t N, size_t N_active, int testparticle_type, int ignore_terms,
                                   double G, double softening, int precision, int compensated,
                                   void* workspace, void* stream);

/* reb_amd_gravity_direct without a workspace and without any allocation:
 * the fast modes' partial sums go to a fixed device array of
 * reb_amd_gravity_noalloc_max_bytes() bytes (32 MiB, one per device, held
 * until the process ends; enough for N up to about 650000 in one call).  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, kernels and results as
 * reb_amd_gravity_direct; returns 4 (with reb_amd_last_error set) when N needs
 * more partial-sum space than the array holds.  Every call on a device shares
 * that device's array, so calls that overlap on one device race. */
REB_API size_t reb_amd_gravity_noalloc_max_bytes(void);
REB_API int reb_amd_gravity_direct_noalloc(const double* x, const double* y, const double* z, const double* m,
                                           double* ax, double* ay, double* az,
                                           size_t N, size_t N_active, int testparticle_type, int ignore_terms,
                                           double G, double softening, int precision, int compensated,
                                           void* stream);

/* Kepler drift of n independent bodies (integrator_whfast.c:241-427): state
 * is n rows of (x,y,z,vx,vy,vz) in host memory, updated in place; mu and dt
 * per body.  Runs on the GPU and waits for it. */
REB_API int reb_amd_kepler_solver_batch(double* state6, const double* mu, const double* dt, size_t n);

/* Builds the octree of the simulation's current particles on the GPU
 * (tree.c:254-271, moments tree.c:147-229) and copies up to `cap` cells into
 * `out` in pre-order, root box by root box, 9 doubles per cell:
 * centre x y z, width, mass, centre of mass x y z, and pt (the body index of a
 * leaf, minus the body count of an internal cell), exactly the fields of
 * struct reb_treecell (tree.h:34-56).  Returns the number of cells or -1. */
REB_API long reb_amd_tree_dump(struct reb_simulation* r, double* out, long cap);
/* Particle-node interactions (accepted cells and leaves, over all ghost boxes)
 * of the last REB_GRAVITY_TREE evaluation; -1 if there was none. */
REB_API long long reb_amd_tree_interactions(struct reb_simulation* r);

/* ---------------------------------------------------------------------------
 * Batched ensembles: many independent planetary systems integrated together
 * with WHFast (default kernel), state resident in device memory: democratic
 * heliocentric coordinates unless reb_amd_ensemble_set_whfast selects Jacobi
 * coordinates and symplectic correctors.  Each system evolves bit-identically
 * to a reb_simulation holding the same particles with integrator "whfast",
 * the same coordinates, correctors and safe_mode, dt and G, driven by
 * reb_simulation_steps (integrator_whfast.c:1051-1259, simulation.c:504-513).
 * The reference has no batched form: its equivalent is a loop over
 * reb_simulation_steps calls.
 * ------------------------------------------------------------------------- */
struct reb_amd_ensemble;

/* n_systems systems; system s has n_bodies[s] bodies (1..65, central body
 * first), stored consecutively in `bodies`; dt[s] its timestep.  gravity is
 * REB_GRAVITY_BASIC or REB_GRAVITY_COMPENSATED.  NULL on error
 * (reb_amd_last_error). */
REB_API struct reb_amd_ensemble* reb_amd_ensemble_create(size_t n_systems, const size_t* n_bodies,
                                                         const struct reb_particle* bodies, const double* dt,
                                                         double G, int safe_mode, int gravity);
REB_API void reb_amd_ensemble_free(struct reb_amd_ensemble* e);
/* Before the first steps: coordinates REB_INTEGRATOR_WHFAST_COORDINATES_JACOBI
 * or _DEMOCRATICHELIOCENTRIC, first corrector order (0, 3, 5, 7, 11, 17) and
 * second corrector (0/1), both Jacobi only; the fields of
 * struct reb_integrator_whfast_state (integrator_whfast.h:28-46). */
REB_API int reb_amd_ensemble_set_whfast(struct reb_amd_ensemble* e, int coordinates, unsigned int corrector,
                                        unsigned int corrector2);
/* Before the first steps, after Jacobi coordinates are selected: the WHFast
 * kernel (field kernel of integrator_whfast.h:28-46): 0 default, 1 modified
 * kick, 2 composition, 3 lazy implementer's; 1 and 3 use the Jacobi gravity
 * (gravity.c:108-163).  Replaces setting the field on each reb_simulation. */
REB_API int reb_amd_ensemble_set_whfast_kernel(struct reb_amd_ensemble* e, unsigned int kernel);
/* Before the first steps (Jacobi, default kernel, safe_mode 1, BASIC): MEGNO
 * for every system, with one first-order variational set per system whose
 * initial particles var0 gives in body order (what reb_simulation_init_megno
 * draws for that system); each system's time starts at 0.  Replaces
 * reb_simulation_init_megno (tools.c:1431-1470) on each reb_simulation; the
 * sets and sums evolve on the device as integrator_whfast.c:392-490,
 * 1250-1258 and tools.c:1504-1522. */
REB_API int reb_amd_ensemble_init_megno(struct reb_amd_ensemble* e, const struct reb_particle* var0);
/* reb_simulation_megno and reb_simulation_lyapunov (tools.c:1471-1482) of
 * every system, into megno[n_systems] and lyapunov[n_systems] (either may be
 * NULL). */
REB_API int reb_amd_ensemble_megno(struct reb_amd_ensemble* e, double* megno, double* lyapunov);
/* The variational particles (inertial x y z vx vy vz ax ay az m), body order. */
REB_API int reb_amd_ensemble_get_var_particles(struct reb_amd_ensemble* e, struct reb_particle* out);
/* Queues n steps of every system, ending synchronized; does not wait. */
REB_API int reb_amd_ensemble_steps(struct reb_amd_ensemble* e, uint64_t n_steps);
REB_API int reb_amd_ensemble_wait(struct reb_amd_ensemble* e);
REB_API uint64_t reb_amd_ensemble_steps_done(const struct reb_amd_ensemble* e);
REB_API size_t reb_amd_ensemble_n_systems(const struct reb_amd_ensemble* e);
REB_API size_t reb_amd_ensemble_n_bodies(const struct reb_amd_ensemble* e);
/* Inertial x y z vx vy vz ax ay az m of all bodies, creation order. */
REB_API int reb_amd_ensemble_get_particles(struct reb_amd_ensemble* e, struct reb_particle* out);
/* reb_simulation_energy of every system (tools.c:108-162). */
REB_API int reb_amd_ensemble_energy(struct reb_amd_ensemble* e, double* out);
/* Per system: bit 0 synchronized, bit 1 Kepler step longer than an orbital
 * period, bit 2 particles modified while not synchronized, bit 3 a Kepler
 * solve took the quartic (Laguerre-Conway) iteration, bit 4 a Kepler solve
 * fell back to bisection (integrator_whfast.c:289-369).  Bits 1-4 stay set
 * from launch to launch. */
REB_API int reb_amd_ensemble_flags(struct reb_amd_ensemble* e, unsigned int* out);
/* Launch shape: persistent workgroups (0 = default) and steps per launch
 * (0 = a whole request in one launch). */
REB_API int reb_amd_ensemble_set_launch(struct reb_amd_ensemble* e, int grid, long chunk);
/* Wave packing: about `tasks` wave tasks (0: the densest, 64 lanes filled per
 * wavefront).  More tasks put fewer systems in each wavefront and more
 * wavefronts on the chip; results are unchanged.  Returns the task count or -1. */
REB_API int reb_amd_ensemble_set_packing(struct reb_amd_ensemble* e, int tasks);
REB_API int reb_amd_ensemble_get_launch(const struct reb_amd_ensemble* e, int* grid, long* chunk, int* ntasks,
                                        int* max_resident);
/* Times candidate launch shapes (packing, grid, chunk) for probe_steps steps
 * each (state restored afterwards) and keeps the fastest; best_ms receives
 * its time. */
REB_API int reb_amd_ensemble_tune(struct reb_amd_ensemble* e, long probe_steps, double* best_ms);
REB_API const char* reb_amd_ensemble_error(const struct reb_amd_ensemble* e);

/* ---------------------------------------------------------------------------
 * Serving: a job queue in front of the GPU.  Jobs are admitted in arrival
 * order while their device-memory footprints fit in the budget (later, smaller
 * jobs may pass one that does not fit yet).  Planetary jobs share one device
 * pool stepped in ticks of at most `whfast_tick_steps` steps, each aiming at
 * about 2 ms of device time; up to `max_nbody_jobs` N-body jobs are batched
 * on a second stream, all advancing by the same number of leapfrog steps per
 * round with three launches per step for the whole batch (EXACT-precision
 * jobs run alone on a third stream).  Jobs join and leave at tick and round
 * boundaries (continuous batching).  A job's result equals that of the same
 * simulation run alone through reb_simulation_steps.  No counterpart in the
 * reference, whose unit of work is one reb_simulation per caller.
 * ------------------------------------------------------------------------- */
enum REB_AMD_JOB_STATE {
    REB_AMD_JOB_UNKNOWN = -2,
    REB_AMD_JOB_FAILED = -1,
    REB_AMD_JOB_PENDING = 0,
    REB_AMD_JOB_RUNNING = 1,
    REB_AMD_JOB_DONE = 2,
};
struct reb_amd_server;

/* memory_budget_bytes 0: 8 GiB; max_nbody_jobs 0: 32; whfast_tick_steps 0: 2000. */
REB_API struct reb_amd_server* reb_amd_server_create(size_t memory_budget_bytes, int max_nbody_jobs,
                                                     long whfast_tick_steps);
/* Cancels pending jobs, waits for running ones, frees everything. */
REB_API void reb_amd_server_free(struct reb_amd_server* s);
/* WHFast (democratic heliocentric, default kernel) on a system of 1..65 bodies,
 * central body first, for n_steps steps ending synchronized.  Job id or -1. */
REB_API long reb_amd_server_submit_whfast(struct reb_amd_server* s, size_t n_bodies, const struct reb_particle* bodies,
                                          double dt, double G, int safe_mode, uint64_t n_steps);
/* Leapfrog with direct gravity (REB_GRAVITY_BASIC or _COMPENSATED) in the given
 * REB_AMD_PRECISION for n_steps steps.  Job id or -1. */
REB_API long reb_amd_server_submit_nbody(struct reb_amd_server* s, size_t n, const struct reb_particle* bodies,
                                         double G, double softening, double dt, uint64_t n_steps, int gravity,
                                         int precision);
/* A simulation as a job that continues from its state, time and step count
 * for n_steps steps: WHFast in democratic heliocentric coordinates (1..65
 * bodies, no correctors, default kernel, BASIC gravity, no softening) or
 * order-2 leapfrog with BASIC or COMPENSATED direct gravity in the
 * simulation's REB_AMD_PRECISION.  A WHFast simulation that carries its
 * heliocentric state (p_jh, e.g. one read from a checkpoint) continues from
 * it.  The simulation is not modified.  Job id or -1.  Serves resumption from
 * reb_amd_server_checkpoint; no counterpart in the reference. */
REB_API long reb_amd_server_submit_simulation(struct reb_amd_server* s, struct reb_simulation* r, uint64_t n_steps);
/* Writes the job's current state to `filename` as one Simulationarchive
 * snapshot (appended when the file exists; simulationarchive.c:411-615
 * format, WHFast's p_jh and synchronization flag included): a pending job as
 * submitted, a running one at its next pool tick, batch round or step-segment
 * boundary, a finished one as it ended.  Blocks until written.  Stepping a
 * simulation read from the snapshot (reb_simulation_create_from_file) for the
 * job's remaining steps, or submitting it with
 * reb_amd_server_submit_simulation, reproduces the job's result bit for bit.
 * 0 on success, 1 (message in reb_amd_last_error) for an unknown or failed
 * job or a file that cannot be written.  No counterpart in the reference. */
REB_API int reb_amd_server_checkpoint(struct reb_amd_server* s, long job, const char* filename);
/* Device-memory footprint of a job, as admission counts it. */
REB_API size_t reb_amd_server_whfast_bytes(size_t n_bodies);
REB_API size_t reb_amd_server_nbody_bytes(size_t n, int precision);
REB_API int reb_amd_server_status(struct reb_amd_server* s, long job);
/* Waits for completion (timeout_s < 0: no limit); returns the job state. */
REB_API int reb_amd_server_wait(struct reb_amd_server* s, long job, double timeout_s);
/* Copies a finished job's bodies (n must equal the submitted count), its
 * energies before and after, and its submit/admit/done times (seconds since
 * the server was created), then forgets the job.  0 on success. */
REB_API int reb_amd_server_result(struct reb_amd_server* s, long job, struct reb_particle* out, size_t n,
                                  double* energy2, double* times3);
/* Why a job failed: an error the simulation reported, a device failure of
 * the pool or batch it ran in, or "non-finite coordinates after N steps"
 * (a job whose bodies are not all finite at its end fails; other jobs are
 * not affected). */
REB_API const char* reb_amd_server_job_error(struct reb_amd_server* s, long job);
/* out[0..9]: jobs held, pending, running, done, failed, bytes in use, peak
 * bytes in use, budget, pool ticks, pool busy seconds. */
REB_API int reb_amd_server_stats(struct reb_amd_server* s, double* out, int n);

#ifdef __cplusplus
}
#endif

#endif /* REBOUND_AMD_H */
