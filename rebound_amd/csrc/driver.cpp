// The step driver: reb_simulation_integrate / steps / step / synchronize /
// update_acceleration, with the reference's control flow (simulation.c:241-689)
// and the residency rules of this engine: force evaluation, integrator
// streams, boundary and collision search run on the GPU; user callbacks run on
// the host after the particle state has been brought back.
#include <cmath>
#include <csignal>
#include <cstring>
#include <sys/time.h>
#include <unistd.h>

#include "internal.h"

using namespace rbamd;

static volatile sig_atomic_t g_sigint = 0;
static void sigint_handler(int signum) {
    if (signum == SIGINT) g_sigint += 1;
}

namespace rbamd {

// simulation.c:640-689
void update_acceleration_internal(struct reb_simulation* r) {
    if (r->gravity == REB_GRAVITY_CUSTOM) {
        if (!r->gravity_custom) {
            reb_simulation_error(r, "REB_GRAVITY_CUSTOM selected, but r->gravity_custom function pointer not provided.");
            return;
        }
        run_host_callback(r, r->gravity_custom);
        return;
    }
    if (!ensure_device(r)) return;
    bool ok = true;
    switch (r->gravity) {
        case REB_GRAVITY_NONE:
            ok = gravity_zero(r);
            break;
        case REB_GRAVITY_TREE:
            ok = gravity_tree(r);
            break;
        case REB_GRAVITY_BASIC:
            ok = gravity_direct(r, false);
            break;
        case REB_GRAVITY_COMPENSATED:
            ok = gravity_direct(r, true);
            break;
        case REB_GRAVITY_JACOBI:
            ok = gravity_jacobi(r);
            break;
        default:
            reb_simulation_error(r, "Gravity module not found.");
            return;
    }
    mark_device_newer(r);
    if (!ok) return;
    if (r->N_var) {
        // the WHFast plugin integrates variational sets inside its own kernel
        reb_simulation_error(r, "Variational particles are supported with the WHFast integrator only on this engine.");
        return;
    }
    DeviceState* d = dev_of(r);
    if (d->dev_forces) {
        // the callback enqueues on the simulation's stream: no host round trip
        const struct reb_amd_device_bodies b = {d->field(FX), d->field(FY), d->field(FZ), d->field(FVX),
                                                d->field(FVY), d->field(FVZ), d->field(FAX), d->field(FAY),
                                                d->field(FAZ), d->field(FM), d->field(FR), r->N, (void*)d->stream};
        d->dev_forces(r, &b, d->dev_forces_user);
        mark_device_newer(r);
    }
    if (r->additional_forces) run_host_callback(r, r->additional_forces);
}

}  // namespace rbamd

// Device reductions of the heartbeat's exit checks (simulation.c:241-275).
namespace rbamd {
bool exit_max_distance_device(struct reb_simulation* r, double max2);
bool exit_min_distance_device(struct reb_simulation* r, double min2);
}

static void run_heartbeat(struct reb_simulation* const r) {
    if (r->heartbeat) run_host_callback(r, r->heartbeat);
    if (r->exit_max_distance) {
        const double max2 = r->exit_max_distance * r->exit_max_distance;
        if (ensure_device(r) && exit_max_distance_device(r, max2)) r->status = REB_STATUS_ESCAPE;
    }
    if (r->exit_min_distance) {
        const double min2 = r->exit_min_distance * r->exit_min_distance;
        if (ensure_device(r) && exit_min_distance_device(r, min2)) r->status = REB_STATUS_ENCOUNTER;
    }
}

static int error_message_waiting(struct reb_simulation* const r) {
    if (r->messages)
        for (size_t i = 0; i < 10; i++)
            if (r->messages[i] != nullptr && r->messages[i][0] == 'e') return 1;
    return 0;
}

// A host-side custom integrator (registered with reb_integrator_register)
// works on r->particles, so the state is brought to the host around it.
static bool builtin_device_integrator(const struct reb_simulation* r) {
    const char* n = r->integrator.name;
    return n && (strcmp(n, "leapfrog") == 0 || strcmp(n, "sei") == 0 || strcmp(n, "whfast") == 0);
}

// simulation.c:514-603
static void reb_simulation_step(struct reb_simulation* const r) {
    struct timeval t0;
    gettimeofday(&t0, nullptr);
    DeviceState* d = dev_of(r);
    if (r->pre_timestep_modifications) {
        reb_simulation_synchronize(r);
        run_host_callback(r, r->pre_timestep_modifications);
    }
    if (r->integrator.callbacks.step) {
        if (builtin_device_integrator(r) || r->integrator.callbacks.step == reb_integrator_none.step) {
            r->integrator.callbacks.step(r, r->integrator.state);
        } else {
            ensure_host(r);
            r->integrator.callbacks.step(r, r->integrator.state);
            mark_host_newer(r);
        }
    }
    if (r->N_odes) reb_simulation_error(r, "User ODEs are not supported by this engine.");
    if (r->post_timestep_modifications) {
        reb_simulation_synchronize(r);
        run_host_callback(r, r->post_timestep_modifications);
    }
    r->did_modify_particles = 0;
    if (r->N_var) {
        const char* n = r->integrator.name;
        if (n && strcmp(n, "whfast") == 0)
            reb_simulation_rescale_var(r);
        else
            reb_simulation_error(r, "Variational particles are supported with the WHFast integrator only on this engine.");
    }
    if (r->boundary != REB_BOUNDARY_NONE) boundary_device(r);
    if (r->collision != REB_COLLISION_NONE) {
        if (collision_search_device(r)) collision_resolve_host(r);
    }
    (void)d;
    struct timeval t1;
    gettimeofday(&t1, nullptr);
    r->walltime_last_step = t1.tv_sec - t0.tv_sec + (t1.tv_usec - t0.tv_usec) / 1e6;
    r->walltime_last_steps_sum += r->walltime_last_step;
    r->walltime_last_steps_N += 1;
    if (r->walltime_last_steps_sum > 0.1) {
        r->walltime_last_steps = r->walltime_last_steps_sum / r->walltime_last_steps_N;
        r->walltime_last_steps_sum = 0;
        r->walltime_last_steps_N = 0;
    }
    r->walltime += r->walltime_last_step;
    r->steps_done++;
}

// simulation.c:291-369
static int reb_check_exit(struct reb_simulation* const r, const double tmax, double* last_full_dt) {
    if (r->status <= REB_STATUS_SINGLE_STEP) {
        if (r->status == REB_STATUS_SINGLE_STEP)
            r->status = REB_STATUS_PAUSED;
        else
            r->status = (enum REB_STATUS)(r->status + 1);
    }
    while (r->status == REB_STATUS_PAUSED || r->status == REB_STATUS_SCREENSHOT) {
        usleep(1000);
        if (g_sigint) r->status = REB_STATUS_SIGINT;
    }
    const double dtsign = copysign(1., r->dt);
    if (error_message_waiting(r)) r->status = REB_STATUS_GENERIC_ERROR;
    if (r->status >= 0) {
    } else if (tmax != INFINITY) {
        if (r->exact_finish_time == 1) {
            if ((r->t + r->dt) * dtsign >= tmax * dtsign) {
                if (r->t == tmax) {
                    r->status = REB_STATUS_SUCCESS;
                } else if (r->status == REB_STATUS_LAST_STEP) {
                    double tscale = 1e-12 * fabs(tmax);
                    if (tscale < 1e-200) tscale = 1e-12;
                    if (fabs(r->t - tmax) < tscale) {
                        r->status = REB_STATUS_SUCCESS;
                    } else {
                        reb_simulation_synchronize(r);
                        r->dt = tmax - r->t;
                    }
                } else {
                    r->status = REB_STATUS_LAST_STEP;
                    reb_simulation_synchronize(r);
                    if (r->dt_last_done != 0.) *last_full_dt = r->dt_last_done;
                    r->dt = tmax - r->t;
                }
            } else {
                if (r->status == REB_STATUS_LAST_STEP) r->status = REB_STATUS_RUNNING;
            }
        } else {
            if (r->t * dtsign >= tmax * dtsign) r->status = REB_STATUS_SUCCESS;
        }
    }
    if (!r->N && !r->N_odes) {
        reb_simulation_warning(r, "No particles in simulation. Will exit.");
        r->status = REB_STATUS_NO_PARTICLES;
    }
    return r->status;
}

static void begin_call(struct reb_simulation* r) {
    // The caller may have written r->particles since the last call (unless the
    // simulation was declared device-resident).
    DeviceState* d = dev_of(r);
    if (!d->resident) mark_host_newer(r);
}

static void end_call(struct reb_simulation* r) {
    check_device_flags(r);
    if (!dev_of(r)->resident) ensure_host(r);
}

extern "C" {

REB_API void reb_simulation_update_acceleration(struct reb_simulation* r) {
    begin_call(r);
    update_acceleration_internal(r);
    end_call(r);
}

REB_API void reb_simulation_synchronize(struct reb_simulation* r) {
    if (r->integrator.callbacks.synchronize) r->integrator.callbacks.synchronize(r, r->integrator.state);
    ensure_host(r);
}

// simulation.c:504-513
REB_API void reb_simulation_steps(struct reb_simulation* const r, size_t N_steps) {
    begin_call(r);
    run_heartbeat(r);
    for (size_t i = 0; i < N_steps; i++) {
        if (r->simulationarchive_filename) archive_heartbeat(r);
        reb_simulation_step(r);
        run_heartbeat(r);
    }
    reb_simulation_synchronize(r);
    if (r->simulationarchive_filename) archive_heartbeat(r);
    end_call(r);
}

// simulation.c:372-502 (no visualization threads in this engine)
REB_API enum REB_STATUS reb_simulation_integrate(struct reb_simulation* const r, double tmax) {
    begin_call(r);
    g_sigint = 0;
    void (*old)(int) = signal(SIGINT, sigint_handler);
    if (tmax != r->t) {
        int dt_sign = (tmax > r->t) ? 1.0 : -1.0;
        r->dt = copysign(r->dt, dt_sign);
    }
    double last_full_dt = r->dt;
    r->dt_last_done = 0.;
    if (r->testparticle_hidewarnings == 0 && r->N_active != SIZE_MAX && r->N_active != r->N &&
        r->testparticle_type == 0) {
        for (size_t i = r->N_active; i < r->N; i++)
            if (r->particles[i].m != 0.) {
                reb_simulation_warning(r, "At least one test particle (type 0) has finite mass. This might lead to "
                                          "unexpected behaviour. Set testparticle_hidewarnings=1 to hide this warning.");
                break;
            }
    }
    if (r->status != REB_STATUS_PAUSED && r->status != REB_STATUS_SCREENSHOT) r->status = REB_STATUS_RUNNING;
    run_heartbeat(r);
    while (reb_check_exit(r, tmax, &last_full_dt) < 0) {
        if (r->simulationarchive_filename) archive_heartbeat(r);
        reb_simulation_step(r);
        run_heartbeat(r);
        if (g_sigint) r->status = REB_STATUS_SIGINT;
        if (r->usleep > 0) usleep((useconds_t)r->usleep);
    }
    reb_simulation_synchronize(r);
    if (r->exact_finish_time == 1) r->dt = last_full_dt;
    if (r->simulationarchive_filename) archive_heartbeat(r);
    signal(SIGINT, old);
    end_call(r);
    return r->status;
}

REB_API double reb_simulation_energy(struct reb_simulation* const r) {
    begin_call(r);
    if (!ensure_device(r)) return nan("");
    double e = energy_device(r);
    end_call(r);
    return e;
}

REB_API void reb_amd_set_profiling(struct reb_simulation* r, int on) {
    DeviceState* d = dev_of(r);
    stream_sync(r);
    d->profiling = on != 0;
    d->kev_used = 0;
}

REB_API int reb_amd_kernel_time(struct reb_simulation* r, double* total_ms, int* launches) {
    DeviceState* d = dev_of(r);
    if (!stream_sync(r)) return 1;
    double tot = 0.;
    int n = 0;
    for (size_t k = 0; k + 1 < d->kev_used; k += 2) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, d->kev[k], d->kev[k + 1]) != hipSuccess) return 2;
        tot += ms;
        n++;
    }
    d->kev_used = 0;
    if (total_ms) *total_ms = tot;
    if (launches) *launches = n;
    return 0;
}

REB_API double reb_amd_kernel_pairs(struct reb_simulation* r) {
    DeviceState* d = dev_of(r);
    const double p = d->kpairs;
    d->kpairs = 0.;
    return p;
}

REB_API void reb_amd_set_resident(struct reb_simulation* r, int on) {
    DeviceState* d = dev_of(r);
    if (!on && d->resident) ensure_host(r);
    d->resident = on != 0;
}

REB_API int reb_amd_sync_host(struct reb_simulation* r) { return ensure_host(r) ? 0 : 1; }

REB_API void reb_amd_mark_host_modified(struct reb_simulation* r) { mark_host_newer(r); }

REB_API int reb_amd_synchronize_device(struct reb_simulation* r) { return stream_sync(r) ? 0 : 1; }

}  // extern "C"
