// Streaming kernels of the integrators and the boundary, and the device
// implementations of the "leapfrog" and "sei" integrator plugins.
//
//   drift / kick          integrator_leapfrog.c:72-93
//   leapfrog step         integrator_leapfrog.c:97-209 (orders 2, 4, 6, 8)
//   SEI step              integrator_sei.c:86-117
//   H0 (epicycle) + phi1  integrator_sei.c:126-174
//   boundary wrap         boundary.c:35-141 (OPEN, PERIODIC, SHEAR)
//
// Each kernel is one pass over the SoA particle block (x, v, a: 48 B read and
// 24-48 B written per body), so they run at HBM bandwidth; the expressions are
// the reference's, term for term, and the library is built without FP
// contraction, so the results are bit-identical to the CPU loops.
#include <cmath>
#include <csignal>

#include "internal.h"

namespace rbamd {

// ---------------------------------------------------------------- kernels
__global__ void k_drift(double* __restrict__ x, double* __restrict__ y, double* __restrict__ z,
                        const double* __restrict__ vx, const double* __restrict__ vy, const double* __restrict__ vz,
                        double dt, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] += dt * vx[i];
    y[i] += dt * vy[i];
    z[i] += dt * vz[i];
}

__global__ void k_kick(double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
                       const double* __restrict__ ax, const double* __restrict__ ay, const double* __restrict__ az,
                       double dt, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    vx[i] += dt * ax[i];
    vy[i] += dt * ay[i];
    vz[i] += dt * az[i];
}

struct SeiConst {
    double dt, OMEGA, OMEGAZ, sindt, tandt, sindtz, tandtz;
};

// Exact solution of Hill's equations over dt as three shears (integrator_sei.c:126-157),
// optionally preceded by the phi1 kick v += a dt (integrator_sei.c:168-174).
__global__ void k_sei(double* __restrict__ x, double* __restrict__ y, double* __restrict__ z,
                      double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
                      const double* __restrict__ ax, const double* __restrict__ ay, const double* __restrict__ az,
                      SeiConst c, int with_kick, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double px = x[i], py = y[i], pz = z[i], pvx = vx[i], pvy = vy[i], pvz = vz[i];
    if (with_kick) {
        pvx += ax[i] * c.dt;
        pvy += ay[i] * c.dt;
        pvz += az[i] * c.dt;
    }
    // vertical
    const double zx = pz * c.OMEGAZ;
    const double zy = pvz;
    const double zt1 = zx - c.tandtz * zy;
    const double zyt = c.sindtz * zt1 + zy;
    const double zxt = zt1 - c.tandtz * zyt;
    pz = zxt / c.OMEGAZ;
    pvz = zyt;
    // in-plane: epicycle centre (aO, bO) and epicycle vector (xs, ys)
    const double aO = 2. * pvy + 4. * px * c.OMEGA;
    const double bO = py * c.OMEGA - 2. * pvx;
    const double ys = (py * c.OMEGA - bO) / 2.;
    const double xs = (px * c.OMEGA - aO);
    const double xst1 = xs - c.tandt * ys;
    const double yst = c.sindt * xst1 + ys;
    const double xst = xst1 - c.tandt * yst;
    px = (xst + aO) / c.OMEGA;
    py = (yst * 2. + bO) / c.OMEGA - 3. / 4. * aO * c.dt;
    pvx = yst;
    pvy = -xst * 2. - 3. / 2. * aO;
    x[i] = px;
    y[i] = py;
    z[i] = pz;
    vx[i] = pvx;
    vy[i] = pvy;
    vz[i] = pvz;
}

struct BoxConst {
    double bx, by, bz;       // box size
    double offp1, offm1;     // shear y offsets for a radial wrap (boundary.c:82-83)
    double dvy;              // 3/2 OMEGA bx
};

static constexpr int kWrapMax = 4096;

__global__ void k_wrap_shear(double* __restrict__ x, double* __restrict__ y, double* __restrict__ z,
                             double* __restrict__ vy, BoxConst b, int n, unsigned int* __restrict__ err) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double px = x[i], py = y[i], pz = z[i], pvy = vy[i];
    // A body wraps once per crossing; kWrapMax bounds the loops for bodies
    // that are non-finite or absurdly far away (reported as an error).
    int k = 0;
    while (px > b.bx / 2. && k < kWrapMax) {
        px -= b.bx;
        py += b.offp1;
        pvy += b.dvy;
        k++;
    }
    while (px < -b.bx / 2. && k < kWrapMax) {
        px += b.bx;
        py += b.offm1;
        pvy -= b.dvy;
        k++;
    }
    while (py > b.by / 2. && k < kWrapMax) { py -= b.by; k++; }
    while (py < -b.by / 2. && k < kWrapMax) { py += b.by; k++; }
    while (pz > b.bz / 2. && k < kWrapMax) { pz -= b.bz; k++; }
    while (pz < -b.bz / 2. && k < kWrapMax) { pz += b.bz; k++; }
    if (k >= kWrapMax) atomicOr(err, 1u);
    x[i] = px;
    y[i] = py;
    z[i] = pz;
    vy[i] = pvy;
}

__global__ void k_wrap_periodic(double* __restrict__ x, double* __restrict__ y, double* __restrict__ z, BoxConst b,
                                int n, unsigned int* __restrict__ err) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double px = x[i], py = y[i], pz = z[i];
    int k = 0;
    while (px > b.bx / 2. && k < kWrapMax) { px -= b.bx; k++; }
    while (px < -b.bx / 2. && k < kWrapMax) { px += b.bx; k++; }
    while (py > b.by / 2. && k < kWrapMax) { py -= b.by; k++; }
    while (py < -b.by / 2. && k < kWrapMax) { py += b.by; k++; }
    while (pz > b.bz / 2. && k < kWrapMax) { pz -= b.bz; k++; }
    while (pz < -b.bz / 2. && k < kWrapMax) { pz += b.bz; k++; }
    if (k >= kWrapMax) atomicOr(err, 1u);
    x[i] = px;
    y[i] = py;
    z[i] = pz;
}

// OPEN boundary: count bodies outside the box; the (rare) removals are then
// applied on the host in index order, as boundary.c:43-76 does.
__global__ void k_count_outside(const double* __restrict__ x, const double* __restrict__ y,
                                const double* __restrict__ z, BoxConst b, int n, unsigned int* __restrict__ count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double px = x[i], py = y[i], pz = z[i];
    const bool out = px > b.bx / 2. || px < -b.bx / 2. || py > b.by / 2. || py < -b.by / 2. || pz > b.bz / 2. ||
                     pz < -b.bz / 2.;
    if (out) atomicAdd(count, 1u);
}

static inline dim3 grid_for(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

// ---------------------------------------------------------------- launchers
bool launch_drift(struct reb_simulation* r, double dt) {
    DeviceState* d = dev_of(r);
    const int N = (int)r->N;
    if (N) {
        hipLaunchKernelGGL(k_drift, grid_for(N), dim3(256), 0, d->stream, d->field(FX), d->field(FY), d->field(FZ),
                           d->field(FVX), d->field(FVY), d->field(FVZ), dt, N);
        if (!RB_HIP(r, hipGetLastError())) return false;
    }
    r->t += dt;  // integrator_leapfrog.c:81
    mark_device_newer(r);
    return true;
}

bool launch_kick(struct reb_simulation* r, double dt) {
    DeviceState* d = dev_of(r);
    const int N = (int)r->N;
    if (N) {
        hipLaunchKernelGGL(k_kick, grid_for(N), dim3(256), 0, d->stream, d->field(FVX), d->field(FVY), d->field(FVZ),
                           d->field(FAX), d->field(FAY), d->field(FAZ), dt, N);
        if (!RB_HIP(r, hipGetLastError())) return false;
    }
    mark_device_newer(r);
    return true;
}

bool launch_sei_h012(struct reb_simulation* r, double dt, double sindt, double tandt, double sindtz, double tandtz,
                     int with_kick) {
    DeviceState* d = dev_of(r);
    const int N = (int)r->N;
    if (!N) return true;
    SeiConst c{dt, r->OMEGA, r->OMEGAZ, sindt, tandt, sindtz, tandtz};
    hipLaunchKernelGGL(k_sei, grid_for(N), dim3(256), 0, d->stream, d->field(FX), d->field(FY), d->field(FZ),
                       d->field(FVX), d->field(FVY), d->field(FVZ), d->field(FAX), d->field(FAY), d->field(FAZ), c,
                       with_kick, N);
    mark_device_newer(r);
    return RB_HIP(r, hipGetLastError());
}

// A device word that the wrap kernels set when a body is too far outside the
// box; checked (and cleared) at the end of each stepping call.
unsigned int* wrap_error_word(struct reb_simulation* r) {
    DeviceState* d = dev_of(r);
    if (!d->flags.p) {
        if (!d->flags.ensure(256)) return nullptr;
        (void)hipMemsetAsync(d->flags.p, 0, 256, d->stream);
    }
    return d->flags.as<unsigned int>();
}

bool check_device_flags(struct reb_simulation* r) {
    DeviceState* d = dev_of(r);
    if (!d->flags.p) return true;
    unsigned int h = 0;
    if (!RB_HIP(r, hipMemcpyAsync(&h, d->flags.p, sizeof(h), hipMemcpyDeviceToHost, d->stream))) return false;
    if (!stream_sync(r)) return false;
    if (h) {
        reb_simulation_error(r, "A particle is non-finite or thousands of box widths outside the box; it was not wrapped.");
        (void)hipMemsetAsync(d->flags.p, 0, sizeof(unsigned int), d->stream);
        return false;
    }
    return true;
}

bool boundary_device(struct reb_simulation* r) {
    DeviceState* d = dev_of(r);
    const int N = (int)r->N;
    if (r->boundary == REB_BOUNDARY_NONE || N == 0) return true;
    if (!ensure_device(r)) return false;
    BoxConst b;
    b.bx = r->root_size * (double)r->N_root_x;
    b.by = r->root_size * (double)r->N_root_y;
    b.bz = r->root_size * (double)r->N_root_z;
    b.offp1 = b.offm1 = b.dvy = 0.;
    switch (r->boundary) {
        case REB_BOUNDARY_SHEAR: {
            const double OMEGA = r->OMEGA;
            b.offp1 = -fmod(-1.5 * OMEGA * b.bx * r->t + b.by / 2., b.by) - b.by / 2.;
            b.offm1 = -fmod(1.5 * OMEGA * b.bx * r->t - b.by / 2., b.by) + b.by / 2.;
            b.dvy = 3. / 2. * OMEGA * b.bx;
            hipLaunchKernelGGL(k_wrap_shear, grid_for(N), dim3(256), 0, d->stream, d->field(FX), d->field(FY),
                               d->field(FZ), d->field(FVY), b, N, wrap_error_word(r));
            mark_device_newer(r);
            return RB_HIP(r, hipGetLastError());
        }
        case REB_BOUNDARY_PERIODIC:
            hipLaunchKernelGGL(k_wrap_periodic, grid_for(N), dim3(256), 0, d->stream, d->field(FX), d->field(FY),
                               d->field(FZ), b, N, wrap_error_word(r));
            mark_device_newer(r);
            return RB_HIP(r, hipGetLastError());
        case REB_BOUNDARY_OPEN: {
            if (!d->scratch.ensure(256)) return false;
            unsigned int* cnt = d->scratch.as<unsigned int>();
            if (!RB_HIP(r, hipMemsetAsync(cnt, 0, sizeof(unsigned int), d->stream))) return false;
            hipLaunchKernelGGL(k_count_outside, grid_for(N), dim3(256), 0, d->stream, d->field(FX), d->field(FY),
                               d->field(FZ), b, N, cnt);
            unsigned int h = 0;
            if (!RB_HIP(r, hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, d->stream))) return false;
            if (!stream_sync(r)) return false;
            if (h == 0) return true;
            // Removal shifts the array and may need the energy (track_energy_offset):
            // done on the host copy, in the reference's order.
            if (!ensure_host(r)) return false;
            int n = (int)r->N;
            for (int i = 0; i < n; i++) {
                const struct reb_particle p = r->particles[i];
                const bool out = p.x > b.bx / 2. || p.x < -b.bx / 2. || p.y > b.by / 2. || p.y < -b.by / 2. ||
                                 p.z > b.bz / 2. || p.z < -b.bz / 2.;
                if (!out) continue;
                if (r->track_energy_offset) {
                    double Ei = reb_simulation_energy(r);
                    reb_simulation_remove_particle(r, i);
                    r->energy_offset += Ei - reb_simulation_energy(r);
                } else {
                    reb_simulation_remove_particle(r, i);
                }
                i--;
                n--;
            }
            mark_host_newer(r);
            return ensure_device(r);
        }
        default:
            return true;
    }
}

// ---------------------------------------------------------------- leapfrog plugin
// Composition coefficients of the 4th (Yoshida), 6th and 8th order schemes
// (integrator_leapfrog.c:66-68); identical values are needed for identical
// trajectories.
static const double lf4_a = 0.675603595979828817023843904485;
static const double lf6_a[5] = {0.1867, 0.5554970237124784, 0.1294669489134754, -0.843265623387734,
                                0.9432033015235604};
static const double lf8_a[9] = {0.128865979381443,   0.581514087105251,  -0.410175371469850,
                                0.1851469357165877,  -0.4095523434208514, 0.1444059410800120,
                                0.2783355003936797,  0.3149566839162949, -0.6269948254051343979};

void update_acceleration_internal(struct reb_simulation* r);

}  // namespace rbamd

using namespace rbamd;

extern "C" {

void* rb_leapfrog_create(void) {
    auto* s = (struct reb_integrator_leapfrog_state*)calloc(1, sizeof(struct reb_integrator_leapfrog_state));
    s->order = 2;
    return s;
}

// Drift-kick-drift, composed to the requested order (integrator_leapfrog.c:97-209).
void rb_leapfrog_step(struct reb_simulation* r, void* state) {
    r->gravity_ignore_terms = REB_GRAVITY_IGNORE_TERMS_NONE;
    const double dt = r->dt;
    auto* lf = (struct reb_integrator_leapfrog_state*)state;
    if (!ensure_device(r)) return;
    auto D = [&](double h) { launch_drift(r, h); };
    auto K = [&](double h) { launch_kick(r, h); };
    auto F = [&]() { update_acceleration_internal(r); ensure_device(r); };
    // Direct gravity with nothing between the force and the kick: the kick and
    // the closing drift run in the pass that finishes the accelerations (same
    // arithmetic, two launches fewer per step).
    const bool fuse = (r->gravity == REB_GRAVITY_BASIC || r->gravity == REB_GRAVITY_COMPENSATED) &&
                      !r->additional_forces && !dev_of(r)->dev_forces && !r->N_var && r->N > 0;
    switch (lf->order) {
        case 2:
            D(dt * 0.5);
            if (fuse) {
                const double kd[2] = {dt, dt * 0.5};
                if (!gravity_direct(r, r->gravity == REB_GRAVITY_COMPENSATED, kd)) return;
                r->t += dt * 0.5;  // integrator_leapfrog.c:81
                mark_device_newer(r);
            } else {
                F(); K(dt); D(dt * 0.5);
            }
            break;
        case 4:
            D(dt * lf4_a); F(); K(dt * 2. * lf4_a);
            D(dt * (0.5 - lf4_a)); F(); K(dt * (1. - 4. * lf4_a));
            D(dt * (0.5 - lf4_a)); F(); K(dt * 2. * lf4_a);
            D(dt * lf4_a);
            break;
        case 6: {
            const double* a = lf6_a;
            D(dt * a[0] * 0.5); F(); K(dt * a[0]);
            D(dt * (a[0] + a[1]) * 0.5); F(); K(dt * a[1]);
            D(dt * (a[1] + a[2]) * 0.5); F(); K(dt * a[2]);
            D(dt * (a[2] + a[3]) * 0.5); F(); K(dt * a[3]);
            D(dt * (a[3] + a[4]) * 0.5); F(); K(dt * a[4]);
            D(dt * (a[3] + a[4]) * 0.5); F(); K(dt * a[3]);
            D(dt * (a[2] + a[3]) * 0.5); F(); K(dt * a[2]);
            D(dt * (a[1] + a[2]) * 0.5); F(); K(dt * a[1]);
            D(dt * (a[0] + a[1]) * 0.5); F(); K(dt * a[0]);
            D(dt * a[0] * 0.5);
            break;
        }
        case 8: {
            const double* a = lf8_a;
            D(dt * a[0] * 0.5); F(); K(dt * a[0]);
            for (int k = 1; k <= 8; k++) {
                D(dt * (a[k - 1] + a[k]) * 0.5); F(); K(dt * a[k]);
            }
            D(dt * (a[7] + a[8]) * 0.5); F(); K(dt * a[7]);
            for (int k = 7; k >= 1; k--) {
                D(dt * (a[k - 1] + a[k]) * 0.5); F(); K(dt * a[k - 1]);
            }
            D(dt * a[0] * 0.5);
            break;
        }
        default:
            reb_simulation_error(r, "Leapfrog order not supported.");
            return;
    }
    r->dt_last_done = dt;
}

// ---------------------------------------------------------------- SEI plugin
// State of integrator_sei.c:33-39 (cached sin/tan of the current dt).
struct rb_sei_state {
    double lastdt, sindt, tandt, sindtz, tandtz;
};

void* rb_sei_create(void) { return calloc(1, sizeof(struct rb_sei_state)); }

void rb_sei_step(struct reb_simulation* r, void* state) {
    auto* sei = (struct rb_sei_state*)state;
    r->gravity_ignore_terms = REB_GRAVITY_IGNORE_TERMS_NONE;
    if (sei->lastdt != r->dt) {
        // sin/tan once per dt on the host (libm), integrator_sei.c:91-101
        if (r->OMEGAZ == -1) r->OMEGAZ = r->OMEGA;
        sei->sindt = sin(r->OMEGA * (-r->dt / 2.));
        sei->tandt = tan(r->OMEGA * (-r->dt / 4.));
        sei->sindtz = sin(r->OMEGAZ * (-r->dt / 2.));
        sei->tandtz = tan(r->OMEGAZ * (-r->dt / 4.));
        sei->lastdt = r->dt;
    }
    if (!ensure_device(r)) return;
    launch_sei_h012(r, r->dt, sei->sindt, sei->tandt, sei->sindtz, sei->tandtz, 0);
    r->t += r->dt / 2.;
    update_acceleration_internal(r);
    if (!ensure_device(r)) return;
    launch_sei_h012(r, r->dt, sei->sindt, sei->tandt, sei->sindtz, sei->tandtz, 1);
    r->t += r->dt / 2.;
    r->dt_last_done = r->dt;
}

}  // extern "C"

// ---------------------------------------------------------------- heartbeat exit checks
// simulation.c:243-274: escape (|x| > exit_max_distance) and close encounter
// (any pair closer than exit_min_distance).  Only a yes/no flag is needed.
namespace rbamd {

__global__ void k_escape(const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
                         double max2, int n, unsigned int* __restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double r2 = x[i] * x[i] + y[i] * y[i] + z[i] * z[i];
    if (r2 > max2) atomicOr(flag, 1u);
}

__global__ __launch_bounds__(256) void k_encounter(const double* __restrict__ x, const double* __restrict__ y,
                                                   const double* __restrict__ z, double min2, int n,
                                                   unsigned int* __restrict__ flag) {
    __shared__ double3 tile[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const bool live = i < n;
    const double xi = live ? x[i] : 0., yi = live ? y[i] : 0., zi = live ? z[i] : 0.;
    bool hit = false;
    for (int t = 0; t < n; t += 256) {
        __syncthreads();
        const int jj = t + threadIdx.x;
        if (jj < n) tile[threadIdx.x] = make_double3(x[jj], y[jj], z[jj]);
        __syncthreads();
        const int m = min(256, n - t);
        if (!live) continue;
        for (int k = 0; k < m; k++) {
            const int j = t + k;
            if (j >= i) break;
            const double dx = xi - tile[k].x, dy = yi - tile[k].y, dz = zi - tile[k].z;
            if (dx * dx + dy * dy + dz * dz < min2) hit = true;
        }
    }
    if (hit) atomicOr(flag, 1u);
}

static bool flag_kernel(struct reb_simulation* r, bool encounter, double lim2) {
    DeviceState* d = dev_of(r);
    const int N = (int)r->N;
    if (N == 0) return false;
    if (!d->scratch.ensure(256)) return false;
    unsigned int* f = d->scratch.as<unsigned int>() + 8;
    if (!RB_HIP(r, hipMemsetAsync(f, 0, sizeof(unsigned int), d->stream))) return false;
    if (encounter)
        hipLaunchKernelGGL(k_encounter, grid_for(N), dim3(256), 0, d->stream, d->field(FX), d->field(FY),
                           d->field(FZ), lim2, N, f);
    else
        hipLaunchKernelGGL(k_escape, grid_for(N), dim3(256), 0, d->stream, d->field(FX), d->field(FY), d->field(FZ),
                           lim2, N, f);
    unsigned int h = 0;
    if (!RB_HIP(r, hipMemcpyAsync(&h, f, sizeof(h), hipMemcpyDeviceToHost, d->stream))) return false;
    if (!stream_sync(r)) return false;
    return h != 0;
}

bool exit_max_distance_device(struct reb_simulation* r, double max2) { return flag_kernel(r, false, max2); }
bool exit_min_distance_device(struct reb_simulation* r, double min2) { return flag_kernel(r, true, min2); }

}  // namespace rbamd

namespace rbamd {
__global__ void k_drag(const double* __restrict__ vx, const double* __restrict__ vy, const double* __restrict__ vz,
                       double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az, double gamma,
                       int N) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    ax[i] -= gamma * vx[i];
    ay[i] -= gamma * vy[i];
    az[i] -= gamma * vz[i];
}
}  // namespace rbamd

extern "C" REB_API void reb_amd_device_force_drag(struct reb_simulation* r, const struct reb_amd_device_bodies* b,
                                                  void* user) {
    using namespace rbamd;
    if (!user) {
        reb_simulation_error(r, "reb_amd_device_force_drag needs a pointer to gamma.");
        return;
    }
    if (b->N == 0) return;
    const double gamma = *static_cast<const double*>(user);
    hipLaunchKernelGGL(k_drag, grid_for(b->N), dim3(256), 0, (hipStream_t)b->stream, b->vx, b->vy, b->vz, b->ax,
                       b->ay, b->az, gamma, (int)b->N);
}
