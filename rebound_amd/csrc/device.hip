// Device residency of a simulation's particles.
//
// The public struct keeps REBOUND's array-of-structs `r->particles` (112 B per
// body).  On the GPU the same state lives as a structure of arrays, one
// contiguous block of NFIELD * cap doubles (x y z vx vy vz ax ay az m r), so
// every kernel reads and writes fully coalesced 8-byte lanes.  `loc` records
// which side is newer; every entry point that computes on the device first
// uploads when the host is newer, and every point where host code can run
// (callbacks, return to the caller) downloads when the device is newer.
#include <chrono>
#include <cstring>
#include <mutex>
#include <string>

#include "internal.h"

namespace rbamd {

static std::mutex g_err_mutex;
static std::string g_last_error;

void set_last_error(const std::string& msg) {
    std::lock_guard<std::mutex> g(g_err_mutex);
    g_last_error = msg;
}

bool hip_check(struct reb_simulation* r, hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    char msg[REB_STRING_SIZE_MAX];
    snprintf(msg, sizeof(msg), "HIP error %s in %s", hipGetErrorString(e), what);
    if (r) {
        reb_simulation_error(r, msg);
    } else {
        set_last_error(msg);
        fprintf(stderr, "\nError! %s\n", msg);
    }
    return false;
}

double now_seconds() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

bool DevBuf::ensure(size_t need) {
    if (need <= bytes && p) return true;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    size_t want = need < 256 ? 256 : need;
    if (hipMalloc(&p, want) != hipSuccess) {
        p = nullptr;
        return false;
    }
    bytes = want;
    return true;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

static int g_device_count = -1;
static std::once_flag g_count_once;

static int device_count() {
    std::call_once(g_count_once, [] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
        g_device_count = n;
    });
    return g_device_count;
}

bool device_ready(struct reb_simulation* r) {
    DeviceState* d = dev_of(r);
    if (d->stream) return RB_HIP(r, hipSetDevice(d->device));
    if (device_count() <= 0) {
        reb_simulation_error(r, "No HIP device is available: the MI355X engine cannot step this simulation.");
        return false;
    }
    int dev = 0;
    if (!RB_HIP(r, hipGetDevice(&dev))) return false;
    d->device = dev;
    if (!RB_HIP(r, hipSetDevice(dev))) return false;
    if (!RB_HIP(r, hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking))) return false;
    d->owns_stream = true;
    return true;
}

bool stream_sync(struct reb_simulation* r) {
    DeviceState* d = dev_of(r);
    if (!d->stream) return true;
    return RB_HIP(r, hipStreamSynchronize(d->stream));
}

void mark_host_newer(struct reb_simulation* r) { dev_of(r)->loc = LOC_HOST; }
void mark_device_newer(struct reb_simulation* r) { dev_of(r)->loc = LOC_DEVICE; }

// Capacity grows by doubling, like the host array (particle.c:53-58).
static bool ensure_capacity(struct reb_simulation* r, size_t n) {
    DeviceState* d = dev_of(r);
    if (n <= d->cap && d->soa.p) return true;
    size_t cap = d->cap ? d->cap : 64;
    while (cap < n) cap *= 2;
    if (!d->soa.ensure(sizeof(double) * NFIELD * cap)) {
        reb_simulation_error(r, "Out of device memory for the particle state.");
        return false;
    }
    d->cap = cap;
    if (d->host_stage) (void)hipHostFree(d->host_stage);
    d->host_stage = nullptr;
    if (!RB_HIP(r, hipHostMalloc((void**)&d->host_stage, sizeof(double) * NFIELD * cap))) return false;
    d->host_cap = cap;
    return true;
}

bool ensure_device(struct reb_simulation* r) {
    DeviceState* d = dev_of(r);
    if (!device_ready(r)) return false;
    if (d->loc != LOC_HOST && d->cap >= r->N) return true;
    const size_t N = r->N;
    if (!ensure_capacity(r, N ? N : 1)) return false;
    if (N == 0) {
        d->loc = LOC_BOTH;
        return true;
    }
    // Pack AoS -> SoA in pinned memory, then one DMA per field block.
    double* s = d->host_stage;
    const size_t cap = d->cap;
    const struct reb_particle* p = r->particles;
    for (size_t i = 0; i < N; i++) {
        s[FX * cap + i] = p[i].x;
        s[FY * cap + i] = p[i].y;
        s[FZ * cap + i] = p[i].z;
        s[FVX * cap + i] = p[i].vx;
        s[FVY * cap + i] = p[i].vy;
        s[FVZ * cap + i] = p[i].vz;
        s[FAX * cap + i] = p[i].ax;
        s[FAY * cap + i] = p[i].ay;
        s[FAZ * cap + i] = p[i].az;
        s[FM * cap + i] = p[i].m;
        s[FR * cap + i] = p[i].r;
    }
    for (int f = 0; f < NFIELD; f++) {
        if (!RB_HIP(r, hipMemcpyAsync(d->field(f), s + (size_t)f * cap, sizeof(double) * N, hipMemcpyHostToDevice,
                                      d->stream)))
            return false;
    }
    if (!stream_sync(r)) return false;
    d->loc = LOC_BOTH;
    return true;
}

bool ensure_host(struct reb_simulation* r) {
    DeviceState* d = dev_of(r);
    if (d->loc != LOC_DEVICE) return true;
    const size_t N = r->N;
    double* s = d->host_stage;
    const size_t cap = d->cap;
    for (int f = 0; f < NFIELD; f++) {
        if (!RB_HIP(r, hipMemcpyAsync(s + (size_t)f * cap, d->field(f), sizeof(double) * N, hipMemcpyDeviceToHost,
                                      d->stream)))
            return false;
    }
    if (!stream_sync(r)) return false;
    struct reb_particle* p = r->particles;
    for (size_t i = 0; i < N; i++) {
        p[i].x = s[FX * cap + i];
        p[i].y = s[FY * cap + i];
        p[i].z = s[FZ * cap + i];
        p[i].vx = s[FVX * cap + i];
        p[i].vy = s[FVY * cap + i];
        p[i].vz = s[FVZ * cap + i];
        p[i].ax = s[FAX * cap + i];
        p[i].ay = s[FAY * cap + i];
        p[i].az = s[FAZ * cap + i];
        p[i].m = s[FM * cap + i];
        p[i].r = s[FR * cap + i];
    }
    d->loc = LOC_BOTH;
    return true;
}

void run_host_callback(struct reb_simulation* r, void (*cb)(struct reb_simulation* const)) {
    ensure_host(r);
    cb(r);
    mark_host_newer(r);
}

void device_release(struct reb_simulation* r) {
    DeviceState* d = dev_of(r);
    if (d->stream) {
        (void)hipSetDevice(d->device);
        (void)hipStreamSynchronize(d->stream);
    }
    whfast_release(r);
    tree_release(r);
    resolve_release(r);
    coll_release(r);
    d->soa.release();
    d->partial.release();
    d->scratch.release();
    d->aos.release();
    d->flags.release();
    if (d->host_stage) (void)hipHostFree(d->host_stage);
    d->host_stage = nullptr;
    for (auto& e : d->kev)
        if (e) (void)hipEventDestroy(e);
    d->kev.clear();
    d->kev_used = 0;
    if (d->stream && d->owns_stream) (void)hipStreamDestroy(d->stream);
    d->stream = nullptr;
    d->cap = 0;
}

}  // namespace rbamd

using namespace rbamd;

extern "C" {

REB_API int reb_amd_device_count(void) { return device_count(); }

REB_API int reb_amd_device_name(char* buf, size_t len) {
    if (device_count() <= 0 || len == 0) return 1;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 1;
    snprintf(buf, len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return 0;
}

REB_API const char* reb_amd_version(void) { return "rebound-amd 0.1 (REBOUND 5.0.0 ABI, gfx950)"; }

REB_API const char* reb_amd_last_error(void) {
    static thread_local std::string copy;
    std::lock_guard<std::mutex> g(g_err_mutex);
    copy = g_last_error;
    return copy.c_str();
}

REB_API int reb_amd_set_precision(struct reb_simulation* r, int precision) {
    if (precision < REB_AMD_PRECISION_FP64 || precision > REB_AMD_PRECISION_SPLIT) {
        reb_simulation_error(r, "Unknown precision mode.");
        return 1;
    }
    dev_of(r)->precision = precision;
    return 0;
}

REB_API int reb_amd_get_precision(const struct reb_simulation* r) { return dev_of(r)->precision; }

}  // extern "C"
