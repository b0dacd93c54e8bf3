// Serving loop: a job queue in front of the GPU with admission by memory
// footprint and continuous batching.
//
// Jobs.  Two kinds, each the reference's own entry point applied to one
// simulation: a planetary system integrated by WHFast (democratic
// heliocentric, reb_simulation_steps: simulation.c:504-513,
// integrator_whfast.c:1113-1259) and an N-body cluster integrated by leapfrog
// with direct gravity (integrator_leapfrog.c:97-209, gravity.c:167-531) in a
// chosen precision.  A job's result is what the reference returns for that
// simulation alone; it does not depend on which other jobs share the GPU.
//
// Admission.  Every job has a device-memory footprint (its state, scratch and
// metadata).  The scheduler admits pending jobs in arrival order while their
// footprints fit in the budget; a job that does not fit waits, and smaller
// later jobs may pass it.  Memory is returned when a job completes.
//
// Continuous batching.  Planetary systems live in one device pool (a
// structure-of-arrays arena, see ensemble.hip for the layout).  The scheduler
// runs the pool in ticks: each tick is one persistent-kernel launch that
// advances every active system by min(remaining, tick) steps; systems that
// reach their horizon stop inside the launch and synchronize at its end.
// Jobs admitted while a tick runs join at the next tick; finished systems
// leave, and their slots and bodies are reused.  N-body jobs are batched the
// same way on a second stream: every scheduler round advances all of them by
// the same number of leapfrog steps, three launches per step for the whole
// batch (gravity.hip k_seg_*), each job with the chunking it would use alone;
// they join and leave between rounds.  Jobs in EXACT precision, and jobs
// large enough for the pair-symmetric force path (gravity.hip), run as single
// simulations from worker threads on a third stream, so that every job's bits
// are those of stepping it alone.  The streams overlap on the device.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <limits>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "internal.h"
#include "whfast_device.h"

namespace rbamd {

enum JobKind { JOB_WHFAST = 0, JOB_NBODY = 1 };

// A request to write a running job's state to an archive file; served by the
// thread that steps the job, at the next point where its state is whole.
struct Checkpoint {
    std::string file;
    bool done = false;   // Server::mu
    int rc = 0;
    std::string error;
};

struct Job {
    long id = 0;
    int kind = JOB_WHFAST;
    std::vector<struct reb_particle> bodies;   // input, then result
    double dt = 0., G = 1., softening = 0.;
    int safe_mode = 0, gravity = REB_GRAVITY_BASIC, precision = REB_AMD_PRECISION_FP64;
    uint64_t n_steps = 0;
    size_t footprint = 0;
    int state = REB_AMD_JOB_PENDING;
    std::string error;
    double t_submit = 0., t_admit = 0., t_done = 0.;
    double e0 = 0., e1 = 0.;
    // where the job starts: time and step count of the submitted simulation,
    // and for WHFast its heliocentric state (x y z vx vy vz m, rows of n;
    // empty: computed from the bodies) with the synchronization flag
    double t0 = 0.;
    uint64_t steps0 = 0;
    bool resumed = false;
    std::vector<double> pjh;
    int synchronized = 1;
    std::vector<std::shared_ptr<Checkpoint>> ckpt;   // requests not yet served (Server::mu)
    // pool bookkeeping (scheduler thread only)
    int slot = -1;
    size_t body0 = 0;
    uint64_t remaining = 0;
    bool fresh = false;
    // N-body batch: the job's device arrays (x y z vx vy vz ax ay az m, n apart)
    DevBuf soa, partial, scratch;
    int chunk = 0, used = 0;
};

// Pinned staging buffer that grows.
struct Pinned {
    void* p = nullptr;
    size_t bytes = 0;
    bool ensure(size_t need) {
        if (need <= bytes && p) return true;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
        size_t want = std::max<size_t>(need, 1 << 16);
        if (hipHostMalloc(&p, want) != hipSuccess) {
            p = nullptr;
            return false;
        }
        bytes = want;
        return true;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
};

__global__ void k_pool_scatter(const double* __restrict__ stage, const int* __restrict__ dest, int nb,
                               double* __restrict__ body, double* __restrict__ ph, size_t stride) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    const size_t b = (size_t)dest[k];
    for (int f = 0; f < 10; f++) body[f * stride + b] = stage[10 * (size_t)k + f];
    for (int f = 0; f < 7; f++) ph[f * stride + b] = 0.;
}

__global__ void k_pool_set_flags(const int* __restrict__ slots, int ns, unsigned int* __restrict__ flags,
                                 unsigned int value) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < ns) flags[slots[k]] = value;
}

// Bodies (10 fields) and heliocentric state (7 fields) of finished systems.
__global__ void k_pool_gather(const int* __restrict__ src, int nb, const double* __restrict__ body,
                              const double* __restrict__ ph, size_t stride, double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    const size_t b = (size_t)src[k];
    for (int f = 0; f < 10; f++) out[17 * (size_t)k + f] = body[f * stride + b];
    for (int f = 0; f < 7; f++) out[17 * (size_t)k + 10 + f] = ph[f * stride + b];
}

// Device arena of planetary systems.
struct Pool {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t cap = 0;                  // body capacity (= field stride)
    size_t nslots = 0;               // system-slot capacity
    DevBuf body, ph;                 // 10 and 7 fields of cap doubles
    DevBuf sys, flags, energy;       // per slot
    DevBuf order, tasks, counter;    // per tick
    DevBuf idx;                      // scatter/gather indices
    DevBuf stage;                    // device side of scatter/gather data
    Pinned h_stage, h_meta;
    std::vector<WHSys> hsys;
    std::vector<Job*> slot_job;
    std::vector<int> free_slots;
    std::map<size_t, size_t> free_ranges;   // start -> length
    int max_resident = 0;
    unsigned long long ticks = 0;
    double busy_s = 0.;
};

struct Server {
    int device = 0;
    size_t budget = 0, used = 0, peak_used = 0;
    int max_nbody = 4, nbody_running = 0;
    long tick_steps = 2000;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job*> pending;
    std::unordered_map<long, std::unique_ptr<Job>> jobs;
    long next_id = 1;
    unsigned long long generation = 0;   // bumped on submit and on completion
    bool stop = false;
    std::thread sched;
    std::vector<std::thread> workers;    // N-body workers, max_nbody of them
    std::vector<hipStream_t> streams;    // N-body streams (worker w uses w % size)
    std::deque<Job*> nbody_ready;        // admitted N-body jobs waiting for a worker
    double tick_target_s = 0.002;        // wall time a pool tick aims at
    double step_s = 0.;                  // measured pool time per step
    std::chrono::steady_clock::time_point t0;
    std::vector<Job*> pool_active;       // scheduler thread only
    Pool pool;
    // N-body batch (scheduler thread only)
    hipStream_t nb_stream = nullptr;
    std::vector<Job*> nb_active;
    bool nb_dirty = true;
    int nb_ranges = 0, nb_tiles = 0;
    DevBuf nb_jobs, nb_rng, nb_til;
    Pinned nb_meta;
    double nb_step_s = 0.;               // device time per batch step (events)
    hipEvent_t nb_ev[2] = {nullptr, nullptr};
    unsigned long long nb_rounds = 0;
    std::string err;
    unsigned long long done = 0, failed = 0;
};

static double since(const Server* S) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - S->t0).count();
}

static size_t whfast_footprint(size_t n) { return n * 17 * sizeof(double) + sizeof(WHSys) + 64; }

static size_t nbody_footprint(size_t n, int precision, int gravity) {
    size_t cap = 64;
    while (cap < n) cap *= 2;
    size_t ws = reb_amd_gravity_workspace_bytes(n, precision);
    if (gravity_symmetric_path(n, precision, gravity == REB_GRAVITY_COMPENSATED))
        ws = std::max(ws, gravity_symmetric_bytes(n));
    return NFIELD * cap * sizeof(double) + ws + 16 * n + 4096;
}

// Jobs that run as single simulations on the worker threads: EXACT, and those
// whose force takes the pair-symmetric path when stepped alone (the batch
// kernels are one-sided, so a job's bits would otherwise depend on the route).
static bool nbody_single(const Job* j) {
    return j->precision == REB_AMD_PRECISION_EXACT ||
           gravity_symmetric_path(j->bodies.size(), j->precision, j->gravity == REB_GRAVITY_COMPENSATED);
}

// ------------------------------------------------------------------ pool memory
static bool pool_fail(Server* S, const char* what, hipError_t e) {
    S->err = std::string("serving pool: ") + what + ": " + hipGetErrorString(e);
    set_last_error(S->err);
    return false;
}
#define POOL_HIP(S, call)                                          \
    do {                                                           \
        hipError_t _e = (call);                                    \
        if (_e != hipSuccess) return pool_fail((S), #call, _e);    \
    } while (0)

static bool pool_grow_bodies(Server* S, size_t need) {
    Pool& P = S->pool;
    size_t cap = std::max<size_t>(P.cap ? P.cap * 2 : 4096, 64);
    while (cap < need) cap *= 2;
    DevBuf nb, np;
    if (!nb.ensure(sizeof(double) * 10 * cap) || !np.ensure(sizeof(double) * 7 * cap))
        return pool_fail(S, "out of device memory for the system pool", hipErrorOutOfMemory);
    if (P.cap) {
        for (int f = 0; f < 10; f++)
            POOL_HIP(S, hipMemcpyAsync(nb.as<double>() + f * cap, P.body.as<double>() + f * P.cap,
                                       sizeof(double) * P.cap, hipMemcpyDeviceToDevice, P.stream));
        for (int f = 0; f < 7; f++)
            POOL_HIP(S, hipMemcpyAsync(np.as<double>() + f * cap, P.ph.as<double>() + f * P.cap,
                                       sizeof(double) * P.cap, hipMemcpyDeviceToDevice, P.stream));
        POOL_HIP(S, hipStreamSynchronize(P.stream));
    }
    P.body.release();
    P.ph.release();
    P.body = nb;
    P.ph = np;
    nb.p = np.p = nullptr;
    // the new tail is free; merge with a free range ending at the old capacity
    size_t start = P.cap, len = cap - P.cap;
    auto it = P.free_ranges.lower_bound(start);
    if (it != P.free_ranges.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == start) {
            start = pv->first;
            len += pv->second;
            P.free_ranges.erase(pv);
        }
    }
    P.free_ranges[start] = len;
    P.cap = cap;
    return true;
}

static bool pool_grow_slots(Server* S, size_t need) {
    Pool& P = S->pool;
    size_t n = P.nslots ? P.nslots * 2 : 1024;
    while (n < need) n *= 2;
    DevBuf nf;
    if (!nf.ensure(sizeof(unsigned int) * n)) return pool_fail(S, "out of device memory", hipErrorOutOfMemory);
    if (P.nslots)
        POOL_HIP(S, hipMemcpyAsync(nf.p, P.flags.p, sizeof(unsigned int) * P.nslots, hipMemcpyDeviceToDevice,
                                   P.stream));
    POOL_HIP(S, hipStreamSynchronize(P.stream));
    P.flags.release();
    P.flags = nf;
    nf.p = nullptr;
    if (!P.sys.ensure(sizeof(WHSys) * n) || !P.energy.ensure(sizeof(double) * n) ||
        !P.order.ensure(sizeof(int) * n) || !P.tasks.ensure(sizeof(WHTask) * n))
        return pool_fail(S, "out of device memory", hipErrorOutOfMemory);
    WHSys empty{};
    empty.steps = 0;
    P.hsys.resize(n, empty);
    P.slot_job.resize(n, nullptr);
    for (size_t s = n; s-- > P.nslots;) P.free_slots.push_back((int)s);
    P.nslots = n;
    return true;
}

static bool alloc_range(Pool& P, size_t n, size_t* start) {
    for (auto it = P.free_ranges.begin(); it != P.free_ranges.end(); ++it) {
        if (it->second >= n) {
            *start = it->first;
            const size_t rest = it->second - n, nstart = it->first + n;
            P.free_ranges.erase(it);
            if (rest) P.free_ranges[nstart] = rest;
            return true;
        }
    }
    return false;
}

static void free_range(Pool& P, size_t start, size_t n) {
    auto it = P.free_ranges.lower_bound(start);
    if (it != P.free_ranges.end() && start + n == it->first) {
        n += it->second;
        it = P.free_ranges.erase(it);
    }
    if (it != P.free_ranges.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == start) {
            pv->second += n;
            return;
        }
    }
    P.free_ranges[start] = n;
}

static WHParams pool_params(Server* S, int ntasks, long nsteps) {
    Pool& P = S->pool;
    WHParams W;
    W.sys = P.sys.as<WHSys>();
    W.tasks = P.tasks.as<WHTask>();
    W.order = P.order.as<int>();
    W.flags = P.flags.as<unsigned int>();
    W.counter = P.counter.as<unsigned int>();
    W.ntasks = ntasks;
    W.grid = std::max(1, std::min(ntasks, P.max_resident));
    W.body = P.body.as<double>();
    W.ph = P.ph.as<double>();
    W.stride = P.cap;
    W.G = 1.;   // per system: WHSys::G
    W.soft2 = 0.;
    W.compensated = 0;
    W.nsteps = nsteps;
    W.safe_mode = 0;
    W.did_modify = 0;
    W.sync_at_end = 0;
    return W;
}

// Places newly admitted planetary systems into the pool.
static bool pool_admit(Server* S, const std::vector<Job*>& fresh) {
    Pool& P = S->pool;
    size_t nb = 0;
    for (Job* j : fresh) {
        const size_t n = j->bodies.size();
        size_t start;
        if (!alloc_range(P, n, &start)) {
            size_t need = P.cap + n;
            if (!pool_grow_bodies(S, need) || !alloc_range(P, n, &start)) return false;
        }
        if (P.free_slots.empty() && !pool_grow_slots(S, P.nslots + 1)) return false;
        const int slot = P.free_slots.back();
        P.free_slots.pop_back();
        j->slot = slot;
        j->body0 = start;
        j->remaining = j->n_steps;
        j->fresh = true;
        WHSys& w = P.hsys[slot];
        w.body0 = (int)start;
        w.N = (int)n;
        w.Nact = (int)n;
        w.Nact_grav = (int)n;
        w.tp_type = 0;
        w.steps = 0;
        w.req = 0;
        w.pad = 0;
        w.dt = j->dt;
        w.G = j->G;
        P.slot_job[slot] = j;
        nb += n;
    }
    // stage bodies (10 doubles each) and destinations, then scatter on the device
    const size_t data_bytes = sizeof(double) * 10 * nb;
    const size_t idx_bytes = sizeof(int) * (nb + fresh.size());
    if (!P.h_stage.ensure(data_bytes + idx_bytes + 256) || !P.stage.ensure(data_bytes + 256) ||
        !P.idx.ensure(idx_bytes + 256))
        return pool_fail(S, "staging", hipErrorOutOfMemory);
    double* hd = static_cast<double*>(P.h_stage.p);
    int* hi = reinterpret_cast<int*>(static_cast<char*>(P.h_stage.p) + data_bytes);
    size_t k = 0;
    for (Job* j : fresh) {
        for (size_t b = 0; b < j->bodies.size(); b++, k++) {
            const struct reb_particle& p = j->bodies[b];
            const double f[10] = {p.x, p.y, p.z, p.vx, p.vy, p.vz, p.ax, p.ay, p.az, p.m};
            memcpy(hd + 10 * k, f, sizeof(f));
            hi[k] = (int)(j->body0 + b);
        }
    }
    for (size_t q = 0; q < fresh.size(); q++) hi[nb + q] = fresh[q]->slot;
    POOL_HIP(S, hipMemcpyAsync(P.stage.p, hd, data_bytes, hipMemcpyHostToDevice, P.stream));
    POOL_HIP(S, hipMemcpyAsync(P.idx.p, hi, idx_bytes, hipMemcpyHostToDevice, P.stream));
    hipLaunchKernelGGL(k_pool_scatter, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, P.stream,
                       P.stage.as<double>(), P.idx.as<int>(), (int)nb, P.body.as<double>(), P.ph.as<double>(), P.cap);
    hipLaunchKernelGGL(k_pool_set_flags, dim3((unsigned)((fresh.size() + 255) / 256)), dim3(256), 0, P.stream,
                       P.idx.as<int>() + nb, (int)fresh.size(), P.flags.as<unsigned int>(), 1u);
    POOL_HIP(S, hipGetLastError());
    // systems resumed from a simulation that carries its heliocentric state
    // (integrator_whfast.c:71-99, p_jh) continue from it, synchronized or not
    std::vector<unsigned int> sync_flag(fresh.size());
    for (size_t q = 0; q < fresh.size(); q++) {
        const Job* j = fresh[q];
        if (j->pjh.empty()) continue;
        const size_t n = j->bodies.size();
        for (int f = 0; f < 7; f++)
            POOL_HIP(S, hipMemcpyAsync(P.ph.as<double>() + f * P.cap + j->body0, j->pjh.data() + f * n,
                                       sizeof(double) * n, hipMemcpyHostToDevice, P.stream));
        sync_flag[q] = j->synchronized ? 1u : 0u;
        POOL_HIP(S, hipMemcpyAsync(P.flags.as<unsigned int>() + j->slot, &sync_flag[q], sizeof(unsigned int),
                                   hipMemcpyHostToDevice, P.stream));
    }
    // initial energies (tools.c:108-162) of the new systems
    POOL_HIP(S, hipMemcpyAsync(P.sys.p, P.hsys.data(), sizeof(WHSys) * P.nslots, hipMemcpyHostToDevice, P.stream));
    WHParams W = pool_params(S, 0, 0);
    if (whfast_energy_launch(W, P.energy.as<double>(), (int)P.nslots, P.stream))
        return pool_fail(S, "energy launch", hipGetLastError());
    if (!P.h_meta.ensure(sizeof(double) * P.nslots)) return pool_fail(S, "staging", hipErrorOutOfMemory);
    POOL_HIP(S, hipMemcpyAsync(P.h_meta.p, P.energy.p, sizeof(double) * P.nslots, hipMemcpyDeviceToHost, P.stream));
    POOL_HIP(S, hipStreamSynchronize(P.stream));
    for (Job* j : fresh) j->e0 = static_cast<double*>(P.h_meta.p)[j->slot];
    return true;
}

// ------------------------------------------------------------------ checkpoints
// One Simulationarchive snapshot (simulationarchive.c:411-615) of a job after
// `steps` of its steps: the bodies, time, step count and the integrator with
// its settings.  `pjh` (7 rows of n) is WHFast's heliocentric state
// (integrator_whfast.c:71-99, p_jh), written with the synchronization flag so
// that a simulation read from the file continues from it.  A submission that
// has not stepped is written as freshly added particles.
static bool write_snapshot(const Job& j, const struct reb_particle* bodies, const double* pjh, int synchronized,
                           uint64_t steps, const std::string& file, std::string* err) {
    const size_t n = j.bodies.size();
    struct reb_simulation* r = reb_simulation_create();
    if (!r) {
        *err = "could not create a simulation";
        return false;
    }
    r->save_messages = 1;
    r->G = j.G;
    r->dt = j.dt;
    r->softening = j.softening;
    if (j.kind == JOB_WHFAST) {
        reb_simulation_set_integrator(r, "whfast");
        auto* s = (struct reb_integrator_whfast_state*)r->integrator.state;
        s->safe_mode = (unsigned int)j.safe_mode;
        s->coordinates = REB_INTEGRATOR_WHFAST_COORDINATES_DEMOCRATICHELIOCENTRIC;
    } else {
        reb_simulation_set_integrator(r, "leapfrog");
        r->gravity = static_cast<decltype(r->gravity)>(j.gravity);
        reb_amd_set_precision(r, j.precision);
    }
    for (size_t i = 0; i < n; i++) reb_simulation_add(r, bodies[i]);
    // the time as reb_simulation_steps sums it: two half steps per step
    // (integrator_leapfrog.c:81, integrator_whfast.c:1193-1257)
    double t = j.t0;
    for (uint64_t k = 0; k < steps; k++) {
        t += j.dt / 2.;
        t += j.dt / 2.;
    }
    r->t = t;
    r->steps_done = j.steps0 + steps;
    r->did_modify_particles = (steps == 0 && !j.resumed) ? 1 : 0;
    r->is_synchronized = synchronized ? 1 : 0;
    if (pjh && j.kind == JOB_WHFAST) {
        auto* s = (struct reb_integrator_whfast_state*)r->integrator.state;
        s->p_jh = (struct reb_particle*)calloc(n, sizeof(struct reb_particle));
        if (!s->p_jh) {
            reb_simulation_free(r);
            *err = "out of host memory";
            return false;
        }
        s->N_allocated = n;
        for (size_t i = 0; i < n; i++) {
            struct reb_particle& q = s->p_jh[i];
            q.x = pjh[0 * n + i]; q.y = pjh[1 * n + i]; q.z = pjh[2 * n + i];
            q.vx = pjh[3 * n + i]; q.vy = pjh[4 * n + i]; q.vz = pjh[5 * n + i];
            q.m = pjh[6 * n + i];
        }
    }
    reb_simulation_save_to_file(r, file.c_str());
    bool ok = true;
    char buf[REB_STRING_SIZE_MAX];
    while (reb_pop_message(r->messages, buf)) {
        if (buf[0] == 'e' && ok) {
            ok = false;
            *err = buf + 1;
        }
    }
    reb_simulation_free(r);
    return ok;
}

// The state a job holds on the host: as submitted (nothing stepped) or final.
static bool write_host_state(const Job& j, bool final, const std::string& file, std::string* err) {
    return write_snapshot(j, j.bodies.data(), j.pjh.empty() ? nullptr : j.pjh.data(), j.synchronized,
                          final ? j.n_steps : 0, file, err);
}

// Takes a job's outstanding checkpoint requests (Server::mu held by nobody).
static std::vector<std::shared_ptr<Checkpoint>> take_requests(Server* S, Job* j) {
    std::vector<std::shared_ptr<Checkpoint>> req;
    std::lock_guard<std::mutex> g(S->mu);
    req.swap(j->ckpt);
    return req;
}

// Writes one snapshot per request with `write(file, err)` and wakes the callers.
template <class W>
static void serve_requests(Server* S, std::vector<std::shared_ptr<Checkpoint>>& req, W write) {
    if (req.empty()) return;
    for (auto& c : req) {
        std::string e;
        c->rc = write(c->file, &e) ? 0 : 1;
        c->error = e;
    }
    std::lock_guard<std::mutex> g(S->mu);
    for (auto& c : req) c->done = true;
    S->cv_done.notify_all();
}

static bool bodies_finite(const std::vector<struct reb_particle>& b) {
    for (const auto& p : b)
        if (!std::isfinite(p.x) || !std::isfinite(p.y) || !std::isfinite(p.z) || !std::isfinite(p.vx) ||
            !std::isfinite(p.vy) || !std::isfinite(p.vz) || !std::isfinite(p.m))
            return false;
    return true;
}

// A job leaves the server: done, or failed with a message.  A job whose
// bodies are no longer finite fails; the other jobs are not affected.
// Checkpoint requests made while it ran are served from its final state.
static void finish_job(Server* S, Job* j, bool ok, const std::string& msg) {
    std::string why = msg;
    if (ok && !bodies_finite(j->bodies)) {
        ok = false;
        why = "non-finite coordinates after " + std::to_string(j->n_steps) + " steps";
    }
    for (;;) {
        std::vector<std::shared_ptr<Checkpoint>> req;
        {
            std::lock_guard<std::mutex> g(S->mu);
            req.swap(j->ckpt);
            if (req.empty()) {
                j->t_done = since(S);
                j->state = ok ? REB_AMD_JOB_DONE : REB_AMD_JOB_FAILED;
                if (!ok) j->error = why;
                S->used -= j->footprint;
                if (ok)
                    S->done++;
                else
                    S->failed++;
                if (j->kind == JOB_NBODY) S->nbody_running--;
                S->generation++;
                S->cv_done.notify_all();
                S->cv_work.notify_all();
                return;
            }
        }
        serve_requests(S, req, [&](const std::string& file, std::string* e) {
            if (!ok) {
                *e = "job failed: " + why;
                return false;
            }
            return write_host_state(*j, true, file, e);
        });
    }
}

// Checkpoints of the systems in the pool, between ticks: the bodies and the
// heliocentric state as the next tick would continue from them.
static void pool_checkpoints(Server* S) {
    Pool& P = S->pool;
    for (Job* j : S->pool_active) {
        auto req = take_requests(S, j);
        if (req.empty()) continue;
        if (j->fresh) {   // admitted, not yet stepped
            serve_requests(S, req, [&](const std::string& f, std::string* e) { return write_host_state(*j, false, f, e); });
            continue;
        }
        const size_t n = j->bodies.size();
        std::vector<double> rows(17 * n);
        unsigned int flag = 0;
        bool got = true;
        for (int f = 0; f < 10 && got; f++)
            got = hipMemcpyAsync(rows.data() + f * n, P.body.as<double>() + f * P.cap + j->body0, sizeof(double) * n,
                                 hipMemcpyDeviceToHost, P.stream) == hipSuccess;
        for (int f = 0; f < 7 && got; f++)
            got = hipMemcpyAsync(rows.data() + (10 + f) * n, P.ph.as<double>() + f * P.cap + j->body0,
                                 sizeof(double) * n, hipMemcpyDeviceToHost, P.stream) == hipSuccess;
        if (got)
            got = hipMemcpyAsync(&flag, P.flags.as<unsigned int>() + j->slot, sizeof(flag), hipMemcpyDeviceToHost,
                                 P.stream) == hipSuccess &&
                  hipStreamSynchronize(P.stream) == hipSuccess;
        std::vector<struct reb_particle> cur = j->bodies;
        for (size_t b = 0; b < n; b++) {
            double* f[10] = {&cur[b].x, &cur[b].y, &cur[b].z, &cur[b].vx, &cur[b].vy,
                             &cur[b].vz, &cur[b].ax, &cur[b].ay, &cur[b].az, &cur[b].m};
            for (int q = 0; q < 10; q++) *f[q] = rows[q * n + b];
        }
        const uint64_t steps = j->n_steps - j->remaining;
        serve_requests(S, req, [&](const std::string& f, std::string* e) {
            if (!got) {
                *e = "device copy failed";
                return false;
            }
            return write_snapshot(*j, cur.data(), rows.data() + 10 * n, (int)(flag & 1u), steps, f, e);
        });
    }
}

// One tick of the planetary pool: every active system advances by
// min(remaining, tick_steps) steps; finished systems are collected.
static bool pool_tick(Server* S) {
    Pool& P = S->pool;
    auto& act = S->pool_active;
    if (act.empty()) return true;
    // Ticks aim at ~2 ms of device time, so jobs join and leave promptly.
    long tick_max = S->tick_steps;
    if (S->step_s > 0.) tick_max = std::max(16L, std::min(tick_max, (long)(S->tick_target_s / S->step_s)));
    long tick = 0;
    for (Job* j : act) {
        WHSys& w = P.hsys[j->slot];
        const uint64_t k = std::min<uint64_t>(j->remaining, (uint64_t)tick_max);
        w.steps = (int)k;
        w.req = (k == j->remaining ? 1 : 0) | (j->fresh && j->pjh.empty() ? 2 : 0) | (j->safe_mode ? 4 : 0);
        tick = std::max<long>(tick, (long)k);
    }
    // tasks over the active systems, bucketed by planet count (largest first)
    std::vector<int> order(act.size());
    for (size_t q = 0; q < act.size(); q++) order[q] = act[q]->slot;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return P.hsys[a].N > P.hsys[b].N; });
    std::vector<WHTask> tasks;
    for (size_t k = 0; k < order.size();) {
        const int Gs = whfast_group_size(P.hsys[order[k]].N - 1);
        const int per = 64 / Gs;
        size_t m = 0;
        while (k + m < order.size() && (int)m < per && P.hsys[order[k + m]].N == P.hsys[order[k]].N) m++;
        tasks.push_back(WHTask{Gs, (int)k, (int)m});
        k += m;
    }
    const size_t sys_bytes = sizeof(WHSys) * P.nslots, ord_bytes = sizeof(int) * order.size(),
                 task_bytes = sizeof(WHTask) * tasks.size();
    if (!P.h_meta.ensure(sys_bytes + ord_bytes + task_bytes + 64)) return pool_fail(S, "staging", hipErrorOutOfMemory);
    char* hm = static_cast<char*>(P.h_meta.p);
    memcpy(hm, P.hsys.data(), sys_bytes);
    memcpy(hm + sys_bytes, order.data(), ord_bytes);
    memcpy(hm + sys_bytes + ord_bytes, tasks.data(), task_bytes);
    POOL_HIP(S, hipMemcpyAsync(P.sys.p, hm, sys_bytes, hipMemcpyHostToDevice, P.stream));
    POOL_HIP(S, hipMemcpyAsync(P.order.p, hm + sys_bytes, ord_bytes, hipMemcpyHostToDevice, P.stream));
    POOL_HIP(S, hipMemcpyAsync(P.tasks.p, hm + sys_bytes + ord_bytes, task_bytes, hipMemcpyHostToDevice, P.stream));
    const double t_start = since(S);
    WHParams W = pool_params(S, (int)tasks.size(), tick);
    if (whfast_launch(W, P.stream)) return pool_fail(S, "tick launch", hipGetLastError());
    // collect the systems that reached their horizon
    std::vector<Job*> fin, keep;
    for (Job* j : act) {
        j->remaining -= (uint64_t)P.hsys[j->slot].steps;
        j->fresh = false;
        (j->remaining == 0 ? fin : keep).push_back(j);
    }
    if (!fin.empty()) {
        size_t nb = 0;
        for (Job* j : fin) nb += j->bodies.size();
        const size_t out_bytes = sizeof(double) * 17 * nb;
        const size_t idx_bytes = sizeof(int) * nb;
        if (!P.h_stage.ensure(out_bytes + idx_bytes + sizeof(double) * P.nslots + 256) ||
            !P.stage.ensure(out_bytes + 256) || !P.idx.ensure(idx_bytes + 256))
            return pool_fail(S, "staging", hipErrorOutOfMemory);
        int* hi = reinterpret_cast<int*>(static_cast<char*>(P.h_stage.p) + out_bytes);
        size_t k = 0;
        for (Job* j : fin)
            for (size_t b = 0; b < j->bodies.size(); b++) hi[k++] = (int)(j->body0 + b);
        POOL_HIP(S, hipMemcpyAsync(P.idx.p, hi, idx_bytes, hipMemcpyHostToDevice, P.stream));
        hipLaunchKernelGGL(k_pool_gather, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, P.stream,
                           P.idx.as<int>(), (int)nb, P.body.as<double>(), P.ph.as<double>(), P.cap,
                           P.stage.as<double>());
        if (whfast_energy_launch(W, P.energy.as<double>(), (int)P.nslots, P.stream))
            return pool_fail(S, "energy launch", hipGetLastError());
        double* he = reinterpret_cast<double*>(static_cast<char*>(P.h_stage.p) + out_bytes + idx_bytes + 64);
        POOL_HIP(S, hipMemcpyAsync(P.h_stage.p, P.stage.p, out_bytes, hipMemcpyDeviceToHost, P.stream));
        POOL_HIP(S, hipMemcpyAsync(he, P.energy.p, sizeof(double) * P.nslots, hipMemcpyDeviceToHost, P.stream));
        POOL_HIP(S, hipStreamSynchronize(P.stream));
        const double* hd = static_cast<const double*>(P.h_stage.p);
        k = 0;
        for (Job* j : fin) {
            const size_t n = j->bodies.size();
            j->pjh.assign(7 * n, 0.);
            j->synchronized = 1;
            for (size_t b = 0; b < n; b++) {
                struct reb_particle& p = j->bodies[b];
                const double* f = hd + 17 * k++;
                p.x = f[0]; p.y = f[1]; p.z = f[2];
                p.vx = f[3]; p.vy = f[4]; p.vz = f[5];
                p.ax = f[6]; p.ay = f[7]; p.az = f[8];
                p.m = f[9];
                for (int q = 0; q < 7; q++) j->pjh[q * n + b] = f[10 + q];
            }
            j->e1 = he[j->slot];
            WHSys& w = P.hsys[j->slot];
            w = WHSys{};
            P.slot_job[j->slot] = nullptr;
            P.free_slots.push_back(j->slot);
            free_range(P, j->body0, j->bodies.size());
        }
    } else {
        POOL_HIP(S, hipStreamSynchronize(P.stream));
    }
    const double took = since(S) - t_start;
    P.busy_s += took;
    P.ticks++;
    if (tick >= 16) S->step_s = S->step_s > 0. ? 0.7 * S->step_s + 0.3 * took / (double)tick : took / (double)tick;
    act.swap(keep);
    for (Job* j : fin) finish_job(S, j, true, "");
    return true;
}

static void run_nbody(Server* S, Job* j, hipStream_t stream) {
    (void)hipSetDevice(S->device);
    struct reb_simulation* r = reb_simulation_create();
    DeviceState* d = dev_of(r);
    d->device = S->device;
    d->stream = stream;        // the worker's stream, kept by the server
    d->owns_stream = false;
    r->save_messages = 1;
    r->G = j->G;
    r->softening = j->softening;
    r->dt = j->dt;
    reb_simulation_set_integrator(r, "leapfrog");
    r->gravity = static_cast<decltype(r->gravity)>(j->gravity);
    for (const auto& p : j->bodies) reb_simulation_add(r, p);
    reb_amd_set_precision(r, j->precision);
    reb_amd_set_resident(r, 1);
    r->t = j->t0;
    r->steps_done = j->steps0;
    j->e0 = reb_simulation_energy(r);
    std::string msg;
    bool ok = true;
    auto drain = [&]() {   // the first error the simulation reported fails the job
        char buf[REB_STRING_SIZE_MAX];
        while (reb_pop_message(r->messages, buf)) {
            if (buf[0] == 'e' && ok) {
                ok = false;
                msg = buf + 1;
            }
        }
    };
    // Stepped in segments of ~10 ms or more, between which checkpoint requests
    // are served from the simulation itself (the steps are those of one call).
    uint64_t left = j->n_steps, seg = 16;
    while (left > 0 && ok) {
        const uint64_t k = std::min(left, seg);
        const auto t_seg = std::chrono::steady_clock::now();
        reb_simulation_steps(r, (size_t)k);
        left -= k;
        drain();
        const double took = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_seg).count();
        if (took < 0.01) seg = std::min<uint64_t>(seg * 2, (uint64_t)1 << 20);
        if (left == 0 || !ok) break;
        auto req = take_requests(S, j);
        serve_requests(S, req, [&](const std::string& f, std::string* e) {
            reb_simulation_save_to_file(r, f.c_str());
            char buf[REB_STRING_SIZE_MAX];
            bool fine = true;
            while (reb_pop_message(r->messages, buf)) {
                if (buf[0] == 'e' && fine) {
                    fine = false;
                    *e = buf + 1;
                }
            }
            return fine;
        });
    }
    reb_amd_set_resident(r, 0);
    j->e1 = reb_simulation_energy(r);
    drain();
    if (ok)
        for (size_t i = 0; i < j->bodies.size() && i < r->N; i++) j->bodies[i] = r->particles[i];
    reb_simulation_free(r);
    finish_job(S, j, ok, msg);
}

// ------------------------------------------------------------------ N-body batch
static bool nb_fail(Server* S, const char* what) {
    S->err = std::string("N-body batch: ") + what;
    set_last_error(S->err);
    return false;
}

static SegJob seg_of(const Job* j) {
    SegJob J;
    const size_t n = j->bodies.size();
    double* b = j->soa.as<double>();
    J.x = b; J.y = b + n; J.z = b + 2 * n;
    J.vx = b + 3 * n; J.vy = b + 4 * n; J.vz = b + 5 * n;
    J.ax = b + 6 * n; J.ay = b + 7 * n; J.az = b + 8 * n;
    J.m = b + 9 * n;
    J.partial = j->partial.as<double>();
    J.n = (int)n;
    J.chunk = j->chunk;
    J.nchunk = j->used;
    J.mode = seg_fast_mode(j->precision, j->gravity == REB_GRAVITY_COMPENSATED ? 1 : 0);
    J.negG = -j->G;
    J.soft2 = j->softening * j->softening;
    J.kick = j->dt;
    J.drift = j->dt * 0.5;
    return J;
}

// Energy (tools.c:108-162) of a batched job; waits for the stream.
static bool nb_energy(Server* S, Job* j, double* e) {
    const SegJob J = seg_of(j);
    const int n = J.n;
    if (energy_launch(J.x, J.y, J.z, J.vx, J.vy, J.vz, J.m, n, n, n, j->G, false, j->scratch.as<double>(),
                      S->nb_stream))
        return nb_fail(S, "energy launch");
    double h[2];
    if (hipMemcpyAsync(h, j->scratch.p, sizeof(h), hipMemcpyDeviceToHost, S->nb_stream) != hipSuccess ||
        hipStreamSynchronize(S->nb_stream) != hipSuccess)
        return nb_fail(S, "energy copy");
    *e = h[0] + h[1];
    return true;
}

static bool nb_admit(Server* S, Job* j) {
    const size_t n = j->bodies.size();
    size_t ws = 0;
    seg_plan((int)n, &j->chunk, &j->used, &ws);
    if (!j->soa.ensure(sizeof(double) * 10 * n) || !j->partial.ensure(ws) ||
        !j->scratch.ensure(sizeof(double) * (2 * n + 16)))
        return nb_fail(S, "out of device memory");
    if (!S->nb_meta.ensure(sizeof(double) * 10 * n + 64)) return nb_fail(S, "staging");
    double* h = static_cast<double*>(S->nb_meta.p);
    for (size_t i = 0; i < n; i++) {
        const struct reb_particle& p = j->bodies[i];
        const double f[10] = {p.x, p.y, p.z, p.vx, p.vy, p.vz, p.ax, p.ay, p.az, p.m};
        for (int k = 0; k < 10; k++) h[k * n + i] = f[k];
    }
    if (hipMemcpyAsync(j->soa.p, h, sizeof(double) * 10 * n, hipMemcpyHostToDevice, S->nb_stream) != hipSuccess)
        return nb_fail(S, "upload");
    if (!nb_energy(S, j, &j->e0)) return false;
    j->remaining = j->n_steps;
    S->nb_dirty = true;
    return true;
}

// Uploads the job, range and tile tables of the active batch.
static bool nb_tables(Server* S) {
    std::vector<SegJob> jobs;
    std::vector<SegRange> rng;
    std::vector<SegTile> til;
    for (size_t q = 0; q < S->nb_active.size(); q++) {
        const Job* j = S->nb_active[q];
        jobs.push_back(seg_of(j));
        const int n = (int)j->bodies.size();
        for (int f = 0; f < n; f += 256) rng.push_back(SegRange{(int)q, f});
        const int bxs = seg_tiles_of(n, 1);
        for (int by = 0; by < j->used; by++)
            for (int bx = 0; bx < bxs; bx++) til.push_back(SegTile{(int)q, bx, by});
    }
    const size_t bj = sizeof(SegJob) * jobs.size(), br = sizeof(SegRange) * rng.size(),
                 bt = sizeof(SegTile) * til.size();
    // the tables go through pinned memory and replace tables in use: wait for
    // the stream first
    if (hipStreamSynchronize(S->nb_stream) != hipSuccess) return nb_fail(S, "stream");
    if (!S->nb_jobs.ensure(bj) || !S->nb_rng.ensure(br) || !S->nb_til.ensure(bt) ||
        !S->nb_meta.ensure(bj + br + bt + 64))
        return nb_fail(S, "out of device memory for the batch tables");
    char* h = static_cast<char*>(S->nb_meta.p);
    memcpy(h, jobs.data(), bj);
    memcpy(h + bj, rng.data(), br);
    memcpy(h + bj + br, til.data(), bt);
    if (hipMemcpyAsync(S->nb_jobs.p, h, bj, hipMemcpyHostToDevice, S->nb_stream) != hipSuccess ||
        hipMemcpyAsync(S->nb_rng.p, h + bj, br, hipMemcpyHostToDevice, S->nb_stream) != hipSuccess ||
        hipMemcpyAsync(S->nb_til.p, h + bj + br, bt, hipMemcpyHostToDevice, S->nb_stream) != hipSuccess)
        return nb_fail(S, "table upload");
    S->nb_ranges = (int)rng.size();
    S->nb_tiles = (int)til.size();
    S->nb_dirty = false;
    return true;
}

// Queues one round of the batch: every job advances by the same number of
// steps (no job past its horizon); returns the step count.
static long nb_launch_round(Server* S) {
    if (S->nb_active.empty()) return 0;
    if (S->nb_dirty && !nb_tables(S)) return -1;
    uint64_t k = std::numeric_limits<uint64_t>::max();
    for (Job* j : S->nb_active) k = std::min(k, j->remaining);
    long target = S->nb_step_s > 0. ? (long)(S->tick_target_s / S->nb_step_s) : 8;
    target = std::max(1L, std::min(target, 4096L));
    const long steps = (long)std::min<uint64_t>(k, (uint64_t)target);
    if (steps > 0 && S->nb_ev[0]) (void)hipEventRecord(S->nb_ev[0], S->nb_stream);
    for (long q = 0; q < steps; q++)
        if (seg_step_launch(S->nb_jobs.as<SegJob>(), S->nb_rng.as<SegRange>(), S->nb_ranges,
                            S->nb_til.as<SegTile>(), S->nb_tiles, S->nb_stream))
            return -1;
    if (steps > 0 && S->nb_ev[1]) (void)hipEventRecord(S->nb_ev[1], S->nb_stream);
    return steps;
}

// After the round's kernels: collects the jobs that reached their horizon.
static bool nb_collect(Server* S, long steps) {
    if (hipStreamSynchronize(S->nb_stream) != hipSuccess) return nb_fail(S, "stream");
    float ms = 0.f;
    if (steps > 0 && S->nb_ev[0] && hipEventElapsedTime(&ms, S->nb_ev[0], S->nb_ev[1]) == hipSuccess) {
        const double per = 1e-3 * ms / (double)steps;
        S->nb_step_s = S->nb_step_s > 0. ? 0.7 * S->nb_step_s + 0.3 * per : per;
        S->nb_rounds++;
    }
    std::vector<Job*> keep, fin;
    for (Job* j : S->nb_active) {
        j->remaining -= (uint64_t)steps;
        (j->remaining == 0 ? fin : keep).push_back(j);
    }
    for (Job* j : fin) {
        const size_t n = j->bodies.size();
        if (!nb_energy(S, j, &j->e1)) return false;
        if (!S->nb_meta.ensure(sizeof(double) * 10 * n + 64)) return nb_fail(S, "staging");
        double* h = static_cast<double*>(S->nb_meta.p);
        if (hipMemcpyAsync(h, j->soa.p, sizeof(double) * 10 * n, hipMemcpyDeviceToHost, S->nb_stream) != hipSuccess ||
            hipStreamSynchronize(S->nb_stream) != hipSuccess)
            return nb_fail(S, "download");
        for (size_t i = 0; i < n; i++) {
            struct reb_particle& p = j->bodies[i];
            double* f[10] = {&p.x, &p.y, &p.z, &p.vx, &p.vy, &p.vz, &p.ax, &p.ay, &p.az, &p.m};
            for (int k = 0; k < 10; k++) *f[k] = h[k * n + i];
        }
        j->soa.release();
        j->partial.release();
        j->scratch.release();
    }
    if (!fin.empty()) S->nb_dirty = true;
    S->nb_active.swap(keep);
    for (Job* j : fin) finish_job(S, j, true, "");
    return true;
}

// Checkpoints of the batched N-body jobs between rounds.
static void nb_checkpoints(Server* S) {
    for (Job* j : S->nb_active) {
        auto req = take_requests(S, j);
        if (req.empty()) continue;
        const size_t n = j->bodies.size();
        std::vector<double> h(10 * n);
        const bool got = hipMemcpyAsync(h.data(), j->soa.p, sizeof(double) * 10 * n, hipMemcpyDeviceToHost,
                                        S->nb_stream) == hipSuccess &&
                         hipStreamSynchronize(S->nb_stream) == hipSuccess;
        std::vector<struct reb_particle> cur = j->bodies;
        for (size_t i = 0; i < n; i++) {
            double* f[10] = {&cur[i].x, &cur[i].y, &cur[i].z, &cur[i].vx, &cur[i].vy,
                             &cur[i].vz, &cur[i].ax, &cur[i].ay, &cur[i].az, &cur[i].m};
            for (int k = 0; k < 10; k++) *f[k] = h[k * n + i];
        }
        const uint64_t steps = j->n_steps - j->remaining;
        serve_requests(S, req, [&](const std::string& f, std::string* e) {
            if (!got) {
                *e = "device copy failed";
                return false;
            }
            return write_snapshot(*j, cur.data(), nullptr, 1, steps, f, e);
        });
    }
}

static void nbody_worker(Server* S, hipStream_t stream) {
    (void)hipSetDevice(S->device);
    for (;;) {
        Job* j = nullptr;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] { return S->stop || !S->nbody_ready.empty(); });
            if (S->nbody_ready.empty()) break;   // stop requested and nothing left
            j = S->nbody_ready.front();
            S->nbody_ready.pop_front();
        }
        run_nbody(S, j, stream);
    }
}

static void scheduler(Server* S) {
    (void)hipSetDevice(S->device);
    unsigned long long seen = ~0ull;
    for (;;) {
        std::vector<Job*> fresh, nb_fresh;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            if (S->pool_active.empty() && S->nb_active.empty())
                S->cv_work.wait(lk, [&] { return S->stop || S->generation != seen; });
            if (S->stop) break;
            seen = S->generation;
            // admission: arrival order, skipping jobs that do not fit yet
            for (auto it = S->pending.begin(); it != S->pending.end();) {
                Job* j = *it;
                const bool fits = S->used + j->footprint <= S->budget;
                const bool slot_ok = j->kind != JOB_NBODY || S->nbody_running < S->max_nbody;
                if (!fits || !slot_ok) {
                    ++it;
                    continue;
                }
                it = S->pending.erase(it);
                S->used += j->footprint;
                S->peak_used = std::max(S->peak_used, S->used);
                j->state = REB_AMD_JOB_RUNNING;
                j->t_admit = since(S);
                if (j->kind == JOB_NBODY) {
                    S->nbody_running++;
                    if (nbody_single(j)) {
                        S->nbody_ready.push_back(j);
                        S->cv_work.notify_all();
                    } else {
                        nb_fresh.push_back(j);
                    }
                } else {
                    fresh.push_back(j);
                }
            }
        }
        for (Job* j : nb_fresh) {
            if (nb_admit(S, j)) {
                S->nb_active.push_back(j);
            } else {
                j->soa.release();
                j->partial.release();
                j->scratch.release();
                finish_job(S, j, false, S->err);
            }
        }
        if (!fresh.empty()) {
            if (pool_admit(S, fresh)) {
                for (Job* j : fresh) S->pool_active.push_back(j);
            } else {
                for (Job* j : fresh) finish_job(S, j, false, S->err);
            }
        }
        // the N-body round is queued first so that it runs beside the pool tick
        const long nb_steps = nb_launch_round(S);
        if (!S->pool_active.empty() && !pool_tick(S)) {
            std::vector<Job*> act;
            act.swap(S->pool_active);
            for (Job* j : act) finish_job(S, j, false, S->err);
        }
        pool_checkpoints(S);
        if (nb_steps < 0 || (!S->nb_active.empty() && !nb_collect(S, nb_steps))) {
            std::vector<Job*> act;
            act.swap(S->nb_active);
            (void)hipStreamSynchronize(S->nb_stream);
            for (Job* j : act) {
                j->soa.release();
                j->partial.release();
                j->scratch.release();
                finish_job(S, j, false, S->err);
            }
            S->nb_dirty = true;
        }
        nb_checkpoints(S);
    }
}

}  // namespace rbamd

using namespace rbamd;

struct reb_amd_server {
    Server s;
};

extern "C" {

REB_API struct reb_amd_server* reb_amd_server_create(size_t memory_budget_bytes, int max_nbody_jobs,
                                                     long whfast_tick_steps) {
    if (reb_amd_device_count() <= 0) {
        set_last_error("reb_amd_server_create: no HIP device is available");
        return nullptr;
    }
    auto* h = new reb_amd_server();
    Server* S = &h->s;
    (void)hipGetDevice(&S->device);
    S->budget = memory_budget_bytes ? memory_budget_bytes : ((size_t)8 << 30);
    S->max_nbody = max_nbody_jobs > 0 ? max_nbody_jobs : 32;
    S->tick_steps = whfast_tick_steps > 0 ? whfast_tick_steps : 2000;
    S->t0 = std::chrono::steady_clock::now();
    S->pool.device = S->device;
    S->pool.max_resident = whfast_max_resident(S->device);
    if (hipStreamCreateWithFlags(&S->pool.stream, hipStreamNonBlocking) != hipSuccess ||
        !S->pool.counter.ensure(256) || hipMemsetAsync(S->pool.counter.p, 0, 256, S->pool.stream) != hipSuccess ||
        hipStreamSynchronize(S->pool.stream) != hipSuccess) {
        set_last_error("reb_amd_server_create: could not set up the device pool");
        delete h;
        return nullptr;
    }
    S->sched = std::thread(scheduler, S);
    // Three streams (planetary pool, N-body batch, single N-body jobs): within
    // the four hardware queues a process gets by default, so no kernel waits
    // behind another stream's launch.
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&S->nb_stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&S->nb_ev[0]) != hipSuccess || hipEventCreate(&S->nb_ev[1]) != hipSuccess) {
        set_last_error("reb_amd_server_create: could not create a stream");
    } else {
        S->streams.push_back(st);
        for (int w = 0; w < 2; w++) S->workers.emplace_back(nbody_worker, S, st);
    }
    return h;
}

// Cancels pending jobs, lets running jobs finish, releases everything.
REB_API void reb_amd_server_free(struct reb_amd_server* h) {
    if (!h) return;
    Server* S = &h->s;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        for (Job* j : S->pending) {
            j->state = REB_AMD_JOB_FAILED;
            j->error = "server shut down";
        }
        S->pending.clear();
    }
    // drain the pool and the workers
    for (;;) {
        std::unique_lock<std::mutex> lk(S->mu);
        bool busy = S->nbody_running > 0;
        for (auto& kv : S->jobs)
            if (kv.second->state == REB_AMD_JOB_RUNNING) busy = true;
        if (!busy) break;
        S->cv_done.wait_for(lk, std::chrono::milliseconds(50));
    }
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
        S->cv_work.notify_all();
    }
    S->sched.join();
    for (auto& t : S->workers)
        if (t.joinable()) t.join();
    Pool& P = S->pool;
    (void)hipSetDevice(S->device);
    (void)hipStreamSynchronize(P.stream);
    for (DevBuf* b : {&P.body, &P.ph, &P.sys, &P.flags, &P.energy, &P.order, &P.tasks, &P.counter, &P.idx, &P.stage})
        b->release();
    P.h_stage.release();
    P.h_meta.release();
    (void)hipStreamDestroy(P.stream);
    if (S->nb_stream) {
        (void)hipStreamSynchronize(S->nb_stream);
        (void)hipStreamDestroy(S->nb_stream);
    }
    for (DevBuf* b : {&S->nb_jobs, &S->nb_rng, &S->nb_til}) b->release();
    for (hipEvent_t e : S->nb_ev)
        if (e) (void)hipEventDestroy(e);
    S->nb_meta.release();
    for (hipStream_t st : S->streams) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    delete h;
}

static long submit(Server* S, std::unique_ptr<Job> j) {
    std::lock_guard<std::mutex> g(S->mu);
    if (j->footprint > S->budget) {
        S->err = "job footprint (" + std::to_string(j->footprint) + " bytes) exceeds the memory budget";
        set_last_error(S->err);
        return -1;
    }
    j->id = S->next_id++;
    j->t_submit = since(S);
    Job* raw = j.get();
    S->jobs[raw->id] = std::move(j);
    S->pending.push_back(raw);
    S->generation++;
    S->cv_work.notify_all();
    return raw->id;
}

// A planetary system (central body first, 1..65 bodies) for n_steps WHFast
// steps in democratic heliocentric coordinates with the given G and safe_mode,
// ending synchronized.
REB_API long reb_amd_server_submit_whfast(struct reb_amd_server* h, size_t n_bodies, const struct reb_particle* bodies,
                                          double dt, double G, int safe_mode, uint64_t n_steps) {
    if (n_bodies < 1 || n_bodies > 65 || !bodies || !std::isfinite(dt) || !std::isfinite(G)) {
        set_last_error("reb_amd_server_submit_whfast: 1 to 65 bodies, a finite dt and a finite G required");
        return -1;
    }
    auto j = std::make_unique<Job>();
    j->kind = JOB_WHFAST;
    j->bodies.assign(bodies, bodies + n_bodies);
    j->dt = dt;
    j->G = G;
    j->safe_mode = safe_mode ? 1 : 0;
    j->n_steps = n_steps;
    j->footprint = whfast_footprint(n_bodies);
    return submit(&h->s, std::move(j));
}

// An N-body simulation: leapfrog (order 2) for n_steps steps with direct
// gravity (REB_GRAVITY_BASIC or _COMPENSATED) in the given precision.
REB_API long reb_amd_server_submit_nbody(struct reb_amd_server* h, size_t n, const struct reb_particle* bodies,
                                         double G, double softening, double dt, uint64_t n_steps, int gravity,
                                         int precision) {
    if (n < 1 || !bodies || (gravity != REB_GRAVITY_BASIC && gravity != REB_GRAVITY_COMPENSATED) ||
        precision < REB_AMD_PRECISION_FP64 || precision > REB_AMD_PRECISION_SPLIT || n > (size_t)1 << 24) {
        set_last_error("reb_amd_server_submit_nbody: invalid arguments");
        return -1;
    }
    auto j = std::make_unique<Job>();
    j->kind = JOB_NBODY;
    j->bodies.assign(bodies, bodies + n);
    j->G = G;
    j->softening = softening;
    j->dt = dt;
    j->n_steps = n_steps;
    j->gravity = gravity;
    j->precision = precision;
    j->footprint = nbody_footprint(n, precision, gravity);
    return submit(&h->s, std::move(j));
}

// A simulation as a job that continues from its state, time and step count
// for n_steps steps: WHFast in democratic heliocentric coordinates (1..65
// bodies, no correctors, default kernel, no softening) or order-2 leapfrog
// with BASIC or COMPENSATED direct gravity in the simulation's precision.  A
// WHFast simulation that carries its heliocentric state (one read from a
// checkpoint, or one that has stepped) continues from it as
// reb_simulation_steps would.  The simulation is not modified.
REB_API long reb_amd_server_submit_simulation(struct reb_amd_server* h, struct reb_simulation* r,
                                              uint64_t n_steps) {
    const char* fn = "reb_amd_server_submit_simulation";
    auto fail = [&](const std::string& why) {
        set_last_error(std::string(fn) + ": " + why);
        return -1L;
    };
    if (!h || !r) return fail("server and simulation required");
    if (!r->integrator.name) return fail("the simulation has no integrator");
    if (r->N_var || r->additional_forces || dev_of(r)->dev_forces || r->N_ghost_x || r->N_ghost_y || r->N_ghost_z ||
        r->collision != REB_COLLISION_NONE || r->boundary != REB_BOUNDARY_NONE ||
        (r->N_active != SIZE_MAX && r->N_active != r->N))
        return fail("variational particles, additional forces, ghost boxes, collisions, boundaries and test "
                    "particles are not served");
    if (!std::isfinite(r->dt) || !std::isfinite(r->G)) return fail("a finite dt and a finite G required");
    if (!prepare_host(r)) return fail("could not read the simulation's state from the device");
    auto j = std::make_unique<Job>();
    j->bodies.assign(r->particles, r->particles + r->N);
    j->dt = r->dt;
    j->G = r->G;
    j->softening = r->softening;
    j->n_steps = n_steps;
    j->t0 = r->t;
    j->steps0 = r->steps_done;
    j->resumed = true;
    if (strcmp(r->integrator.name, "whfast") == 0) {
        auto* s = (const struct reb_integrator_whfast_state*)r->integrator.state;
        if (r->N < 1 || r->N > 65) return fail("WHFast jobs hold 1 to 65 bodies");
        if (!s || s->coordinates != REB_INTEGRATOR_WHFAST_COORDINATES_DEMOCRATICHELIOCENTRIC || s->corrector ||
            s->corrector2 || s->kernel || s->keep_unsynchronized)
            return fail("WHFast jobs use democratic heliocentric coordinates without correctors");
        if (r->gravity != REB_GRAVITY_BASIC || r->softening != 0.)
            return fail("WHFast jobs use BASIC gravity without softening");
        j->kind = JOB_WHFAST;
        j->safe_mode = s->safe_mode ? 1 : 0;
        // the heliocentric state the simulation would continue from
        // (whfast.hip: a simulation read from a file steps from its p_jh)
        if (s->p_jh && s->N_allocated == r->N && !r->did_modify_particles) {
            const size_t n = r->N;
            j->pjh.assign(7 * n, 0.);
            for (size_t i = 0; i < n; i++) {
                const struct reb_particle& q = s->p_jh[i];
                const double f[7] = {q.x, q.y, q.z, q.vx, q.vy, q.vz, q.m};
                for (int k = 0; k < 7; k++) j->pjh[k * n + i] = f[k];
            }
            j->synchronized = r->is_synchronized ? 1 : 0;
        }
        j->footprint = whfast_footprint(r->N);
    } else if (strcmp(r->integrator.name, "leapfrog") == 0) {
        auto* s = (const struct reb_integrator_leapfrog_state*)r->integrator.state;
        if (!s || s->order != 2) return fail("leapfrog jobs are of order 2");
        if (r->gravity != REB_GRAVITY_BASIC && r->gravity != REB_GRAVITY_COMPENSATED)
            return fail("N-body jobs use BASIC or COMPENSATED direct gravity");
        if (r->N < 1 || r->N > ((size_t)1 << 24)) return fail("N-body jobs hold 1 to 2^24 bodies");
        j->kind = JOB_NBODY;
        j->gravity = r->gravity;
        j->precision = reb_amd_get_precision(r);
        j->footprint = nbody_footprint(r->N, j->precision, j->gravity);
    } else {
        return fail("only WHFast and leapfrog simulations are served");
    }
    return submit(&h->s, std::move(j));
}

// Writes a job's current state to `filename` as one Simulationarchive
// snapshot (appended when the file exists): a pending job as submitted, a
// running job at the next point where its state is whole (between pool
// ticks, batch rounds or step segments), a finished job as it ended.  Blocks
// until the file is written.  Stepping a simulation read from the snapshot
// for the job's remaining steps gives the job's result bit for bit, and so
// does submitting it with reb_amd_server_submit_simulation.  0 on success.
REB_API int reb_amd_server_checkpoint(struct reb_amd_server* h, long job, const char* filename) {
    if (!h || !filename) {
        set_last_error("reb_amd_server_checkpoint: server and file name required");
        return 1;
    }
    Server* S = &h->s;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->jobs.find(job);
    if (it == S->jobs.end()) {
        set_last_error("reb_amd_server_checkpoint: unknown job");
        return 1;
    }
    Job* j = it->second.get();
    std::string err;
    if (j->state == REB_AMD_JOB_FAILED) {
        set_last_error("reb_amd_server_checkpoint: job failed: " + j->error);
        return 1;
    }
    if (j->state == REB_AMD_JOB_PENDING || j->state == REB_AMD_JOB_DONE) {
        // the host copy is the state; a pending job is only read by admission
        // and a finished one not at all, so the copy can be made unlocked
        Job copy;
        copy.kind = j->kind;
        copy.bodies = j->bodies;
        copy.dt = j->dt; copy.G = j->G; copy.softening = j->softening;
        copy.safe_mode = j->safe_mode; copy.gravity = j->gravity; copy.precision = j->precision;
        copy.n_steps = j->n_steps; copy.t0 = j->t0; copy.steps0 = j->steps0; copy.resumed = j->resumed;
        copy.pjh = j->pjh;
        copy.synchronized = j->synchronized;
        const bool final = j->state == REB_AMD_JOB_DONE;
        lk.unlock();
        if (!write_host_state(copy, final, filename, &err)) {
            set_last_error("reb_amd_server_checkpoint: " + err);
            return 1;
        }
        return 0;
    }
    auto c = std::make_shared<Checkpoint>();
    c->file = filename;
    j->ckpt.push_back(c);
    S->generation++;
    S->cv_work.notify_all();
    S->cv_done.wait(lk, [&] { return c->done; });
    if (c->rc) set_last_error("reb_amd_server_checkpoint: " + c->error);
    return c->rc;
}

// Device-memory footprints used for admission.
REB_API size_t reb_amd_server_whfast_bytes(size_t n_bodies) { return whfast_footprint(n_bodies); }
REB_API size_t reb_amd_server_nbody_bytes(size_t n, int precision) {
    return nbody_footprint(n, precision, REB_GRAVITY_BASIC);
}

REB_API int reb_amd_server_status(struct reb_amd_server* h, long job) {
    Server* S = &h->s;
    std::lock_guard<std::mutex> g(S->mu);
    auto it = S->jobs.find(job);
    return it == S->jobs.end() ? REB_AMD_JOB_UNKNOWN : it->second->state;
}

// Waits until the job is done or failed (timeout_s < 0: no limit).  Returns
// its state (REB_AMD_JOB_RUNNING or _PENDING after a timeout).
REB_API int reb_amd_server_wait(struct reb_amd_server* h, long job, double timeout_s) {
    Server* S = &h->s;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->jobs.find(job);
    if (it == S->jobs.end()) return REB_AMD_JOB_UNKNOWN;
    Job* j = it->second.get();
    auto done = [&] { return j->state == REB_AMD_JOB_DONE || j->state == REB_AMD_JOB_FAILED; };
    if (timeout_s < 0)
        S->cv_done.wait(lk, done);
    else
        S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done);
    return j->state;
}

// Copies a finished job's bodies (x y z vx vy vz ax ay az m updated; other
// fields as submitted), energies before and after (tools.c:108-162) and its
// submit / admit / done times in seconds since the server was created, then
// forgets the job.  0 on success.
REB_API int reb_amd_server_result(struct reb_amd_server* h, long job, struct reb_particle* out, size_t n,
                                  double* energy2, double* times3) {
    Server* S = &h->s;
    std::lock_guard<std::mutex> g(S->mu);
    auto it = S->jobs.find(job);
    if (it == S->jobs.end()) {
        set_last_error("reb_amd_server_result: unknown job");
        return 1;
    }
    Job* j = it->second.get();
    if (j->state != REB_AMD_JOB_DONE) {
        set_last_error(j->state == REB_AMD_JOB_FAILED ? "job failed: " + j->error : "job not finished");
        return 1;
    }
    if (out) {
        if (n != j->bodies.size()) {
            set_last_error("reb_amd_server_result: wrong body count");
            return 1;
        }
        std::copy(j->bodies.begin(), j->bodies.end(), out);
    }
    if (energy2) {
        energy2[0] = j->e0;
        energy2[1] = j->e1;
    }
    if (times3) {
        times3[0] = j->t_submit;
        times3[1] = j->t_admit;
        times3[2] = j->t_done;
    }
    S->jobs.erase(it);
    return 0;
}

REB_API const char* reb_amd_server_job_error(struct reb_amd_server* h, long job) {
    Server* S = &h->s;
    static thread_local std::string copy;
    std::lock_guard<std::mutex> g(S->mu);
    auto it = S->jobs.find(job);
    copy = it == S->jobs.end() ? std::string("unknown job") : it->second->error;
    return copy.c_str();
}

// out[0..9]: submitted-but-unreported jobs, pending, running, done, failed,
// bytes in use, peak bytes in use, budget, pool ticks, pool busy seconds.
REB_API int reb_amd_server_stats(struct reb_amd_server* h, double* out, int n) {
    Server* S = &h->s;
    std::lock_guard<std::mutex> g(S->mu);
    size_t running = 0;
    for (auto& kv : S->jobs)
        if (kv.second->state == REB_AMD_JOB_RUNNING) running++;
    const double v[10] = {(double)S->jobs.size(), (double)S->pending.size(), (double)running, (double)S->done,
                          (double)S->failed, (double)S->used, (double)S->peak_used, (double)S->budget,
                          (double)S->pool.ticks, S->pool.busy_s};
    for (int i = 0; i < n && i < 10; i++) out[i] = v[i];
    return 0;
}

}  // extern "C"
