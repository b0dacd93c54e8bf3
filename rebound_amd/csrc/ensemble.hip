// Batched ensembles of independent planetary systems, integrated with WHFast
// in democratic heliocentric coordinates (integrator_whfast.c:1113-1259,
// transformations.c:388-477) entirely on the GPU.
//
// Layout.  All bodies of all systems live in one structure-of-arrays block
// (x y z vx vy vz ax ay az m, `stride` doubles per field) plus the heliocentric
// copy (x y z vx vy vz m).  Systems are bucketed by planet count K; a wave task
// holds floor(64/K) systems of one bucket, one lane per planet.  The persistent
// kernel (whfast.hip) pulls tasks from a work counter, keeps a task's whole
// state in registers for all steps of the launch and writes it back once.
//
// Step loop.  Every launch is a replay of a captured hipGraph of
// [reset work counter, persistent kernel for `nsteps` steps]; the counter reset
// is a kernel (not a memset) so a graph can be replayed any number of times.
// A request of n steps is split into launches of `chunk` steps (0: one launch),
// the last one ending with the synchronization that reb_simulation_steps does
// (simulation.c:504-513).  Grid and chunk come from reb_amd_ensemble_tune or
// the defaults (one resident wavefront per task, one launch per request).
//
// Each system's arithmetic is exactly the per-simulation WHFast plugin's, so
// every system evolves bit-identically to the reference integrating it alone.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"
#include "whfast_device.h"

namespace rbamd {

struct GraphEntry {
    long nsteps = -1;
    int did_modify = 0, sync_at_end = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    unsigned long long last_use = 0;
};

struct Ensemble {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t nsys = 0, nbody = 0, stride = 0;
    std::vector<size_t> n_bodies, first;
    double G = 1.;
    int safe_mode = 0;
    int compensated = 0;
    int coords = WH_DH;
    int corrector = 0, corrector2 = 0;
    int kernel = 0;
    DevBuf body, ph, sys, tasks, order, flags, counter, energy;
    int ntasks = 0;
    int ntasks_dense = 0;            // wave tasks at the densest packing
    int packing = 0;                 // requested task count (0: densest)
    std::vector<int> order_h, nb_h;  // host copies: task order, bodies per system
    int grid = 0;
    long chunk = 0;
    bool first_step_done = false;
    uint64_t steps_done = 0;
    int max_resident = 0;
    GraphEntry graphs[4];
    unsigned long long use_clock = 0;
    // MEGNO (one first-order set per system): pv, jv (10 * stride doubles
    // each), the systems' MEGNO fields, the set offset {0}
    bool megno = false;
    DevBuf var;
    size_t var_bytes = 0;
    std::string err;
};

static bool ens_ok(Ensemble* e, hipError_t rc, const char* what) {
    if (rc == hipSuccess) return true;
    e->err = std::string("HIP error ") + hipGetErrorString(rc) + " in " + what;
    set_last_error(e->err);
    return false;
}
#define ENS_HIP(e, call) ens_ok((e), (call), #call)

static WHParams ens_params(const Ensemble* e) {
    WHParams P;
    P.sys = e->sys.as<WHSys>();
    P.tasks = e->tasks.as<WHTask>();
    P.order = e->order.as<int>();
    P.flags = e->flags.as<unsigned int>();
    P.counter = e->counter.as<unsigned int>();
    P.ntasks = e->ntasks;
    P.grid = e->grid;
    P.body = e->body.as<double>();
    P.ph = e->ph.as<double>();
    P.stride = e->stride;
    P.G = e->G;
    P.soft2 = 0.;
    P.compensated = e->compensated;
    P.nsteps = 0;
    P.safe_mode = e->safe_mode;
    P.did_modify = 0;
    P.sync_at_end = 0;
    P.coords = e->coords;
    P.corrector = e->corrector;
    P.corrector2 = e->corrector2;
    P.kernel = e->kernel;
    P.jacgrav = e->kernel == 1 || e->kernel == 3;
    if (e->megno) {
        double* base = e->var.as<double>();
        P.nvar = 1;
        P.pv = base;
        P.jv = base + 10 * e->stride;
        P.vstride = e->stride;
        P.megno = 2;
        P.megno_sys = base + 20 * e->stride;
        P.var_index = reinterpret_cast<const int*>(base + 20 * e->stride + (size_t)kMegnoFields * e->nsys);
    }
    return P;
}

static void drop_graph(GraphEntry& g) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g = GraphEntry();
}

static void drop_graphs(Ensemble* e) {
    for (auto& g : e->graphs) drop_graph(g);
}

// The graph of one launch shape, captured on first use (a small LRU cache).
static GraphEntry* graph_for(Ensemble* e, long nsteps, int did_modify, int sync_at_end) {
    GraphEntry* slot = nullptr;
    for (auto& g : e->graphs) {
        if (g.exec && g.nsteps == nsteps && g.did_modify == did_modify && g.sync_at_end == sync_at_end) {
            g.last_use = ++e->use_clock;
            return &g;
        }
        if (!slot || !g.exec || (slot->exec && g.last_use < slot->last_use)) slot = &g;
    }
    drop_graph(*slot);
    WHParams P = ens_params(e);
    P.nsteps = nsteps;
    P.did_modify = did_modify;
    P.sync_at_end = sync_at_end;
    if (!ENS_HIP(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal))) return nullptr;
    const int rc = whfast_launch(P, e->stream);
    hipGraph_t graph = nullptr;
    const hipError_t end = hipStreamEndCapture(e->stream, &graph);
    if (rc || end != hipSuccess || !graph) {
        if (graph) (void)hipGraphDestroy(graph);
        e->err = "capture of the ensemble step launch failed";
        set_last_error(e->err);
        return nullptr;
    }
    slot->graph = graph;
    if (!ENS_HIP(e, hipGraphInstantiate(&slot->exec, graph, nullptr, nullptr, 0))) {
        drop_graph(*slot);
        return nullptr;
    }
    slot->nsteps = nsteps;
    slot->did_modify = did_modify;
    slot->sync_at_end = sync_at_end;
    slot->last_use = ++e->use_clock;
    return slot;
}

// Wave tasks over the systems in `order` (bucketed by body count).  At the
// densest packing a task holds 64 / (lanes per system) systems of one bucket;
// with target > 0 every bucket is cut into about target / densest times as
// many tasks, fewer systems each, so that more wavefronts share the chip
// (latency hiding when there are few wavefronts per SIMD).  The systems and
// their arithmetic are the same either way.
static std::vector<WHTask> make_tasks(const std::vector<int>& order, const std::vector<int>& nb, int target,
                                      int* dense_count) {
    struct Bucket {
        size_t k0, n;
        int Gs;
    };
    std::vector<Bucket> buckets;
    for (size_t k = 0; k < order.size();) {
        size_t m = 0;
        while (k + m < order.size() && nb[order[k + m]] == nb[order[k]]) m++;
        buckets.push_back(Bucket{k, m, whfast_group_size(nb[order[k]] - 1)});
        k += m;
    }
    size_t dense = 0;
    for (const Bucket& b : buckets) dense += (b.n + (64 / b.Gs) - 1) / (64 / b.Gs);
    if (dense_count) *dense_count = (int)dense;
    std::vector<WHTask> tasks;
    for (const Bucket& b : buckets) {
        const size_t per = (size_t)(64 / b.Gs);
        size_t nt = (b.n + per - 1) / per;
        if (target > 0 && (size_t)target > dense) nt = std::min(b.n, (nt * (size_t)target + dense - 1) / dense);
        const size_t each = (b.n + nt - 1) / nt;
        for (size_t q = 0; q < b.n; q += each)
            tasks.push_back(WHTask{b.Gs, (int)(b.k0 + q), (int)std::min(each, b.n - q)});
    }
    return tasks;
}

// Re-packs the wave tasks (target 0: densest); drops the captured graphs.
static bool set_packing(Ensemble* e, int target) {
    if (target == e->packing && e->ntasks > 0) return true;
    const std::vector<WHTask> tasks = make_tasks(e->order_h, e->nb_h, target, &e->ntasks_dense);
    if (!ENS_HIP(e, hipMemcpyAsync(e->tasks.p, tasks.data(), sizeof(WHTask) * tasks.size(), hipMemcpyHostToDevice,
                                   e->stream)) ||
        !ENS_HIP(e, hipStreamSynchronize(e->stream)))
        return false;
    drop_graphs(e);
    e->packing = target;
    e->ntasks = (int)tasks.size();
    e->grid = std::min(e->ntasks, e->max_resident);
    return true;
}

static bool launch(Ensemble* e, long nsteps, int did_modify, int sync_at_end) {
    GraphEntry* g = graph_for(e, nsteps, did_modify, sync_at_end);
    return g && ENS_HIP(e, hipGraphLaunch(g->exec, e->stream));
}

// n steps, ending synchronized like reb_simulation_steps.
static bool ens_steps(Ensemble* e, uint64_t n) {
    if (!ENS_HIP(e, hipSetDevice(e->device))) return false;
    if (n == 0) return true;
    const uint64_t chunk = e->chunk > 0 ? (uint64_t)e->chunk : n;
    uint64_t left = n;
    while (left > 0) {
        const uint64_t k = std::min(chunk, left);
        // The first launch derives the heliocentric state from the inertial one
        // (particles were added: integrator_whfast.c:1120-1135).
        const int did_modify = e->first_step_done ? 0 : 1;
        if (!launch(e, (long)k, did_modify, k == left ? 1 : 0)) return false;
        e->first_step_done = true;
        left -= k;
    }
    e->steps_done += n;
    return true;
}

}  // namespace rbamd

using namespace rbamd;

struct reb_amd_ensemble {
    Ensemble e;
};

extern "C" {

REB_API void reb_amd_ensemble_free(struct reb_amd_ensemble* h);

REB_API struct reb_amd_ensemble* reb_amd_ensemble_create(size_t n_systems, const size_t* n_bodies,
                                                         const struct reb_particle* bodies, const double* dt,
                                                         double G, int safe_mode, int gravity) {
    if (reb_amd_device_count() <= 0) {
        set_last_error("reb_amd_ensemble_create: no HIP device is available");
        return nullptr;
    }
    if (n_systems == 0 || !n_bodies || !bodies || !dt) {
        set_last_error("reb_amd_ensemble_create: empty ensemble or null argument");
        return nullptr;
    }
    if (gravity != REB_GRAVITY_BASIC && gravity != REB_GRAVITY_COMPENSATED) {
        set_last_error("reb_amd_ensemble_create: gravity must be REB_GRAVITY_BASIC or REB_GRAVITY_COMPENSATED");
        return nullptr;
    }
    if (n_systems > (size_t)INT32_MAX / 65) {
        set_last_error("reb_amd_ensemble_create: too many systems");
        return nullptr;
    }
    size_t tot = 0;
    for (size_t s = 0; s < n_systems; s++) {
        if (n_bodies[s] < 1 || n_bodies[s] > 65) {
            set_last_error("reb_amd_ensemble_create: each system needs 1 to 65 bodies, the central body first");
            return nullptr;
        }
        tot += n_bodies[s];
    }
    auto* h = new reb_amd_ensemble();
    Ensemble* e = &h->e;
    (void)hipGetDevice(&e->device);
    e->nsys = n_systems;
    e->G = G;
    e->safe_mode = safe_mode ? 1 : 0;
    e->compensated = gravity == REB_GRAVITY_COMPENSATED ? 1 : 0;
    e->n_bodies.assign(n_bodies, n_bodies + n_systems);
    e->first.resize(n_systems);
    for (size_t s = 0, b = 0; s < n_systems; s++) {
        e->first[s] = b;
        b += n_bodies[s];
    }
    e->nbody = tot;
    e->stride = (tot + 63) / 64 * 64;
    if (!ENS_HIP(e, hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking))) {
        delete h;
        return nullptr;
    }
    std::vector<double> hb(10 * e->stride, 0.), hp(7 * e->stride, 0.);
    for (size_t b = 0; b < tot; b++) {
        const struct reb_particle& p = bodies[b];
        const double f[10] = {p.x, p.y, p.z, p.vx, p.vy, p.vz, p.ax, p.ay, p.az, p.m};
        for (int k = 0; k < 10; k++) hb[k * e->stride + b] = f[k];
    }
    std::vector<WHSys> sys(n_systems);
    std::vector<int> order(n_systems);
    for (size_t s = 0; s < n_systems; s++) {
        sys[s].body0 = (int)e->first[s];
        sys[s].N = (int)n_bodies[s];
        sys[s].Nact = (int)n_bodies[s];
        sys[s].Nact_grav = (int)n_bodies[s];
        sys[s].tp_type = 0;
        sys[s].steps = -1;
        sys[s].req = 0;
        sys[s].pad = 0;
        sys[s].G = G;
        sys[s].dt = dt[s];
        order[s] = (int)s;
    }
    // Bucket by planet count, largest first (the longest tasks start first).
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sys[a].N > sys[b].N; });
    e->order_h = order;
    e->nb_h.resize(n_systems);
    for (size_t q = 0; q < n_systems; q++) e->nb_h[q] = sys[q].N;
    const std::vector<WHTask> tasks = make_tasks(order, e->nb_h, 0, &e->ntasks_dense);
    e->ntasks = (int)tasks.size();
    e->packing = 0;
    std::vector<unsigned int> flags(n_systems, 1u);   // freshly added: synchronized
    bool ok = e->body.ensure(sizeof(double) * hb.size()) && e->ph.ensure(sizeof(double) * hp.size()) &&
              e->sys.ensure(sizeof(WHSys) * n_systems) && e->tasks.ensure(sizeof(WHTask) * n_systems) &&
              e->order.ensure(sizeof(int) * n_systems) && e->flags.ensure(sizeof(unsigned int) * n_systems) &&
              e->counter.ensure(256) && e->energy.ensure(sizeof(double) * n_systems);
    if (!ok) {
        set_last_error("reb_amd_ensemble_create: out of device memory");
        reb_amd_ensemble_free(h);
        return nullptr;
    }
    ok = ENS_HIP(e, hipMemcpy(e->body.p, hb.data(), sizeof(double) * hb.size(), hipMemcpyHostToDevice)) &&
         ENS_HIP(e, hipMemcpy(e->ph.p, hp.data(), sizeof(double) * hp.size(), hipMemcpyHostToDevice)) &&
         ENS_HIP(e, hipMemcpy(e->sys.p, sys.data(), sizeof(WHSys) * n_systems, hipMemcpyHostToDevice)) &&
         ENS_HIP(e, hipMemcpy(e->tasks.p, tasks.data(), sizeof(WHTask) * tasks.size(), hipMemcpyHostToDevice)) &&
         ENS_HIP(e, hipMemcpy(e->order.p, order.data(), sizeof(int) * n_systems, hipMemcpyHostToDevice)) &&
         ENS_HIP(e, hipMemcpy(e->flags.p, flags.data(), sizeof(unsigned int) * n_systems, hipMemcpyHostToDevice)) &&
         ENS_HIP(e, hipMemset(e->counter.p, 0, 256));
    if (!ok) {
        reb_amd_ensemble_free(h);
        return nullptr;
    }
    e->max_resident = whfast_max_resident(e->device);
    e->grid = std::min(e->ntasks, e->max_resident);
    e->chunk = 0;
    return h;
}

REB_API void reb_amd_ensemble_free(struct reb_amd_ensemble* h) {
    if (!h) return;
    Ensemble* e = &h->e;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graphs(e);
    for (DevBuf* b : {&e->body, &e->ph, &e->sys, &e->tasks, &e->order, &e->flags, &e->counter, &e->energy, &e->var})
        b->release();
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete h;
}

// Coordinates and symplectic correctors of every system (the fields
// coordinates, corrector and corrector2 of integrator_whfast.h:28-46), before
// the first steps; democratic heliocentric without correctors by default.
REB_API int reb_amd_ensemble_set_whfast(struct reb_amd_ensemble* h, int coordinates, unsigned int corrector,
                                        unsigned int corrector2) {
    Ensemble* e = &h->e;
    if (e->first_step_done) {
        set_last_error("reb_amd_ensemble_set_whfast: the ensemble has already been stepped");
        return 1;
    }
    if (coordinates != REB_INTEGRATOR_WHFAST_COORDINATES_JACOBI &&
        coordinates != REB_INTEGRATOR_WHFAST_COORDINATES_DEMOCRATICHELIOCENTRIC) {
        set_last_error("reb_amd_ensemble_set_whfast: coordinates must be JACOBI or DEMOCRATICHELIOCENTRIC");
        return 1;
    }
    const bool jac = coordinates == REB_INTEGRATOR_WHFAST_COORDINATES_JACOBI;
    if ((corrector || corrector2) && !jac) {
        set_last_error("reb_amd_ensemble_set_whfast: symplectic correctors are only compatible with Jacobi coordinates");
        return 1;
    }
    if (corrector != 0 && corrector != 3 && corrector != 5 && corrector != 7 && corrector != 11 && corrector != 17) {
        set_last_error("reb_amd_ensemble_set_whfast: first symplectic correctors are only available in the "
                       "following orders: 0, 3, 5, 7, 11, 17");
        return 1;
    }
    if (!jac && e->kernel != 0) {
        set_last_error("reb_amd_ensemble_set_whfast: non-standard kernel requires Jacobi coordinates");
        return 1;
    }
    drop_graphs(e);
    e->coords = jac ? WH_JACOBI : WH_DH;
    e->corrector = (int)corrector;
    e->corrector2 = corrector2 ? 1 : 0;
    return 0;
}

// WHFast kernel of every system (the field kernel of integrator_whfast.h:28-46:
// 0 default, 1 modified kick, 2 composition, 3 lazy implementer's), before
// the first steps, with Jacobi coordinates.  Kernels 1 and 3 evaluate the
// Jacobi gravity (gravity.c:108-163), as the reference switches r->gravity to
// it (integrator_whfast.c:929-931).
REB_API int reb_amd_ensemble_set_whfast_kernel(struct reb_amd_ensemble* h, unsigned int kernel) {
    Ensemble* e = &h->e;
    if (e->first_step_done) {
        set_last_error("reb_amd_ensemble_set_whfast_kernel: the ensemble has already been stepped");
        return 1;
    }
    if (kernel > 3) {
        set_last_error("reb_amd_ensemble_set_whfast_kernel: kernel method must be 0 (default), 1 (exact modified "
                       "kick), 2 (composition kernel), or 3 (lazy implementer's modified kick)");
        return 1;
    }
    if (kernel != 0 && e->coords != WH_JACOBI) {
        set_last_error("reb_amd_ensemble_set_whfast_kernel: non-standard kernel requires Jacobi coordinates");
        return 1;
    }
    drop_graphs(e);
    e->kernel = (int)kernel;
    return 0;
}

// MEGNO for every system (tools.c:1431-1470 with the variational particles
// given): var0 holds one variational particle per body, in body order (x y z
// vx vy vz m are used), e.g. those reb_simulation_init_megno draws for the
// system.  Jacobi coordinates, the default kernel, no second corrector, safe
// mode and BASIC gravity, as integrator_whfast.c:877-910 requires; before
// the first steps.  Each system's time starts at 0.
REB_API int reb_amd_ensemble_init_megno(struct reb_amd_ensemble* h, const struct reb_particle* var0) {
    Ensemble* e = &h->e;
    const char* why = nullptr;
    if (e->first_step_done) why = "the ensemble has already been stepped";
    else if (e->coords != WH_JACOBI) why = "variational particles are only compatible with Jacobi coordinates";
    else if (e->kernel != 0) why = "variational particles are only compatible with the standard kernel";
    else if (e->corrector2) why = "variational particles not compatible with 2nd corrector";
    else if (!e->safe_mode) why = "MEGNO is not compatible with WHFast's safe_mode=0";
    else if (e->compensated) why = "variational gravity needs REB_GRAVITY_BASIC";
    else if (!var0) why = "null argument";
    if (why) {
        set_last_error(std::string("reb_amd_ensemble_init_megno: ") + why);
        return 1;
    }
    if (!ENS_HIP(e, hipSetDevice(e->device))) return 1;
    const size_t nd = 20 * e->stride + (size_t)kMegnoFields * e->nsys + 1;
    if (!e->var.ensure(sizeof(double) * nd)) {
        set_last_error("reb_amd_ensemble_init_megno: out of device memory");
        return 1;
    }
    std::vector<double> hb(nd, 0.);
    for (size_t b = 0; b < e->nbody; b++) {
        const struct reb_particle& q = var0[b];
        const double f[6] = {q.x, q.y, q.z, q.vx, q.vy, q.vz};
        for (int k = 0; k < 6; k++) hb[k * e->stride + b] = f[k];
        hb[9 * e->stride + b] = q.m;
    }
    // the set offset 0 in the last double (its first four bytes)
    if (!ENS_HIP(e, hipMemcpyAsync(e->var.p, hb.data(), sizeof(double) * nd, hipMemcpyHostToDevice, e->stream)) ||
        !ENS_HIP(e, hipStreamSynchronize(e->stream)))
        return 1;
    e->var_bytes = sizeof(double) * nd;
    e->megno = true;
    drop_graphs(e);
    return 0;
}

// Each system's MEGNO <Y> and Lyapunov number (tools.c:1471-1482); either
// pointer may be null.
REB_API int reb_amd_ensemble_megno(struct reb_amd_ensemble* h, double* megno, double* lyapunov) {
    Ensemble* e = &h->e;
    if (!e->megno) {
        set_last_error("reb_amd_ensemble_megno: MEGNO is not enabled (reb_amd_ensemble_init_megno)");
        return 1;
    }
    std::vector<double> M((size_t)kMegnoFields * e->nsys);
    if (!ENS_HIP(e, hipSetDevice(e->device)) ||
        !ENS_HIP(e, hipMemcpyAsync(M.data(), e->var.as<double>() + 20 * e->stride, sizeof(double) * M.size(),
                                   hipMemcpyDeviceToHost, e->stream)) ||
        !ENS_HIP(e, hipStreamSynchronize(e->stream)))
        return 1;
    for (size_t i = 0; i < e->nsys; i++) {
        const double* m = &M[i * kMegnoFields];
        if (megno) megno[i] = m[0] == m[1] ? 0. : m[3] / (m[0] - m[1]);
        if (lyapunov) lyapunov[i] = m[5] == 0.0 ? 0. : m[4] / m[5];
    }
    return 0;
}

// The variational particles (inertial, x y z vx vy vz ax ay az m) in body order.
REB_API int reb_amd_ensemble_get_var_particles(struct reb_amd_ensemble* h, struct reb_particle* out) {
    Ensemble* e = &h->e;
    if (!e->megno) {
        set_last_error("reb_amd_ensemble_get_var_particles: MEGNO is not enabled");
        return 1;
    }
    std::vector<double> hb(10 * e->stride);
    if (!ENS_HIP(e, hipSetDevice(e->device)) ||
        !ENS_HIP(e, hipMemcpyAsync(hb.data(), e->var.p, sizeof(double) * hb.size(), hipMemcpyDeviceToHost,
                                   e->stream)) ||
        !ENS_HIP(e, hipStreamSynchronize(e->stream)))
        return 1;
    for (size_t b = 0; b < e->nbody; b++) {
        struct reb_particle& p = out[b];
        double* f[10] = {&p.x, &p.y, &p.z, &p.vx, &p.vy, &p.vz, &p.ax, &p.ay, &p.az, &p.m};
        for (int k = 0; k < 10; k++) *f[k] = hb[k * e->stride + b];
    }
    return 0;
}

// Queues n steps of every system (ending synchronized); returns without
// waiting for the device.  0 on success.
REB_API int reb_amd_ensemble_steps(struct reb_amd_ensemble* h, uint64_t n_steps) {
    return ens_steps(&h->e, n_steps) ? 0 : 1;
}

REB_API int reb_amd_ensemble_wait(struct reb_amd_ensemble* h) {
    return ENS_HIP(&h->e, hipStreamSynchronize(h->e.stream)) ? 0 : 1;
}

REB_API uint64_t reb_amd_ensemble_steps_done(const struct reb_amd_ensemble* h) { return h->e.steps_done; }

REB_API size_t reb_amd_ensemble_n_systems(const struct reb_amd_ensemble* h) { return h->e.nsys; }

REB_API size_t reb_amd_ensemble_n_bodies(const struct reb_amd_ensemble* h) { return h->e.nbody; }

// Inertial state of all bodies in creation order (x y z vx vy vz ax ay az m
// filled; the remaining fields of `out` are left as they are).
REB_API int reb_amd_ensemble_get_particles(struct reb_amd_ensemble* h, struct reb_particle* out) {
    Ensemble* e = &h->e;
    std::vector<double> hb(10 * e->stride);
    if (!ENS_HIP(e, hipSetDevice(e->device)) ||
        !ENS_HIP(e, hipMemcpyAsync(hb.data(), e->body.p, sizeof(double) * hb.size(), hipMemcpyDeviceToHost,
                                   e->stream)) ||
        !ENS_HIP(e, hipStreamSynchronize(e->stream)))
        return 1;
    for (size_t b = 0; b < e->nbody; b++) {
        struct reb_particle& p = out[b];
        double* f[10] = {&p.x, &p.y, &p.z, &p.vx, &p.vy, &p.vz, &p.ax, &p.ay, &p.az, &p.m};
        for (int k = 0; k < 10; k++) *f[k] = hb[k * e->stride + b];
    }
    return 0;
}

// Total energy of every system (tools.c:108-162 order), into out[n_systems].
REB_API int reb_amd_ensemble_energy(struct reb_amd_ensemble* h, double* out) {
    Ensemble* e = &h->e;
    if (!ENS_HIP(e, hipSetDevice(e->device))) return 1;
    const WHParams P = ens_params(e);
    if (whfast_energy_launch(P, e->energy.as<double>(), (int)e->nsys, e->stream)) {
        ENS_HIP(e, hipGetLastError());
        return 1;
    }
    return ENS_HIP(e, hipMemcpyAsync(out, e->energy.p, sizeof(double) * e->nsys, hipMemcpyDeviceToHost, e->stream)) &&
                   ENS_HIP(e, hipStreamSynchronize(e->stream))
               ? 0
               : 1;
}

// Per-system status bits: 1 synchronized, 2 a Kepler step longer than an
// orbital period was taken, 4 particles modified while unsynchronized.
REB_API int reb_amd_ensemble_flags(struct reb_amd_ensemble* h, unsigned int* out) {
    Ensemble* e = &h->e;
    return ENS_HIP(e, hipSetDevice(e->device)) &&
                   ENS_HIP(e, hipMemcpyAsync(out, e->flags.p, sizeof(unsigned int) * e->nsys, hipMemcpyDeviceToHost,
                                             e->stream)) &&
                   ENS_HIP(e, hipStreamSynchronize(e->stream))
               ? 0
               : 1;
}

// Launch shape: persistent workgroups (0: one per task, capped at what is
// resident at once) and steps per launch (0: a whole request in one launch).
REB_API int reb_amd_ensemble_set_launch(struct reb_amd_ensemble* h, int grid, long chunk) {
    Ensemble* e = &h->e;
    const int g = grid > 0 ? std::min(grid, e->ntasks) : std::min(e->ntasks, e->max_resident);
    const long c = chunk > 0 ? chunk : 0;
    if (g != e->grid) drop_graphs(e);
    e->grid = g;
    e->chunk = c;
    return 0;
}

// Wave packing: about `tasks` wave tasks (0: the densest, every wavefront
// filled).  Returns the resulting task count, or -1.
REB_API int reb_amd_ensemble_set_packing(struct reb_amd_ensemble* h, int tasks) {
    Ensemble* e = &h->e;
    if (tasks < 0 || !ENS_HIP(e, hipSetDevice(e->device)) || !set_packing(e, tasks)) return -1;
    return e->ntasks;
}

REB_API int reb_amd_ensemble_get_launch(const struct reb_amd_ensemble* h, int* grid, long* chunk, int* ntasks,
                                        int* max_resident) {
    if (grid) *grid = h->e.grid;
    if (chunk) *chunk = h->e.chunk;
    if (ntasks) *ntasks = h->e.ntasks;
    if (max_resident) *max_resident = h->e.max_resident;
    return 0;
}

// Autotuner: times `probe_steps` steps of every candidate (packing, grid,
// chunk) on the ensemble's own state, restoring the state after each probe,
// and keeps the fastest shape.  The ensemble is left exactly as it was before the call.
REB_API int reb_amd_ensemble_tune(struct reb_amd_ensemble* h, long probe_steps, double* best_ms) {
    Ensemble* e = &h->e;
    if (probe_steps < 4) probe_steps = 4;
    if (!ENS_HIP(e, hipSetDevice(e->device))) return 1;
    DevBuf sb, sp, sf, sv;
    const size_t bb = sizeof(double) * 10 * e->stride, pb = sizeof(double) * 7 * e->stride,
                 fb = sizeof(unsigned int) * e->nsys, vb = e->megno ? e->var_bytes : 0;
    if (!sb.ensure(bb) || !sp.ensure(pb) || !sf.ensure(fb) || (vb && !sv.ensure(vb))) {
        set_last_error("reb_amd_ensemble_tune: out of device memory");
        return 1;
    }
    // the bodies, heliocentric state and flags (and the MEGNO state) to or from the backup
    auto copy_state = [&](bool save) {
        void* bk[4] = {sb.p, sp.p, sf.p, sv.p};
        void* live[4] = {e->body.p, e->ph.p, e->flags.p, e->var.p};
        const size_t n[4] = {bb, pb, fb, vb};
        for (int k = 0; k < 4; k++)
            if (n[k] && !ENS_HIP(e, hipMemcpyAsync(save ? bk[k] : live[k], save ? live[k] : bk[k], n[k],
                                                   hipMemcpyDeviceToDevice, e->stream)))
                return false;
        return true;
    };
    bool ok = copy_state(true);
    const bool first_done = e->first_step_done;
    const uint64_t done = e->steps_done;
    const int grid0 = e->grid;
    const long chunk0 = e->chunk;
    const int pack0 = e->packing;
    // packings: the densest, and as many tasks as wavefronts fit at once
    std::vector<int> packs = {0};
    if (e->max_resident > e->ntasks_dense + e->ntasks_dense / 8) packs.push_back(e->max_resident);
    const std::vector<long> chunks = {0, std::max(1L, probe_steps / 4), std::max(1L, probe_steps / 16)};
    hipEvent_t a = nullptr, b = nullptr;
    ok = ok && ENS_HIP(e, hipEventCreate(&a)) && ENS_HIP(e, hipEventCreate(&b));
    double best = 1e300, best_raw = 0.;
    int bg = grid0, bp = pack0;
    long bc = chunk0;
    // one candidate: warm-up run (captures the graphs), restore, timed run, restore
    auto probe = [&](int g, long c, float* ms) {
        reb_amd_ensemble_set_launch(h, g, c);
        bool fine = ens_steps(e, (uint64_t)probe_steps);
        fine = fine && copy_state(false);
        e->first_step_done = first_done;
        fine = fine && ENS_HIP(e, hipEventRecord(a, e->stream)) && ens_steps(e, (uint64_t)probe_steps) &&
               ENS_HIP(e, hipEventRecord(b, e->stream)) && ENS_HIP(e, hipEventSynchronize(b)) &&
               ENS_HIP(e, hipEventElapsedTime(ms, a, b));
        fine = fine && copy_state(false);
        e->first_step_done = first_done;
        e->steps_done = done;
        return fine;
    };
    for (int pk : packs) {
        ok = ok && set_packing(e, pk);
        std::vector<int> grids = {std::min(e->ntasks, e->max_resident)};
        for (int g : {e->ntasks / 2, e->ntasks / 4, 1024, 512})
            if (g > 0 && g < grids[0] && std::find(grids.begin(), grids.end(), g) == grids.end()) grids.push_back(g);
        for (int g : grids) {
            for (long c : chunks) {
                float ms = 0.f;
                ok = ok && probe(g, c, &ms);
                // A chunked request waits at every chunk for its slowest system;
                // over a long request the slow systems differ from chunk to
                // chunk, which a short probe cannot see, so chunking must win
                // clearly.
                const double score = c > 0 ? ms / 0.85 : ms;
                if (ok && score < best) {
                    best = score;
                    best_raw = ms;
                    bg = g;
                    bc = c;
                    bp = pk;
                }
            }
        }
    }
    ok = ENS_HIP(e, hipStreamSynchronize(e->stream)) && ok;
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    sb.release();
    sp.release();
    sf.release();
    ok = set_packing(e, bp) && ok;
    reb_amd_ensemble_set_launch(h, bg, bc);
    if (best_ms) *best_ms = best_raw;
    return ok ? 0 : 1;
}

REB_API const char* reb_amd_ensemble_error(const struct reb_amd_ensemble* h) { return h->e.err.c_str(); }

}  // extern "C"
