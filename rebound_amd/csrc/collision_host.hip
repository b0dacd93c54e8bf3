// Collision resolution (collision.c:336-404 and the built-in resolvers
// :573-737).
//
// The search runs on the GPU (tree.hip) and hands over the collisions in the
// reference's detection order.  Resolution is sequential by definition (a
// body may collide twice in one step, each resolution sees the velocities the
// previous one produced) and calls user callbacks (collision_resolve,
// coefficient_of_restitution), so it runs on the host: the shuffle draws from
// r->rand_seed with rand_r as the reference does, and the built-in
// hard-sphere resolver uses the host libm (atan2/sin/cos) like the reference.
// Only the bodies named in the collision list travel between device and host
// when the resolver is one of the built-ins that touch p1 and p2 alone.
//
// The built-in hard-sphere resolver with a constant (or no) coefficient of
// restitution touches nothing but p1, p2 and the two logging fields, so the
// shuffled list is split into levels: a collision's level is one more than the
// last level that touched either of its bodies.  The collisions of one level
// touch disjoint bodies and are resolved in parallel on host threads; every
// body still sees its collisions in list order, and the pressure log is summed
// in list order afterwards, so the result is the sequential one, bit for bit.
// Right after the search the device numbers the colliding bodies and packs
// the eight fields the resolver reads into 64-byte records, which travel to
// pinned host memory while the host draws the shuffle.  The shuffle goes to
// the device, which writes the shuffled list twice: in compact body numbers
// for the resolve, and as the reference's struct list straight into
// r->collisions (pinned), a copy that runs while the host resolves.  The
// resolver works on the body records in place; they travel back for a
// scatter of the velocities.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>

#include "internal.h"

namespace rbamd {

// Exclusive prefix sum of n ints with rocPRIM's device scan; with tmp ==
// nullptr it only reports the temporary storage in `bytes`.
static hipError_t exclusive_sum(void* tmp, size_t& bytes, const int* in, int* out, size_t n, hipStream_t st) {
    return rocprim::exclusive_scan(tmp, bytes, in, out, 0, n, rocprim::plus<int>(), st);
}

// A small persistent pool of host threads for the level-parallel resolve.
class HostPool {
  public:
    static HostPool& get() {
        static HostPool pool;
        return pool;
    }
    int size() const { return (int)threads_.size() + 1; }
    // Runs fn(i) for i in [0, n) on the pool and the calling thread.  The
    // pool serves one caller at a time: a caller that finds it busy (another
    // simulation resolving on another host thread) runs its items itself.
    void run(size_t n, const std::function<void(size_t)>& fn) {
        std::unique_lock<std::mutex> owner(dispatch_, std::try_to_lock);
        if (n < 256 || threads_.empty() || !owner.owns_lock()) {
            for (size_t i = 0; i < n; i++) fn(i);
            return;
        }
        {
            std::lock_guard<std::mutex> g(mu_);
            fn_ = &fn;
            all_ = nullptr;
            n_ = n;
            next_.store(0);
            busy_ = (int)threads_.size();
            gen_++;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return busy_ == 0; });
        fn_ = nullptr;
    }
    // Runs fn(t) once on every thread of the pool, t = 0 .. size()-1 (0 is
    // the calling thread), and returns when all have finished.  The caller
    // holds dispatch_.
    void run_all(const std::function<void(int)>& fn) {
        if (threads_.empty()) {
            fn(0);
            return;
        }
        {
            std::lock_guard<std::mutex> g(mu_);
            fn_ = nullptr;
            all_ = &fn;
            busy_ = (int)threads_.size();
            gen_++;
        }
        cv_.notify_all();
        fn(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return busy_ == 0; });
        all_ = nullptr;
    }
    // fn(order[q]) for the items q of level 1, then of level 2, ... (level l
    // holds q in [start[l], start[l+1])): items of one level in parallel, a
    // barrier between levels, all levels in one dispatch of the pool.  The
    // last levels, once fewer than kTail items remain, run on one thread in
    // level order (a barrier costs more than their items).
    void run_levels(int nlev, const int* start, const int* order, const std::function<void(size_t)>& fn) {
        constexpr int kTail = 1024;
        const int total = start[nlev + 1] - start[1];
        // levels in order on this thread give the same result, so a busy
        // pool (another caller) is not waited for
        std::unique_lock<std::mutex> owner(dispatch_, std::try_to_lock);
        if (threads_.empty() || total < 2 * kTail || !owner.owns_lock()) {
            for (int q = start[1]; q < start[nlev + 1]; q++) fn((size_t)order[q]);
            return;
        }
        int par = nlev;   // levels 1..par in parallel
        while (par > 1 && start[nlev + 1] - start[par] < kTail) par--;
        const int P = size();
        std::vector<std::atomic<int>> next(par + 2);
        for (int l = 1; l <= par; l++) next[l].store(start[l]);
        std::atomic<long> arrived{0};
        run_all([&](int t) {
            for (int l = 1; l <= par; l++) {
                const int e = start[l + 1];
                const int chunk = (e - start[l]) / (4 * P) + 1;
                for (;;) {
                    const int b = next[l].fetch_add(chunk);
                    if (b >= e) break;
                    const int end = std::min(e, b + chunk);
                    for (int q = b; q < end; q++) fn((size_t)order[q]);
                }
                // barrier: every thread has finished level l
                arrived.fetch_add(1);
                const long want = (long)l * P;
                // spin briefly, then yield: a thread of the pool that the OS
                // has preempted must get its CPU back quickly
                for (long spin = 0; arrived.load(std::memory_order_acquire) < want; spin++) {
                    if (spin < 2000)
                        __builtin_ia32_pause();
                    else
                        std::this_thread::yield();
                }
            }
            if (t == 0)
                for (int q = start[par + 1]; q < start[nlev + 1]; q++) fn((size_t)order[q]);
        });
    }

  private:
    // One thread fewer than the CPUs (at most 16) besides the caller: the
    // HIP runtime's own threads keep a CPU, so the level barriers do not wait
    // for a preempted pool thread.
    HostPool() {
        unsigned hw = std::thread::hardware_concurrency();
        int k = (int)std::min<unsigned>(hw ? hw : 1, 16u) - 2;
        if (k < 0) k = 0;
        for (int t = 0; t < k; t++) threads_.emplace_back([this, t] { loop(t + 1); });
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
            gen_++;
        }
        cv_.notify_all();
        for (auto& t : threads_) t.join();
    }
    void work() {
        const size_t chunk = 64;
        for (;;) {
            const size_t b = next_.fetch_add(chunk);
            if (b >= n_) break;
            const size_t e = std::min(n_, b + chunk);
            for (size_t i = b; i < e; i++) (*fn_)(i);
        }
    }
    void loop(int index) {
        unsigned long long seen = 0;
        for (;;) {
            const std::function<void(int)>* all = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                all = all_;
            }
            if (all)
                (*all)(index);
            else
                work();
            std::lock_guard<std::mutex> g(mu_);
            if (--busy_ == 0) done_.notify_all();
        }
    }
    std::vector<std::thread> threads_;
    std::mutex dispatch_;   // held by the caller that owns the pool's threads
    std::mutex mu_;
    std::condition_variable cv_, done_;
    const std::function<void(size_t)>* fn_ = nullptr;
    const std::function<void(int)>* all_ = nullptr;
    size_t n_ = 0;
    std::atomic<size_t> next_{0};
    int busy_ = 0;
    unsigned long long gen_ = 0;
    bool stop_ = false;
};

// collision.c:573-665 on a host copy: applies one hard-sphere collision to
// bodies i1 and i2 (reb_particle or HSBody records) and returns 1 with the
// pressure-log term in *plog if it did.
template <class Body>
static int hardsphere_apply(const struct reb_simulation* r, Body* particles, size_t i1, size_t i2,
                            const struct reb_vec6d& gb, double* plog);

// Gathers / scatters the full state of a list of bodies.
__global__ void k_gather_bodies(const double* __restrict__ soa, size_t cap, const int* __restrict__ ids, int n,
                                double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = ids[k];
    for (int f = 0; f < NFIELD; f++) out[(size_t)k * NFIELD + f] = soa[(size_t)f * cap + i];
}

__global__ void k_scatter_bodies(double* __restrict__ soa, size_t cap, const int* __restrict__ ids, int n,
                                 const double* __restrict__ in) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = ids[k];
    for (int f = 0; f < NFIELD; f++) soa[(size_t)f * cap + i] = in[(size_t)k * NFIELD + f];
}

// The eight fields the hard-sphere resolver reads, one 64-byte record per
// colliding body; it writes back vx, vy, vz.
struct HSBody {
    double x, y, z, vx, vy, vz, m, r;
};
// A collision in compact body numbers, with its ghost-box index.
struct HSPair {
    int p1, p2, gb;
};

// flag[p] = 1 for every body named by a record.
__global__ void k_coll_mark(const CollRecH* __restrict__ rec, int nc, int* __restrict__ flag) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc) return;
    flag[rec[k].p1] = 1;
    flag[rec[k].p2] = 1;
}

// Compact number cidx[p] (exclusive scan of flag) -> the body's fields.
__global__ void k_coll_gather(const double* __restrict__ soa, size_t cap, const int* __restrict__ flag,
                              const int* __restrict__ cidx, int N, int* __restrict__ ids, HSBody* __restrict__ out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N || !flag[p]) return;
    const int c = cidx[p];
    ids[c] = p;
    out[c] = HSBody{soa[(size_t)FX * cap + p],  soa[(size_t)FY * cap + p],  soa[(size_t)FZ * cap + p],
                    soa[(size_t)FVX * cap + p], soa[(size_t)FVY * cap + p], soa[(size_t)FVZ * cap + p],
                    soa[(size_t)FM * cap + p],  soa[(size_t)FR * cap + p]};
}

// Ghost boxes of a collision search (the inner ring: at most 3 x 3 x 3).
static constexpr int kCollGB = 27;
struct CollGB {
    int n;
    struct reb_vec6d g[kCollGB];
};

// The shuffled list: entry k is detection record perm[k], in compact body
// numbers and as the reference's struct.
__global__ void k_coll_shuffled(const CollRecH* __restrict__ rec, const int* __restrict__ perm, int nc,
                                const int* __restrict__ cidx, CollGB gb, HSPair* __restrict__ cl,
                                struct reb_collision* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc) return;
    const CollRecH h = rec[perm[k]];
    cl[k] = HSPair{cidx[h.p1], cidx[h.p2], h.gb};
    struct reb_collision c;
    c.p1 = (size_t)h.p1;
    c.p2 = (size_t)h.p2;
    c.gb = gb.g[h.gb];
    c.ri = (size_t)h.ri;
    out[k] = c;
}

__global__ void k_coll_scatter_v(double* __restrict__ soa, size_t cap, const int* __restrict__ ids, int n,
                                 const HSBody* __restrict__ in) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const size_t i = (size_t)ids[k];
    soa[(size_t)FVX * cap + i] = in[k].vx;
    soa[(size_t)FVY * cap + i] = in[k].vy;
    soa[(size_t)FVZ * cap + i] = in[k].vz;
}

// Levels of the shuffled list on the device.  The level of collision k is one
// more than the largest level of the collisions before it that share a body.
// prev holds, for each of the two bodies of k, the nearest earlier collision
// of that body (from each body's list of collisions); rounds then give a
// level to every collision whose predecessors had one after the round
// before.  Each round reads one buffer and writes the other, so it sees only
// what earlier launches wrote: round l settles exactly the collisions of
// level l.  Levels are at most kMaxLevels, so level starts fit kLevBins.
static constexpr int kLevBins = 1024;
static constexpr int kMaxLevels = kLevBins - 2;
static constexpr int kLevBlocks = 64;   // blocks of the histogram / placement passes

// Collisions per body (both bodies of every collision).
__global__ void k_lev_body_count(const HSPair* __restrict__ cl, int nc, int* __restrict__ cnt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc) return;
    atomicAdd(&cnt[cl[k].p1], 1);
    atomicAdd(&cnt[cl[k].p2], 1);
}

// Each body's collisions at its offset (in any order).
__global__ void k_lev_body_fill(const HSPair* __restrict__ cl, int nc, const int* __restrict__ off,
                                int* __restrict__ cursor, int* __restrict__ list) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc) return;
    const int a = cl[k].p1, b = cl[k].p2;
    list[off[a] + atomicAdd(&cursor[a], 1)] = k;
    list[off[b] + atomicAdd(&cursor[b], 1)] = k;
}

// prev[2k + s]: the latest collision before k of body s of k (-1: none).
__global__ void k_lev_links(const HSPair* __restrict__ cl, int nc, const int* __restrict__ off,
                            const int* __restrict__ list, int* __restrict__ prev) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 2 * nc) return;
    const int k = e >> 1;
    const int body = (e & 1) ? cl[k].p2 : cl[k].p1;
    int best = -1;
    for (int j = off[body]; j < off[body + 1]; j++) {
        const int q = list[j];
        if (q < k && q > best) best = q;
    }
    prev[e] = best;
}

__global__ void k_lev_round(const int* __restrict__ prev, const int* __restrict__ in, int* __restrict__ out, int nc,
                            int* __restrict__ pending) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc) return;
    int v = in[k];
    if (v == 0) {
        const int q1 = prev[2 * k], q2 = prev[2 * k + 1];
        const int a = q1 < 0 ? 0 : in[q1];
        const int b = q2 < 0 ? 0 : in[q2];
        if ((q1 < 0 || a > 0) && (q2 < 0 || b > 0))
            v = max(a, b) + 1;
        else
            *pending = 1;
    }
    out[k] = v;
}

// Block b of kLevBlocks takes collisions [b*per, (b+1)*per): its count per
// level into hist[b][*].
__global__ __launch_bounds__(256) void k_lev_hist(const int* __restrict__ level, int nc, int per,
                                                  int* __restrict__ hist) {
    __shared__ int h[kLevBins];
    for (int i = threadIdx.x; i < kLevBins; i += 256) h[i] = 0;
    __syncthreads();
    const int k0 = blockIdx.x * per, k1 = min(nc, k0 + per);
    for (int k = k0 + (int)threadIdx.x; k < k1; k += 256) atomicAdd(&h[level[k]], 1);
    __syncthreads();
    for (int i = threadIdx.x; i < kLevBins; i += 256) hist[blockIdx.x * kLevBins + i] = h[i];
}

// One block of kLevBins threads: start[l] (first slot of level l) and the
// first slot of each histogram block in each level, offs[b][l].
__global__ __launch_bounds__(kLevBins) void k_lev_offsets(const int* __restrict__ hist, int* __restrict__ start,
                                                          int* __restrict__ offs) {
    __shared__ int sc[kLevBins];
    const int l = threadIdx.x;
    int tot = 0;
    for (int b = 0; b < kLevBlocks; b++) tot += hist[b * kLevBins + l];
    sc[l] = tot;
    __syncthreads();
    for (int d = 1; d < kLevBins; d <<= 1) {   // inclusive scan
        const int v = l >= d ? sc[l - d] : 0;
        __syncthreads();
        sc[l] += v;
        __syncthreads();
    }
    int o = sc[l] - tot;   // exclusive
    start[l] = o;
    for (int b = 0; b < kLevBlocks; b++) {
        offs[b * kLevBins + l] = o;
        o += hist[b * kLevBins + l];
    }
}

// Collision k to a slot of its level (any order within a level: the
// collisions of one level touch disjoint bodies).
__global__ __launch_bounds__(256) void k_lev_place(const int* __restrict__ level, int nc, int per,
                                                   const int* __restrict__ offs, int* __restrict__ order) {
    __shared__ int cur[kLevBins];
    for (int i = threadIdx.x; i < kLevBins; i += 256) cur[i] = offs[blockIdx.x * kLevBins + i];
    __syncthreads();
    const int k0 = blockIdx.x * per, k1 = min(nc, k0 + per);
    for (int k = k0 + (int)threadIdx.x; k < k1; k += 256) order[atomicAdd(&cur[level[k]], 1)] = k;
}

static void copy_body_in(struct reb_particle& p, const double* s) {
    p.x = s[FX]; p.y = s[FY]; p.z = s[FZ];
    p.vx = s[FVX]; p.vy = s[FVY]; p.vz = s[FVZ];
    p.ax = s[FAX]; p.ay = s[FAY]; p.az = s[FAZ];
    p.m = s[FM]; p.r = s[FR];
}
static void copy_body_out(const struct reb_particle& p, double* s) {
    s[FX] = p.x; s[FY] = p.y; s[FZ] = p.z;
    s[FVX] = p.vx; s[FVY] = p.vy; s[FVZ] = p.vz;
    s[FAX] = p.ax; s[FAY] = p.ay; s[FAZ] = p.az;
    s[FM] = p.m; s[FR] = p.r;
}

// Brings the listed bodies' device state into r->particles (host), or back.
static bool move_bodies(struct reb_simulation* r, const std::vector<int>& ids, bool to_host) {
    DeviceState* d = dev_of(r);
    const int n = (int)ids.size();
    if (n == 0) return true;
    const size_t idbytes = sizeof(int) * n, databytes = sizeof(double) * NFIELD * n;
    if (!d->aos.ensure(idbytes + databytes + 256)) return false;
    int* dids = d->aos.as<int>();
    double* ddata = reinterpret_cast<double*>(d->aos.as<char>() + ((idbytes + 255) / 256) * 256);
    // pinned staging: the simulation's host stage holds NFIELD * cap doubles (n <= cap)
    std::vector<double> fallback;
    double* hd = d->host_stage;
    if (!hd || (size_t)n > d->host_cap) {
        fallback.resize((size_t)NFIELD * n);
        hd = fallback.data();
    }
    struct {
        double* p;
        double* data() { return p; }
    } h{hd};
    if (!RB_HIP(r, hipMemcpyAsync(dids, ids.data(), idbytes, hipMemcpyHostToDevice, d->stream))) return false;
    if (to_host) {
        hipLaunchKernelGGL(k_gather_bodies, dim3((n + 127) / 128), dim3(128), 0, d->stream, d->soa.as<double>(),
                           d->cap, dids, n, ddata);
        if (!RB_HIP(r, hipMemcpyAsync(h.data(), ddata, databytes, hipMemcpyDeviceToHost, d->stream))) return false;
        if (!stream_sync(r)) return false;
        for (int k = 0; k < n; k++) copy_body_in(r->particles[ids[k]], h.data() + (size_t)k * NFIELD);
    } else {
        for (int k = 0; k < n; k++) copy_body_out(r->particles[ids[k]], h.data() + (size_t)k * NFIELD);
        if (!RB_HIP(r, hipMemcpyAsync(ddata, h.data(), databytes, hipMemcpyHostToDevice, d->stream))) return false;
        hipLaunchKernelGGL(k_scatter_bodies, dim3((n + 127) / 128), dim3(128), 0, d->stream, d->soa.as<double>(),
                           d->cap, dids, n, ddata);
        if (!stream_sync(r)) return false;
    }
    return RB_HIP(r, hipGetLastError());
}

static int rand_r_step(unsigned int* seed) {
    unsigned int next = *seed;
    int result;
    next *= 1103515245u;
    next += 12345u;
    result = (unsigned int)(next / 65536) % 2048;
    next *= 1103515245u;
    next += 12345u;
    result <<= 10;
    result ^= (unsigned int)(next / 65536) % 1024;
    next *= 1103515245u;
    next += 12345u;
    result <<= 10;
    result ^= (unsigned int)(next / 65536) % 1024;
    *seed = next;
    return result;
}

}  // namespace rbamd

// Buffers of the compact resolve, reused from step to step (sized to the
// largest step so far).  `ready` is set when the device has prepared the
// compact records of the last search.
struct rbamd::ResolveScratch {
    DevBuf flag, cidx, ids, body, pair, perm, coll, prim_tmp;   // device
    DevBuf bcnt, boff, bcur, blist, prev, level_d, lmisc, lprim_tmp;
    HSBody* body_h = nullptr;                              // pinned host
    HSPair* pair_h = nullptr;
    int* perm_h = nullptr;    // the permutation, then the level order
    int* n_h = nullptr;
    int* lev_h = nullptr;     // round flags, then level starts (kLevBins)
    int last_nlev = 0;
    size_t body_cap = 0, pair_cap = 0, perm_cap = 0;
    hipEvent_t ev_pairs = nullptr, ev_coll = nullptr;
    const CollRecH* drec = nullptr;   // device records of the prepared search
    int n = 0;
    bool ready = false;
    std::vector<int> level, last, hstart, horder;
    // rand_r draws made ahead from the seed `draw_seed` while the device
    // searched: draws[i] is the i-th result, seeds[i] the seed after it
    std::vector<int> draws;
    std::vector<unsigned int> seeds;
    unsigned int draw_seed = 0;
    // the shuffle made while the device gathers the bodies: perm_h holds it
    // for perm_nc collisions, and r->rand_seed was seed_before
    bool perm_ready = false;
    int perm_nc = 0;
    unsigned int seed_before = 0;
    std::vector<double> plog;
    std::vector<unsigned char> hit;
    void release() {
        for (DevBuf* b : {&flag, &cidx, &ids, &body, &pair, &perm, &coll, &prim_tmp, &bcnt, &boff, &bcur, &blist,
                          &prev, &level_d, &lmisc, &lprim_tmp})
            b->release();
        for (void* p : {(void*)body_h, (void*)pair_h, (void*)perm_h, (void*)n_h, (void*)lev_h})
            if (p) (void)hipHostFree(p);
        lev_h = nullptr;
        if (ev_pairs) (void)hipEventDestroy(ev_pairs);
        if (ev_coll) (void)hipEventDestroy(ev_coll);
        body_h = nullptr;
        pair_h = nullptr;
        perm_h = nullptr;
        n_h = nullptr;
        ev_pairs = ev_coll = nullptr;
        body_cap = pair_cap = perm_cap = 0;
    }
};

namespace rbamd {

static void shuffle_perm(struct reb_simulation* r, ResolveScratch& W, int nc);

void resolve_release(struct reb_simulation* r) {
    DeviceState* d = dev_of(r);
    if (d->resolve) d->resolve->release();
    delete d->resolve;
    d->resolve = nullptr;
}

void compact_resolve_draw_ahead(struct reb_simulation* r, size_t expect) {
    DeviceState* d = dev_of(r);
    if (!d->resolve) d->resolve = new ResolveScratch();
    ResolveScratch& W = *d->resolve;
    W.draws.resize(expect);
    W.seeds.resize(expect);
    unsigned int seed = r->rand_seed;
    W.draw_seed = seed;
    for (size_t i = 0; i < expect; i++) {
        W.draws[i] = rand_r_step(&seed);
        W.seeds[i] = seed;
    }
}

bool compact_resolve_eligible(const struct reb_simulation* r, int n_ghostboxes) {
    const DeviceState* d = dev_of(r);
    return n_ghostboxes <= kCollGB && r->collision_resolve == reb_collision_resolve_hardsphere &&
           (r->coefficient_of_restitution == nullptr ||
            r->coefficient_of_restitution == reb_amd_coefficient_of_restitution_constant) &&
           d->loc != LOC_HOST;
}

// Grows a pinned host array of T to at least `need` elements (contents lost).
template <class T>
static bool pinned_ensure(struct reb_simulation* r, T*& p, size_t& cap, size_t need) {
    if (need <= cap && p) return true;
    size_t c = cap ? cap : 4096;
    while (c < need) c *= 2;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    if (!RB_HIP(r, hipHostMalloc((void**)&p, sizeof(T) * c, hipHostMallocDefault))) return false;
    cap = c;
    return true;
}

bool compact_resolve_mark(struct reb_simulation* r, const CollRecH* rec, size_t nc) {
    DeviceState* d = dev_of(r);
    if (!d->resolve) d->resolve = new ResolveScratch();
    ResolveScratch& W = *d->resolve;
    W.ready = false;
    const int N = (int)r->N;
    size_t tmp = 0;
    (void)exclusive_sum(nullptr, tmp, W.flag.as<int>(), W.cidx.as<int>(), N + 1, d->stream);
    if (!W.flag.ensure(sizeof(int) * (N + 1)) || !W.cidx.ensure(sizeof(int) * (N + 1)) ||
        !W.prim_tmp.ensure(tmp + 256)) {
        reb_simulation_error(r, "Out of device memory for the collision resolve.");
        return false;
    }
    if (!W.n_h && !RB_HIP(r, hipHostMalloc((void**)&W.n_h, sizeof(int), hipHostMallocDefault))) return false;
    if (!RB_HIP(r, hipMemsetAsync(W.flag.p, 0, sizeof(int) * (N + 1), d->stream))) return false;
    hipLaunchKernelGGL(k_coll_mark, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, d->stream, rec, (int)nc,
                       W.flag.as<int>());
    if (!RB_HIP(r, exclusive_sum(W.prim_tmp.p, tmp, W.flag.as<int>(), W.cidx.as<int>(), N + 1, d->stream)))
        return false;
    return RB_HIP(r, hipMemcpyAsync(W.n_h, W.cidx.as<int>() + N, sizeof(int), hipMemcpyDeviceToHost, d->stream));
}

bool compact_resolve_gather(struct reb_simulation* r, const CollRecH* rec, size_t nc) {
    DeviceState* d = dev_of(r);
    ResolveScratch& W = *d->resolve;
    const int N = (int)r->N;
    const int n = *W.n_h;   // the stream has been synchronized since compact_resolve_mark
    if (!W.ids.ensure(sizeof(int) * n + 256) || !W.body.ensure(sizeof(HSBody) * n + 256) ||
        !W.pair.ensure(sizeof(HSPair) * nc + 256) || !W.perm.ensure(sizeof(int) * nc + 256) ||
        !W.coll.ensure(sizeof(struct reb_collision) * nc + 256)) {
        reb_simulation_error(r, "Out of device memory for the collision resolve.");
        return false;
    }
    if (!pinned_ensure(r, W.body_h, W.body_cap, (size_t)n) || !pinned_ensure(r, W.pair_h, W.pair_cap, nc) ||
        !pinned_ensure(r, W.perm_h, W.perm_cap, nc) || !collisions_reserve(r, nc))
        return false;
    if (!W.ev_pairs && !RB_HIP(r, hipEventCreateWithFlags(&W.ev_pairs, hipEventDisableTiming))) return false;
    if (!W.ev_coll && !RB_HIP(r, hipEventCreateWithFlags(&W.ev_coll, hipEventDisableTiming))) return false;
    hipLaunchKernelGGL(k_coll_gather, dim3((N + 255) / 256), dim3(256), 0, d->stream, d->soa.as<double>(), d->cap,
                       W.flag.as<int>(), W.cidx.as<int>(), N, W.ids.as<int>(), W.body.as<HSBody>());
    if (!RB_HIP(r, hipGetLastError())) return false;
    if (!RB_HIP(r, hipMemcpyAsync(W.body_h, W.body.p, sizeof(HSBody) * n, hipMemcpyDeviceToHost, d->stream)))
        return false;
    W.drec = rec;
    W.n = n;
    W.ready = true;
    // the resolve follows this search at once (its shuffle is the next use
    // of the seed): shuffle now, while the device gathers the bodies
    W.seed_before = r->rand_seed;
    shuffle_perm(r, W, (int)nc);
    W.perm_ready = true;
    W.perm_nc = (int)nc;
    return true;
}

bool collisions_reserve(struct reb_simulation* r, size_t nc) {
    if (r->N_allocated_collisions >= nc && r->collisions) return true;
    size_t cap = r->N_allocated_collisions ? r->N_allocated_collisions : 32;
    while (cap < nc) cap *= 2;
    struct reb_collision* p = nullptr;
    if (hipHostMalloc((void**)&p, sizeof(struct reb_collision) * cap, hipHostMallocDefault) != hipSuccess || !p) {
        reb_simulation_error(r, "Out of memory for the collision list.");
        return false;
    }
    if (r->collisions) (void)hipHostFree(r->collisions);
    r->collisions = p;
    r->N_allocated_collisions = cap;
    return true;
}

void collisions_free(struct reb_simulation* r) {
    if (r->collisions) (void)hipHostFree(r->collisions);
    r->collisions = nullptr;
    r->N_allocated_collisions = 0;
}

// Levels of the shuffled list in W.pair on the device (see k_lev_round).
// Returns 1 with the level starts (kLevBins entries) and the collisions
// ordered by level enqueued for copying into W.lev_h and W.perm_h and *nlev
// an upper bound of the levels, 0 if the list needs more than kMaxLevels
// levels, -1 on a HIP error.
static int device_levels(struct reb_simulation* r, ResolveScratch& W, int nc, int n, int* nlev) {
    DeviceState* d = dev_of(r);
    const size_t m = 2 * (size_t)nc;
    // lmisc: round flags | level starts | per-block histograms | per-block offsets
    const size_t misc = (size_t)kLevBins * (2 + 2 * kLevBlocks);
    if (!W.bcnt.ensure(sizeof(int) * (n + 1)) || !W.boff.ensure(sizeof(int) * (n + 1)) ||
        !W.bcur.ensure(sizeof(int) * n) || !W.blist.ensure(sizeof(int) * m) || !W.prev.ensure(sizeof(int) * m) ||
        !W.level_d.ensure(sizeof(int) * 2 * (size_t)nc) || !W.lmisc.ensure(sizeof(int) * misc)) {
        reb_simulation_error(r, "Out of device memory for the collision resolve.");
        return -1;
    }
    if (!W.lev_h && !RB_HIP(r, hipHostMalloc((void**)&W.lev_h, sizeof(int) * kLevBins, hipHostMallocDefault)))
        return -1;
    int* flags = W.lmisc.as<int>();
    int* start = flags + kLevBins;
    int* hist = start + kLevBins;
    int* offs = hist + (size_t)kLevBins * kLevBlocks;
    size_t t_scan = 0;
    (void)exclusive_sum(nullptr, t_scan, W.bcnt.as<int>(), W.boff.as<int>(), n + 1, d->stream);
    if (!W.lprim_tmp.ensure(t_scan + 256)) {
        reb_simulation_error(r, "Out of device memory for the collision resolve.");
        return -1;
    }
    const dim3 gk((unsigned)((nc + 255) / 256)), gm((unsigned)((m + 255) / 256)), blk(256);
    const HSPair* cl = W.pair.as<HSPair>();
    // each body's collisions, then the predecessors
    if (!RB_HIP(r, hipMemsetAsync(W.bcnt.p, 0, sizeof(int) * (n + 1), d->stream)) ||
        !RB_HIP(r, hipMemsetAsync(W.bcur.p, 0, sizeof(int) * n, d->stream)))
        return -1;
    hipLaunchKernelGGL(k_lev_body_count, gk, blk, 0, d->stream, cl, nc, W.bcnt.as<int>());
    if (!RB_HIP(r, exclusive_sum(W.lprim_tmp.p, t_scan, W.bcnt.as<int>(), W.boff.as<int>(), n + 1, d->stream)))
        return -1;
    hipLaunchKernelGGL(k_lev_body_fill, gk, blk, 0, d->stream, cl, nc, W.boff.as<int>(), W.bcur.as<int>(),
                       W.blist.as<int>());
    hipLaunchKernelGGL(k_lev_links, gm, blk, 0, d->stream, cl, nc, W.boff.as<int>(), W.blist.as<int>(),
                       W.prev.as<int>());
    if (!RB_HIP(r, hipMemsetAsync(W.level_d.p, 0, sizeof(int) * nc, d->stream)) ||
        !RB_HIP(r, hipMemsetAsync(flags, 0, sizeof(int) * kLevBins, d->stream)))
        return -1;
    // rounds in batches; after each batch the last round's flag says whether
    // any collision still had no level when it ran.  The levels alternate
    // between the two halves of level_d; `level` is the latest.
    int* lv_buf[2] = {W.level_d.as<int>(), W.level_d.as<int>() + nc};
    int* level = lv_buf[0];
    // the first batch has a margin over the last list's level count: an empty
    // round costs a few microseconds, another batch a host round trip
    int rounds = 0, batch = std::max(8, W.last_nlev + 6);
    for (;;) {
        const int b = std::min(batch, kMaxLevels - rounds);
        for (int i = 0; i < b; i++, rounds++) {
            int* next = lv_buf[(rounds + 1) & 1];
            hipLaunchKernelGGL(k_lev_round, gk, blk, 0, d->stream, W.prev.as<int>(), level, next, nc,
                               flags + rounds);
            level = next;
        }
        if (!RB_HIP(r, hipGetLastError()) ||
            !RB_HIP(r, hipMemcpyAsync(W.lev_h, flags + rounds - 1, sizeof(int), hipMemcpyDeviceToHost, d->stream)))
            return -1;
        if (!stream_sync(r)) return -1;
        if (W.lev_h[0] == 0) break;
        if (rounds >= kMaxLevels) return 0;
        batch = 8;
    }
    // collisions ordered by level
    const int per = (nc + kLevBlocks - 1) / kLevBlocks;
    hipLaunchKernelGGL(k_lev_hist, dim3(kLevBlocks), blk, 0, d->stream, level, nc, per, hist);
    hipLaunchKernelGGL(k_lev_offsets, dim3(1), dim3(kLevBins), 0, d->stream, hist, start, offs);
    hipLaunchKernelGGL(k_lev_place, dim3(kLevBlocks), blk, 0, d->stream, level, nc, per, offs, W.perm.as<int>());
    if (!RB_HIP(r, hipGetLastError()) ||
        !RB_HIP(r, hipMemcpyAsync(W.lev_h, start, sizeof(int) * kLevBins, hipMemcpyDeviceToHost, d->stream)) ||
        !RB_HIP(r, hipMemcpyAsync(W.perm_h, W.perm.p, sizeof(int) * nc, hipMemcpyDeviceToHost, d->stream)))
        return -1;
    *nlev = rounds;
    return 1;
}

// The same levels on the host (a list with more than kMaxLevels levels).
static void host_levels(ResolveScratch& W, const HSPair* cl, int nc, int n, int* nlev) {
    std::vector<int>& level = W.level;
    std::vector<int>& last = W.last;
    std::vector<int>& start = W.hstart;
    level.resize(nc);
    last.assign(n, 0);
    start.assign(64, 0);
    int top = 0;
    for (int k = 0; k < nc; k++) {
        const HSPair q = cl[k];
        const int l = std::max(last[q.p1], last[q.p2]) + 1;
        level[k] = l;
        last[q.p1] = l;
        last[q.p2] = l;
        top = std::max(top, l);
        if ((size_t)l + 1 >= start.size()) start.resize(2 * ((size_t)l + 1), 0);
        start[l + 1]++;
    }
    start.resize(top + 2);
    for (int l = 1; l <= top + 1; l++) start[l] += start[l - 1];
    std::vector<int>& order = W.horder;
    order.resize(nc);
    std::vector<int> fill(start.begin(), start.end());
    for (int k = 0; k < nc; k++) order[fill[level[k]]++] = k;
    *nlev = top;
}

// Shuffle (collision.c:336-342) as a permutation of the detection order in
// W.perm_h: the same swaps applied to indices instead of 72-byte records,
// with the rand_r draws made ahead while the device searched when they start
// from the current seed.  Advances r->rand_seed as the reference does.
static void shuffle_perm(struct reb_simulation* r, ResolveScratch& W, int nc) {
    int* perm = W.perm_h;
    for (int k = 0; k < nc; k++) perm[k] = k;
    const int ahead = W.draw_seed == r->rand_seed ? std::min(nc, (int)W.draws.size()) : 0;
    for (int i = 0; i < ahead; i++) {
        const int nw = (int)((size_t)W.draws[i] % (size_t)nc);
        std::swap(perm[i], perm[nw]);
    }
    if (ahead > 0) r->rand_seed = W.seeds[ahead - 1];
    for (int i = ahead; i < nc; i++) {
        const int nw = (int)((size_t)rand_r_step(&r->rand_seed) % (size_t)nc);
        std::swap(perm[i], perm[nw]);
    }
    W.draws.clear();
}

// The built-in hard-sphere resolver with a constant restitution, on the
// compact records prepared by compact_resolve_gather: only the eight fields it
// reads come over, only the velocities go back, and collisions of one level
// run in parallel.
static void resolve_hardsphere_compact(struct reb_simulation* r) {
    DeviceState* d = dev_of(r);
    ResolveScratch& W = *d->resolve;
    W.ready = false;
    const CollList& L = *d->coll;
    const int nc = (int)L.n;
    const int n = W.n;
    if (W.perm_ready && W.perm_nc != nc) r->rand_seed = W.seed_before;   // (not expected) shuffle again
    if (!W.perm_ready || W.perm_nc != nc) shuffle_perm(r, W, nc);
    W.perm_ready = false;
    int* perm = W.perm_h;
    CollGB gb;
    gb.n = (int)L.gb.size();
    for (int k = 0; k < gb.n; k++) gb.g[k] = L.gb[k];
    if (!RB_HIP(r, hipMemcpyAsync(W.perm.p, perm, sizeof(int) * nc, hipMemcpyHostToDevice, d->stream))) return;
    hipLaunchKernelGGL(k_coll_shuffled, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, d->stream, W.drec,
                       W.perm.as<int>(), nc, W.cidx.as<int>(), gb, W.pair.as<HSPair>(),
                       W.coll.as<struct reb_collision>());
    if (!RB_HIP(r, hipGetLastError())) return;
    // every entry of plog and hit is written by the resolve below
    std::vector<double>& plog = W.plog;
    std::vector<unsigned char>& hit = W.hit;
    plog.resize(nc);
    hit.resize(nc);
    int nlev = 0;
    const int lv = device_levels(r, W, nc, n, &nlev);
    if (lv < 0) return;
    if (!RB_HIP(r, hipMemcpyAsync(W.pair_h, W.pair.p, sizeof(HSPair) * nc, hipMemcpyDeviceToHost, d->stream)) ||
        !RB_HIP(r, hipEventRecord(W.ev_pairs, d->stream)))
        return;
    // r->collisions as the reference leaves it (the shuffled list), copied
    // while the host resolves
    if (!RB_HIP(r, hipMemcpyAsync(r->collisions, W.coll.p, sizeof(struct reb_collision) * nc, hipMemcpyDeviceToHost,
                                  d->stream)) ||
        !RB_HIP(r, hipEventRecord(W.ev_coll, d->stream)))
        return;
    if (!RB_HIP(r, hipEventSynchronize(W.ev_pairs))) return;
    const HSPair* cl = W.pair_h;
    const int* start;
    const int* order;
    if (lv == 1) {
        start = W.lev_h;
        order = W.perm_h;
        while (nlev > 0 && start[nlev + 1] == start[nlev]) nlev--;
    } else {
        host_levels(W, cl, nc, n, &nlev);
        start = W.hstart.data();
        order = W.horder.data();
    }
    W.last_nlev = nlev;
    HSBody* B = W.body_h;
    // a collision that does not bounce logs -0.0, the exact identity of the
    // sum below, so the sum needs no branch
    HostPool::get().run_levels(nlev, start, order, [&](size_t k) {
        double pl = -0.;
        hit[k] = (unsigned char)hardsphere_apply(r, B, (size_t)cl[k].p1, (size_t)cl[k].p2, L.gb[cl[k].gb], &pl);
        plog[k] = pl;
    });
    // the pressure log in list order (collision.c:655-663)
    double sum = r->collisions_plog;
    long nhit = 0;
    for (int k = 0; k < nc; k++) {
        sum += plog[k];
        nhit += hit[k];
    }
    r->collisions_plog = sum;
    r->collisions_log_n += nhit;
    // the records go back whole; the scatter takes the velocities.  The next
    // search's copy into body_h is ordered after this one on the stream.
    if (!RB_HIP(r, hipMemcpyAsync(W.body.p, B, sizeof(HSBody) * n, hipMemcpyHostToDevice, d->stream))) return;
    hipLaunchKernelGGL(k_coll_scatter_v, dim3((n + 255) / 256), dim3(256), 0, d->stream, d->soa.as<double>(), d->cap,
                       W.ids.as<int>(), n, W.body.as<HSBody>());
    (void)RB_HIP(r, hipGetLastError());
    (void)RB_HIP(r, hipEventSynchronize(W.ev_coll));   // r->collisions is complete
    d->coll->pending = false;
    mark_device_newer(r);
}

void collision_resolve_host(struct reb_simulation* r) {
    if (r->N_collisions == 0) return;
    enum REB_COLLISION_RESOLVE_OUTCOME (*resolve)(struct reb_simulation* const, struct reb_collision) =
        r->collision_resolve;
    if (resolve == nullptr) resolve = reb_collision_resolve_halt;
    DeviceState* d = dev_of(r);
    if (d->resolve && d->resolve->ready && d->coll && d->coll->pending && d->coll->n == r->N_collisions &&
        compact_resolve_eligible(r, (int)d->coll->gb.size())) {
        resolve_hardsphere_compact(r);   // shuffles, resolves and writes r->collisions
        return;
    }
    if (d->resolve) {
        d->resolve->ready = false;
        if (d->resolve->perm_ready) {   // a shuffle made for the compact path: undo its draws
            r->rand_seed = d->resolve->seed_before;
            d->resolve->perm_ready = false;
        }
    }
    if (!materialize_collisions(r)) return;
    // shuffle (collision.c:336-342)
    for (size_t i = 0; i < r->N_collisions; i++) {
        size_t nw = (size_t)rand_r_step(&r->rand_seed) % r->N_collisions;
        struct reb_collision c1 = r->collisions[i];
        r->collisions[i] = r->collisions[nw];
        r->collisions[nw] = c1;
    }
    // The built-in hard-sphere and halt resolvers touch only p1 and p2: move
    // just those bodies.  Anything else sees the whole host state.
    const bool local = (resolve == reb_collision_resolve_hardsphere || resolve == reb_collision_resolve_halt);
    std::vector<int> ids;
    if (local) {
        std::vector<unsigned char> seen(r->N, 0);
        for (size_t k = 0; k < r->N_collisions; k++) {
            for (size_t p : {r->collisions[k].p1, r->collisions[k].p2})
                if (p < r->N && !seen[p]) {
                    seen[p] = 1;
                    ids.push_back((int)p);
                }
        }
        if (!move_bodies(r, ids, true)) return;
    } else {
        ensure_host(r);
    }
    // resolve in order (collision.c:345-404)
    bool removed_any = false;
    for (size_t i = 0; i < r->N_collisions; i++) {
        struct reb_collision c = r->collisions[i];
        if (c.p1 == SIZE_MAX || c.p2 == SIZE_MAX) continue;
        enum REB_COLLISION_RESOLVE_OUTCOME outcome = resolve(r, c);
        if (outcome & REB_COLLISION_RESOLVE_OUTCOME_REMOVE_P1) {
            const int removedp1 = !reb_simulation_remove_particle(r, c.p1);
            removed_any = true;
            if (removedp1) {
                if (c.p2 > c.p1 && c.p2 != SIZE_MAX) c.p2--;
                for (size_t j = i + 1; j < r->N_collisions; j++) {
                    struct reb_collision* cp = &(r->collisions[j]);
                    if (cp->p1 == c.p1 || cp->p2 == c.p1) {
                        cp->p1 = SIZE_MAX;
                        cp->p2 = SIZE_MAX;
                    }
                    if (cp->p1 > c.p1 && cp->p1 != SIZE_MAX) cp->p1--;
                    if (cp->p2 > c.p1 && cp->p2 != SIZE_MAX) cp->p2--;
                }
            }
        }
        if (outcome & REB_COLLISION_RESOLVE_OUTCOME_REMOVE_P2) {
            const int removedp2 = !reb_simulation_remove_particle(r, c.p2);
            removed_any = true;
            if (removedp2) {
                for (size_t j = i + 1; j < r->N_collisions; j++) {
                    struct reb_collision* cp = &(r->collisions[j]);
                    if (cp->p1 == c.p2 || cp->p2 == c.p2) {
                        cp->p1 = SIZE_MAX;
                        cp->p2 = SIZE_MAX;
                    }
                    if (cp->p1 > c.p2 && cp->p1 != SIZE_MAX) cp->p1--;
                    if (cp->p2 > c.p2 && cp->p2 != SIZE_MAX) cp->p2--;
                }
            }
        }
    }
    if (local && !removed_any) {
        move_bodies(r, ids, false);
    } else {
        mark_host_newer(r);
    }
}

}  // namespace rbamd

using namespace rbamd;

extern "C" {

// collision.c:573-665
REB_API enum REB_COLLISION_RESOLVE_OUTCOME reb_collision_resolve_hardsphere(struct reb_simulation* const r,
                                                                            struct reb_collision c) {
    double plog = 0.;
    if (rbamd::hardsphere_apply(r, r->particles, c.p1, c.p2, c.gb, &plog)) {
        r->collisions_plog += plog;
        r->collisions_log_n++;
    }
    return REB_COLLISION_RESOLVE_OUTCOME_REMOVE_NONE;
}

}  // extern "C"

namespace rbamd {

template <class Body>
static int hardsphere_apply(const struct reb_simulation* r, Body* particles, size_t i1, size_t i2,
                            const struct reb_vec6d& gb, double* plog) {
    const Body p1 = particles[i1];
    const Body p2 = particles[i2];
    double x21 = p1.x + gb.x - p2.x;
    double y21 = p1.y + gb.y - p2.y;
    double z21 = p1.z + gb.z - p2.z;
    double rp = p1.r + p2.r;
    double oldvyouter;
    if (x21 > 0) {
        oldvyouter = p1.vy;
    } else {
        oldvyouter = p2.vy;
    }
    if (rp * rp < x21 * x21 + y21 * y21 + z21 * z21) return 0;
    double vx21 = p1.vx + gb.vx - p2.vx;
    double vy21 = p1.vy + gb.vy - p2.vy;
    double vz21 = p1.vz + gb.vz - p2.vz;
    if (vx21 * x21 + vy21 * y21 + vz21 * z21 > 0) return 0;
    // rotate the pair into the xy plane, then onto the x axis.  The reference
    // library, built by gcc, evaluates each sin/cos pair with one glibc
    // sincos() call (its results differ from separate sin and cos calls in
    // the last bit for ~0.1% of angles), so the same call is made here.
    double theta = atan2(z21, y21);
    double stheta, ctheta;
    sincos(theta, &stheta, &ctheta);
    double vy21n = ctheta * vy21 + stheta * vz21;
    double y21n = ctheta * y21 + stheta * z21;
    double phi = atan2(y21n, x21);
    double sphi, cphi;
    sincos(phi, &sphi, &cphi);
    double vx21nn = cphi * vx21 + sphi * vy21n;
    double eps = 1;
    if (r->coefficient_of_restitution) eps = r->coefficient_of_restitution(r, vx21nn);
    double dvx2 = -(1.0 + eps) * vx21nn;
    double minr = (p1.r > p2.r) ? p2.r : p1.r;
    double maxr = (p1.r < p2.r) ? p2.r : p1.r;
    double mindv = minr * r->minimum_collision_velocity;
    double _r = sqrt(x21 * x21 + y21 * y21 + z21 * z21);
    mindv *= 1. - (_r - maxr) / minr;
    if (mindv > maxr * r->minimum_collision_velocity) mindv = maxr * r->minimum_collision_velocity;
    if (dvx2 < mindv) dvx2 = mindv;
    double dvx2n = cphi * dvx2;
    double dvy2n = sphi * dvx2;
    double dvy2nn = ctheta * dvy2n;
    double dvz2nn = stheta * dvy2n;
    const double p2pf = p1.m / (p1.m + p2.m);
    particles[i2].vx -= p2pf * dvx2n;
    particles[i2].vy -= p2pf * dvy2nn;
    particles[i2].vz -= p2pf * dvz2nn;
    const double p1pf = p2.m / (p1.m + p2.m);
    particles[i1].vx += p1pf * dvx2n;
    particles[i1].vy += p1pf * dvy2nn;
    particles[i1].vz += p1pf * dvz2nn;
    if (x21 > 0) {
        *plog = -fabs(x21) * (oldvyouter - particles[i1].vy) * p1.m;
    } else {
        *plog = -fabs(x21) * (oldvyouter - particles[i2].vy) * p2.m;
    }
    return 1;
}

}  // namespace rbamd

extern "C" {

// collision.c:667-672
REB_API enum REB_COLLISION_RESOLVE_OUTCOME reb_collision_resolve_halt(struct reb_simulation* const r,
                                                                      struct reb_collision c) {
    (void)c;
    r->status = REB_STATUS_COLLISION;
    return REB_COLLISION_RESOLVE_OUTCOME_REMOVE_NONE;
}

// collision.c:674-737: merge into the lower index, conserving mass, momentum
// and volume; the energy change is tracked when track_energy_offset is set.
REB_API enum REB_COLLISION_RESOLVE_OUTCOME reb_collision_resolve_merge(struct reb_simulation* const r,
                                                                       struct reb_collision c) {
    int swap = 0;
    size_t i = c.p1;
    size_t j = c.p2;
    if (j < i) {
        swap = 1;
        i = c.p2;
        j = c.p1;
    }
    struct reb_particle* pi = &(r->particles[i]);
    struct reb_particle* pj = &(r->particles[j]);
    double invmass = 1.0 / (pi->m + pj->m);
    double Ei = 0, Ef = 0;
    if (r->track_energy_offset) {
        double vix = pi->vx, viy = pi->vy, viz = pi->vz;
        Ei += 0.5 * pi->m * (vix * vix + viy * viy + viz * viz);
        double vjx = pj->vx, vjy = pj->vy, vjz = pj->vz;
        Ei += 0.5 * pj->m * (vjx * vjx + vjy * vjy + vjz * vjz);
        const size_t N_active = ((r->N_active == SIZE_MAX) ? r->N : r->N_active);
        if (i < N_active || j < N_active) {
            double x = pi->x - pj->x, y = pi->y - pj->y, z = pi->z - pj->z;
            double _r = sqrt(x * x + y * y + z * z);
            Ei += -r->G * pi->m * pj->m / _r;
        }
    }
    pi->vx = (pi->vx * pi->m + pj->vx * pj->m) * invmass;
    pi->vy = (pi->vy * pi->m + pj->vy * pj->m) * invmass;
    pi->vz = (pi->vz * pi->m + pj->vz * pj->m) * invmass;
    pi->x = (pi->x * pi->m + pj->x * pj->m) * invmass;
    pi->y = (pi->y * pi->m + pj->y * pj->m) * invmass;
    pi->z = (pi->z * pi->m + pj->z * pj->m) * invmass;
    pi->m = pi->m + pj->m;
    pi->r = cbrt(pi->r * pi->r * pi->r + pj->r * pj->r * pj->r);
    if (r->track_energy_offset) {
        double vx = pi->vx, vy = pi->vy, vz = pi->vz;
        Ef += 0.5 * pi->m * (vx * vx + vy * vy + vz * vz);
        r->energy_offset += Ei - Ef;
    }
    return swap ? REB_COLLISION_RESOLVE_OUTCOME_REMOVE_P1 : REB_COLLISION_RESOLVE_OUTCOME_REMOVE_P2;
}

}  // extern "C"
