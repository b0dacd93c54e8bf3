// Barnes-Hut octree on the GPU: construction from sorted keys, monopole
// moments, tree gravity, and the collision search that uses the same tree.
//
// Reference: tree.c:64-328 (insertion octree, one body per leaf, moments,
// recursive walk), gravity.c:47-106 (tree gravity driver), collision.c:197-269
// and :422-503 (tree collision search), simulation.c:718-799 (two largest radii).
//
// Construction.  The reference inserts bodies one by one; a cell is split as
// soon as it holds two bodies, so the finished tree is the canonical point-
// region octree: a cell is internal iff it holds >= 2 bodies, children are
// visited in octant order 0..7.  Here each body's octant digits are derived
// with the reference's own arithmetic (octant bit = x < cx, child centre =
// parent centre +/- (w/2)/2, tree.c:98-103 and :136-142), packed into a key
// (root box index first), and sorted.  In sorted order the pre-order of the
// octree is "by first body, then by level": a body s starts the internal cells
// at levels lcp(s-1,s)+1 .. lcp(s,s+1) and its own leaf at level
// max(lcp(s-1,s), lcp(s,s+1)) + 1, where lcp is the number of common leading
// digits of neighbouring keys.  A scan of those counts places every cell; each
// internal cell finds the end of its body range by a binary search on keys,
// which gives its pre-order successor (`next`), the skip pointer of the
// stackless walks.  Moments are accumulated bottom-up, level by level, in
// child (octant) order with the reference's expressions, so cell masses and
// centres of mass are bit-identical to tree.c:147-229.
//
// Walks.  One wavefront walks one pre-order sequence for 64 bodies (in key
// order, so neighbours share most of the path): at cell c a lane that accepts
// the cell (or has accepted an ancestor) skips c's subtree; the wave descends
// into c if any lane opens it.  Each lane therefore sees its own accepted
// cells in the reference's depth-first order and sums them in that order.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.h"

namespace rbamd {

// Exclusive prefix sum of n ints with rocPRIM's device scan; with tmp ==
// nullptr it only reports the temporary storage in `bytes`.
static hipError_t exclusive_sum(void* tmp, size_t& bytes, const int* in, int* out, size_t n, hipStream_t st) {
    return rocprim::exclusive_scan(tmp, bytes, in, out, 0, n, rocprim::plus<int>(), st);
}

static constexpr int kMaxGBT = 64;   // keeps the by-value argument under 4 KB
// `info` block: ints 0..7 (build results and flags), bytes 32..39 the
// interaction counter, bytes 64.. the per-block top-two radii of k_top2
// (at most kTop2Blocks blocks, two doubles each), then the second largest
// radius of all bodies.
static constexpr int kTop2Blocks = 64;
static constexpr size_t kInfoBytes = 64 + 2 * sizeof(double) * kTop2Blocks + 16;
struct GBT {
    int n;
    double g[kMaxGBT][6];
};

struct TreeBuffers {
    size_t last_total = 0;                        // collisions found by the last tree search
    DevBuf key0, key1, idx, key0b, key1b, idxb;   // keys and body permutation
    DevBuf lcp, cnt, base;                        // per body
    DevBuf prim_tmp;                              // rocPRIM temporary storage
    DevBuf node_d;                                // 8 doubles per cell: cx cy cz w m mx my mz
    DevBuf node_i;                                // 4 ints per cell: next pt level start
    DevBuf info;                                  // total, maxlevel, flags
    DevBuf coll_cnt, coll_off, coll_rec;          // collision search
    DevBuf coll_slot;                             // kCollSlots records per body (single-pass search)
    DevBuf grec, crec;                            // 64-byte walk records per cell (gravity, collisions)
    size_t body_cap = 0, node_cap = 0, rec_cap = 0;
    int nnodes = 0;
    int maxlevel = 0;
    int L1 = 0, L2 = 0, rbits = 0;
    bool two_words = false;
    int nroot = 0;
};

// ---------------------------------------------------------------- key construction
struct TreeGeom {
    double root_size;
    int nrx, nry, nrz;
    int L1, L2;     // digits in word 0 (below the root index) and word 1
};

__device__ __forceinline__ int root_index(const TreeGeom& g, double x, double y, double z) {
    // particle.c:119-126
    const int i = ((int)floor((x + g.root_size * (double)g.nrx / 2.) / g.root_size) + g.nrx) % g.nrx;
    const int j = ((int)floor((y + g.root_size * (double)g.nry / 2.) / g.root_size) + g.nry) % g.nry;
    const int k = ((int)floor((z + g.root_size * (double)g.nrz / 2.) / g.root_size) + g.nrz) % g.nrz;
    return (k * g.nry + j) * g.nrx + i;
}

// Centre of root box `ri` exactly as tree.c:91-97 computes it from a member body.
__device__ __forceinline__ void root_center(const TreeGeom& g, int ri, double& cx, double& cy, double& cz) {
    const int i = ri % g.nrx;
    const int j = (ri / g.nrx) % g.nry;
    const int k = ri / (g.nrx * g.nry);
    const double bx = g.root_size * (double)g.nrx, by = g.root_size * (double)g.nry, bz = g.root_size * (double)g.nrz;
    cx = -bx / 2. + g.root_size * (0.5 + (double)i);
    cy = -by / 2. + g.root_size * (0.5 + (double)j);
    cz = -bz / 2. + g.root_size * (0.5 + (double)k);
}

// One level down (tree.c:98-103): child w = w/2, centre +/- child w / 2.
__device__ __forceinline__ void descend(int o, double& cx, double& cy, double& cz, double& w) {
    w = w / 2.;
    cx = cx + w / 2. * (((o >> 0) % 2 == 0) ? 1. : -1);
    cy = cy + w / 2. * (((o >> 1) % 2 == 0) ? 1. : -1);
    cz = cz + w / 2. * (((o >> 2) % 2 == 0) ? 1. : -1);
}

// flags: bit0 outside box, bit1 non-finite, bit2 unresolved at max depth, bit3 identical positions
__global__ void k_tree_keys(const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
                            int n, TreeGeom g, int with_word1, unsigned long long* __restrict__ key0,
                            unsigned long long* __restrict__ key1, int* __restrict__ idx, int* __restrict__ info) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double px = x[i], py = y[i], pz = z[i];
    idx[i] = i;
    if (fabs(px) > g.root_size * (double)g.nrx / 2. || fabs(py) > g.root_size * (double)g.nry / 2. ||
        fabs(pz) > g.root_size * (double)g.nrz / 2.) {
        atomicOr(&info[2], 1);
        key0[i] = 0;
        if (with_word1) key1[i] = 0;
        return;
    }
    if (!isfinite(px) || !isfinite(py) || !isfinite(pz)) {
        atomicOr(&info[2], 2);
        key0[i] = 0;
        if (with_word1) key1[i] = 0;
        return;
    }
    const int ri = root_index(g, px, py, pz);
    double cx, cy, cz, w = g.root_size;
    root_center(g, ri, cx, cy, cz);
    unsigned long long k0 = (unsigned long long)ri;
    for (int l = 0; l < g.L1; l++) {
        int o = 0;
        if (px < cx) o += 1;
        if (py < cy) o += 2;
        if (pz < cz) o += 4;
        k0 = (k0 << 3) | (unsigned long long)o;
        descend(o, cx, cy, cz, w);
    }
    key0[i] = k0;
    if (with_word1) {
        unsigned long long k1 = 0;
        for (int l = 0; l < g.L2; l++) {
            int o = 0;
            if (px < cx) o += 1;
            if (py < cy) o += 2;
            if (pz < cz) o += 4;
            k1 = (k1 << 3) | (unsigned long long)o;
            descend(o, cx, cy, cz, w);
        }
        key1[i] = k1;
    }
}

// Number of common leading octant digits of two keys (-1: different root box).
__device__ __forceinline__ int key_lcp(unsigned long long a0, unsigned long long a1, unsigned long long b0,
                                       unsigned long long b1, int L1, int L2, int two) {
    const unsigned long long x0 = a0 ^ b0;
    if ((x0 >> (3 * L1)) != 0ull) return -1;
    if (x0 != 0ull) return (__clzll((long long)x0) - (64 - 3 * L1)) / 3;
    if (!two) return L1;
    const unsigned long long x1 = a1 ^ b1;
    if (x1 != 0ull) return L1 + (__clzll((long long)x1) - (64 - 3 * L2)) / 3;
    return L1 + L2;
}

__global__ void k_gather_u64(const unsigned long long* __restrict__ src, const int* __restrict__ perm,
                             unsigned long long* __restrict__ dst, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

// lcp[s] = common digits of sorted keys s and s+1 (lcp[n-1] = -1); counts of
// the cells each body starts; the deepest level.
__global__ void k_tree_lcp(const unsigned long long* __restrict__ k0, const unsigned long long* __restrict__ k1,
                           int n, int L1, int L2, int two, const int* __restrict__ idx, const double* __restrict__ x,
                           const double* __restrict__ y, const double* __restrict__ z, int* __restrict__ lcp,
                           int* __restrict__ info) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int l = -1;
    if (s + 1 < n) {
        l = key_lcp(k0[s], two ? k1[s] : 0ull, k0[s + 1], two ? k1[s + 1] : 0ull, L1, L2, two);
        if (l == L1 + (two ? L2 : 0)) {
            // both bodies share the deepest representable cell
            const int a = idx[s], b = idx[s + 1];
            if (x[a] == x[b] && y[a] == y[b] && z[a] == z[b])
                atomicOr(&info[2], 8);
            else
                atomicOr(&info[2], 4);
        }
    }
    lcp[s] = l;
}

__global__ void k_tree_count(const int* __restrict__ lcp, int n, int* __restrict__ cnt, int* __restrict__ info) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int prev = s > 0 ? lcp[s - 1] : -1;
    const int cur = lcp[s];
    const int nint = cur > prev ? cur - prev : 0;
    cnt[s] = nint + 1;
    atomicMax(&info[1], (prev > cur ? prev : cur) + 1);
}

struct NodeArrays {
    double *cx, *cy, *cz, *w, *m, *mx, *my, *mz;
    int *next, *pt, *level, *start;
};

__device__ __forceinline__ NodeArrays node_arrays(double* d, int* ii, size_t cap) {
    NodeArrays A;
    A.cx = d;
    A.cy = d + cap;
    A.cz = d + 2 * cap;
    A.w = d + 3 * cap;
    A.m = d + 4 * cap;
    A.mx = d + 5 * cap;
    A.my = d + 6 * cap;
    A.mz = d + 7 * cap;
    A.next = ii;
    A.pt = ii + cap;
    A.level = ii + 2 * cap;
    A.start = ii + 3 * cap;
    return A;
}

// Geometry of the cell at `level` on the path of key (k0,k1): level 0 = root box.
__device__ __forceinline__ void cell_geometry(const TreeGeom& g, unsigned long long k0, unsigned long long k1,
                                              int level, double& cx, double& cy, double& cz, double& w) {
    const int ri = (int)(k0 >> (3 * g.L1));
    root_center(g, ri, cx, cy, cz);
    w = g.root_size;
    for (int l = 1; l <= level; l++) {
        int o;
        if (l <= g.L1)
            o = (int)((k0 >> (3 * (g.L1 - l))) & 7ull);
        else
            o = (int)((k1 >> (3 * (g.L2 - (l - g.L1)))) & 7ull);
        descend(o, cx, cy, cz, w);
    }
}

__global__ void k_tree_nodes(const unsigned long long* __restrict__ k0, const unsigned long long* __restrict__ k1,
                             const int* __restrict__ idx, const int* __restrict__ lcp, const int* __restrict__ base,
                             int n, int total, TreeGeom g, int two, double* __restrict__ nd, int* __restrict__ ni,
                             size_t cap, const double* __restrict__ pm, const double* __restrict__ px,
                             const double* __restrict__ py, const double* __restrict__ pz) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    NodeArrays A = node_arrays(nd, ni, cap);
    const int prev = s > 0 ? lcp[s - 1] : -1;
    const int cur = lcp[s];
    const unsigned long long a0 = k0[s], a1 = two ? k1[s] : 0ull;
    const int b = base[s];
    int id = b;
    for (int d = prev + 1; d <= cur; d++, id++) {
        // end of the body range of this cell: largest e with lcp(key_s, key_e) >= d
        int lo = s + 1, step = 1, hi;
        while (true) {
            const int j = s + step;
            if (j >= n) {
                hi = n;
                break;
            }
            const int l = key_lcp(a0, a1, k0[j], two ? k1[j] : 0ull, g.L1, g.L2, two);
            if (l < d) {
                hi = j;
                break;
            }
            lo = j + 1;
            step *= 2;
        }
        // first index in [lo, hi) whose lcp < d (hi if none)
        while (lo < hi) {
            const int mid = lo + (hi - lo) / 2;
            const int l = key_lcp(a0, a1, k0[mid], two ? k1[mid] : 0ull, g.L1, g.L2, two);
            if (l >= d)
                lo = mid + 1;
            else
                hi = mid;
        }
        const int e = lo - 1;
        A.next[id] = (e + 1 < n) ? base[e + 1] : total;
        A.pt[id] = -(e - s + 1);
        A.level[id] = d;
        A.start[id] = s;
        double cx, cy, cz, w;
        cell_geometry(g, a0, a1, d, cx, cy, cz, w);
        A.cx[id] = cx;
        A.cy[id] = cy;
        A.cz[id] = cz;
        A.w[id] = w;
    }
    // the leaf of body s
    const int leaf_level = (prev > cur ? prev : cur) + 1;
    const int p = idx[s];
    A.next[id] = id + 1;
    A.pt[id] = p;
    A.level[id] = leaf_level;
    A.start[id] = s;
    double cx, cy, cz, w;
    cell_geometry(g, a0, a1, leaf_level, cx, cy, cz, w);
    A.cx[id] = cx;
    A.cy[id] = cy;
    A.cz[id] = cz;
    A.w[id] = w;
    // leaf moments (tree.c:220-226)
    A.m[id] = pm[p];
    A.mx[id] = px[p];
    A.my[id] = py[p];
    A.mz[id] = pz[p];
}

// Internal cells of one level: children in octant order (tree.c:160-177).
__global__ void k_tree_moments(double* __restrict__ nd, int* __restrict__ ni, size_t cap, int total, int level) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= total) return;
    NodeArrays A = node_arrays(nd, ni, cap);
    if (A.pt[c] >= 0 || A.level[c] != level) return;
    double m = 0, mx = 0, my = 0, mz = 0;
    const int end = A.next[c];
    for (int d = c + 1; d < end; d = A.next[d]) {
        const double d_m = A.m[d];
        mx += A.mx[d] * d_m;
        my += A.my[d] * d_m;
        mz += A.mz[d] * d_m;
        m += d_m;
    }
    const double m_tot = m;
    if (m_tot > 0) {
        mx /= m_tot;
        my /= m_tot;
        mz /= m_tot;
    }
    A.m[c] = m;
    A.mx[c] = mx;
    A.my[c] = my;
    A.mz[c] = mz;
}

// One cell as the walks read it: 64 bytes, so that a wavefront's (uniform)
// cell fetch is one scalar load, and the pre-order successor and the skip
// target can both be fetched while the current cell is evaluated.
struct __align__(64) GravRec {
    double mx, my, mz, m, w;
    int next, pt;
    double pad;
};
// A cell's record for the collision walk.  Inner cell (pt < 0): its
// geometric centre and width.  Leaf (pt >= 0): its body's x, y, z and radius
// in cx, cy, cz, w and its velocity in v, copied when the tree is built (the
// search runs on the state the tree was built from), so the walk reads one
// record per visited cell and never the scattered body arrays.
struct __align__(64) CollCell {
    double cx, cy, cz, w;
    int next, pt;
    double v[3];
};

__global__ void k_tree_pack(const double* __restrict__ nd, const int* __restrict__ ni, size_t cap, int total,
                            GravRec* __restrict__ g, CollCell* __restrict__ cc, const double* __restrict__ x,
                            const double* __restrict__ y, const double* __restrict__ z,
                            const double* __restrict__ vx, const double* __restrict__ vy,
                            const double* __restrict__ vz, const double* __restrict__ rad) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= total) return;
    GravRec a;
    a.mx = nd[5 * cap + c];
    a.my = nd[6 * cap + c];
    a.mz = nd[7 * cap + c];
    a.m = nd[4 * cap + c];
    a.w = nd[3 * cap + c];
    a.next = ni[c];
    a.pt = ni[cap + c];
    a.pad = 0.;
    g[c] = a;
    CollCell b;
    b.next = a.next;
    b.pt = a.pt;
    if (a.pt >= 0) {
        const int p = a.pt;
        b.cx = x[p];
        b.cy = y[p];
        b.cz = z[p];
        b.w = rad[p];
        b.v[0] = vx[p];
        b.v[1] = vy[p];
        b.v[2] = vz[p];
    } else {
        b.cx = nd[c];
        b.cy = nd[cap + c];
        b.cz = nd[2 * cap + c];
        b.w = a.w;
        b.v[0] = b.v[1] = b.v[2] = 0.;
    }
    cc[c] = b;
}

// ---------------------------------------------------------------- host: build
static bool tree_alloc(struct reb_simulation* r, TreeBuffers* T, size_t n) {
    if (n <= T->body_cap && T->key0.p) return true;
    size_t cap = T->body_cap ? T->body_cap : 1024;
    while (cap < n) cap *= 2;
    bool ok = T->key0.ensure(8 * cap) && T->key1.ensure(8 * cap) && T->idx.ensure(4 * cap) &&
              T->key0b.ensure(8 * cap) && T->key1b.ensure(8 * cap) && T->idxb.ensure(4 * cap) &&
              T->lcp.ensure(4 * cap) && T->cnt.ensure(4 * cap) && T->base.ensure(4 * cap) && T->info.ensure(kInfoBytes);
    if (!ok) {
        reb_simulation_error(r, "Out of device memory for the tree.");
        return false;
    }
    T->body_cap = cap;
    return true;
}

static TreeGeom tree_geom(struct reb_simulation* r, TreeBuffers* T) {
    TreeGeom g;
    g.root_size = r->root_size;
    g.nrx = (int)r->N_root_x;
    g.nry = (int)r->N_root_y;
    g.nrz = (int)r->N_root_z;
    int nroot = g.nrx * g.nry * g.nrz;
    int rbits = 0;
    while ((1 << rbits) < nroot) rbits++;
    g.L1 = (63 - rbits) / 3;
    g.L2 = 21;
    T->L1 = g.L1;
    T->L2 = g.L2;
    T->rbits = rbits;
    T->nroot = nroot;
    return g;
}

// Sort (key0[,key1]) with the body permutation; the result is in key0/key1/idx.
static bool tree_sort(struct reb_simulation* r, TreeBuffers* T, int n, bool two) {
    DeviceState* d = dev_of(r);
    auto* K0 = T->key0.as<unsigned long long>();
    auto* K1 = T->key1.as<unsigned long long>();
    auto* K0b = T->key0b.as<unsigned long long>();
    auto* K1b = T->key1b.as<unsigned long long>();
    int* I = T->idx.as<int>();
    int* Ib = T->idxb.as<int>();
    const int keybits = 3 * T->L1 + T->rbits;
    size_t tmp = 0;
    (void)rocprim::radix_sort_pairs(nullptr, tmp, K0, K0b, I, Ib, n, 0, keybits, d->stream);
    size_t tmp2 = 0;
    (void)rocprim::radix_sort_pairs(nullptr, tmp2, K1, K1b, I, Ib, n, 0, 3 * T->L2, d->stream);
    tmp = std::max(tmp, tmp2);
    if (!T->prim_tmp.ensure(tmp + 256)) return false;
    if (!two) {
        if (rocprim::radix_sort_pairs(T->prim_tmp.p, tmp, K0, K0b, I, Ib, n, 0, keybits, d->stream) !=
            hipSuccess)
            return RB_HIP(r, hipGetLastError());
        std::swap(T->key0, T->key0b);
        std::swap(T->idx, T->idxb);
        return true;
    }
    // least significant word first, then a stable sort on the most significant
    if (rocprim::radix_sort_pairs(T->prim_tmp.p, tmp, K1, K1b, I, Ib, n, 0, 3 * T->L2, d->stream) != hipSuccess)
        return RB_HIP(r, hipGetLastError());
    hipLaunchKernelGGL(k_gather_u64, dim3((n + 255) / 256), dim3(256), 0, d->stream, K0, Ib, K0b, n);
    // K0b: key0 in key1 order; Ib: permutation; sort (K0b, Ib) -> (K0, I)
    if (rocprim::radix_sort_pairs(T->prim_tmp.p, tmp, K0b, K0, Ib, I, n, 0, keybits, d->stream) != hipSuccess)
        return RB_HIP(r, hipGetLastError());
    hipLaunchKernelGGL(k_gather_u64, dim3((n + 255) / 256), dim3(256), 0, d->stream, K1, I, K1b, n);
    std::swap(T->key1, T->key1b);   // key1 now in final order
    return RB_HIP(r, hipGetLastError());
}

// Builds the tree of the current device positions.  tree.c:254-271 + :147-229.
bool tree_build(struct reb_simulation* r) {
    DeviceState* d = dev_of(r);
    if (r->root_size <= 0.0) {
        reb_simulation_error(r, "Set root_size to a finite value to use a tree based gravity or collision solver.");
        return false;
    }
    if (!d->tree) d->tree = new TreeBuffers();
    TreeBuffers* T = d->tree;
    const int n = (int)r->N;
    if (!tree_alloc(r, T, n ? n : 1)) return false;
    TreeGeom g = tree_geom(r, T);
    int* info = T->info.as<int>();
    bool two = false;
    int hinfo[4] = {0, 0, 0, 0};
    int last[2] = {0, 0};
    for (int attempt = 0; attempt < 2; attempt++) {
        two = attempt == 1;
        if (!RB_HIP(r, hipMemsetAsync(info, 0, 16, d->stream))) return false;
        if (n == 0) break;
        hipLaunchKernelGGL(k_tree_keys, dim3((n + 255) / 256), dim3(256), 0, d->stream, d->field(FX), d->field(FY),
                           d->field(FZ), n, g, two ? 1 : 0, T->key0.as<unsigned long long>(),
                           T->key1.as<unsigned long long>(), T->idx.as<int>(), info);
        if (!tree_sort(r, T, n, two)) return false;
        hipLaunchKernelGGL(k_tree_lcp, dim3((n + 255) / 256), dim3(256), 0, d->stream,
                           T->key0.as<unsigned long long>(), T->key1.as<unsigned long long>(), n, g.L1, g.L2,
                           two ? 1 : 0, T->idx.as<int>(), d->field(FX), d->field(FY), d->field(FZ),
                           T->lcp.as<int>(), info);
        // cells per body and their offsets, read back with the build flags in
        // one synchronization (used only when the flags report no error)
        hipLaunchKernelGGL(k_tree_count, dim3((n + 255) / 256), dim3(256), 0, d->stream, T->lcp.as<int>(), n,
                           T->cnt.as<int>(), info);
        size_t tmp = 0;
        (void)exclusive_sum(nullptr, tmp, T->cnt.as<int>(), T->base.as<int>(), n, d->stream);
        if (!T->prim_tmp.ensure(tmp + 256)) return false;
        if (!RB_HIP(r, exclusive_sum(T->prim_tmp.p, tmp, T->cnt.as<int>(), T->base.as<int>(), n, d->stream)))
            return false;
        if (!RB_HIP(r, hipMemcpyAsync(&last[0], T->base.as<int>() + (n - 1), 4, hipMemcpyDeviceToHost, d->stream)) ||
            !RB_HIP(r, hipMemcpyAsync(&last[1], T->cnt.as<int>() + (n - 1), 4, hipMemcpyDeviceToHost, d->stream)) ||
            !RB_HIP(r, hipMemcpyAsync(hinfo, info, 16, hipMemcpyDeviceToHost, d->stream)))
            return false;
        if (!stream_sync(r)) return false;
        if (hinfo[2] & 1) {
            reb_simulation_error(r, "Particle is outside of simulation box. Cannot add to tree.");
            return false;
        }
        if (hinfo[2] & 2) {
            reb_simulation_error(r, "Particle has non-finite coordinates. Cannot add to tree.");
            return false;
        }
        if (hinfo[2] & 8) {
            reb_simulation_error(r, "Cannot add two particles with the same coordinates to the tree.");
            return false;
        }
        if (!(hinfo[2] & 4)) break;
        if (two) {
            reb_simulation_error(r, "Tree deeper than 63 levels (two bodies closer than root_size/2^62).");
            return false;
        }
    }
    T->two_words = two;
    T->nnodes = 0;
    if (n == 0) return true;
    const int total = last[0] + last[1];
    T->maxlevel = hinfo[1];
    if ((size_t)total > T->node_cap) {
        size_t cap = T->node_cap ? T->node_cap : 4096;
        while (cap < (size_t)total) cap *= 2;
        if (!T->node_d.ensure(sizeof(double) * 8 * cap) || !T->node_i.ensure(sizeof(int) * 4 * cap)) {
            reb_simulation_error(r, "Out of device memory for the tree cells.");
            return false;
        }
        T->node_cap = cap;
    }
    T->nnodes = total;
    hipLaunchKernelGGL(k_tree_nodes, dim3((n + 255) / 256), dim3(256), 0, d->stream,
                       T->key0.as<unsigned long long>(), T->key1.as<unsigned long long>(), T->idx.as<int>(),
                       T->lcp.as<int>(), T->base.as<int>(), n, total, g, two ? 1 : 0, T->node_d.as<double>(),
                       T->node_i.as<int>(), T->node_cap, d->field(FM), d->field(FX), d->field(FY), d->field(FZ));
    for (int lev = T->maxlevel; lev >= 0; lev--)
        hipLaunchKernelGGL(k_tree_moments, dim3((total + 255) / 256), dim3(256), 0, d->stream,
                           T->node_d.as<double>(), T->node_i.as<int>(), T->node_cap, total, lev);
    if (!T->grec.ensure(sizeof(GravRec) * (size_t)total) || !T->crec.ensure(sizeof(CollCell) * (size_t)total)) {
        reb_simulation_error(r, "Out of device memory for the tree cells.");
        return false;
    }
    hipLaunchKernelGGL(k_tree_pack, dim3((total + 255) / 256), dim3(256), 0, d->stream, T->node_d.as<double>(),
                       T->node_i.as<int>(), T->node_cap, total, T->grec.as<GravRec>(), T->crec.as<CollCell>(),
                       d->field(FX), d->field(FY), d->field(FZ), d->field(FVX), d->field(FVY), d->field(FVZ),
                       d->field(FR));
    return RB_HIP(r, hipGetLastError());
}

// ---------------------------------------------------------------- tree gravity
// tree.c:275-328 for every ghost box (gravity.c:84-102), in the reference's
// expressions; 64 bodies (in key order) per wavefront.
__global__ __launch_bounds__(64) void k_tree_gravity(const GravRec* __restrict__ rec, int total,
                                                     const int* __restrict__ idx, int n,
                                                     const double* __restrict__ x, const double* __restrict__ y,
                                                     const double* __restrict__ z, GBT gl, double G, double soft2,
                                                     double theta2, double* __restrict__ ax,
                                                     double* __restrict__ ay, double* __restrict__ az,
                                                     unsigned long long* __restrict__ counter) {
    const int s = blockIdx.x * 64 + threadIdx.x;
    const bool live = s < n;
    const int me = live ? idx[s] : -1;
    const double xi = live ? x[me] : 0., yi = live ? y[me] : 0., zi = live ? z[me] : 0.;
    double a_x = 0., a_y = 0., a_z = 0.;
    unsigned long long inter = 0;
    for (int g = 0; g < gl.n; g++) {
        const double gx = gl.g[g][0] + xi, gy = gl.g[g][1] + yi, gz = gl.g[g][2] + zi;
        int skip = live ? 0 : 0x7fffffff;
        int c = 0;
        GravRec cur = rec[0];
        while (c < total) {
            // both possible successors are fetched before this cell is evaluated
            const int c_open = c + 1, c_skip = cur.next;
            const GravRec r_open = rec[c_open < total ? c_open : total - 1];
            const GravRec r_skip = rec[c_skip < total ? c_skip : total - 1];
            const int pt = cur.pt;
            bool open = false;
            if (c >= skip) {
                const double dx = gx - cur.mx;
                const double dy = gy - cur.my;
                const double dz = gz - cur.mz;
                const double r2 = dx * dx + dy * dy + dz * dz;
                bool use = true;
                if (pt < 0) {
                    const double w = cur.w;
                    if (w * w > theta2 * r2) {
                        open = true;
                        use = false;
                    } else {
                        skip = c_skip;
                    }
                } else if (pt == me) {
                    use = false;
                }
                if (use) {
                    const double _r = sqrt(r2 + soft2);
                    const double prefact = -G / (_r * _r * _r) * cur.m;
                    a_x += prefact * dx;
                    a_y += prefact * dy;
                    a_z += prefact * dz;
                    inter++;
                }
            }
            const bool any = __ballot(open) != 0ull;
            c = any ? c_open : c_skip;
            cur = any ? r_open : r_skip;
        }
    }
    if (live) {
        ax[me] = a_x;
        ay[me] = a_y;
        az[me] = a_z;
    }
    if (counter) {
        // per-wave reduction of the interaction count
        for (int off = 32; off > 0; off >>= 1) inter += __shfl_down(inter, off);
        if (threadIdx.x == 0) atomicAdd(counter, inter);
    }
}

bool gravity_tree(struct reb_simulation* r) {
    DeviceState* d = dev_of(r);
    // gravity.c:54-56: wrap into the box before building
    if (!boundary_device(r)) return false;
    if (!tree_build(r)) return false;
    TreeBuffers* T = d->tree;
    const int n = (int)r->N;
    if (n == 0) return true;
    auto gbv = ghostboxes(r, r->N_ghost_x, r->N_ghost_y, r->N_ghost_z);
    if ((int)gbv.size() > kMaxGBT) {
        reb_simulation_error(r, "Too many ghost boxes for tree gravity (at most 64).");
        return false;
    }
    GBT gl;
    gl.n = (int)gbv.size();
    for (int k = 0; k < gl.n; k++) {
        gl.g[k][0] = gbv[k].x;
        gl.g[k][1] = gbv[k].y;
        gl.g[k][2] = gbv[k].z;
        gl.g[k][3] = gbv[k].vx;
        gl.g[k][4] = gbv[k].vy;
        gl.g[k][5] = gbv[k].vz;
    }
    unsigned long long* counter = reinterpret_cast<unsigned long long*>(T->info.as<int>() + 8);
    if (!RB_HIP(r, hipMemsetAsync(counter, 0, 8, d->stream))) return false;
    hipLaunchKernelGGL(k_tree_gravity, dim3((n + 63) / 64), dim3(64), 0, d->stream, T->grec.as<GravRec>(),
                       T->nnodes, T->idx.as<int>(), n, d->field(FX),
                       d->field(FY), d->field(FZ), gl, r->G, r->softening * r->softening, r->opening_angle2,
                       d->field(FAX), d->field(FAY), d->field(FAZ), counter);
    mark_device_newer(r);
    return RB_HIP(r, hipGetLastError());
}

// ---------------------------------------------------------------- collision search
// Largest and second-largest radius (simulation.c:718-799); only the second
// value is used by the search.  Value-based, so independent of the order.
__global__ __launch_bounds__(256) void k_top2(const double* __restrict__ rad, int n, double* __restrict__ out) {
    __shared__ double s1[256], s2[256];
    double a1 = -1.0, a2 = -1.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const double v = rad[i];
        if (v > a1) {
            a2 = a1;
            a1 = v;
        } else if (v > a2) {
            a2 = v;
        }
    }
    s1[threadIdx.x] = a1;
    s2[threadIdx.x] = a2;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b1 = -1.0, b2 = -1.0;
        for (int k = 0; k < 256; k++) {
            const double vs[2] = {s1[k], s2[k]};
            for (double v : vs) {
                if (v > b1) {
                    b2 = b1;
                    b1 = v;
                } else if (v > b2) {
                    b2 = v;
                }
            }
        }
        out[2 * blockIdx.x] = b1;
        out[2 * blockIdx.x + 1] = b2;
    }
}

// The second largest radius over the blocks' top two (simulation.c:718-799),
// left on the device for the search (0 for fewer than two bodies).
__global__ void k_top2_final(const double* __restrict__ top, int nb, int n, double* __restrict__ second) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double b1 = -1.0, b2 = -1.0;
    for (int k = 0; k < 2 * nb; k++) {
        const double v = top[k];
        if (v > b1) {
            b2 = b1;
            b1 = v;
        } else if (v > b2) {
            b2 = v;
        }
    }
    *second = n >= 2 ? b2 : 0.;
}

// Swept-sphere test of LINE and LINETREE (collision.c:150-171, :510-534):
// the closest distance of two bodies moving on straight lines over the last
// step (dt_last_done), from the separation d1 and relative velocity dv now,
// against the sum of their radii.  MIN as collision.c:43 spells it.
__device__ __forceinline__ bool line_overlap(double dx1, double dy1, double dz1, double dvx1, double dvy1,
                                             double dvz1, double dt, double rsum) {
    const double r1 = (dx1 * dx1 + dy1 * dy1 + dz1 * dz1);
    const double dx2 = dx1 - dt * dvx1;
    const double dy2 = dy1 - dt * dvy1;
    const double dz2 = dz1 - dt * dvz1;
    const double r2 = (dx2 * dx2 + dy2 * dy2 + dz2 * dz2);
    const double t_closest = (dx1 * dvx1 + dy1 * dvy1 + dz1 * dvz1) / (dvx1 * dvx1 + dvy1 * dvy1 + dvz1 * dvz1);
    double rmin2_ab = r1 > r2 ? r2 : r1;
    if (t_closest / dt >= 0. && t_closest / dt <= 1.) {
        const double dx3 = dx1 - t_closest * dvx1;
        const double dy3 = dy1 - t_closest * dvy1;
        const double dz3 = dz1 - t_closest * dvz1;
        const double r3 = (dx3 * dx3 + dy3 * dy3 + dz3 * dz3);
        rmin2_ab = rmin2_ab > r3 ? r3 : rmin2_ab;
    }
    return !(rmin2_ab > rsum * rsum);
}

// Largest |v|^2 of each block's bodies (collision.c:273-277); the host takes
// the maximum over blocks (a maximum is exact in any order).
__global__ __launch_bounds__(256) void k_vmax2(const double* __restrict__ vx, const double* __restrict__ vy,
                                               const double* __restrict__ vz, int n, double* __restrict__ out) {
    __shared__ double red[256];
    double best = 0.;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const double v2 = vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i];
        best = best > v2 ? best : v2;
    }
    red[threadIdx.x] = best;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const double o = red[threadIdx.x + w];
            red[threadIdx.x] = red[threadIdx.x] > o ? red[threadIdx.x] : o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// Record of one detected collision: p1, p2, ghost box index, root box index.
struct CollRec {
    int p1, p2, gb, ri;
};

// Records kept per body by the single-pass search; the bodies with more
// collisions are searched again by a pass that walks only them.
static constexpr int kCollSlots = 32;

// Tree search of collision.c:197-269 + :422-503 (TREE) or :270-331 +
// :506-566 (LINETREE, LINE = true), in detection order per body (ghost box,
// root box, depth-first).  MODE 0 counts the collisions of each body; MODE 1
// stores them at the body's offset; MODE 2 counts them and keeps the first
// kCollSlots in the body's slots, raising *overflow when a body has more;
// MODE 3 stores at their offsets the collisions of the bodies whose count
// exceeds kCollSlots, and a wave without such a body returns at once.
// TREE passes the second largest radius in device memory (`reach_d`); LINETREE
// passes `reach`, the largest
// drift of the last step, dt_last_done*|v|max.
template <int MODE, bool LINE = false>
__global__ __launch_bounds__(64) void k_tree_collide(const CollCell* __restrict__ cells, int total,
                                                     const int* __restrict__ idx, int n,
                                                     const double* __restrict__ x, const double* __restrict__ y,
                                                     const double* __restrict__ z, const double* __restrict__ vx,
                                                     const double* __restrict__ vy, const double* __restrict__ vz,
                                                     const double* __restrict__ rad, GBT gl, double reach,
                                                     int* __restrict__ counts, const int* __restrict__ offsets,
                                                     CollRec* __restrict__ rec, double dt_last = 0.,
                                                     int* __restrict__ overflow = nullptr,
                                                     const double* __restrict__ reach_d = nullptr) {
    if (reach_d) reach = *reach_d;
    constexpr bool WRITE = MODE == 1 || MODE == 3;
    const int s = blockIdx.x * 64 + threadIdx.x;
    const int me = s < n ? idx[s] : -1;
    const bool live = s < n && (MODE != 3 || counts[me] > kCollSlots);
    if (MODE == 3 && __ballot(live) == 0ull) return;   // one wave per block: uniform exit
    const double px = live ? x[me] : 0., py = live ? y[me] : 0., pz = live ? z[me] : 0.;
    const double pvx = live ? vx[me] : 0., pvy = live ? vy[me] : 0., pvz = live ? vz[me] : 0.;
    const double p1_r = live ? rad[me] : 0.;
    // LINETREE: the radius grown by the body's own drift (collision.c:293)
    const double p1_r_plus_dtv = LINE ? p1_r + dt_last * sqrt(pvx * pvx + pvy * pvy + pvz * pvz) : 0.;
    int found = 0;
    int out = (WRITE && live) ? offsets[me] : 0;
    CollRec* const slot = MODE == 2 && live ? rec + (size_t)me * kCollSlots : nullptr;
    auto emit = [&](int pt, int g, int ri) {
        if (WRITE) rec[out++] = CollRec{me, pt, g, ri};
        if (MODE == 2 && found < kCollSlots) slot[found] = CollRec{me, pt, g, ri};
        found++;
    };
    for (int g = 0; g < gl.n; g++) {
        const double gx = gl.g[g][0] + px, gy = gl.g[g][1] + py, gz = gl.g[g][2] + pz;
        const double gvx = gl.g[g][3] + pvx, gvy = gl.g[g][4] + pvy, gvz = gl.g[g][5] + pvz;
        int skip = live ? 0 : 0x7fffffff;
        int c = 0;
        int ri = -1;
        CollCell cur = cells[0];
        while (c < total) {
            // both possible successors are fetched before this cell is evaluated
            const int c_open = c + 1, nx = cur.next;
            const CollCell r_open = cells[c_open < total ? c_open : total - 1];
            const CollCell r_skip = cells[nx < total ? nx : total - 1];
            const int pt = cur.pt;
            bool open = false;
            if (c >= skip) {
                if (pt >= 0) {
                    if (pt != me) {   // the leaf's record holds its body (k_tree_pack)
                        const double dx = gx - cur.cx;
                        const double dy = gy - cur.cy;
                        const double dz = gz - cur.cz;
                        if (LINE) {
                            if (line_overlap(dx, dy, dz, gvx - cur.v[0], gvy - cur.v[1], gvz - cur.v[2], dt_last,
                                             p1_r + cur.w))
                                emit(pt, g, ri);
                        } else {
                            const double r2 = dx * dx + dy * dy + dz * dz;
                            const double rp = p1_r + cur.w;
                            if (!(r2 > rp * rp)) {
                                const double dvx = gvx - cur.v[0];
                                const double dvy = gvy - cur.v[1];
                                const double dvz = gvz - cur.v[2];
                                if (!(dvx * dx + dvy * dy + dvz * dz > 0)) emit(pt, g, ri);
                            }
                        }
                    }
                } else {
                    const double dx = gx - cur.cx;
                    const double dy = gy - cur.cy;
                    const double dz = gz - cur.cz;
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    const double rp = LINE ? p1_r_plus_dtv + reach + 0.86602540378443 * cur.w
                                           : p1_r + reach + 0.86602540378443 * cur.w;
                    if (r2 < rp * rp)
                        open = true;
                    else
                        skip = nx;
                }
            }
            const bool any = __ballot(open) != 0ull;
            c = any ? c_open : nx;
            cur = any ? r_open : r_skip;
        }
    }
    if ((MODE == 0 || MODE == 2) && live) counts[me] = found;
    if (MODE == 2 && live && found > kCollSlots) *overflow = 1;   // same value from every writer
}

// The slot records of every body with at most kCollSlots collisions to its
// offset; the others are written by k_tree_collide<3>.
__global__ void k_coll_compact(const int* __restrict__ counts, const int* __restrict__ offsets,
                               const CollRec* __restrict__ slots, int n, CollRec* __restrict__ rec) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = counts[i], o = offsets[i];
    if (c > kCollSlots) return;
    for (int k = 0; k < c; k++) rec[o + k] = slots[(size_t)i * kCollSlots + k];
}

// Root box of each record: the reference stores the root box being searched.
// Recomputed on the host from p2's position is not exact for wrapped bodies;
// the search kernel cannot cheaply know which root subtree it is in, so the
// root index comes from the leaf's key (same value: the leaf lies in that root).
__global__ void k_coll_root(CollRec* __restrict__ rec, int nrec, const unsigned long long* __restrict__ key0,
                            const int* __restrict__ pos_of, int L1) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nrec) return;
    rec[k].ri = (int)(key0[pos_of[rec[k].p2]] >> (3 * L1));
}

__global__ void k_inverse_perm(const int* __restrict__ idx, int* __restrict__ pos_of, int n) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) pos_of[idx[s]] = s;
}

// ---------------------------------------------------------------- direct search
// collision.c:64-124 (DIRECT): ghost box outer, projectile i, target j
// ascending.  collision.c:125-196 (LINE = true): the same loops over the
// pairs j > i with the swept-sphere test.
template <bool WRITE, bool LINE = false>
__global__ __launch_bounds__(256) void k_direct_collide(const double* __restrict__ x, const double* __restrict__ y,
                                                        const double* __restrict__ z, const double* __restrict__ vx,
                                                        const double* __restrict__ vy, const double* __restrict__ vz,
                                                        const double* __restrict__ rad, int n, int n_targets,
                                                        GBT gl, int* __restrict__ counts,
                                                        const int* __restrict__ offsets, CollRec* __restrict__ rec,
                                                        double dt_last = 0.) {
    __shared__ double4 tp[256];
    __shared__ double4 tv[256];
    const int g = blockIdx.y;
    const int i = blockIdx.x * 256 + threadIdx.x;
    const bool live = i < n;
    const double gx = live ? gl.g[g][0] + x[i] : 0., gy = live ? gl.g[g][1] + y[i] : 0.,
                 gz = live ? gl.g[g][2] + z[i] : 0.;
    const double gvx = live ? gl.g[g][3] + vx[i] : 0., gvy = live ? gl.g[g][4] + vy[i] : 0.,
                 gvz = live ? gl.g[g][5] + vz[i] : 0.;
    const double p1_r = live ? rad[i] : 0.;
    int found = 0;
    int out = (WRITE && live) ? offsets[(size_t)g * n + i] : 0;
    for (int t = 0; t < n_targets; t += 256) {
        __syncthreads();
        const int jj = t + threadIdx.x;
        if (jj < n_targets) {
            tp[threadIdx.x] = make_double4(x[jj], y[jj], z[jj], rad[jj]);
            tv[threadIdx.x] = make_double4(vx[jj], vy[jj], vz[jj], 0.);
        }
        __syncthreads();
        if (!live) continue;
        const int m = min(256, n_targets - t);
        for (int k = 0; k < m; k++) {
            const int j = t + k;
            if (LINE ? j <= i : i == j) continue;
            const double4 q = tp[k];
            const double dx = gx - q.x, dy = gy - q.y, dz = gz - q.z;
            if (LINE) {
                const double4 qv = tv[k];
                if (line_overlap(dx, dy, dz, gvx - qv.x, gvy - qv.y, gvz - qv.z, dt_last, p1_r + q.w)) {
                    if (WRITE) rec[out++] = CollRec{i, j, g, 0};
                    found++;
                }
                continue;
            }
            const double sr = p1_r + q.w;
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > sr * sr) continue;
            const double4 qv = tv[k];
            const double dvx = gvx - qv.x, dvy = gvy - qv.y, dvz = gvz - qv.z;
            if (dvx * dx + dvy * dy + dvz * dz > 0) continue;
            if (WRITE) rec[out++] = CollRec{i, j, g, 0};
            found++;
        }
    }
    if (!WRITE && live) counts[(size_t)g * n + i] = found;
}

static bool ensure_coll(struct reb_simulation* r, TreeBuffers* T, size_t nslots, size_t nrec) {
    if (!T->coll_cnt.ensure(4 * (nslots + 1)) || !T->coll_off.ensure(4 * (nslots + 1))) return false;
    if (nrec > T->rec_cap || !T->coll_rec.p) {
        size_t cap = T->rec_cap ? T->rec_cap : 4096;
        while (cap < nrec) cap *= 2;
        if (!T->coll_rec.ensure(sizeof(CollRec) * cap)) {
            reb_simulation_error(r, "Out of device memory for the collision list.");
            return false;
        }
        T->rec_cap = cap;
    }
    return true;
}

// Exclusive scan of counts -> offsets; returns the total.
static long scan_counts(struct reb_simulation* r, TreeBuffers* T, int nslots) {
    DeviceState* d = dev_of(r);
    size_t tmp = 0;
    (void)exclusive_sum(nullptr, tmp, T->coll_cnt.as<int>(), T->coll_off.as<int>(), nslots + 1, d->stream);
    if (!T->prim_tmp.ensure(tmp + 256)) return -1;
    // counts[nslots] = 0 so that offsets[nslots] is the total
    if (!RB_HIP(r, hipMemsetAsync(T->coll_cnt.as<int>() + nslots, 0, 4, d->stream))) return -1;
    if (!RB_HIP(r, exclusive_sum(T->prim_tmp.p, tmp, T->coll_cnt.as<int>(), T->coll_off.as<int>(), nslots + 1,
                                 d->stream)))
        return -1;
    int total = 0;
    if (!RB_HIP(r, hipMemcpyAsync(&total, T->coll_off.as<int>() + nslots, 4, hipMemcpyDeviceToHost, d->stream)))
        return -1;
    if (!stream_sync(r)) return -1;
    return total;
}

// The device records of a search into pinned host memory; r->collisions is
// written from them by materialize_collisions when a resolve needs it.
static bool fetch_collisions(struct reb_simulation* r, TreeBuffers* T, long total, const GBT& gl) {
    DeviceState* d = dev_of(r);
    r->N_collisions = 0;
    if (!d->coll) d->coll = new CollList();
    CollList& L = *d->coll;
    L.n = 0;
    L.pending = false;
    if (total <= 0) return true;
    if ((size_t)total > L.cap) {
        size_t cap = L.cap ? L.cap : 4096;
        while (cap < (size_t)total) cap *= 2;
        if (L.rec) (void)hipHostFree(L.rec);
        L.rec = nullptr;
        L.cap = 0;
        if (!RB_HIP(r, hipHostMalloc((void**)&L.rec, sizeof(CollRecH) * cap, hipHostMallocDefault))) return false;
        L.cap = cap;
    }
    static_assert(sizeof(CollRecH) == sizeof(CollRec), "host and device collision records");
    // the compact resolve reads the records on the device; any other resolve
    // reads them on the host
    const CollRecH* drec = T->coll_rec.as<CollRecH>();
    const bool compact = compact_resolve_eligible(r, gl.n);
    L.rec_host = !compact;
    if (!compact &&
        !RB_HIP(r, hipMemcpyAsync(L.rec, T->coll_rec.p, sizeof(CollRec) * total, hipMemcpyDeviceToHost, d->stream)))
        return false;
    if (compact && !compact_resolve_mark(r, drec, (size_t)total)) return false;
    L.gb.resize(gl.n);
    for (int k = 0; k < gl.n; k++)
        L.gb[k] = reb_vec6d{gl.g[k][0], gl.g[k][1], gl.g[k][2], gl.g[k][3], gl.g[k][4], gl.g[k][5]};
    if (!stream_sync(r)) return false;
    L.n = (size_t)total;
    L.pending = true;
    r->N_collisions = (size_t)total;
    if (compact && !compact_resolve_gather(r, drec, (size_t)total)) return false;
    return true;
}

bool materialize_collisions(struct reb_simulation* r) {
    DeviceState* d = dev_of(r);
    if (!d->coll || !d->coll->pending) return true;
    CollList& L = *d->coll;
    if (!L.rec_host) {   // prepared for the compact resolve: the records are still on the device
        if (!RB_HIP(r, hipMemcpyAsync(L.rec, d->tree->coll_rec.p, sizeof(CollRecH) * L.n, hipMemcpyDeviceToHost,
                                      d->stream)) ||
            !stream_sync(r))
            return false;
        L.rec_host = true;
    }
    if (!collisions_reserve(r, L.n)) return false;
    for (size_t k = 0; k < L.n; k++) {
        const CollRecH& h = L.rec[k];
        struct reb_collision c;
        c.p1 = (size_t)h.p1;
        c.p2 = (size_t)h.p2;
        c.gb = L.gb[h.gb];
        c.ri = (size_t)h.ri;
        r->collisions[k] = c;
    }
    L.pending = false;
    return true;
}

void coll_release(struct reb_simulation* r) {
    DeviceState* d = dev_of(r);
    if (!d->coll) return;
    if (d->coll->rec) (void)hipHostFree(d->coll->rec);
    delete d->coll;
    d->coll = nullptr;
}

// TREE / LINETREE search into T->coll_rec in detection order; returns the
// number of collisions or -1.  One pass keeps up to kCollSlots records per
// body; when a body has more, a second pass searches again for those bodies.
template <bool LINE>
static long tree_collide_run(struct reb_simulation* r, TreeBuffers* T, int n, const GBT& gl, double reach,
                             double dt_last, const double* reach_d = nullptr) {
    DeviceState* d = dev_of(r);
    if (!ensure_coll(r, T, n, 1)) return -1;
    if (!T->coll_slot.ensure(sizeof(CollRec) * kCollSlots * (size_t)n)) {
        reb_simulation_error(r, "Out of device memory for the collision search.");
        return -1;
    }
    int* overflow = T->info.as<int>() + 12;
    if (!RB_HIP(r, hipMemsetAsync(overflow, 0, sizeof(int), d->stream))) return -1;
    const dim3 grid((n + 63) / 64);
    hipLaunchKernelGGL((k_tree_collide<2, LINE>), grid, dim3(64), 0, d->stream, T->crec.as<CollCell>(), T->nnodes,
                       T->idx.as<int>(), n, d->field(FX), d->field(FY), d->field(FZ), d->field(FVX), d->field(FVY),
                       d->field(FVZ), d->field(FR), gl, reach, T->coll_cnt.as<int>(), nullptr,
                       T->coll_slot.as<CollRec>(), dt_last, overflow, reach_d);
    // while the device searches: the shuffle's random draws, sized by the
    // last search (the resolve draws any further ones itself)
    if (compact_resolve_eligible(r, gl.n)) compact_resolve_draw_ahead(r, T->last_total + T->last_total / 4 + 64);
    int hover = 0;
    if (!RB_HIP(r, hipMemcpyAsync(&hover, overflow, sizeof(int), hipMemcpyDeviceToHost, d->stream))) return -1;
    const long total = scan_counts(r, T, n);   // synchronizes: hover is current
    if (total < 0) return -1;
    T->last_total = (size_t)total;
    if (!ensure_coll(r, T, n, (size_t)total + 1)) return -1;
    if (total == 0) return 0;
    hipLaunchKernelGGL(k_coll_compact, dim3((n + 255) / 256), dim3(256), 0, d->stream, T->coll_cnt.as<int>(),
                       T->coll_off.as<int>(), T->coll_slot.as<CollRec>(), n, T->coll_rec.as<CollRec>());
    if (hover)
        hipLaunchKernelGGL((k_tree_collide<3, LINE>), grid, dim3(64), 0, d->stream, T->crec.as<CollCell>(),
                           T->nnodes, T->idx.as<int>(), n, d->field(FX), d->field(FY), d->field(FZ), d->field(FVX),
                           d->field(FVY), d->field(FVZ), d->field(FR), gl, reach, T->coll_cnt.as<int>(),
                           T->coll_off.as<int>(), T->coll_rec.as<CollRec>(), dt_last, nullptr, reach_d);
    // root index of each record (the lcp buffer is free now: reuse as pos_of)
    int* pos_of = T->lcp.as<int>();
    hipLaunchKernelGGL(k_inverse_perm, dim3((n + 255) / 256), dim3(256), 0, d->stream, T->idx.as<int>(), pos_of, n);
    hipLaunchKernelGGL(k_coll_root, dim3((total + 255) / 256), dim3(256), 0, d->stream, T->coll_rec.as<CollRec>(),
                       (int)total, T->key0.as<unsigned long long>(), pos_of, T->L1);
    return total;
}

bool collision_search_device(struct reb_simulation* r) {
    DeviceState* d = dev_of(r);
    r->N_collisions = 0;
    if (!ensure_device(r)) return false;
    const int n = (int)r->N;
    if (n == 0) return true;
    if (r->map) {
        reb_simulation_error(r, "r->map is not supported by the collision search of this engine.");
        return false;
    }
    // inner ring of ghost boxes (collision.c:66-69, :204-206)
    const int gx = r->N_ghost_x > 1 ? 1 : r->N_ghost_x;
    const int gy = r->N_ghost_y > 1 ? 1 : r->N_ghost_y;
    const int gz = r->N_ghost_z > 1 ? 1 : r->N_ghost_z;
    auto gbv = ghostboxes(r, gx, gy, gz);
    GBT gl;
    gl.n = (int)gbv.size();
    for (int k = 0; k < gl.n; k++) {
        gl.g[k][0] = gbv[k].x;
        gl.g[k][1] = gbv[k].y;
        gl.g[k][2] = gbv[k].z;
        gl.g[k][3] = gbv[k].vx;
        gl.g[k][4] = gbv[k].vy;
        gl.g[k][5] = gbv[k].vz;
    }
    if (!d->tree) d->tree = new TreeBuffers();
    TreeBuffers* T = d->tree;
    switch (r->collision) {
        case REB_COLLISION_DIRECT: {
            const int n_targets = (r->N_targets != SIZE_MAX) ? (int)r->N_targets : n;
            const int nslots = gl.n * n;
            if (!ensure_coll(r, T, nslots, 1)) return false;
            dim3 grid((n + 255) / 256, gl.n);
            hipLaunchKernelGGL(k_direct_collide<false>, grid, dim3(256), 0, d->stream, d->field(FX), d->field(FY),
                               d->field(FZ), d->field(FVX), d->field(FVY), d->field(FVZ), d->field(FR), n, n_targets,
                               gl, T->coll_cnt.as<int>(), nullptr, nullptr);
            const long total = scan_counts(r, T, nslots);
            if (total < 0) return false;
            if (!ensure_coll(r, T, nslots, (size_t)total + 1)) return false;
            if (total > 0)
                hipLaunchKernelGGL(k_direct_collide<true>, grid, dim3(256), 0, d->stream, d->field(FX), d->field(FY),
                                   d->field(FZ), d->field(FVX), d->field(FVY), d->field(FVZ), d->field(FR), n,
                                   n_targets, gl, nullptr, T->coll_off.as<int>(), T->coll_rec.as<CollRec>());
            return fetch_collisions(r, T, total, gl);
        }
        case REB_COLLISION_TREE: {
            if (!tree_build(r)) return false;
            // second largest radius, left on the device for the search
            double* top = reinterpret_cast<double*>(T->info.as<int>() + 16);
            double* second_r = top + 2 * kTop2Blocks;
            const int nb = std::min(kTop2Blocks, (n + 255) / 256);
            hipLaunchKernelGGL(k_top2, dim3(nb), dim3(256), 0, d->stream, d->field(FR), n, top);
            hipLaunchKernelGGL(k_top2_final, dim3(1), dim3(64), 0, d->stream, top, nb, n, second_r);
            const long total = tree_collide_run<false>(r, T, n, gl, 0., 0., second_r);
            if (total < 0) return false;
            return fetch_collisions(r, T, total, gl);
        }
        case REB_COLLISION_LINE: {
            // collision.c:125-196: pairs j > i of all bodies (no N_targets)
            const int nslots = gl.n * n;
            const double dt_last = r->dt_last_done;
            if (!ensure_coll(r, T, nslots, 1)) return false;
            dim3 grid((n + 255) / 256, gl.n);
            hipLaunchKernelGGL((k_direct_collide<false, true>), grid, dim3(256), 0, d->stream, d->field(FX),
                               d->field(FY), d->field(FZ), d->field(FVX), d->field(FVY), d->field(FVZ), d->field(FR), n,
                               n, gl, T->coll_cnt.as<int>(), nullptr, nullptr, dt_last);
            const long total = scan_counts(r, T, nslots);
            if (total < 0) return false;
            if (!ensure_coll(r, T, nslots, (size_t)total + 1)) return false;
            if (total > 0)
                hipLaunchKernelGGL((k_direct_collide<true, true>), grid, dim3(256), 0, d->stream, d->field(FX),
                                   d->field(FY), d->field(FZ), d->field(FVX), d->field(FVY), d->field(FVZ),
                                   d->field(FR), n, n, gl, nullptr, T->coll_off.as<int>(), T->coll_rec.as<CollRec>(),
                                   dt_last);
            return fetch_collisions(r, T, total, gl);
        }
        case REB_COLLISION_LINETREE: {
            // collision.c:270-331: the largest drift of the last step widens
            // every cell test
            const double dt_last = r->dt_last_done;
            const int nb = std::min(kTop2Blocks, (n + 255) / 256);
            if (!T->info.ensure(kInfoBytes)) return false;
            double* top = reinterpret_cast<double*>(T->info.as<int>() + 16);
            hipLaunchKernelGGL(k_vmax2, dim3(nb), dim3(256), 0, d->stream, d->field(FVX), d->field(FVY),
                               d->field(FVZ), n, top);
            std::vector<double> h(nb);
            if (!RB_HIP(r, hipMemcpyAsync(h.data(), top, sizeof(double) * nb, hipMemcpyDeviceToHost, d->stream)))
                return false;
            if (!stream_sync(r)) return false;
            double vmax2 = 0.;
            for (double v : h) vmax2 = vmax2 > v ? vmax2 : v;
            const double maxdrift = dt_last * sqrt(vmax2);
            if (!tree_build(r)) return false;
            const long total = tree_collide_run<true>(r, T, n, gl, maxdrift, dt_last);
            if (total < 0) return false;
            return fetch_collisions(r, T, total, gl);
        }
        default:
            return true;
    }
}

void tree_release(struct reb_simulation* r) {
    DeviceState* d = dev_of(r);
    if (!d->tree) return;
    TreeBuffers* T = d->tree;
    for (DevBuf* b : {&T->key0, &T->key1, &T->idx, &T->key0b, &T->key1b, &T->idxb, &T->lcp, &T->cnt, &T->base,
                      &T->prim_tmp, &T->node_d, &T->node_i, &T->info, &T->coll_cnt, &T->coll_off, &T->coll_rec,
                      &T->coll_slot, &T->grec, &T->crec})
        b->release();
    delete T;
    d->tree = nullptr;
}

// Tree dump for tests (pre-order rows x y z w m mx my mz pt, root order).
int tree_dump_device(struct reb_simulation* r, double* out, long cap) {
    DeviceState* d = dev_of(r);
    TreeBuffers* T = d->tree;
    if (!T) return -1;
    const int total = T->nnodes;
    if (!out) return total;
    std::vector<double> hd(8 * (size_t)total);
    std::vector<int> hi(2 * (size_t)total);
    for (int f = 0; f < 8; f++)
        (void)hipMemcpyAsync(hd.data() + (size_t)f * total, T->node_d.as<double>() + (size_t)f * T->node_cap,
                             sizeof(double) * total, hipMemcpyDeviceToHost, d->stream);
    (void)hipMemcpyAsync(hi.data(), T->node_i.as<int>() + T->node_cap, sizeof(int) * total, hipMemcpyDeviceToHost,
                         d->stream);
    (void)hipStreamSynchronize(d->stream);
    const long nrow = std::min<long>(cap, total);
    for (long k = 0; k < nrow; k++) {
        for (int f = 0; f < 8; f++) out[9 * k + f] = hd[(size_t)f * total + k];
        out[9 * k + 8] = (double)hi[k];
    }
    return total;
}

}  // namespace rbamd

extern "C" {

// Test hook: builds the tree of the simulation's current particles on the GPU
// and copies the cells out in pre-order.  Returns the number of cells, or -1.
REB_API long reb_amd_tree_dump(struct reb_simulation* r, double* out, long cap) {
    using namespace rbamd;
    mark_host_newer(r);
    if (!ensure_device(r)) return -1;
    if (!tree_build(r)) return -1;
    return tree_dump_device(r, out, cap);
}

// Particle-cell and particle-particle interactions summed by the last tree
// gravity evaluation (over all ghost boxes), or -1 if there was none.
REB_API long long reb_amd_tree_interactions(struct reb_simulation* r) {
    using namespace rbamd;
    DeviceState* d = dev_of(r);
    if (!d->tree || !d->stream) return -1;
    unsigned long long v = 0;
    const auto* counter = reinterpret_cast<const unsigned long long*>(d->tree->info.as<int>() + 8);
    if (!RB_HIP(r, hipMemcpyAsync(&v, counter, sizeof(v), hipMemcpyDeviceToHost, d->stream))) return -1;
    if (!stream_sync(r)) return -1;
    return (long long)v;
}

}  // extern "C"
