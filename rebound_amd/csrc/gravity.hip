// Direct O(N^2) gravity on the GPU (reference: gravity.c:167-531) and the
// energy diagnostic (tools.c:108-162).
//
// Three precisions (rebound_amd.h):
//   EXACT  one lane per target i, sources j streamed through LDS tiles in
//          ascending order, IEEE sqrt/div and no contraction (the library is
//          built with -ffp-contract=off): every term and every partial sum is
//          the reference's OpenMP loop (gravity.c:215-232, 308-343), so the
//          result is bit-identical.
//   FP64   r^-3 from v_rsq_f64 refined to full double accuracy, fused
//          multiply-adds, each LDS tile's sum folded into the running sum
//          with Kahan compensation, and the source range split over
//          blockIdx.y so that N=65536 fills every SIMD; the partial sums are
//          added in a fixed order by a second kernel (deterministic).
//   SPLIT  displacement and r^2 in FP64, r^-1 from the FP32 reciprocal
//          square root (v_rsq_f32, about 2^-23 relative), cubed and
//          multiplied by m_j in FP64 (no FP32 exponent limit on r^-3 or m),
//          accumulation in FP64; each LDS tile's FP64 partial is folded into
//          the running sum with Kahan compensation.
//
// Data layout: structure of arrays (x, y, z, m as separate double arrays).
// Sources are staged per 256-wide tile into LDS as double4 (x, y, z, m) and
// read back as broadcasts (every lane of a wavefront reads the same entry).
#include <cmath>
#include <mutex>
#include <string>

#include "internal.h"

namespace rbamd {

static constexpr int kBlock = 256;     // targets per workgroup = sources per LDS tile (exact kernels)

// Shape of the fast kernels (tuning knobs, fixed at build time): targets per
// workgroup (= sources per LDS tile), unroll of the pair loop, and wavefronts
// per SIMD the register budget is sized for.
#ifndef RB_FAST_BLOCK
#define RB_FAST_BLOCK 256
#endif
#ifndef RB_FAST_UNROLL
#define RB_FAST_UNROLL 4
#endif
#ifndef RB_FAST_WAVES
#define RB_FAST_WAVES 4
#endif
#define RB_STR(x) #x
#define RB_UNROLL(n) _Pragma(RB_STR(unroll n))
static constexpr int kFast = RB_FAST_BLOCK;

struct GB3 { double x, y, z; };
// Ghost-box offsets travel by value in the kernel arguments (kernarg segment,
// read with scalar loads), so no copy or synchronisation is needed per call.
static constexpr int kMaxGB = 125;       // N_ghost up to 2 in all three axes
struct GBList {
    int n;
    GB3 g[kMaxGB];
};

// The leapfrog's kick and closing drift, applied by the pass that finishes the
// accelerations (on == 0: nothing).
struct KickDrift {
    double *x, *y, *z, *vx, *vy, *vz;
    double kick, drift;
    int on;
};

__device__ __forceinline__ bool skip_pair(int i, int j, int ignore) {
    if (ignore == 1 && ((j == 1 && i == 0) || (i == 1 && j == 0))) return true;
    if (ignore == 2 && (j == 0 || i == 0)) return true;
    return i == j;
}

__device__ __forceinline__ void kahan_add(double term, double& a, double& c) {
    const double yv = term - c;
    const double tv = a + yv;
    c = (tv - a) - yv;
    a = tv;
}

// ---------------------------------------------------------------- EXACT
// BASIC: gravity.c:217-231 (targets 0..N-1, sources 0..N_active-1 for every
// ghost box), then :259-276 (active targets also feel test particles when
// testparticle_type is set).  COMPENSATED: gravity.c:308-343 (active
// targets), :346-378 (test-particle targets), :379-410 (active targets feel
// test particles), no ghost boxes, Kahan sums.
//
// The terms are independent: all 256 threads of a workgroup compute them for
// TGT targets x TILE sources at a time (IEEE sqrt and division, no
// contraction: the reference's expressions) into LDS rows, one row per
// (target, component), and one lane per row then adds its row in ascending
// source order: the reference's sequential sum, each component its own
// chain.  The rows are double-buffered, so one barrier per tile suffices and
// the summing lanes' chains overlap the other waves' terms of the next tile;
// each thread loads its sources one tile ahead.
static constexpr int kExThreads = 256;
static constexpr int kExactRowwiseN = 49152;   // from here on the row-wise kernels

template <bool COMP, int TGT, int TILE>
__global__ __launch_bounds__(kExThreads) void k_exact(const double* __restrict__ x, const double* __restrict__ y,
                                                       const double* __restrict__ z, const double* __restrict__ m,
                                                       double* __restrict__ ax, double* __restrict__ ay,
                                                       double* __restrict__ az, int N, int N_active, int tp_type,
                                                       int ignore, double G, double soft2, GBList gl) {
    constexpr int PER = TGT * TILE / kExThreads;   // terms per thread and tile
    constexpr int ROWS = 3 * TGT;
    constexpr int PITCH = TILE + 2;                // row pitch (doubles): 16-byte aligned rows, spread banks
    static_assert(PER >= 1 && PER * kExThreads == TGT * TILE, "tile shape");
    static_assert(ROWS <= kExThreads, "one summing lane per row");
    __shared__ __align__(16) double T[2][ROWS][PITCH];
    const int tid = threadIdx.x;
    const int i0 = blockIdx.x * TGT;
    // term phase: thread -> (target ti, sources tj0 .. tj0 + PER - 1 of a tile)
    const int ti = tid % TGT;
    const int tj0 = (tid / TGT) * PER;
    const int i = i0 + ti;
    const bool live = i < N;
    const double xi = live ? x[i] : 0., yi = live ? y[i] : 0., zi = live ? z[i] : 0.;
    // sum phase: lanes 0 .. ROWS-1, lane -> (target si, component sc)
    const int si = i0 + tid / 3;
    const int sc = tid % 3;
    const bool summer = tid < ROWS && si < N;
    double acc = 0., comp = 0.;
    int buf = 0;
    const int ngb = COMP ? 1 : gl.n;
    for (int g = 0; g < ngb; g++) {
        const GB3 gb = COMP ? GB3{0., 0., 0.} : gl.g[g];
        for (int pass = 0; pass < 2; pass++) {
            const int j0 = pass == 0 ? 0 : N_active;
            const int j1 = pass == 0 ? N_active : N;
            if (pass == 1 && !tp_type) break;
            // the second pass only reaches active targets
            const bool tgt = live && !(pass == 1 && i >= N_active);
            const bool stgt = summer && !(pass == 1 && si >= N_active);
            auto skipped = [&](int ii, int j) {
                if (COMP || pass == 0) return skip_pair(ii, j, ignore);
                return (ignore == 1 && j == 1 && ii == 0) || (ignore == 2 && (j == 0 || ii == 0));
            };
            // this thread's sources of the next tile, loaded one tile ahead
            double nx[PER], ny[PER], nz[PER], nm[PER];
            auto fetch = [&](int t) {
#pragma unroll
                for (int q = 0; q < PER; q++) {
                    const int j = t + tj0 + q;
                    if (j < j1) {
                        nx[q] = x[j]; ny[q] = y[j]; nz[q] = z[j]; nm[q] = m[j];
                    }
                }
            };
            if (j0 < j1) fetch(j0);
            for (int t = j0; t < j1; t += TILE, buf ^= 1) {
                const int n = min(TILE, j1 - t);
                double cx[PER], cy[PER], cz[PER], cm[PER];
#pragma unroll
                for (int q = 0; q < PER; q++) {
                    cx[q] = nx[q]; cy[q] = ny[q]; cz[q] = nz[q]; cm[q] = nm[q];
                }
                if (t + TILE < j1) fetch(t + TILE);
#pragma unroll
                for (int q = 0; q < PER; q++) {
                    const int k = tj0 + q;
                    const int j = t + k;
                    if (!tgt || k >= n || skipped(i, j)) continue;
                    double tx, ty, tz;
                    if (COMP) {
                        const double dx = xi - cx[q];
                        const double dy = yi - cy[q];
                        const double dz = zi - cz[q];
                        const double r2 = dx * dx + dy * dy + dz * dz + soft2;
                        const double rr = sqrt(r2);
                        const double prefact = G / (r2 * rr);
                        const double prefactj = -prefact * cm[q];
                        tx = prefactj * dx;
                        ty = prefactj * dy;
                        tz = prefactj * dz;
                    } else {
                        const double dx = (gb.x + xi) - cx[q];
                        const double dy = (gb.y + yi) - cy[q];
                        const double dz = (gb.z + zi) - cz[q];
                        const double _r = sqrt(dx * dx + dy * dy + dz * dz + soft2);
                        const double prefact = -G / (_r * _r * _r) * cm[q];
                        tx = prefact * dx;
                        ty = prefact * dy;
                        tz = prefact * dz;
                    }
                    T[buf][3 * ti + 0][k] = tx;
                    T[buf][3 * ti + 1][k] = ty;
                    T[buf][3 * ti + 2][k] = tz;
                }
                // After this barrier the summers read T[buf]; the next writes
                // to T[buf] come two tiles later, after the next barrier, which
                // the summers reach only once this sum is done.
                __syncthreads();
                if (stgt) {
                    const double* row = T[buf][tid];
                    auto add = [&](double term) {
                        if (COMP) {
                            const double yv = term - comp;
                            const double tv = acc + yv;
                            comp = (tv - acc) - yv;
                            acc = tv;
                        } else {
                            acc += term;
                        }
                    };
                    // Pairs the reference leaves out: the target itself and,
                    // with ignored terms, pairs with bodies 0 and 1.
                    const bool check = (t <= si && si < t + n) || (ignore != 0 && (t < 2 || si < 2));
                    if (check) {
                        for (int k = 0; k < n; k++)
                            if (!skipped(si, t + k)) add(row[k]);
                    } else if (n == TILE) {
#pragma unroll 16
                        for (int k = 0; k < TILE; k++) add(row[k]);
                    } else {
                        for (int k = 0; k < n; k++) add(row[k]);
                    }
                }
            }
        }
    }
    if (summer) {
        double* out = sc == 0 ? ax : (sc == 1 ? ay : az);
        out[si] = acc;
    }
}

// ---------------------------------------------------------------- EXACT, row-wise (large N)
// One thread per target adds its terms in ascending source order; the terms
// of tiles away from the target are independent and overlap in flight.  At
// large N this beats the LDS-term kernel above, whose per-term LDS traffic
// and barriers cost more than the serial chain hides.
//
// gravity.c:217-231 (targets 0..N-1, sources 0..N_active-1 for every ghost
// box), then :259-276 (active targets also feel test particles when
// testparticle_type is set).
__global__ __launch_bounds__(kBlock) void k_basic_exact(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ m, double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az,
    int N, int N_active, int tp_type, int ignore, double G, double soft2, GBList gl) {
    __shared__ double4 tile[kBlock];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const bool live = i < N;
    const int blk_lo = blockIdx.x * kBlock, blk_hi = blk_lo + kBlock;
    const double xi = live ? x[i] : 0., yi = live ? y[i] : 0., zi = live ? z[i] : 0.;
    double a_x = 0., a_y = 0., a_z = 0.;
    for (int g = 0; g < gl.n; g++) {
        const GB3 gb = gl.g[g];
        for (int pass = 0; pass < 2; pass++) {
            int j0 = pass == 0 ? 0 : N_active;
            int j1 = pass == 0 ? N_active : N;
            if (pass == 1 && !tp_type) break;
            for (int t = j0; t < j1; t += kBlock) {
                __syncthreads();
                const int jj = t + (int)threadIdx.x;
                if (jj < j1) tile[threadIdx.x] = make_double4(x[jj], y[jj], z[jj], m[jj]);
                __syncthreads();
                const int n = min(kBlock, j1 - t);
                if (!live || (pass == 1 && i >= N_active)) continue;
                auto term = [&](const double4 s) {
                    const double dx = (gb.x + xi) - s.x;
                    const double dy = (gb.y + yi) - s.y;
                    const double dz = (gb.z + zi) - s.z;
                    const double _r = sqrt(dx * dx + dy * dy + dz * dz + soft2);
                    const double prefact = -G / (_r * _r * _r) * s.w;
                    a_x += prefact * dx;
                    a_y += prefact * dy;
                    a_z += prefact * dz;
                };
                // Tiles without this block's targets (and, with ignored terms,
                // without bodies 0 and 1 as sources or among the block's
                // targets) need no per-pair test: the terms are then
                // independent work the compiler can overlap, while the sums
                // still add them in ascending j.
                const bool check =
                    (pass == 0 && t < blk_hi && t + n > blk_lo) || (ignore != 0 && (t < 2 || blk_lo < 2));
                if (!check) {
#pragma unroll 4
                    for (int k = 0; k < n; k++) term(tile[k]);
                    continue;
                }
                for (int k = 0; k < n; k++) {
                    const int j = t + k;
                    if (pass == 0) {
                        if (skip_pair(i, j, ignore)) continue;
                    } else {
                        if (ignore == 1 && j == 1 && i == 0) continue;
                        if (ignore == 2 && (j == 0 || i == 0)) continue;
                    }
                    term(tile[k]);
                }
            }
        }
    }
    if (live) {
        ax[i] = a_x;
        ay[i] = a_y;
        az[i] = a_z;
    }
}

// COMPENSATED, row-wise
// gravity.c:308-343 (active targets), :346-378 (test-particle targets), :379-410
// (active targets feel test particles).  No ghost boxes in the reference.

__global__ __launch_bounds__(kBlock) void k_comp_exact(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ m, double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az,
    int N, int N_active, int tp_type, int ignore, double G, double soft2) {
    __shared__ double4 tile[kBlock];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const bool live = i < N;
    const int blk_lo = blockIdx.x * kBlock, blk_hi = blk_lo + kBlock;
    const double xi = live ? x[i] : 0., yi = live ? y[i] : 0., zi = live ? z[i] : 0.;
    double a_x = 0., a_y = 0., a_z = 0., c_x = 0., c_y = 0., c_z = 0.;
    for (int pass = 0; pass < 2; pass++) {
        int j0 = pass == 0 ? 0 : N_active;
        int j1 = pass == 0 ? N_active : N;
        if (pass == 1 && !tp_type) break;
        for (int t = j0; t < j1; t += kBlock) {
            __syncthreads();
            const int jj = t + (int)threadIdx.x;
            if (jj < j1) tile[threadIdx.x] = make_double4(x[jj], y[jj], z[jj], m[jj]);
            __syncthreads();
            const int n = min(kBlock, j1 - t);
            if (!live || (pass == 1 && i >= N_active)) continue;
            auto term = [&](const double4 s) {
                const double dx = xi - s.x;
                const double dy = yi - s.y;
                const double dz = zi - s.z;
                const double r2 = dx * dx + dy * dy + dz * dz + soft2;
                const double rr = sqrt(r2);
                const double prefact = G / (r2 * rr);
                const double prefactj = -prefact * s.w;
                kahan_add(prefactj * dx, a_x, c_x);
                kahan_add(prefactj * dy, a_y, c_y);
                kahan_add(prefactj * dz, a_z, c_z);
            };
            const bool check = (t < blk_hi && t + n > blk_lo) || (ignore != 0 && (t < 2 || blk_lo < 2));
            if (!check) {
#pragma unroll 4
                for (int k = 0; k < n; k++) term(tile[k]);
                continue;
            }
            for (int k = 0; k < n; k++) {
                const int j = t + k;
                if (skip_pair(i, j, ignore)) continue;
                term(tile[k]);
            }
        }
    }
    if (live) {
        ax[i] = a_x;
        ay[i] = a_y;
        az[i] = a_z;
    }
}

template <bool COMP>
static void launch_exact(hipStream_t st, const double* x, const double* y, const double* z, const double* m,
                         double* ax, double* ay, double* az, int N, int N_active, int tp_type, int ignore, double G,
                         double soft2, const GBList& gl) {
    // Measured on gfx950 (shapes 16x16 .. 1x256 targets x sources, N = 256
    // .. 65536): 8 targets x 64 sources is within 10 % of the best shape at
    // every N below ~50k; the row-wise kernel wins above.
#define RB_EXACT(TGT, TILE)                                                                                           \
    hipLaunchKernelGGL((k_exact<COMP, TGT, TILE>), dim3((N + TGT - 1) / TGT), dim3(kExThreads), 0, st, x, y, z, m,   \
                       ax, ay, az, N, N_active, tp_type, ignore, G, soft2, gl)
    if (N >= kExactRowwiseN) {
        if (COMP)
            hipLaunchKernelGGL(k_comp_exact, dim3((N + kBlock - 1) / kBlock), dim3(kBlock), 0, st, x, y, z, m, ax,
                               ay, az, N, N_active, tp_type, ignore, G, soft2);
        else
            hipLaunchKernelGGL(k_basic_exact, dim3((N + kBlock - 1) / kBlock), dim3(kBlock), 0, st, x, y, z, m, ax,
                               ay, az, N, N_active, tp_type, ignore, G, soft2, gl);
    } else {
        RB_EXACT(8, 64);
    }
#undef RB_EXACT
}

// ---------------------------------------------------------------- FP64 / SPLIT / COMP-FP64
// Pair kernel of the fast modes.  Accumulates S_i = sum_j m_j r_ij^-3 d_ij over
// one chunk of sources (blockIdx.y) and writes it (and for the compensated
// variant its Kahan carry) to the partial buffer; k_reduce applies -G.
enum FastMode { FM_FP64 = 0, FM_SPLIT = 1, FM_COMP = 2 };

// r^-3 in FP64.  v_rsq_f64 is good to 2^-24 (measured on gfx950); it is
// refined through e = 1 - r2*y^2 (|e| < 2^-23):
//   r^-3 = y^3 (1 - e)^(-3/2) = y^3 (1 + 3/2 e + 15/8 e^2 + O(e^3)),
// leaving a relative error of order 2^-69, far below FP64 rounding.
// `c15` is 1.5 held in a register: a GFX9 VOP3 instruction takes at most one
// scalar or literal operand, so fma(e, 1.875, 1.5) with both constants would
// be rebuilt from two v_mov per pair.
__device__ __forceinline__ double rinv3_fp64(double r2, double c15, double c1875) {
    const double yv = __builtin_amdgcn_rsq(r2);
    const double h = yv * yv;
    const double e = fma(-r2, h, 1.0);
    const double y3 = h * yv;
    const double p = fma(e, c1875, c15);
    return fma(y3, e * p, y3);
}

// A loop-invariant constant kept in a VGPR (opaque to constant folding).
__device__ __forceinline__ double vreg_const(double v) {
    asm volatile("" : "+v"(v));
    return v;
}

template <int MODE, bool CHECK>
__device__ __forceinline__ void pair_tile(const double4* __restrict__ tile, int n,
                                          int jbase, int i, int ignore, double gx, double gy, double gz,
                                          double soft2, double& sx, double& sy, double& sz, double& cx, double& cy,
                                          double& cz, double c15, double c1875) {
    RB_UNROLL(RB_FAST_UNROLL)
    for (int k = 0; k < n; k++) {
        const double4 s = tile[k];
        const double dx = gx - s.x;
        const double dy = gy - s.y;
        const double dz = gz - s.z;
        const double r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, soft2)));
        double w;
        if (MODE == FM_SPLIT) {
            // r^-1 from the FP32 reciprocal square root; its cube and the
            // mass factor in FP64, so only r^2 itself must lie in FP32's
            // normal range (about 1e-19 < r < 1.8e19 in simulation units)
            const double yv = (double)__builtin_amdgcn_rsqf((float)r2);
            w = yv * yv * yv * s.w;
        } else {
            w = s.w * rinv3_fp64(r2, c15, c1875);
        }
        if (CHECK) {
            const int j = jbase + k;
            if (skip_pair(i, j, ignore)) w = 0.;
        }
        if (MODE == FM_COMP) {
            kahan_add(w * dx, sx, cx);
            kahan_add(w * dy, sy, cy);
            kahan_add(w * dz, sz, cz);
        } else {
            sx = fma(w, dx, sx);
            sy = fma(w, dy, sy);
            sz = fma(w, dz, sz);
        }
    }
}

// Ghost-box offsets of a pass: the list of the simulation (kernel argument) or
// the single zero offset of a simulation without ghost boxes.
struct GBFromList {
    const GBList& gl;
    __device__ int n() const { return gl.n; }
    __device__ GB3 get(int g) const { return gl.g[g]; }
};
struct GBZero {
    __device__ int n() const { return 1; }
    __device__ GB3 get(int) const { return GB3{0., 0., 0.}; }
};

// One workgroup of the pair kernel: targets bx*kFast.. of n_recv, sources of
// chunk `by` of [j_begin, j_end).  Writes the chunk's sums and carries to
// `partial` (pstride doubles per component).
template <int MODE, class GBS>
__device__ __forceinline__ void direct_fast_block(const double* __restrict__ x, const double* __restrict__ y,
                                                  const double* __restrict__ z, const double* __restrict__ m,
                                                  int n_recv, int j_begin, int j_end, int chunk, int ignore,
                                                  double soft2, const GBS& gbs, double* __restrict__ partial,
                                                  int pstride, int bx, int by, double4* tile) {
    const int i = bx * kFast + threadIdx.x;
    const bool live = i < n_recv;
    const double xi = live ? x[i] : 0., yi = live ? y[i] : 0., zi = live ? z[i] : 0.;
    const int j0 = j_begin + by * chunk;
    const int j1 = min(j_end, j0 + chunk);
    const int blk_lo = bx * kFast;
    const int blk_hi = blk_lo + kFast;
    double sx = 0., sy = 0., sz = 0.;      // running sums
    double cx = 0., cy = 0., cz = 0.;      // Kahan carries (COMP per pair, FP64/SPLIT per tile)
    const double c15 = vreg_const(1.5);
    const double c1875 = vreg_const(1.875);
    for (int g = 0; g < gbs.n(); g++) {
        const GB3 gb = gbs.get(g);
        const double gx = gb.x + xi, gy = gb.y + yi, gz = gb.z + zi;
        for (int t = j0; t < j1; t += kFast) {
            __syncthreads();
            const int jj = t + (int)threadIdx.x;
            if (jj < j1) {
                tile[threadIdx.x] = make_double4(x[jj], y[jj], z[jj], m[jj]);
            }
            __syncthreads();
            const int n = min(kFast, j1 - t);
            // Only tiles holding a target of this block, or sources 0/1 when
            // terms are ignored, need the per-pair skip test (target 0 with
            // ignore == 2 is zeroed by the reduction).
            const bool check = (t < blk_hi && t + n > blk_lo) || (ignore != 0 && t < 2);
            if (MODE != FM_COMP) {
                // plain FP64 sum over the tile, folded into the running sum with Kahan
                double tx = 0., ty = 0., tz = 0., d0 = 0., d1 = 0., d2 = 0.;
                if (check)
                    pair_tile<MODE, true>(tile, n, t, i, ignore, gx, gy, gz, soft2, tx, ty, tz, d0, d1, d2,
                                          c15, c1875);
                else
                    pair_tile<MODE, false>(tile, n, t, i, ignore, gx, gy, gz, soft2, tx, ty, tz, d0, d1, d2,
                                           c15, c1875);
                kahan_add(tx, sx, cx);
                kahan_add(ty, sy, cy);
                kahan_add(tz, sz, cz);
            } else {
                if (check)
                    pair_tile<MODE, true>(tile, n, t, i, ignore, gx, gy, gz, soft2, sx, sy, sz, cx, cy, cz,
                                          c15, c1875);
                else
                    pair_tile<MODE, false>(tile, n, t, i, ignore, gx, gy, gz, soft2, sx, sy, sz, cx, cy, cz,
                                           c15, c1875);
            }
        }
    }
    if (live) {
        double* p = partial + (size_t)by * 6 * pstride;
        p[0 * pstride + i] = sx;
        p[1 * pstride + i] = sy;
        p[2 * pstride + i] = sz;
        p[3 * pstride + i] = cx;
        p[4 * pstride + i] = cy;
        p[5 * pstride + i] = cz;
    }
}

// Four wavefronts per SIMD (128 VGPRs, no spills in the pair loop) measured
// faster than eight on gfx950; at N=65536 the grid is 256 target blocks x 8
// source chunks.
template <int MODE>
__global__ __launch_bounds__(kFast) __attribute__((amdgpu_waves_per_eu(RB_FAST_WAVES, RB_FAST_WAVES))) void k_direct_fast(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ m, int n_recv, int j_begin, int j_end, int chunk, int ignore, double soft2,
    GBList gl, double* __restrict__ partial, int pstride) {
    __shared__ double4 tile[kFast];
    direct_fast_block<MODE>(x, y, z, m, n_recv, j_begin, j_end, chunk, ignore, soft2, GBFromList{gl}, partial,
                            pstride, (int)blockIdx.x, (int)blockIdx.y, tile);
}

// Sum the chunk partials in chunk order (Kahan), apply -G and write (or add to)
// the acceleration.  ignore==2 zeroes particle 0 (it feels nothing).  With
// kd.on, the leapfrog's kick and closing drift follow in the same pass
// (integrator_leapfrog.c:84-93 then :72-82: v += dt*a, then x += dt/2*v),
// which is safe because every read of the positions has finished.
__device__ __forceinline__ void reduce_one(const double* __restrict__ partial, int pstride, int nchunk, int i,
                                           double negG, int accumulate, int ignore, double* __restrict__ ax,
                                           double* __restrict__ ay, double* __restrict__ az, const KickDrift& kd) {
    // Per component, chunk by chunk: the chunk's sum, then its carry with the
    // sign flipped, Kahan-added in chunk order.  The loads of eight chunks are
    // issued before their additions, so their latencies overlap.
    double s[3] = {0., 0., 0.}, k[3] = {0., 0., 0.};
    constexpr int B = 8;
    int q = 0;
    for (; q + B <= nchunk; q += B) {
        double v[B][6];
#pragma unroll
        for (int u = 0; u < B; u++) {
            const double* p = partial + (size_t)(q + u) * 6 * pstride + i;
#pragma unroll
            for (int f = 0; f < 6; f++) v[u][f] = p[(size_t)f * pstride];
        }
#pragma unroll
        for (int u = 0; u < B; u++) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                kahan_add(v[u][c], s[c], k[c]);
                kahan_add(-v[u][c + 3], s[c], k[c]);
            }
        }
    }
    for (; q < nchunk; q++) {
        const double* p = partial + (size_t)q * 6 * pstride + i;
        for (int c = 0; c < 3; c++) {
            kahan_add(p[(size_t)c * pstride], s[c], k[c]);
            kahan_add(-p[(size_t)(c + 3) * pstride], s[c], k[c]);
        }
    }
    double out[3];
    for (int c = 0; c < 3; c++) out[c] = negG * s[c];
    if (ignore == 2 && i == 0) out[0] = out[1] = out[2] = 0.;
    if (accumulate) {
        out[0] = ax[i] + out[0];
        out[1] = ay[i] + out[1];
        out[2] = az[i] + out[2];
    }
    ax[i] = out[0];
    ay[i] = out[1];
    az[i] = out[2];
    if (kd.on) {
        const double vx = kd.vx[i] + kd.kick * out[0];
        const double vy = kd.vy[i] + kd.kick * out[1];
        const double vz = kd.vz[i] + kd.kick * out[2];
        kd.vx[i] = vx;
        kd.vy[i] = vy;
        kd.vz[i] = vz;
        kd.x[i] = kd.x[i] + kd.drift * vx;
        kd.y[i] = kd.y[i] + kd.drift * vy;
        kd.z[i] = kd.z[i] + kd.drift * vz;
    }
}

__global__ void k_reduce(const double* __restrict__ partial, int pstride, int nchunk, int n_recv, double negG,
                         int accumulate, int ignore, double* __restrict__ ax, double* __restrict__ ay,
                         double* __restrict__ az, KickDrift kd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_recv) return;
    reduce_one(partial, pstride, nchunk, i, negG, accumulate, ignore, ax, ay, az, kd);
}

// The same kick and drift after a kernel that writes the accelerations itself.
__global__ void k_kick_drift(const double* __restrict__ ax, const double* __restrict__ ay,
                             const double* __restrict__ az, int n, KickDrift kd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double vx = kd.vx[i] + kd.kick * ax[i];
    const double vy = kd.vy[i] + kd.kick * ay[i];
    const double vz = kd.vz[i] + kd.kick * az[i];
    kd.vx[i] = vx;
    kd.vy[i] = vy;
    kd.vz[i] = vz;
    kd.x[i] = kd.x[i] + kd.drift * vx;
    kd.y[i] = kd.y[i] + kd.drift * vy;
    kd.z[i] = kd.z[i] + kd.drift * vz;
}

__global__ void k_zero3(double* __restrict__ a, double* __restrict__ b, double* __restrict__ c, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = b[i] = c[i] = 0.;
}

// ---------------------------------------------------------------- many simulations per launch
// One order-2 leapfrog step with direct gravity (no ghost boxes, no test
// particles, nothing ignored) for a batch of independent simulations in three
// launches: drift, pair workgroups of every simulation, then reduction with
// the kick and closing drift.  Each simulation keeps its own chunking
// (pick_chunks of its own N), so its bits are those of stepping it alone.
__global__ __launch_bounds__(256) void k_seg_drift(const SegJob* __restrict__ jobs, const SegRange* __restrict__ rg) {
    const SegRange R = rg[blockIdx.x];
    const SegJob J = jobs[R.job];
    const int i = R.first + (int)threadIdx.x;
    if (i >= J.n) return;
    J.x[i] += J.drift * J.vx[i];
    J.y[i] += J.drift * J.vy[i];
    J.z[i] += J.drift * J.vz[i];
}

__global__ __launch_bounds__(kFast) __attribute__((amdgpu_waves_per_eu(RB_FAST_WAVES, RB_FAST_WAVES))) void k_seg_direct(
    const SegJob* __restrict__ jobs, const SegTile* __restrict__ tiles) {
    __shared__ double4 tile[kFast];
    const SegTile T = tiles[blockIdx.x];
    const SegJob J = jobs[T.job];
    switch (J.mode) {
        case FM_FP64:
            direct_fast_block<FM_FP64>(J.x, J.y, J.z, J.m, J.n, 0, J.n, J.chunk, 0, J.soft2, GBZero{}, J.partial, J.n,
                                       T.bx, T.by, tile);
            break;
        case FM_SPLIT:
            direct_fast_block<FM_SPLIT>(J.x, J.y, J.z, J.m, J.n, 0, J.n, J.chunk, 0, J.soft2, GBZero{}, J.partial,
                                        J.n, T.bx, T.by, tile);
            break;
        default:
            direct_fast_block<FM_COMP>(J.x, J.y, J.z, J.m, J.n, 0, J.n, J.chunk, 0, J.soft2, GBZero{}, J.partial, J.n,
                                       T.bx, T.by, tile);
            break;
    }
}

__global__ __launch_bounds__(256) void k_seg_reduce(const SegJob* __restrict__ jobs, const SegRange* __restrict__ rg) {
    const SegRange R = rg[blockIdx.x];
    const SegJob J = jobs[R.job];
    const int i = R.first + (int)threadIdx.x;
    if (i >= J.n) return;
    const KickDrift kd{J.x, J.y, J.z, J.vx, J.vy, J.vz, J.kick, J.drift, 1};
    reduce_one(J.partial, J.n, J.nchunk, i, J.negG, 0, 0, J.ax, J.ay, J.az, kd);
}

// ---------------------------------------------------------------- symmetric pair kernel
// Newton's third law, as the reference's serial build uses it (gravity.c:195-214
// visits each pair once and adds the term to both bodies): each unordered pair
// of bodies in different superblocks is evaluated once.  The bodies are cut
// into superblocks of kSymSB; a workgroup takes one pair of superblocks A < B
// (a task).  Its 4 wavefronts hold the targets of A, SYM_TPL per lane, in
// registers; the sources of B stream through LDS 64 at a time (a J-block),
// and lane k evaluates source (k - s) & 63 at step s, so the 64 lanes see 64
// different sources at once.  The source-side sum travels with its source:
// after each step the three accumulators are rotated one lane over the whole
// wavefront (DPP wave_ror:1), and after 64 steps every lane again holds the
// sum of the source it started with.  Per unordered pair: the FP64 r^-3 of
// the fast mode (or the split-precision one) once, two mass products, three
// fused multiply-adds on each side; per step the rotation adds six moves.
//
// Outputs, all in fixed order (the results are deterministic):
//   ipart[t][c][SB]  target-side sums of task t (targets of A, sources of B)
//   jpart[t][c][SB]  source-side sums of task t (bodies of B, all targets of A)
//                    with the sign of the term for A's bodies: S_j = -jpart
//   dpart[6][N]      the one-sided sums within each superblock (direct_fast_block)
// and k_sym_reduce adds, per body, the partials of every superblock in
// ascending order with Kahan compensation.
#ifndef RB_SYM_TPL
#define RB_SYM_TPL 4
#endif
static constexpr int kSymTPL = RB_SYM_TPL;           // targets per lane
static constexpr int kSymSB = 4 * 64 * kSymTPL;       // bodies per superblock
static constexpr int kSymJB = 64;                     // sources per J-block
static constexpr int kSymMinN = 8192;                 // below: the one-sided kernel

// Lane k receives lane k-1's value, lane 0 lane 63's (DPP wave_ror:1).
__device__ __forceinline__ double dpp_ror1(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x13C, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x13C, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

__host__ __device__ __forceinline__ int sym_task(int A, int B, int nsb) {
    return A * nsb - A * (A + 1) / 2 + (B - A - 1);
}

template <int MODE, bool CHECK>
__global__ __launch_bounds__(256) void k_direct_sym(const double* __restrict__ x, const double* __restrict__ y,
                                                    const double* __restrict__ z, const double* __restrict__ m,
                                                    int N, int nsb, double soft2, const int2* __restrict__ tasks,
                                                    int ntasks_done, double* __restrict__ ipart,
                                                    double* __restrict__ jpart) {
    // J-block sources (x, y, z, m), double-buffered, each stored twice in a
    // row so that lane k reads entry k + 64 - s at step s without wrapping
    __shared__ double4 jt[2][2 * kSymJB];
    __shared__ double jred[2][4][3][kSymJB];       // per-wave source sums of a J-block
    const int2 tk = tasks[ntasks_done + blockIdx.x];
    const int A = tk.x, B = tk.y;
    const int t = sym_task(A, B, nsb);
    const int tid = threadIdx.x;
    const int w = tid >> 6, k = tid & 63;
    // targets of this lane
    double tx[kSymTPL], ty[kSymTPL], tz[kSymTPL], tm[kSymTPL];
    double sx[kSymTPL], sy[kSymTPL], sz[kSymTPL];
#pragma unroll
    for (int q = 0; q < kSymTPL; q++) {
        const int i = A * kSymSB + w * 64 * kSymTPL + q * 64 + k;
        const bool ok = !CHECK || i < N;
        tx[q] = ok ? x[i] : 0.;
        ty[q] = ok ? y[i] : 0.;
        tz[q] = ok ? z[i] : 0.;
        tm[q] = ok ? m[i] : 0.;
        sx[q] = sy[q] = sz[q] = 0.;
    }
    const int jbase = B * kSymSB;
    constexpr int NJB = kSymSB / kSymJB;
    // prologue: J-block 0 into LDS, J-block 1 into registers
    double4 nxt = make_double4(0., 0., 0., 0.);
    if (tid < kSymJB) {
        const int j0 = jbase + tid;
        const bool ok0 = !CHECK || j0 < N;
        const double4 v0 = ok0 ? make_double4(x[j0], y[j0], z[j0], m[j0]) : make_double4(0., 0., 0., 0.);
        jt[0][tid] = v0;
        jt[0][tid + kSymJB] = v0;
        const int j1 = jbase + kSymJB + tid;
        const bool ok1 = !CHECK || j1 < N;
        if (ok1) nxt = make_double4(x[j1], y[j1], z[j1], m[j1]);
    }
    const double c15 = vreg_const(1.5);
    const double c1875 = vreg_const(1.875);
    for (int jb = 0; jb < NJB; jb++) {
        __syncthreads();
        if (jb > 0 && w < 3) {
            // source sums of J-block jb-1: the four waves' in order
            const double* r = &jred[(jb - 1) & 1][0][w][k];
            const double v = ((r[0] + r[3 * kSymJB]) + r[6 * kSymJB]) + r[9 * kSymJB];
            jpart[((size_t)t * 3 + w) * kSymSB + (jb - 1) * kSymJB + k] = v;
        }
        if (tid < kSymJB) {
            if (jb + 1 < NJB) {
                jt[(jb + 1) & 1][tid] = nxt;
                jt[(jb + 1) & 1][tid + kSymJB] = nxt;
            }
            const int jn = jbase + (jb + 2) * kSymJB + tid;
            const bool okn = jb + 2 < NJB && (!CHECK || jn < N);
            nxt = okn ? make_double4(x[jn], y[jn], z[jn], m[jn]) : make_double4(0., 0., 0., 0.);
        }
        const double4* src = jt[jb & 1];
        double bx[kSymTPL], by[kSymTPL], bz[kSymTPL];
#pragma unroll
        for (int q = 0; q < kSymTPL; q++) bx[q] = by[q] = bz[q] = 0.;
        double jx = 0., jy = 0., jz = 0.;
        const int jlim = CHECK ? N - (jbase + jb * kSymJB) : kSymJB;   // valid sources of this J-block
#pragma unroll 4
        for (int s = 0; s < kSymJB; s++) {
            const int idx = (k - s) & (kSymJB - 1);     // the source's index in the J-block
            const double4 sj = src[k + kSymJB - s];
#pragma unroll
            for (int q = 0; q < kSymTPL; q++) {
                const double dx = tx[q] - sj.x;
                const double dy = ty[q] - sj.y;
                const double dz = tz[q] - sj.z;
                const double r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, soft2)));
                double wr;
                if (MODE == FM_SPLIT) {
                    const double yv = (double)__builtin_amdgcn_rsqf((float)r2);
                    wr = yv * yv * yv;
                } else {
                    wr = rinv3_fp64(r2, c15, c1875);
                }
                if (CHECK) {
                    const int i = A * kSymSB + w * 64 * kSymTPL + q * 64 + k;
                    if (i >= N || idx >= jlim) wr = 0.;
                }
                const double wj = wr * sj.w;
                const double wi = wr * tm[q];
                bx[q] = fma(wj, dx, bx[q]);
                by[q] = fma(wj, dy, by[q]);
                bz[q] = fma(wj, dz, bz[q]);
                jx = fma(wi, dx, jx);
                jy = fma(wi, dy, jy);
                jz = fma(wi, dz, jz);
            }
            jx = dpp_ror1(jx);
            jy = dpp_ror1(jy);
            jz = dpp_ror1(jz);
        }
#pragma unroll
        for (int q = 0; q < kSymTPL; q++) {
            sx[q] += bx[q];
            sy[q] += by[q];
            sz[q] += bz[q];
        }
        double* jr = &jred[jb & 1][w][0][k];
        jr[0] = jx;
        jr[kSymJB] = jy;
        jr[2 * kSymJB] = jz;
    }
    __syncthreads();
    if (w < 3) {
        const double* r = &jred[(NJB - 1) & 1][0][w][k];
        const double v = ((r[0] + r[3 * kSymJB]) + r[6 * kSymJB]) + r[9 * kSymJB];
        jpart[((size_t)t * 3 + w) * kSymSB + (NJB - 1) * kSymJB + k] = v;
    }
#pragma unroll
    for (int q = 0; q < kSymTPL; q++) {
        const int li = w * 64 * kSymTPL + q * 64 + k;
        double* p = ipart + (size_t)t * 3 * kSymSB + li;
        p[0] = sx[q];
        p[kSymSB] = sy[q];
        p[2 * kSymSB] = sz[q];
    }
}

// The one-sided sums inside each superblock (pairs i != j of one superblock):
// the fast pair block with targets and sources limited to the superblock.
template <int MODE>
__global__ __launch_bounds__(kFast) __attribute__((amdgpu_waves_per_eu(RB_FAST_WAVES, RB_FAST_WAVES))) void k_direct_diag(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ m, int N, double soft2, double* __restrict__ dpart) {
    __shared__ double4 tile[kFast];
    constexpr int per = kSymSB / kFast;                 // target blocks per superblock
    const int A = (int)blockIdx.x / per;
    const int j0 = A * kSymSB;
    const int j1 = min(N, j0 + kSymSB);
    direct_fast_block<MODE>(x, y, z, m, N, j0, j1, kSymSB, 0, soft2, GBZero{}, dpart, N, (int)blockIdx.x, 0, tile);
}

// Per body: the partials of every superblock in ascending order (Kahan), then
// -G and the fused kick and drift as in the one-sided reduction.
__global__ void k_sym_reduce(const double* __restrict__ ipart, const double* __restrict__ jpart,
                             const double* __restrict__ dpart, int N, int nsb, double negG, double* __restrict__ ax,
                             double* __restrict__ ay, double* __restrict__ az, KickDrift kd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int X = i / kSymSB, loc = i % kSymSB;
    double s[3] = {0., 0., 0.}, c[3] = {0., 0., 0.};
    for (int Y = 0; Y < nsb; Y++) {
        if (Y < X) {
            const double* p = jpart + (size_t)sym_task(Y, X, nsb) * 3 * kSymSB + loc;
#pragma unroll
            for (int q = 0; q < 3; q++) kahan_add(-p[(size_t)q * kSymSB], s[q], c[q]);
        } else if (Y == X) {
#pragma unroll
            for (int q = 0; q < 3; q++) {
                kahan_add(dpart[(size_t)q * N + i], s[q], c[q]);
                kahan_add(-dpart[(size_t)(q + 3) * N + i], s[q], c[q]);
            }
        } else {
            const double* p = ipart + (size_t)sym_task(X, Y, nsb) * 3 * kSymSB + loc;
#pragma unroll
            for (int q = 0; q < 3; q++) kahan_add(p[(size_t)q * kSymSB], s[q], c[q]);
        }
    }
    const double o0 = negG * s[0], o1 = negG * s[1], o2 = negG * s[2];
    ax[i] = o0;
    ay[i] = o1;
    az[i] = o2;
    if (kd.on) {
        const double vx = kd.vx[i] + kd.kick * o0;
        const double vy = kd.vy[i] + kd.kick * o1;
        const double vz = kd.vz[i] + kd.kick * o2;
        kd.vx[i] = vx;
        kd.vy[i] = vy;
        kd.vz[i] = vz;
        kd.x[i] = kd.x[i] + kd.drift * vx;
        kd.y[i] = kd.y[i] + kd.drift * vy;
        kd.z[i] = kd.z[i] + kd.drift * vz;
    }
}

// Superblock layout of the symmetric path for N bodies.
struct SymPlan {
    int nsb = 0;          // superblocks
    int ntasks = 0;       // superblock pairs A < B
    int nfull = 0;        // of them, pairs without a partial superblock (launched first)
    size_t ipart_off = 0, jpart_off = 0, dpart_off = 0, tasks_off = 0, bytes = 0;   // workspace layout (bytes)
};

static SymPlan sym_plan(int N) {
    SymPlan p;
    p.nsb = (N + kSymSB - 1) / kSymSB;
    p.ntasks = p.nsb * (p.nsb - 1) / 2;
    const bool ragged = N % kSymSB != 0;
    p.nfull = ragged ? (p.nsb - 1) * (p.nsb - 2) / 2 : p.ntasks;
    const size_t part = sizeof(double) * 3 * (size_t)kSymSB * (size_t)p.ntasks;
    auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
    p.ipart_off = 0;
    p.jpart_off = up(part);
    p.dpart_off = p.jpart_off + up(part);
    p.tasks_off = p.dpart_off + up(sizeof(double) * 6 * (size_t)N);
    p.bytes = p.tasks_off + up(sizeof(int2) * (size_t)(p.ntasks > 0 ? p.ntasks : 1));
    return p;
}

// The task list (A, B), A < B: pairs without the ragged last superblock first,
// then those with it (launched with the bounds checks).
static std::vector<int2> sym_tasks(const SymPlan& p) {
    std::vector<int2> t;
    t.reserve(p.ntasks);
    const int last = (p.nfull == p.ntasks) ? p.nsb : p.nsb - 1;
    for (int A = 0; A < last; A++)
        for (int B = A + 1; B < last; B++) t.push_back(make_int2(A, B));
    if (last < p.nsb)
        for (int A = 0; A < last; A++) t.push_back(make_int2(A, last));
    return t;
}

// Number of source chunks for the fast modes: enough workgroups to fill every
// SIMD of the 256-CU part several times over, chunks no shorter than
// kMinChunk sources (short chunks keep the time of one step small for N of a
// few thousand, where a single workgroup's chunk sets the kernel's duration).
static constexpr int kMinChunk = 32;
static_assert(kFast % kMinChunk == 0, "chunks tile the LDS tile");
static int round_chunk(int c) { return ((c + kMinChunk - 1) / kMinChunk) * kMinChunk; }
static int pick_chunks(int n_recv, int n_src, int ngb) {
    const int blocks_x = (n_recv + kFast - 1) / kFast;
    const int target_blocks = 2048 * 256 / kFast;
    int s = (target_blocks + blocks_x - 1) / blocks_x;
    int max_s = (n_src * ngb) / kMinChunk;
    if (max_s < 1) max_s = 1;
    if (s > max_s) s = max_s;
    if (s > 64) s = 64;
    if (s < 1) s = 1;
    return s;
}

static size_t workspace_bytes(int n, int nchunk) { return sizeof(double) * 6 * (size_t)n * (size_t)nchunk; }

// The chunking a simulation of n bodies uses when stepped alone (launch_fast).
void seg_plan(int n, int* chunk, int* used, size_t* workspace) {
    const int nchunk = pick_chunks(n, n, 1);
    const int c = round_chunk((n + nchunk - 1) / nchunk);
    *chunk = c;
    *used = n > 0 ? (n + c - 1) / c : 1;
    *workspace = workspace_bytes(n, nchunk) + 256;
}

int seg_tiles_of(int n, int used) { return ((n + kFast - 1) / kFast) * used; }

int seg_step_launch(const SegJob* jobs, const SegRange* ranges, int nranges, const SegTile* tiles, int ntiles,
                    hipStream_t st) {
    if (nranges == 0) return 0;
    hipLaunchKernelGGL(k_seg_drift, dim3(nranges), dim3(256), 0, st, jobs, ranges);
    hipLaunchKernelGGL(k_seg_direct, dim3(ntiles), dim3(kFast), 0, st, jobs, tiles);
    hipLaunchKernelGGL(k_seg_reduce, dim3(nranges), dim3(256), 0, st, jobs, ranges);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int seg_fast_mode(int precision, int compensated) {
    return compensated ? FM_COMP : (precision == REB_AMD_PRECISION_SPLIT ? FM_SPLIT : FM_FP64);
}

// Whether a simulation of N bodies, all active, without ghost boxes or
// ignored terms, takes the symmetric path in this precision and gravity, and
// the workspace it then needs.
bool gravity_symmetric_path(size_t N, int precision, bool compensated) {
    return !compensated && precision != REB_AMD_PRECISION_EXACT && N >= (size_t)kSymMinN;
}
size_t gravity_symmetric_bytes(size_t N) { return sym_plan((int)N).bytes; }

// Ordered pairs inside the superblocks (evaluated by k_direct_diag, not by
// the symmetric kernel): sum over superblocks of n_A^2, self pairs included.
static double sym_diag_pairs(int N) {
    const int full = N / kSymSB, rest = N % kSymSB;
    return (double)full * kSymSB * kSymSB + (double)rest * rest;
}

// One force evaluation on the symmetric path: the pairs inside superblocks,
// the superblock pairs (events around these), then the per-body reduction
// with the fused kick and drift.
static int sym_launch(int mode, const double* x, const double* y, const double* z, const double* m, double* ax,
                      double* ay, double* az, int N, double G, double soft2, const SymPlan& plan, char* base,
                      hipStream_t st, hipEvent_t* ev, const KickDrift& kd) {
    double* ipart = reinterpret_cast<double*>(base + plan.ipart_off);
    double* jpart = reinterpret_cast<double*>(base + plan.jpart_off);
    double* dpart = reinterpret_cast<double*>(base + plan.dpart_off);
    const int2* tasks = reinterpret_cast<const int2*>(base + plan.tasks_off);
    const dim3 gdiag(plan.nsb * (kSymSB / kFast));
    if (mode == FM_SPLIT)
        hipLaunchKernelGGL(k_direct_diag<FM_SPLIT>, gdiag, dim3(kFast), 0, st, x, y, z, m, N, soft2, dpart);
    else
        hipLaunchKernelGGL(k_direct_diag<FM_FP64>, gdiag, dim3(kFast), 0, st, x, y, z, m, N, soft2, dpart);
    if (ev) (void)hipEventRecord(ev[0], st);
    const int nrag = plan.ntasks - plan.nfull;
    if (plan.nfull > 0) {
        if (mode == FM_SPLIT)
            hipLaunchKernelGGL((k_direct_sym<FM_SPLIT, false>), dim3(plan.nfull), dim3(256), 0, st, x, y, z, m, N,
                               plan.nsb, soft2, tasks, 0, ipart, jpart);
        else
            hipLaunchKernelGGL((k_direct_sym<FM_FP64, false>), dim3(plan.nfull), dim3(256), 0, st, x, y, z, m, N,
                               plan.nsb, soft2, tasks, 0, ipart, jpart);
    }
    if (nrag > 0) {
        if (mode == FM_SPLIT)
            hipLaunchKernelGGL((k_direct_sym<FM_SPLIT, true>), dim3(nrag), dim3(256), 0, st, x, y, z, m, N,
                               plan.nsb, soft2, tasks, plan.nfull, ipart, jpart);
        else
            hipLaunchKernelGGL((k_direct_sym<FM_FP64, true>), dim3(nrag), dim3(256), 0, st, x, y, z, m, N,
                               plan.nsb, soft2, tasks, plan.nfull, ipart, jpart);
    }
    if (ev) (void)hipEventRecord(ev[1], st);
    hipLaunchKernelGGL(k_sym_reduce, dim3((N + 255) / 256), dim3(256), 0, st, ipart, jpart, dpart, N, plan.nsb, -G,
                       ax, ay, az, kd);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// Launch one fast-mode pass: targets [0,n_recv) feel sources [j_begin,j_end).
static int launch_fast(int mode, const double* x, const double* y, const double* z, const double* m, int n_recv,
                       int j_begin, int j_end, int ignore, double G, double soft2, const GBList& gl,
                       double* partial, int nchunk, int accumulate, double* ax, double* ay, double* az,
                       hipStream_t st, hipEvent_t* ev, const KickDrift& kd) {
    const int nsrc = j_end - j_begin;
    const int chunk = round_chunk((nsrc + nchunk - 1) / nchunk);
    const int used = nsrc > 0 ? (nsrc + chunk - 1) / chunk : 1;
    dim3 grid((n_recv + kFast - 1) / kFast, used);
    if (ev) (void)hipEventRecord(ev[0], st);
    switch (mode) {
        case FM_FP64:
            hipLaunchKernelGGL(k_direct_fast<FM_FP64>, grid, dim3(kFast), 0, st, x, y, z, m, n_recv, j_begin, j_end,
                               chunk, ignore, soft2, gl, partial, n_recv);
            break;
        case FM_SPLIT:
            hipLaunchKernelGGL(k_direct_fast<FM_SPLIT>, grid, dim3(kFast), 0, st, x, y, z, m, n_recv, j_begin,
                               j_end, chunk, ignore, soft2, gl, partial, n_recv);
            break;
        default:
            hipLaunchKernelGGL(k_direct_fast<FM_COMP>, grid, dim3(kFast), 0, st, x, y, z, m, n_recv, j_begin, j_end,
                               chunk, ignore, soft2, gl, partial, n_recv);
            break;
    }
    if (ev) (void)hipEventRecord(ev[1], st);
    hipLaunchKernelGGL(k_reduce, dim3((n_recv + 255) / 256), dim3(256), 0, st, partial, n_recv, used, n_recv, -G,
                       accumulate, ignore, ax, ay, az, kd);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// Shared dispatch for the simulation path and the raw entry point.
static int direct_dispatch(const double* x, const double* y, const double* z, const double* m, double* ax,
                           double* ay, double* az, int N, int N_active, int tp_type, int ignore, double G,
                           double soft2, int precision, int compensated, const GBList& gl,
                           double* workspace, hipStream_t st, hipEvent_t* ev = nullptr,
                           const KickDrift& kd = KickDrift{}) {
    if (N == 0) return 0;
    const KickDrift none{};
    if (precision == REB_AMD_PRECISION_EXACT) {
        if (ev) (void)hipEventRecord(ev[0], st);
        if (compensated)
            launch_exact<true>(st, x, y, z, m, ax, ay, az, N, N_active, tp_type, ignore, G, soft2, gl);
        else
            launch_exact<false>(st, x, y, z, m, ax, ay, az, N, N_active, tp_type, ignore, G, soft2, gl);
        if (ev) (void)hipEventRecord(ev[1], st);
        if (kd.on) hipLaunchKernelGGL(k_kick_drift, dim3((N + 255) / 256), dim3(256), 0, st, ax, ay, az, N, kd);
        return hipGetLastError() == hipSuccess ? 0 : 1;
    }
    const int mode = compensated ? FM_COMP : (precision == REB_AMD_PRECISION_SPLIT ? FM_SPLIT : FM_FP64);
    const int nchunk = pick_chunks(N, N_active, gl.n);
    int rc = 0;
    const bool second = tp_type && N_active < N;
    if (N_active > 0) {
        rc |= launch_fast(mode, x, y, z, m, N, 0, N_active, ignore, G, soft2, gl, workspace, nchunk, 0, ax, ay, az,
                          st, ev, second ? none : kd);
    } else {
        hipLaunchKernelGGL(k_zero3, dim3((N + 255) / 256), dim3(256), 0, st, ax, ay, az, N);
    }
    if (second) {
        // Active targets also feel the test particles (gravity.c:259-276).
        const int nchunk2 = pick_chunks(N_active, N - N_active, gl.n);
        rc |= launch_fast(mode, x, y, z, m, N_active, N_active, N, ignore, G, soft2, gl, workspace,
                          nchunk2 < nchunk ? nchunk2 : nchunk, 1, ax, ay, az, st, nullptr, none);
    }
    if (kd.on && (second || N_active == 0))
        hipLaunchKernelGGL(k_kick_drift, dim3((N + 255) / 256), dim3(256), 0, st, ax, ay, az, N, kd);
    return rc;
}

// ---------------------------------------------------------------- JACOBI
// gravity.c:108-163, the gravity of the WHFast modified-kick and lazy kernels:
// Jacobi and direct terms with every pair visited once (j = 0..N-1, i = 0..j)
// and added to both bodies, without softening.  The order of the additions to
// each body is the whole definition, so one thread walks the pairs; the
// systems it serves have a few bodies (the WHFast group kernel evaluates the
// same sums for its own systems).
__global__ void k_gravity_jacobi(const double* __restrict__ x, const double* __restrict__ y,
                                 const double* __restrict__ z, const double* __restrict__ m, double* ax, double* ay,
                                 double* az, int N, double G) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double Rx = 0., Ry = 0., Rz = 0., Mt = 0.;
    for (int j = 0; j < N; j++) {
        ax[j] = 0.;
        ay[j] = 0.;
        az[j] = 0.;
        const double xj = x[j], yj = y[j], zj = z[j], mj = m[j];
        for (int i = 0; i <= j; i++) {
            if (j > 1) {
                const double qx = xj - Rx / Mt;
                const double qy = yj - Ry / Mt;
                const double qz = zj - Rz / Mt;
                const double dr = sqrt(qx * qx + qy * qy + qz * qz);
                const double w = i < j ? -mj : Mt;
                const double pf = G * w / (dr * dr * dr);
                ax[i] += pf * qx;
                ay[i] += pf * qy;
                az[i] += pf * qz;
            }
            if (i != j && (i != 0 || j != 1)) {
                const double dx = x[i] - xj;
                const double dy = y[i] - yj;
                const double dz = z[i] - zj;
                const double dr = sqrt(dx * dx + dy * dy + dz * dz);
                const double pf = G / (dr * dr * dr);
                const double pfi = pf * m[i];
                const double pfj = pf * mj;
                ax[i] -= pfj * dx;
                ay[i] -= pfj * dy;
                az[i] -= pfj * dz;
                ax[j] += pfi * dx;
                ay[j] += pfi * dy;
                az[j] += pfi * dz;
            }
        }
        Rx += mj * xj;
        Ry += mj * yj;
        Rz += mj * zj;
        Mt += mj;
    }
}

// ---------------------------------------------------------------- simulation path
bool gravity_jacobi(struct reb_simulation* r) {
    DeviceState* d = dev_of(r);
    const int N = (int)r->N;
    if (N == 0) return true;
    hipLaunchKernelGGL(k_gravity_jacobi, dim3(1), dim3(64), 0, d->stream, d->field(FX), d->field(FY), d->field(FZ),
                       d->field(FM), d->field(FAX), d->field(FAY), d->field(FAZ), N, r->G);
    return RB_HIP(r, hipGetLastError());
}

bool gravity_zero(struct reb_simulation* r) {
    DeviceState* d = dev_of(r);
    const int N = (int)r->N;
    if (N == 0) return true;
    hipLaunchKernelGGL(k_zero3, dim3((N + 255) / 256), dim3(256), 0, d->stream, d->field(FAX), d->field(FAY),
                       d->field(FAZ), N);
    return RB_HIP(r, hipGetLastError());
}

// kick_drift (optional): {kick dt, drift dt} of a leapfrog step, applied in the
// pass that finishes the accelerations.
bool gravity_direct(struct reb_simulation* r, bool compensated, const double* kick_drift) {
    DeviceState* d = dev_of(r);
    const int N = (int)r->N;
    if (N == 0) return true;
    const int N_active = (r->N_active == SIZE_MAX) ? N : (int)r->N_active;
    const double soft2 = r->softening * r->softening;
    // Ghost boxes (BASIC only; gravity.c:184-186).  COMPENSATED has none.
    std::vector<struct reb_vec6d> gbv;
    if (compensated)
        gbv.push_back(reb_vec6d{0, 0, 0, 0, 0, 0});
    else
        gbv = ghostboxes(r, r->N_ghost_x, r->N_ghost_y, r->N_ghost_z);
    const int ngb = (int)gbv.size();
    if (ngb > kMaxGB) {
        reb_simulation_error(r, "Too many ghost boxes for direct gravity (at most 125).");
        return false;
    }
    GBList gl;
    gl.n = ngb;
    for (int g = 0; g < ngb; g++) gl.g[g] = GB3{gbv[g].x, gbv[g].y, gbv[g].z};
    hipEvent_t* ev = nullptr;
    if (d->profiling) {
        if (d->kev_used + 2 > d->kev.size()) {
            for (int k = 0; k < 64; k++) {
                hipEvent_t e;
                if (!RB_HIP(r, hipEventCreate(&e))) return false;
                d->kev.push_back(e);
            }
        }
        ev = &d->kev[d->kev_used];
        d->kev_used += 2;
    }
    KickDrift kd{};
    if (kick_drift) {
        kd = KickDrift{d->field(FX), d->field(FY), d->field(FZ), d->field(FVX), d->field(FVY), d->field(FVZ),
                       kick_drift[0], kick_drift[1], 1};
    }
    // Every body active, no ghost boxes, nothing ignored: each pair once
    // (Newton's third law); otherwise the one-sided kernels.
    if (ngb == 1 && N_active == N && r->gravity_ignore_terms == 0 &&
        gravity_symmetric_path((size_t)N, d->precision, compensated)) {
        const SymPlan plan = sym_plan(N);
        if (!d->partial.ensure(plan.bytes)) {
            reb_simulation_error(r, "Out of device memory for the gravity workspace.");
            return false;
        }
        char* base = static_cast<char*>(d->partial.p);
        int2* tasks = reinterpret_cast<int2*>(base + plan.tasks_off);
        if (d->sym_tasks_n != N || d->sym_tasks_dev != (const void*)tasks) {
            d->sym_tasks_host = sym_tasks(plan);
            if (!RB_HIP(r, hipMemcpyAsync(tasks, d->sym_tasks_host.data(), sizeof(int2) * d->sym_tasks_host.size(),
                                          hipMemcpyHostToDevice, d->stream)))
                return false;
            d->sym_tasks_n = N;
            d->sym_tasks_dev = tasks;
        }
        if (d->profiling) d->kpairs += (double)N * (double)N - sym_diag_pairs(N);
        const int mode = d->precision == REB_AMD_PRECISION_SPLIT ? FM_SPLIT : FM_FP64;
        const int rc = sym_launch(mode, d->field(FX), d->field(FY), d->field(FZ), d->field(FM), d->field(FAX),
                                  d->field(FAY), d->field(FAZ), N, r->G, soft2, plan, base, d->stream, ev, kd);
        if (rc) return RB_HIP(r, hipGetLastError());
        return true;
    }
    const int nchunk = pick_chunks(N, N_active, ngb);
    if (!d->partial.ensure(workspace_bytes(N, nchunk) + 256)) {
        reb_simulation_error(r, "Out of device memory for the gravity workspace.");
        return false;
    }
    if (d->profiling) d->kpairs += (double)N * (double)(N - 1);
    int rc = direct_dispatch(d->field(FX), d->field(FY), d->field(FZ), d->field(FM), d->field(FAX), d->field(FAY),
                             d->field(FAZ), N, N_active, r->testparticle_type, (int)r->gravity_ignore_terms, r->G,
                             soft2, d->precision, compensated ? 1 : 0, gl, d->partial.as<double>(), d->stream, ev, kd);
    if (rc) return RB_HIP(r, hipGetLastError());
    return true;
}

// ---------------------------------------------------------------- energy
// tools.c:108-162.  EXACT: the kinetic and potential sums are accumulated in
// the reference's sequential order, so the value is bit-identical: the
// kinetic terms by one wavefront (lanes evaluate 64 terms, lane 0 folds them
// in order); the pair terms of a batch of rows are written in the
// reference's (i, j > i) order by a grid, then one lane folds them in order
// while its wavefront streams the next chunk in.  Otherwise each target's row
// sum over j>i is compensated and rows are folded with Kahan.
__global__ __launch_bounds__(64) void k_energy_kin(const double* __restrict__ vx, const double* __restrict__ vy,
                                                   const double* __restrict__ vz, const double* __restrict__ m,
                                                   int N_interact, double* __restrict__ out) {
    __shared__ double terms[64];
    const int lane = threadIdx.x;
    double e_kin = 0.;
    for (int b = 0; b < N_interact; b += 64) {
        const int i = b + lane;
        if (i < N_interact) {
            const double vv = vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i];
            terms[lane] = 0.5 * m[i] * vv;
        }
        __syncthreads();
        if (lane == 0) {
            const int n = min(64, N_interact - b);
            for (int k = 0; k < n; k++) e_kin += terms[k];
        }
        __syncthreads();
    }
    if (lane == 0) {
        out[0] = e_kin;
        out[1] = 0.;
    }
}

// Pair terms G m_j m_i / r_ij of rows i0 .. i0 + gridDim.x - 1, j = i+1 ..
// N_interact-1, row after row (the reference's loop order).
__device__ __forceinline__ long energy_row_start(long i, long NI) { return i * (NI - 1) - i * (i - 1) / 2; }

__global__ __launch_bounds__(256) void k_energy_pair_terms(const double* __restrict__ x, const double* __restrict__ y,
                                                           const double* __restrict__ z, const double* __restrict__ m,
                                                           int i0, int N_interact, double G,
                                                           double* __restrict__ buf) {
    const int i = i0 + blockIdx.x;
    const long off = energy_row_start(i, N_interact) - energy_row_start(i0, N_interact) - (i + 1);
    const double xi = x[i], yi = y[i], zi = z[i], mi = m[i];
    for (int j = i + 1 + (int)threadIdx.x; j < N_interact; j += blockDim.x) {
        const double dx = xi - x[j];
        const double dy = yi - y[j];
        const double dz = zi - z[j];
        buf[off + j] = G * m[j] * mi / sqrt(dx * dx + dy * dy + dz * dz);
    }
}

// e_pot -= buf[0 .. n-1] in order, continuing out[1].
__global__ __launch_bounds__(64) void k_energy_fold_terms(const double* __restrict__ buf, long n,
                                                          double* __restrict__ out) {
    constexpr int PER = 16;
    constexpr int CH = 64 * PER;
    __shared__ __align__(16) double stage[2][CH];
    const int lane = threadIdx.x;
    double e = out[1];
    const long nch = (n + CH - 1) / CH;
    double r[PER];
    auto load = [&](long c) {
        const long base = c * CH + (long)lane * PER;
#pragma unroll
        for (int q = 0; q < PER; q++) r[q] = base + q < n ? buf[base + q] : 0.;
    };
    auto store = [&](int b) {
#pragma unroll
        for (int q = 0; q < PER; q++) stage[b][lane * PER + q] = r[q];
    };
    if (nch > 0) {
        load(0);
        store(0);
    }
    __syncthreads();
    for (long c = 0; c < nch; c++) {
        const int b = (int)(c & 1);
        if (c + 1 < nch) load(c + 1);   // in flight while lane 0 folds chunk c
        if (lane == 0) {
            const long cnt = n - c * CH;
            if (cnt >= CH) {
#pragma unroll 16
                for (int k = 0; k < CH; k++) e -= stage[b][k];
            } else {
                for (int k = 0; k < (int)cnt; k++) e -= stage[b][k];
            }
        }
        // stage[b ^ 1] was last read in the previous iteration, before its barrier
        if (c + 1 < nch) store(b ^ 1);
        __syncthreads();
    }
    if (lane == 0) out[1] = e;
}

__global__ __launch_bounds__(256) void k_energy_rows(const double* __restrict__ x, const double* __restrict__ y,
                                                     const double* __restrict__ z, const double* __restrict__ vx,
                                                     const double* __restrict__ vy, const double* __restrict__ vz,
                                                     const double* __restrict__ m, int N_active, int N_interact,
                                                     double G, double* __restrict__ rows) {
    __shared__ double4 tile[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const bool live = i < N_interact;
    const double xi = live ? x[i] : 0., yi = live ? y[i] : 0., zi = live ? z[i] : 0., mi = live ? m[i] : 0.;
    double s = 0., c = 0.;
    const int jstart = blockIdx.x * 256;
    for (int t = jstart; t < N_interact; t += 256) {
        __syncthreads();
        const int jj = t + threadIdx.x;
        if (jj < N_interact) tile[threadIdx.x] = make_double4(x[jj], y[jj], z[jj], m[jj]);
        __syncthreads();
        const int n = min(256, N_interact - t);
        if (live && i < N_active) {
            for (int k = 0; k < n; k++) {
                const int j = t + k;
                if (j <= i) continue;
                const double4 q = tile[k];
                const double dx = xi - q.x, dy = yi - q.y, dz = zi - q.z;
                kahan_add(-(G * q.w * mi / sqrt(dx * dx + dy * dy + dz * dz)), s, c);
            }
        }
    }
    if (live) {
        rows[i] = s - c;
        rows[N_interact + i] = 0.5 * mi * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
    }
}

__global__ __launch_bounds__(256) void k_energy_fold(const double* __restrict__ rows, int n, double* __restrict__ out) {
    // One workgroup: each lane folds a strided subset with Kahan, then lane 0
    // folds the 256 partials in lane order.
    __shared__ double ps[256], pk[256];
    for (int which = 0; which < 2; which++) {
        double s = 0., c = 0.;
        for (int i = threadIdx.x; i < n; i += 256) kahan_add(rows[which * n + i], s, c);
        ps[threadIdx.x] = s;
        pk[threadIdx.x] = c;
        __syncthreads();
        if (threadIdx.x == 0) {
            double S = 0., C = 0.;
            for (int k = 0; k < 256; k++) {
                kahan_add(ps[k], S, C);
                kahan_add(-pk[k], S, C);
            }
            out[which == 0 ? 1 : 0] = S;
        }
        __syncthreads();
    }
}

int energy_launch(const double* x, const double* y, const double* z, const double* vx, const double* vy,
                  const double* vz, const double* m, int n, int n_active, int n_interact, double G, bool exact,
                  double* out, hipStream_t st, double* terms, size_t terms_cap) {
    if (exact && !(terms && terms_cap > 0)) return 1;
    if (exact) {
        hipLaunchKernelGGL(k_energy_kin, dim3(1), dim3(64), 0, st, vx, vy, vz, m, n_interact, out);
        // rows in batches whose terms fit the buffer (a row has at most n_interact - 1)
        const long NI = n_interact;
        for (int i0 = 0; i0 < n_active;) {
            int i1 = i0;
            long cnt = 0;
            while (i1 < n_active && cnt + (NI - 1 - i1) <= (long)terms_cap) {
                cnt += NI - 1 - i1;
                i1++;
            }
            if (i1 == i0) return 1;   // a single row larger than the buffer
            if (cnt > 0) {
                hipLaunchKernelGGL(k_energy_pair_terms, dim3(i1 - i0), dim3(256), 0, st, x, y, z, m, i0, n_interact, G,
                                   terms);
                hipLaunchKernelGGL(k_energy_fold_terms, dim3(1), dim3(64), 0, st, terms, cnt, out);
            }
            i0 = i1;
        }
    } else {
        double* rows = out + 16;
        hipLaunchKernelGGL(k_energy_rows, dim3((n_interact + 255) / 256), dim3(256), 0, st, x, y, z, vx, vy, vz, m,
                           n_active, n_interact, G, rows);
        hipLaunchKernelGGL(k_energy_fold, dim3(1), dim3(256), 0, st, rows, n_interact, out);
    }
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

static constexpr size_t kEnergyExactMax = 16384;             // bodies
static constexpr size_t kEnergyTermsMax = size_t(4) << 20;    // pair terms per batch (32 MB)

double energy_device(struct reb_simulation* r) {
    DeviceState* d = dev_of(r);
    const size_t N = r->N;
    const size_t N_active = (r->N_active == SIZE_MAX) ? N : r->N_active;
    const size_t N_interact = (r->testparticle_type == 0) ? N_active : N;
    if (N == 0) return r->energy_offset;
    // EXACT: the reference's sequential sum, bit-identical, up to kEnergyExactMax
    // bodies (a serial chain of N^2/2 additions); above, the compensated sum.
    const bool exact = d->precision == REB_AMD_PRECISION_EXACT && N_interact <= kEnergyExactMax;
    if (d->precision == REB_AMD_PRECISION_EXACT && !exact && !d->warned_energy_exact) {
        d->warned_energy_exact = true;
        reb_simulation_warning(r,
                               "EXACT precision: above 16384 interacting bodies the energy is summed with Kahan "
                               "compensation in parallel, not in the reference's serial order, so its last bits "
                               "can differ from the reference's.");
    }
    const size_t terms_cap = exact ? std::min<size_t>(kEnergyTermsMax, N_active * N_interact / 2 + N_interact) : 0;
    const size_t head = 2 * N + 16;
    if (!d->scratch.ensure(sizeof(double) * (head + terms_cap))) {
        reb_simulation_error(r, "Out of device memory for the energy reduction.");
        return nan("");
    }
    double* out = d->scratch.as<double>();
    if (energy_launch(d->field(FX), d->field(FY), d->field(FZ), d->field(FVX), d->field(FVY), d->field(FVZ),
                      d->field(FM), (int)N, (int)N_active, (int)N_interact, r->G, exact, out, d->stream,
                      exact ? out + head : nullptr, terms_cap)) {
        (void)RB_HIP(r, hipGetLastError());
        return nan("");
    }
    double h[2] = {0, 0};
    if (!RB_HIP(r, hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, d->stream))) return nan("");
    if (!stream_sync(r)) return nan("");
    return h[0] + h[1] + r->energy_offset;
}

}  // namespace rbamd

using namespace rbamd;

extern "C" {

REB_API size_t reb_amd_gravity_workspace_bytes(size_t N, int precision) {
    if (precision == REB_AMD_PRECISION_EXACT) return 256;
    return workspace_bytes((int)N, pick_chunks((int)N, (int)N, 1)) + 256;
}

REB_API int reb_amd_gravity_direct(const double* x, const double* y, const double* z, const double* m, double* ax,
                                   double* ay, double* az, size_t N, size_t N_active, int testparticle_type,
                                   int ignore_terms, double G, double softening, int precision, int compensated,
                                   void* workspace, void* stream) {
    if (N > (size_t)INT32_MAX / 8) {
        set_last_error("reb_amd_gravity_direct: N too large");
        return 2;
    }
    if (N_active == SIZE_MAX || N_active > N) N_active = N;
    if (precision != REB_AMD_PRECISION_EXACT && workspace == nullptr) {
        set_last_error("reb_amd_gravity_direct: workspace required for this precision");
        return 3;
    }
    hipStream_t st = (hipStream_t)stream;
    GBList gl;
    gl.n = 1;
    gl.g[0] = GB3{0., 0., 0.};
    return direct_dispatch(x, y, z, m, ax, ay, az, (int)N, (int)N_active, testparticle_type, ignore_terms, G,
                           softening * softening, precision, compensated, gl, static_cast<double*>(workspace), st);
}

}  // extern "C"

// ---------------------------------------------------------------- no-allocation variant
// The chunk partial sums of the fast modes live in a fixed device array
// instead of a caller workspace.  The name `scratch` is a host-side stand-in
// in host code; its device address is looked up once per device
// (hipGetSymbolAddress with that device current) and only that address is
// passed to kernels.  Every partial-sum entry the reduction reads is written
// by the pair kernel of the same call first, so nothing needs clearing.
// Calls that overlap on one device share the array and race.
namespace rbamd {

static constexpr size_t kScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];

static void* scratch_on_current_device() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_last_error("reb_amd_gravity_direct_noalloc: no current HIP device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            set_last_error(std::string("reb_amd_gravity_direct_noalloc: hipGetSymbolAddress failed: ") +
                           hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

}  // namespace rbamd

extern "C" {

REB_API size_t reb_amd_gravity_noalloc_max_bytes(void) { return kScratchBytes; }

REB_API int reb_amd_gravity_direct_noalloc(const double* x, const double* y, const double* z, const double* m,
                                           double* ax, double* ay, double* az, size_t N, size_t N_active,
                                           int testparticle_type, int ignore_terms, double G, double softening,
                                           int precision, int compensated, void* stream) {
    if (N > (size_t)INT32_MAX / 8) {
        set_last_error("reb_amd_gravity_direct_noalloc: N too large");
        return 2;
    }
    if (N_active == SIZE_MAX || N_active > N) N_active = N;
    const size_t need = reb_amd_gravity_workspace_bytes(N, precision);
    if (need > kScratchBytes) {
        set_last_error("reb_amd_gravity_direct_noalloc: N=" + std::to_string(N) + " needs " + std::to_string(need) +
                       " bytes of partial sums, above the fixed " + std::to_string(kScratchBytes) +
                       "-byte device array; use reb_amd_gravity_direct with a workspace");
        return 4;
    }
    void* ws = scratch_on_current_device();
    if (ws == nullptr) return 5;
    hipStream_t st = (hipStream_t)stream;
    GBList gl;
    gl.n = 1;
    gl.g[0] = GB3{0., 0., 0.};
    return direct_dispatch(x, y, z, m, ax, ay, az, (int)N, (int)N_active, testparticle_type, ignore_terms, G,
                           softening * softening, precision, compensated, gl, static_cast<double*>(ws), st);
}

}  // extern "C"
