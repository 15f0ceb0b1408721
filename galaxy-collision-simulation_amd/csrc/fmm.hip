// fmm.hip — hip-fmm: the reference's octree FMM (cpp/fmm/fmm_solver.cpp,
// cpp/fmm/multipole.cpp) on the resident particle state.
//
// Per force evaluation, all on the context's stream:
//  1. bounding box -> root cube exactly as fmm_solver.cpp:60-85 (FP64);
//  2. per particle a 96-bit octant key: 32 levels of the reference's
//     ">=" tests against child centres recursed as c +- hw/2 in FP64
//     (fmm_solver.cpp:106-134), so the key path equals the insertion path;
//     levels 1-21 in one 63-bit word, levels 22-32 in a 33-bit word;
//  3. stable radix sort of (key, index) (hipCUB) on the first word; only
//     when more than leaf_capacity particles share all 21 levels of it (and
//     max_depth > 21) are both words sorted;
//  4. level-by-level split: a node at depth L is internal iff it holds more
//     than leaf_capacity particles and L < max_depth — the tree the
//     reference's insertion builds (fmm_solver.cpp:136-155), to its depth
//     cap of 32.  Empty children are not materialised (they have zero mass
//     and never enter a list).  The deepest internal level is known before
//     the split (a node at level L holds more than c particles iff sorted
//     particles i and i + c share L levels), and level sizes stay on the
//     device: each level's kernels run over a bound on its node count;
//  5. upward pass, deepest level first: mass, centre of mass, 35 raw
//     moments about the COM in FP64 (P2M at leaves, binomial M2M shift);
//     then per node the far-field coefficient table of far_m2p.hpp (FP64
//     and FP32) from N_alpha = M_alpha / alpha!;
//  6. per target leaf a preorder DFS with the reference's acceptance test
//     2 (h_t + h_s) / |c_s - c_t| < theta on geometric centres
//     (fmm_solver.cpp:228-277): far cells and near leaves, in the
//     reference's list order;
//  7. evaluation, one 64-lane wave per target leaf: every far cell by M2P
//     (the radial-kernel form of far_m2p.hpp over the cell's table), every
//     near leaf by P2P over sources staged in LDS as offsets from the
//     target leaf's centre; self pairs skipped (fmm_solver.cpp:295-310).
// The evaluation is instantiated in FP32 (precision "mixed": FP32 pairs and
// M2P on FP32 offsets, FP64 accumulation) and FP64.
#include "gc_fmm.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace gc {
namespace {

constexpr int kLevels = 21;       // levels in the first key word: 3 bits each, 63 bits
constexpr int kDeepLevels = 11;   // levels 22-32 in the second word: 33 bits
constexpr int kMaxLevels = kLevels + kDeepLevels;  // the reference's depth cap
constexpr int kMom = 35;          // raw Cartesian moments, total degree <= 4
constexpr int kEvalLanes = 64;
// device statistics: far, near, M2P, pairs (FMM) or M2P, pairs, leaves
// (tree), then internal nodes at kStatInternal
constexpr int kStats = 16;
constexpr int kStatInternal = 4;
constexpr int kStatMassMax = 5;  // largest |m| of the build (double bits)
constexpr int kStatMassMin = 6;  // smallest m of the build (double bits)
constexpr int kStatTotals = 8;   // build totals read with the statistics (build_totals_kernel)

__device__ __forceinline__ float gc_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double gc_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float gc_rsq(float s) { return __builtin_amdgcn_rsqf(s); }
__device__ __forceinline__ double gc_rsq(double s) { return 1.0 / __builtin_sqrt(s); }

#define GC_M2P_INLINE __device__ __forceinline__
#define GC_FMA(a, b, c) gc_fma((a), (b), (c))
#define GC_RSQ(s) gc_rsq(s)
#include "far_m2p.hpp"
#undef GC_FMA
#undef GC_RSQ

// ---- multipole tables (graded exponent order of multipole.cpp:17-26) ----
// 1 / alpha! for the 35 moments in the graded order of multipole.cpp:17-26
// (N_alpha = M_alpha / alpha! is what M2P contracts with).
struct Tables {
    double inv_fact[kMom];
};

constexpr Tables make_tables() {
    Tables t{};
    int k = 0;
    const int fact[5] = {1, 1, 2, 6, 24};
    for (int d = 0; d <= 4; ++d)
        for (int a = d; a >= 0; --a)
            for (int b = d - a; b >= 0; --b) {
                t.inv_fact[k] = 1.0 / static_cast<double>(fact[a] * fact[b] * fact[d - a - b]);
                ++k;
            }
    return t;
}

__constant__ Tables c_tables = make_tables();

// ---- 1. bounding box -----------------------------------------------------
__global__ __launch_bounds__(256) void bbox_partial_kernel(int n, const double* __restrict__ x,
                                                           const double* __restrict__ y,
                                                           const double* __restrict__ z,
                                                           double* __restrict__ part) {
    __shared__ double red[6][256];
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const double p[3] = {x[i], y[i], z[i]};
        for (int d = 0; d < 3; ++d) {
            v[d] = fmin(v[d], p[d]);
            v[3 + d] = fmax(v[3 + d], p[d]);
        }
    }
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            for (int d = 0; d < 3; ++d) {
                red[d][threadIdx.x] = fmin(red[d][threadIdx.x], red[d][threadIdx.x + s]);
                red[3 + d][threadIdx.x] = fmax(red[3 + d][threadIdx.x], red[3 + d][threadIdx.x + s]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// Root cube of fmm_solver.cpp:76-85 into root[4] (centre, half width), and
// the first level offsets (level 0 is node 0).  One wave: lanes take the
// block partials in strides, then wavefront min / max (exact, so the order
// does not matter).
__global__ __launch_bounds__(64) void bbox_final_kernel(int nblocks, const double* __restrict__ part,
                                                        double eps, double* __restrict__ root,
                                                        int* __restrict__ lvl) {
#pragma clang fp contract(off)
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < nblocks; b += 64)
        for (int d = 0; d < 3; ++d) {
            lo[d] = fmin(lo[d], part[b * 6 + d]);
            hi[d] = fmax(hi[d], part[b * 6 + 3 + d]);
        }
    for (int d = 0; d < 3; ++d) {
        for (int o = 32; o > 0; o >>= 1) {
            lo[d] = fmin(lo[d], __shfl_xor(lo[d], o, 64));
            hi[d] = fmax(hi[d], __shfl_xor(hi[d], o, 64));
        }
    }
    if (threadIdx.x != 0) return;
    double ext = hi[0] - lo[0];
    if (ext < hi[1] - lo[1]) ext = hi[1] - lo[1];
    if (ext < hi[2] - lo[2]) ext = hi[2] - lo[2];
    double hw = 0.5 * ext;
    if (hw < eps + 1.0e-6) hw = eps + 1.0e-6;
    hw *= 1.0001;
    for (int d = 0; d < 3; ++d) root[d] = (lo[d] + hi[d]) * 0.5;
    root[3] = hw;
    lvl[0] = 0;
    lvl[1] = 1;
}

// Node 0 of the node arrays: the root cube over all n particles.
__global__ void root_kernel(int n, const double* __restrict__ root, double* __restrict__ geom,
                            int* __restrict__ nb, int* __restrict__ ne) {
    for (int d = 0; d < 4; ++d) geom[d] = root[d];
    nb[0] = 0;
    ne[0] = n;
}

// ---- 2. octant keys --------------------------------------------------------
// Levels 1-21 into keys (digit of level l at bits 3 (21 - l)), levels 22-32
// into keys_lo (digit of level l at bits 3 (32 - l)) when keys_lo is given.
__global__ __launch_bounds__(256) void keys_kernel(int n, const double* __restrict__ x,
                                                   const double* __restrict__ y,
                                                   const double* __restrict__ z,
                                                   const double* __restrict__ root,
                                                   unsigned long long* __restrict__ keys,
                                                   unsigned long long* __restrict__ keys_lo,
                                                   int* __restrict__ idx) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double c0 = root[0], c1 = root[1], c2 = root[2], hw = root[3];
    const double px = x[i], py = y[i], pz = z[i];
    unsigned long long key[2] = {0, 0};
    const int levels = keys_lo ? kMaxLevels : kLevels;
    for (int l = 0; l < levels; ++l) {
        const int ox = px >= c0 ? 1 : 0, oy = py >= c1 ? 1 : 0, oz = pz >= c2 ? 1 : 0;
        const int w = l < kLevels ? 0 : 1;
        key[w] = (key[w] << 3) | static_cast<unsigned long long>(ox + 2 * oy + 4 * oz);
        const double chw = hw * 0.5;
        c0 = c0 + (ox ? 1.0 : -1.0) * chw;
        c1 = c1 + (oy ? 1.0 : -1.0) * chw;
        c2 = c2 + (oz ? 1.0 : -1.0) * chw;
        hw = chw;
    }
    keys[i] = key[0];
    if (keys_lo) keys_lo[i] = key[1];
    idx[i] = i;
}

// out[k] = in[perm[k]]
template <typename T>
__global__ __launch_bounds__(256) void permute_kernel(int n, const int* __restrict__ perm,
                                                      const T* __restrict__ in, T* __restrict__ out) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < n) out[k] = in[perm[k]];
}

// Levels two keys share (0..21 on the first word; 21 + 0..11 with the second).
__device__ __forceinline__ int shared_levels(unsigned long long a, unsigned long long b,
                                             const unsigned long long* lo, int i, int j) {
    const unsigned long long d = a ^ b;
    if (d != 0) return kLevels - 1 - (63 - __clzll(static_cast<long long>(d))) / 3;
    if (!lo) return kLevels;
    const unsigned long long e = lo[i] ^ lo[j];
    if (e != 0) return kMaxLevels - 1 - (63 - __clzll(static_cast<long long>(e))) / 3;
    return kMaxLevels;
}

// The deepest level at which some node holds more than cap particles: the
// most levels that sorted particles i and i + cap share (atomic max into
// *out, which starts at -1).
__global__ __launch_bounds__(256) void deepest_kernel(int n, int cap,
                                                      const unsigned long long* __restrict__ keys,
                                                      const unsigned long long* __restrict__ lo,
                                                      int* __restrict__ out) {
    int best = -1;
    for (int i = blockIdx.x * 256 + threadIdx.x; static_cast<long long>(i) + cap < n;
         i += gridDim.x * 256)
        best = max(best, shared_levels(keys[i], keys[i + cap], lo, i, i + cap));
    for (int o = 32; o > 0; o >>= 1) best = max(best, __shfl_xor(best, o, 64));
    if ((threadIdx.x & 63) == 0 && best >= 0) atomicMax(out, best);
}

// ---- 4. level split -------------------------------------------------------
// Digit of level 1..32 from the two key words.
__device__ __forceinline__ int digit_at(const unsigned long long* keys,
                                        const unsigned long long* lo, int i, int level) {
    if (level <= kLevels) return static_cast<int>((keys[i] >> (3 * (kLevels - level))) & 7ull);
    return static_cast<int>((lo[i] >> (3 * (kMaxLevels - level))) & 7ull);
}

// Children of the nodes of level L, nodes [lvl[L], lvl[L + 1]): bounds[9]
// in sorted order and the non-empty count.  Launched over `bound` >= the
// level's node count; entries past it get count 0, so the scan over
// `bound` entries is the scan over the level.  Eight lanes per node: lane d
// finds the first position whose digit is >= d by its own binary search.
constexpr int kSplitLanes = 8;

__global__ __launch_bounds__(256) void split_count_kernel(
    int level, int bound, const int* __restrict__ lvl, int cap, int max_depth,
    const unsigned long long* __restrict__ keys, const unsigned long long* __restrict__ keys_lo,
    const int* __restrict__ nb, const int* __restrict__ ne, int* __restrict__ bounds,
    int* __restrict__ count, int* __restrict__ overflow) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int j = t / kSplitLanes, d = t % kSplitLanes;  // a group's exits are uniform (one j)
    if (j >= bound) return;
    const int k = lvl[level] + j;
    if (k >= lvl[level + 1]) {
        if (d == 0) count[j] = 0;
        return;
    }
    const int b = nb[k], e = ne[k];
    int* bd = bounds + 9 * j;
    const bool internal = (e - b) > cap && level < max_depth;
    if (!internal) {
        if (d == 0) count[j] = 0;
        return;
    }
    if (level >= kMaxLevels || (level >= kLevels && !keys_lo)) {  // excluded by the host
        if (d == 0) {
            atomicOr(overflow, 1);
            count[j] = 0;
        }
        return;
    }
    int lo = b, hi = e;  // first position with digit >= d (b for d = 0)
    while (d > 0 && lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (digit_at(keys, keys_lo, mid, level + 1) < d) lo = mid + 1;
        else hi = mid;
    }
    bd[d] = lo;
    if (d == 0) bd[8] = e;
    const int up = __shfl_down(lo, 1, kSplitLanes);
    int c = (d == kSplitLanes - 1 ? e : up) > lo ? 1 : 0;  // child d is non-empty
    c += __shfl_xor(c, 1, kSplitLanes);
    c += __shfl_xor(c, 2, kSplitLanes);
    c += __shfl_xor(c, 4, kSplitLanes);
    if (d == 0) count[j] = c;
}

// Emits level L + 1 (children in digit order, parents in order) at
// lvl[L + 1] + offset, and its end lvl[L + 2].  Node slots beyond node_cap
// (excluded by the host's bound) set overflow bit 2 instead of writing.
__global__ __launch_bounds__(256) void split_emit_kernel(
    int level, int bound, int* __restrict__ lvl, int node_cap, const int* __restrict__ bounds,
    const int* __restrict__ count, const int* __restrict__ offset, double* __restrict__ geom,
    int* __restrict__ nb, int* __restrict__ ne, int* __restrict__ first_child,
    int* __restrict__ nchild, int* __restrict__ overflow) {
#pragma clang fp contract(off)
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= bound) return;
    const int lb = lvl[level], le = lvl[level + 1];
    const int k = lb + j;
    if (k >= le) {
        if (j == 0) lvl[level + 2] = le;  // empty level
        return;
    }
    const int c = count[j];
    if (k == le - 1) lvl[level + 2] = le + offset[j] + c;
    nchild[k] = c;
    if (c == 0) {
        first_child[k] = -1;
        return;
    }
    const int* bd = bounds + 9 * j;
    int pos = le + offset[j];
    if (pos + c > node_cap) {
        atomicOr(overflow, 2);
        nchild[k] = 0;
        first_child[k] = -1;
        return;
    }
    first_child[k] = pos;
    const double chw = geom[4 * k + 3] * 0.5;
    for (int d = 0; d < 8; ++d) {
        if (bd[d + 1] <= bd[d]) continue;
        geom[4 * pos + 0] = geom[4 * k + 0] + ((d & 1) ? 1.0 : -1.0) * chw;
        geom[4 * pos + 1] = geom[4 * k + 1] + ((d & 2) ? 1.0 : -1.0) * chw;
        geom[4 * pos + 2] = geom[4 * k + 2] + ((d & 4) ? 1.0 : -1.0) * chw;
        geom[4 * pos + 3] = chw;
        nb[pos] = bd[d];
        ne[pos] = bd[d + 1];
        ++pos;
    }
}

// Internal nodes (the reference's node_count is 1 + 8 x this): one atomic
// per wave into *out.
__global__ __launch_bounds__(256) void internal_count_kernel(int nnodes, const int* __restrict__ nchild,
                                                             unsigned long long* __restrict__ out) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    const unsigned long long m = __ballot(k < nnodes && nchild[k] > 0);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(out, static_cast<unsigned long long>(__popcll(m)));
}

// ---- 5. upward pass -------------------------------------------------------
#define GC_UP_INLINE __device__ __forceinline__
#include "upward_generated.inc"
#undef GC_UP_INLINE

// Eight lanes per node of a level: mass and centre of mass, then the raw
// moments about the COM — P2M over the leaf's particles (key order, the
// sorted copies; lane g takes particles g, g + 8, ...) or M2M over the
// children (lane g takes child g) (multipole.cpp:226-261,
// fmm_solver.cpp:157-208).  The lanes' partial sums are added by a fixed
// butterfly over the 8 lanes (every lane ends with the same sums).
// Moments stay in registers (generated code).
constexpr int kUpLanes = 8;

__device__ __forceinline__ double up_sum(double v) {
#pragma clang fp contract(off)
    v += __shfl_xor(v, 1, kUpLanes);
    v += __shfl_xor(v, 2, kUpLanes);
    v += __shfl_xor(v, 4, kUpLanes);
    return v;
}

__global__ __launch_bounds__(64) void upward_kernel(
    int lb, int le, const int* __restrict__ nb, const int* __restrict__ ne,
    const int* __restrict__ first_child, const int* __restrict__ nchild,
    const double* __restrict__ geom, const double* __restrict__ sx, const double* __restrict__ sy,
    const double* __restrict__ sz, const double* __restrict__ sm, double* __restrict__ mcom,
    double* __restrict__ mom) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * 64 + threadIdx.x;
    const int g = t & (kUpLanes - 1);
    // a group past the level's end still runs its shuffles (all lanes of
    // a group take the same path), then writes nothing
    const int k = lb + t / kUpLanes;
    const bool live = k < le;
    const int kk = live ? k : lb;
    double mass = 0.0, w[3] = {0.0, 0.0, 0.0}, M[kMom];
#pragma unroll
    for (int a = 0; a < kMom; ++a) M[a] = 0.0;
    const int nc = nchild[kk];
    const int b = nb[kk], e = ne[kk];
    const int fc = first_child[kk];
    if (nc == 0) {
        for (int j = b + g; j < e; j += kUpLanes) {
            const double mj = sm[j];
            mass += mj;
            w[0] += sx[j] * mj;
            w[1] += sy[j] * mj;
            w[2] += sz[j] * mj;
        }
    } else if (g < nc) {
        const int c = fc + g;
        const double cm = mcom[4 * c];
        mass += cm;
        for (int d = 0; d < 3; ++d) w[d] += mcom[4 * c + 1 + d] * cm;
    }
    mass = up_sum(mass);
    for (int d = 0; d < 3; ++d) w[d] = up_sum(w[d]);
    double com[3];
    for (int d = 0; d < 3; ++d) com[d] = mass > 0.0 ? w[d] / mass : geom[4 * kk + d];
    if (nc == 0) {
        for (int j = b + g; j < e; j += kUpLanes)
            p2m_add(sx[j] - com[0], sy[j] - com[1], sz[j] - com[2], sm[j], M);
    } else if (g < nc) {
        const int c = fc + g;
        double C[kMom];
        const double* cmom = mom + static_cast<size_t>(kMom) * c;
#pragma unroll
        for (int a = 4; a < kMom; ++a) C[a] = cmom[a];
        m2m_add(mcom[4 * c + 1] - com[0], mcom[4 * c + 2] - com[1], mcom[4 * c + 3] - com[2],
                mcom[4 * c], C, M);
    }
#pragma unroll
    for (int a = 0; a < kMom; ++a) M[a] = up_sum(M[a]);
    if (!live) return;
    if (g == 0) {
        mcom[4 * k] = mass;
        for (int d = 0; d < 3; ++d) mcom[4 * k + 1 + d] = com[d];
    }
    double* out = mom + static_cast<size_t>(kMom) * k;
#pragma unroll
    for (int a = 0; a < kMom; ++a)
        if (a % kUpLanes == g) out[a] = M[a];
}

// Per node, the far-field coefficient table of far_m2p.hpp from the FP64
// moments (N_alpha = M_alpha / alpha!, N_0 = mass), in FP64 and FP32.
__global__ __launch_bounds__(64) void far_coef_kernel(int nnodes, const double* __restrict__ mcom,
                                                      const double* __restrict__ mom,
                                                      double* __restrict__ kc64,
                                                      float* __restrict__ kc32) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * 64 + threadIdx.x;
    if (k >= nnodes) return;
    double N[kMom], K[kFarCoef];
    const double* src = mom + static_cast<size_t>(kMom) * k;
    N[0] = mcom[4 * k];
    N[1] = N[2] = N[3] = 0.0;
#pragma unroll
    for (int a = 4; a < kMom; ++a) N[a] = src[a] * c_tables.inv_fact[a];
    far_coefficients(N, K);
    double* o64 = kc64 + static_cast<size_t>(kFarCoef) * k;
    float* o32 = kc32 + static_cast<size_t>(kFarCoef) * k;
#pragma unroll
    for (int a = 0; a < kFarCoef; ++a) {
        o64[a] = K[a];
        o32[a] = static_cast<float>(K[a]);
    }
}

// leaf flags for the scan of target leaves: non-empty leaves with mass > 0
__global__ __launch_bounds__(256) void leaf_flag_kernel(int nnodes, const int* __restrict__ nchild,
                                                        const double* __restrict__ mcom,
                                                        int* __restrict__ flag) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= nnodes) return;
    flag[k] = (nchild[k] == 0 && mcom[4 * k] > 0.0) ? 1 : 0;
}

__global__ __launch_bounds__(256) void leaf_collect_kernel(int nnodes, const int* __restrict__ flag,
                                                           const int* __restrict__ pos,
                                                           int* __restrict__ leaves) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= nnodes) return;
    if (flag[k]) leaves[pos[k]] = k;
}

// ---- 6. interaction lists ----------------------------------------------

// The reference's acceptance test (fmm_solver.cpp:243-277): cell s is far
// from target t when 2 (h_t + h_s) / |c_s - c_t| < theta (never at zero
// distance); h is the half width.
__device__ __forceinline__ bool well_separated_at(double tx, double ty, double tz, double th,
                                                  double sx, double sy, double sz, double sh,
                                                  double theta) {
#pragma clang fp contract(off)
    const double dx = sx - tx, dy = sy - ty, dz = sz - tz;
    const double dist = __builtin_sqrt(dx * dx + dy * dy + dz * dz);
    if (dist == 0.0) return false;
    const double w = 2.0 * (th + sh);
    return w / dist < theta;
}

// Interaction lists, one wave per target leaf.  The wave walks the tree
// breadth first: each round tests up to 64 frontier nodes at once with the
// reference's rules (fmm_solver.cpp:243-277) — a node is skipped if
// massless, a far cell if it is not the target and
// 2 (h_t + h_s) / |c_s - c_t| < theta, a near leaf if it has no children,
// and otherwise opened, its children joining the next frontier in child
// order.  Far cells and near leaves are appended in lane order (ballot
// prefix sums), so the lists are deterministic; they hold the reference's
// cells, in breadth-first rather than depth-first order.  The grid is
// persistent: wave w takes leaves w, w + G, ...; its two frontier buffers of
// `fcap` nodes hold their first kListLds entries in LDS and the rest in
// `scratch`.  A round loads all the fields of its 64 nodes at once.  A frontier that would outgrow `fcap` sets
// `overflow` and the host repeats the pass with larger buffers.
// FILL == false: count far cells, near leaves and near source particles per
// leaf; FILL == true: write the lists at the scanned offsets.
// MODE 0 counts each leaf's far cells, near leaves and near sources;
// MODE 1 writes the lists at far_off / near_off from an earlier count;
// MODE 2 counts and at once stashes the lists in the wave's own scratch
// region (wcap_far / wcap_near entries per wave, the leaf's start there in
// lpos_far / lpos_near), for list_copy_kernel to move them to their places
// once the counts are scanned.  Bit 1 of *overflow flags a wave whose
// lists outgrew its region (counts stay exact; MODE 1 then writes the
// lists), bit 0 a frontier overflow.
constexpr int kListLds = 512;  // frontier entries per buffer held in LDS

template <int MODE>
__global__ __launch_bounds__(64) void lists_kernel(
    int nleaves, const int* __restrict__ leaves, const double* __restrict__ geom,
    const double* __restrict__ mcom, const int* __restrict__ first_child,
    const int* __restrict__ nchild, const int* __restrict__ nb, const int* __restrict__ ne,
    double theta, int* __restrict__ nfar, int* __restrict__ nnear, int* __restrict__ nsrc,
    const long long* __restrict__ far_off, const long long* __restrict__ near_off,
    long long wcap_far, long long wcap_near, long long* __restrict__ lpos_far,
    long long* __restrict__ lpos_near, int* __restrict__ far, int* __restrict__ near,
    int* __restrict__ scratch, int fcap, int* __restrict__ overflow) {
    constexpr bool COUNT = MODE != 1, WRITE = MODE != 0;
    const int lane = threadIdx.x;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int* bufa = scratch + static_cast<size_t>(blockIdx.x) * 2 * fcap;
    int* bufb = bufa + fcap;
    // MODE 2: this wave's stash regions and how much of them is used
    int* const wfar = far + static_cast<size_t>(blockIdx.x) * wcap_far;
    int* const wnear = near + static_cast<size_t>(blockIdx.x) * wcap_near;
    long long used_f = 0, used_n = 0;
    // frontier entries below kListLds live in LDS, the rest in `scratch`
    __shared__ int lfront[2][kListLds];
    for (int li = blockIdx.x; li < nleaves; li += gridDim.x) {
        const int t = leaves[li];
        const double tx = geom[4 * t], ty = geom[4 * t + 1], tz = geom[4 * t + 2], th = geom[4 * t + 3];
        int* cur = bufa;
        int* nxt = bufb;
        int* lcur = lfront[0];
        int* lnxt = lfront[1];
        if (lane == 0) lcur[0] = 0;
        __syncthreads();  // one wave: orders the frontier writes before the reads
        int ncur = 1;
        long long cf = 0, cn = 0, cs = 0;
        long long fo = 0, no = 0, flim = 0, nlim = 0;
        int* fdst = far;
        int* ndst = near;
        if (MODE == 1) {
            fo = far_off[li];
            no = near_off[li];
            flim = nlim = 1ll << 62;
        } else if (MODE == 2) {
            fdst = wfar;
            ndst = wnear;
            fo = used_f;
            no = used_n;
            flim = wcap_far - used_f;
            nlim = wcap_near - used_n;
        }
        bool ovf = false;
        while (ncur > 0 && !ovf) {
            int nnext = 0;
            for (int base = 0; base < ncur; base += 64) {
                const int j = base + lane;
                int node = -1, kind = 3, nc = 0, fc = 0;
                if (j < ncur) {
                    node = j < kListLds ? lcur[j] : cur[j];
                    // the node's fields load together (one memory round trip)
                    const double mass = mcom[4 * node];
                    const double gx = geom[4 * node], gy = geom[4 * node + 1],
                                 gz = geom[4 * node + 2], gh = geom[4 * node + 3];
                    nc = nchild[node];
                    fc = first_child[node];
                    if (mass > 0.0) {
                        if (node != t && well_separated_at(tx, ty, tz, th, gx, gy, gz, gh, theta)) {
                            kind = 0;
                        } else {
                            kind = nc == 0 ? 1 : 2;
                        }
                    }
                }
                const unsigned long long mfar = __ballot(kind == 0);
                const unsigned long long mnear = __ballot(kind == 1);
                if (WRITE) {
                    const long long pf = cf + __popcll(mfar & below);
                    const long long pn = cn + __popcll(mnear & below);
                    if (kind == 0 && pf < flim) fdst[fo + pf] = node;
                    if (kind == 1 && pn < nlim) ndst[no + pn] = node;
                }
                cf += __popcll(mfar);
                cn += __popcll(mnear);
                if (COUNT && kind == 1) cs += ne[node] - nb[node];  // this lane's part
                // children of opened nodes (at most 8) join the next frontier
                // in lane order: prefix sums of the counts' bits by ballot
                const int open = kind == 2 ? nc : 0;
                int excl = 0, total = 0;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const unsigned long long mb = __ballot((open >> b) & 1);
                    excl += __popcll(mb & below) << b;
                    total += __popcll(mb) << b;
                }
                if (nnext + total > fcap) {
                    ovf = true;
                    break;
                }
                if (kind == 2)
                    for (int c = 0; c < nc; ++c) {
                        const int k = nnext + excl + c;
                        if (k < kListLds) lnxt[k] = fc + c;
                        else nxt[k] = fc + c;
                    }
                nnext += total;
            }
            __syncthreads();
            int* tmp = cur;
            cur = nxt;
            nxt = tmp;
            tmp = lcur;
            lcur = lnxt;
            lnxt = tmp;
            ncur = nnext;
        }
        if (ovf) {
            if (lane == 0) atomicOr(overflow, 1);
            continue;
        }
        if (COUNT)
            for (int o = 32; o > 0; o >>= 1) cs += __shfl_xor(cs, o, 64);
        if (COUNT && lane == 0) {
            nfar[li] = static_cast<int>(cf);
            nnear[li] = static_cast<int>(cn);
            nsrc[li] = static_cast<int>(cs);
        }
        if (MODE == 2) {
            if (lane == 0) {
                lpos_far[li] = fo;
                lpos_near[li] = no;
                if (cf > flim || cn > nlim) atomicOr(overflow, 2);
            }
            used_f += min(cf, flim);
            used_n += min(cn, nlim);
        }
    }
}

// Moves leaf li's stashed lists (lists_kernel<2>) to far_off / near_off:
// one block per leaf, coalesced.
__global__ __launch_bounds__(256) void list_copy_kernel(
    int nleaves, int lwaves, const int* __restrict__ nfar, const int* __restrict__ nnear,
    const long long* __restrict__ far_off, const long long* __restrict__ near_off,
    const long long* __restrict__ lpos_far, const long long* __restrict__ lpos_near,
    long long wcap_far, long long wcap_near, const int* __restrict__ wfar,
    const int* __restrict__ wnear, int* __restrict__ far, int* __restrict__ near) {
    const int li = blockIdx.x;
    if (li >= nleaves) return;
    const int w = li % lwaves;
    const int* sf = wfar + static_cast<size_t>(w) * wcap_far + lpos_far[li];
    const int* sn = wnear + static_cast<size_t>(w) * wcap_near + lpos_near[li];
    int* df = far + far_off[li];
    int* dn = near + near_off[li];
    for (int k = threadIdx.x; k < nfar[li]; k += 256) df[k] = sf[k];
    for (int k = threadIdx.x; k < nnear[li]; k += 256) dn[k] = sn[k];
}

// per-wave list totals of lists_kernel's leaf-to-wave assignment (li mod
// lwaves), to size the stash regions
// Persistent waves of the list traversals (and slots of the per-wave
// totals): 8 per SIMD on 256 CUs.  The traversal is latency-bound; at
// N = 2^20 8192 waves build in 5.68 ms, 4096 in 5.96 ms, 16384 in 7.10 ms.
constexpr int kListWaves = 8192;

__global__ __launch_bounds__(256) void wave_totals_kernel(int nleaves, int lwaves,
                                                          const int* __restrict__ nfar,
                                                          const int* __restrict__ nnear,
                                                          long long* __restrict__ wtot) {
    const int w = blockIdx.x * 256 + threadIdx.x;
    if (w >= lwaves) return;
    long long f = 0, n = 0;
    for (int li = w; li < nleaves; li += lwaves) {
        f += nfar[li];
        n += nnear[li];
    }
    wtot[w] = f;
    wtot[lwaves + w] = n;
}

// ---- 7. evaluation ----------------------------------------------------------
template <typename F>
struct Src {
    F x, y, z, gm;
    int id;
};

// Far cells staged per chunk: COM offset from the target leaf's centre and
// the cell's far-field coefficient table (far_m2p.hpp).  32 cells per chunk
// in FP32 (8.7 KB of tables), 16 in FP64.
template <typename F>
constexpr int far_chunk() { return sizeof(F) == sizeof(float) ? 32 : 16; }
constexpr int kNearPiece = 64;

template <typename F>
struct alignas(16) FarStage {
    F c[far_chunk<F>()][4];
    F K[far_chunk<F>()][kFarCoef];
    int id[far_chunk<F>()];
};

// Evaluation, in two kernels per target leaf (one 64-lane wave each),
// particles in key order (sorted FP64 copies sx..sm).  A leaf chunk of
// c <= 64 targets is served by G = floor(64 / c) lane groups; per lane the
// FP32 (or FP64) terms of a chunk are added into FP64 sums, and the groups'
// sums are added in group order through LDS (deterministic).
//  far  in FP64 (eval_far_kernel): chunks of 16 far cells are staged in LDS
//       by one cooperative load (COM offsets, coefficient tables); group g
//       evaluates cells g, g + G, ... of the chunk by M2P; writes G * sum.
//       In FP32 the far field runs with the cells across the lanes instead
//       (eval_far_cells_kernel, below).
//  near (eval_near_kernel): the near leaves' particles, 64 leaves at a time
//       laid end to end, are staged kNearStage sources per round, the next
//       round's sources loaded into registers while the current one is
//       evaluated (each lane finds its sources' leaves by binary search over
//       the leaves' prefix sizes); group g takes sources g, g + G, ...; the
//       self pair is skipped (fmm_solver.cpp:295-310); adds its sum to the
//       far kernel's and applies the closing kick.
// Splitting keeps the register-heavy M2P out of the P2P kernel, which then
// runs at full occupancy.
// Targets per lane for a leaf chunk of c targets: t = 1, 2 or 3, whichever
// keeps the most of the wave's 64 x t target slots busy when the chunk is
// spread over floor(64 / ceil(c / t)) lane groups (ties to the smaller t).
// On a Plummer sphere of 2^20 particles in leaves of <= 64 this serves
// 93.5% of the slots, against 81% for one target per lane.
__host__ __device__ inline int far_targets_per_lane(int c) {
    int best = 1, best_used = 0;
    for (int t = 1; t <= 3; ++t) {
        const int h = (c + t - 1) / t;
        const int used3 = c * (kEvalLanes / h) * (6 / t);  // slots used, scaled by 6/t
        if (used3 > best_used) {
            best_used = used3;
            best = t;
        }
    }
    return best;
}

// One leaf chunk of c targets at TPL targets per lane: lane (g, l) of group
// g holds targets l, l + h, .. (h = ceil(c / TPL) lanes per group) and
// evaluates cells g, g + groups, .. of each staged chunk for all of them,
// reading each cell's table once.  Group sums are added in group order and
// written by group 0.
// Sorted position of target ti of a leaf chunk starting at list entry base:
// the leaf's own particles (tpos null) or its entries of the target list.
__device__ __forceinline__ int target_pos(const int* tpos, int base, int ti) {
    return tpos ? tpos[base + ti] : base + ti;
}

template <typename F, int ORDER, int TPL>
__device__ __forceinline__ void far_leaf_chunk(
    FarStage<F>& fs, double (&red)[3][kEvalLanes], int base, int c, double cx, double cy,
    double cz, F eps2, const double* __restrict__ mcom, const F* __restrict__ kcoef,
    const int* __restrict__ sidx, const double* __restrict__ sx, const double* __restrict__ sy,
    const double* __restrict__ sz, long long fo, int nf, const int* __restrict__ far, double G,
    int begin, int end, const int* __restrict__ tpos, double* __restrict__ ax,
    double* __restrict__ ay, double* __restrict__ az) {
#pragma clang fp contract(off)
    constexpr int kChunk = far_chunk<F>();
    const int lane = threadIdx.x;
    const int h = (c + TPL - 1) / TPL;
    const int groups = kEvalLanes / h;
    const int g = lane / h;
    const int tl = lane - g * h;
    const bool active = g < groups;
    F tx[TPL], ty[TPL], tz[TPL];
    double far_a[TPL][3];
#pragma unroll
    for (int k = 0; k < TPL; ++k) {
        const int ti = min(tl + k * h, c - 1);  // a spare slot repeats the last target
        const int pos = target_pos(tpos, base, ti);
        tx[k] = static_cast<F>(sx[pos] - cx);
        ty[k] = static_cast<F>(sy[pos] - cy);
        tz[k] = static_cast<F>(sz[pos] - cz);
        far_a[k][0] = far_a[k][1] = far_a[k][2] = 0.0;
    }
    for (int f0 = 0; f0 < nf; f0 += kChunk) {
        const int k = min(kChunk, nf - f0);
        __syncthreads();
        if (lane < k) {
            const int cell = far[fo + f0 + lane];
            fs.id[lane] = cell;
            fs.c[lane][0] = static_cast<F>(mcom[4 * cell + 1] - cx);
            fs.c[lane][1] = static_cast<F>(mcom[4 * cell + 2] - cy);
            fs.c[lane][2] = static_cast<F>(mcom[4 * cell + 3] - cz);
        }
        __syncthreads();
        // the chunk's tables, 16-byte units spread over the whole wave
        constexpr int kUnits = kFarCoef * static_cast<int>(sizeof(F)) / 16;
        for (int e = lane; e < k * kUnits; e += kEvalLanes) {
            const int f = e / kUnits, u = e - f * kUnits;
            const uint4 v = reinterpret_cast<const uint4*>(
                kcoef + static_cast<size_t>(kFarCoef) * fs.id[f])[u];
            reinterpret_cast<uint4*>(fs.K[f])[u] = v;
        }
        __syncthreads();
        if (!active) continue;
        F sf[TPL][3];
#pragma unroll
        for (int q = 0; q < TPL; ++q) sf[q][0] = sf[q][1] = sf[q][2] = 0;
        for (int f = g; f < k; f += groups) {
#pragma unroll
            for (int q = 0; q < TPL; ++q) {
                const F yx = fs.c[f][0] - tx[q], yy = fs.c[f][1] - ty[q], yz = fs.c[f][2] - tz[q];
                far_m2p<F, ORDER>(yx, yy, yz, eps2, fs.K[f], sf[q][0], sf[q][1], sf[q][2]);
            }
        }
#pragma unroll
        for (int q = 0; q < TPL; ++q)
#pragma unroll
            for (int d = 0; d < 3; ++d) far_a[q][d] += static_cast<double>(sf[q][d]);
    }
#pragma unroll
    for (int q = 0; q < TPL; ++q) {
        __syncthreads();
        for (int d = 0; d < 3; ++d) red[d][lane] = active ? far_a[q][d] : 0.0;
        __syncthreads();
        const int ti = tl + q * h;
        if (g != 0 || ti >= c) continue;
        double a[3];
        for (int d = 0; d < 3; ++d) {
            double acc = red[d][tl];
            for (int j = 1; j < groups; ++j) acc += red[d][tl + j * h];
            a[d] = G * acc;
        }
        const int me = sidx[target_pos(tpos, base, ti)];
        if (me < begin || me >= end) continue;
        ax[me] = a[0];
        ay[me] = a[1];
        az[me] = a[2];
    }
}

template <typename F, int ORDER>
__global__ __launch_bounds__(kEvalLanes) void eval_far_kernel(
    const int* __restrict__ leaves, const double* __restrict__ geom,
    const double* __restrict__ mcom, const F* __restrict__ kcoef,
    const int* __restrict__ nb, const int* __restrict__ ne, const int* __restrict__ sidx,
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    const long long* __restrict__ far_off, const int* __restrict__ nfar,
    const int* __restrict__ far, double G, double eps, int begin, int end,
    const int* __restrict__ tpos, const int* __restrict__ toff, const int* __restrict__ tcnt,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az) {
    __shared__ FarStage<F> fs;
    __shared__ double red[3][kEvalLanes];
    const int leaf = leaves[blockIdx.x];
    // targets: the leaf's particles, or its entries of the range's target list
    const int tb = tpos ? toff[blockIdx.x] : nb[leaf];
    const int te = tpos ? tb + tcnt[blockIdx.x] : ne[leaf];
    if (tb >= te) return;
    const double cx = geom[4 * leaf], cy = geom[4 * leaf + 1], cz = geom[4 * leaf + 2];
    const F eps2 = static_cast<F>(eps * eps);
    const long long fo = far_off[blockIdx.x];
    const int nf = nfar[blockIdx.x];
    for (int base = tb; base < te; base += kEvalLanes) {
        const int c = min(kEvalLanes, te - base);
#define GC_CHUNK(T)                                                                              \
    far_leaf_chunk<F, ORDER, T>(fs, red, base, c, cx, cy, cz, eps2, mcom, kcoef, sidx, sx, sy, sz, \
                                fo, nf, far, G, begin, end, tpos, ax, ay, az)
        switch (far_targets_per_lane(c)) {  // uniform per block
        case 3: GC_CHUNK(3); break;
        case 2: GC_CHUNK(2); break;
        default: GC_CHUNK(1); break;
        }
#undef GC_CHUNK
    }
}

// Far field in FP32 with the cells across the lanes: each lane holds one far
// cell's coefficient table in registers (64 cells per pass, loaded once per
// leaf chunk) and the chunk's targets are taken four at a time from LDS, the
// same for every lane.  A lane's four contributions are summed over the 64
// lanes by a fixed network (halving over lane bits 0 and 1, then rotations
// by 4 and 8 within each 16-lane row, then the four row sums in row order
// through LDS), so lane t adds target t's FP32 pass sum to its FP64 sum.
// No table is read from LDS, and no lane is idle but in a leaf's last pass.
// Lanes past the list's end hold a zero table at a point 4 half-widths plus
// one away from the leaf's centre, so their terms are exactly zero.
constexpr int kRowQuad = 4;  // targets per network evaluation

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}
constexpr int kDppXor1 = 0xB1;  // quad_perm [1, 0, 3, 2]
constexpr int kDppXor2 = 0x4E;  // quad_perm [2, 3, 0, 1]
constexpr int kDppRor4 = 0x124;  // row_ror:4
constexpr int kDppRor8 = 0x128;  // row_ror:8

// Lane l returns the sum over its 16-lane row of a[l & 3] (every lane active).
__device__ __forceinline__ float row_quad_sum(float a0, float a1, float a2, float a3, int lane) {
#pragma clang fp contract(off)
    const bool b0 = lane & 1, b1 = lane & 2;
    const float n0 = (b0 ? a1 : a0) + dpp_f<kDppXor1>(b0 ? a0 : a1);
    const float n1 = (b0 ? a3 : a2) + dpp_f<kDppXor1>(b0 ? a2 : a3);
    float v = (b1 ? n1 : n0) + dpp_f<kDppXor2>(b1 ? n0 : n1);
    v += dpp_f<kDppRor4>(v);
    v += dpp_f<kDppRor8>(v);
    return v;
}

template <int ORDER>
__global__ __launch_bounds__(kEvalLanes) void eval_far_cells_kernel(
    const int* __restrict__ leaves, const double* __restrict__ geom,
    const double* __restrict__ mcom, const float* __restrict__ kcoef,
    const int* __restrict__ nb, const int* __restrict__ ne, const int* __restrict__ sidx,
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    const long long* __restrict__ far_off, const int* __restrict__ nfar,
    const int* __restrict__ far, double G, double eps, int begin, int end,
    const int* __restrict__ tpos, const int* __restrict__ toff, const int* __restrict__ tcnt,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az) {
#pragma clang fp contract(off)
    constexpr int kVec = ORDER == 4 ? 17 : (ORDER == 2 ? 2 : 1);  // float4s of the table used
    __shared__ float tgt[kEvalLanes][4];
    __shared__ float red[4][kRowQuad][4];
    const int li = blockIdx.x;  // leaf-list position
    const int leaf = leaves[li];
    const int tb = tpos ? toff[li] : nb[leaf];
    const int te = tpos ? tb + tcnt[li] : ne[leaf];
    if (tb >= te) return;
    const double cx = geom[4 * leaf], cy = geom[4 * leaf + 1], cz = geom[4 * leaf + 2];
    const float pad = static_cast<float>(4.0 * geom[4 * leaf + 3] + 1.0);
    const float eps2 = static_cast<float>(eps * eps);
    const long long fo = far_off[li];
    const int nf = nfar[li];
    const int lane = threadIdx.x, q = lane & 3, row = lane >> 4;
    for (int base = tb; base < te; base += kEvalLanes) {
        const int c = min(kEvalLanes, te - base);
        __syncthreads();
        if (lane < c) {
            const int pos = target_pos(tpos, base, lane);
            tgt[lane][0] = static_cast<float>(sx[pos] - cx);
            tgt[lane][1] = static_cast<float>(sy[pos] - cy);
            tgt[lane][2] = static_cast<float>(sz[pos] - cz);
        }
        __syncthreads();
        double a64[3] = {0.0, 0.0, 0.0};
        int next = lane < nf ? far[fo + lane] : -1;  // cell ids run one pass ahead
        for (int f0 = 0; f0 < nf; f0 += kEvalLanes) {
            const int cell = next;
            next = f0 + kEvalLanes + lane < nf ? far[fo + f0 + kEvalLanes + lane] : -1;
            float K[4 * kVec];
            float px = pad, py = pad, pz = pad;
            if (cell >= 0) {
                const float4* src =
                    reinterpret_cast<const float4*>(kcoef + static_cast<size_t>(kFarCoef) * cell);
#pragma unroll
                for (int v = 0; v < kVec; ++v) {
                    const float4 k4 = src[v];
                    K[4 * v] = k4.x;
                    K[4 * v + 1] = k4.y;
                    K[4 * v + 2] = k4.z;
                    K[4 * v + 3] = k4.w;
                }
                px = static_cast<float>(mcom[4 * cell + 1] - cx);
                py = static_cast<float>(mcom[4 * cell + 2] - cy);
                pz = static_cast<float>(mcom[4 * cell + 3] - cz);
            } else {
#pragma unroll
                for (int i = 0; i < 4 * kVec; ++i) K[i] = 0.0f;
            }
            for (int t0 = 0; t0 < c; t0 += kRowQuad) {
                float acc[kRowQuad][3];
#pragma unroll
                for (int u = 0; u < kRowQuad; ++u) {
                    acc[u][0] = acc[u][1] = acc[u][2] = 0.0f;
                    if (t0 + u < c) {  // uniform
                        const float* T = tgt[t0 + u];
                        far_m2p<float, ORDER>(px - T[0], py - T[1], pz - T[2], eps2, K, acc[u][0],
                                              acc[u][1], acc[u][2]);
                    }
                }
                float r[3];
#pragma unroll
                for (int d = 0; d < 3; ++d)
                    r[d] = row_quad_sum(acc[0][d], acc[1][d], acc[2][d], acc[3][d], lane);
                __syncthreads();  // the previous quad's row sums are read
                if ((lane & 15) < kRowQuad)
                    for (int d = 0; d < 3; ++d) red[row][q][d] = r[d];
                __syncthreads();
                if ((lane >> 2) == (t0 >> 2)) {  // lane t0 + q: target t0 + q
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                        a64[d] += static_cast<double>(((red[0][q][d] + red[1][q][d]) + red[2][q][d]) +
                                                      red[3][q][d]);
                }
            }
        }
        if (lane < c) {
            const int me = sidx[target_pos(tpos, base, lane)];
            if (me >= begin && me < end) {
                ax[me] = G * a64[0];
                ay[me] = G * a64[1];
                az[me] = G * a64[2];
            }
        }
    }
}

constexpr int kNearStage = 256;     // sources staged per round
constexpr int kNearPad = kEvalLanes;  // zero-mass entries past the stage
constexpr int kNearPer = kNearStage / kEvalLanes;
constexpr int kNearItem = 256;      // near leaves per work item

// Near-field work items: item k of a leaf covers its near leaves
// [k kNearItem, (k + 1) kNearItem); each item writes FP64 partial sums for
// the leaf's targets at `slot` (rows of 3).  A leaf's items are consecutive.
struct NearItem {
    int leaf;   // index into the leaf list
    int q0, q1;
    int slot;   // first partial row (one row per target of the leaf)
};

__global__ __launch_bounds__(256) void near_item_count_kernel(int nleaves,
                                                              const int* __restrict__ nnear,
                                                              int* __restrict__ icount) {
    const int li = blockIdx.x * 256 + threadIdx.x;
    if (li >= nleaves) return;
    icount[li] = max(1, (nnear[li] + kNearItem - 1) / kNearItem);
}

// Items of each leaf at the leaf's scanned item offset; partial rows at the
// scanned offset of (items x targets).
__global__ __launch_bounds__(256) void near_item_fill_kernel(
    int nleaves, const int* __restrict__ leaves, const int* __restrict__ nb,
    const int* __restrict__ ne, const int* __restrict__ nnear, const int* __restrict__ icount,
    const int* __restrict__ ioff, const long long* __restrict__ row_off,
    NearItem* __restrict__ items) {
    const int li = blockIdx.x * 256 + threadIdx.x;
    if (li >= nleaves) return;
    const int cnt = ne[leaves[li]] - nb[leaves[li]];
    const int k = icount[li];
    const long long r0 = row_off[li];
    for (int j = 0; j < k; ++j) {
        NearItem it;
        it.leaf = li;
        it.q0 = j * kNearItem;
        it.q1 = min(nnear[li], (j + 1) * kNearItem);
        it.slot = static_cast<int>(r0 + static_cast<long long>(j) * cnt);
        items[ioff[li] + j] = it;
    }
}

// Near-field sources staged in LDS: offset from the target leaf's centre
// and G m (zero past the end of the item's sources).
template <typename F>
struct alignas(4 * sizeof(F)) NearSrc {
    F x, y, z, gm;
};

// Pair sums of one staged round for TPL targets per lane: group g takes
// sources g, g + groups, ...; CHECK drops the self pair (source cb + jj is
// the target's own particle when it equals selfpos).
// The round's trip count is the same for every lane group (a scalar loop):
// entries past cnt, up to kNearPad past the stage, hold zero-mass sources.
template <typename F, bool SAFE, int TPL, bool CHECK>
__device__ __forceinline__ void near_round(const NearSrc<F>* tile, int g, int cnt, int groups,
                                           int cb, F eps2, const F* tx, const F* ty, const F* tz,
                                           const int* selfpos, F* px, F* py, F* pz) {
#pragma clang fp contract(off)
    const int steps = (cnt + groups - 1) / groups;
    const NearSrc<F>* src = tile + g;
    int jj = g;
    for (int k = 0; k < steps; ++k, src += groups, jj += groups) {
        const NearSrc<F> sj = *src;
#pragma unroll
        for (int q = 0; q < TPL; ++q) {
            const F dx = sj.x - tx[q], dy = sj.y - ty[q], dz = sj.z - tz[q];
            const F s2 = gc_fma(dz, dz, gc_fma(dy, dy, gc_fma(dx, dx, eps2)));
            F r;
            if (SAFE) {
                r = s2 > F(1e-30) ? gc_rsq(s2) : F(0);
            } else {
                r = gc_rsq(s2);
            }
            if (CHECK) r = cb + jj == selfpos[q] ? F(0) : r;
            const F kk = (sj.gm * r) * (r * r);
            px[q] = gc_fma(dx, kk, px[q]);
            py[q] = gc_fma(dy, kk, py[q]);
            pz[q] = gc_fma(dz, kk, pz[q]);
        }
    }
}

// One leaf chunk of c targets of a near item at TPL targets per lane (lane
// (g, l) holds targets l, l + h, .., h = ceil(c / TPL)); group g takes
// sources g, g + groups, .. of each staged round.  The self pair
// (fmm_solver.cpp:295-310) is dropped without a branch: each target knows
// its own position among the piece's sources (the target leaf is one of
// its near leaves) and zeroes that pair's factor.
template <typename F, bool SAFE, int TPL>
__device__ __forceinline__ void near_leaf_chunk(
    NearSrc<F> (&tile)[kNearStage + kNearPad], int (&qs)[kNearPiece], int (&qpre)[kNearPiece + 1],
    double (&red)[3][kEvalLanes], const NearItem& item, int lb, int lsize, int base, int c,
    double cx, double cy, double cz, F eps2, const int* __restrict__ nb,
    const int* __restrict__ ne, const double* __restrict__ sx, const double* __restrict__ sy,
    const double* __restrict__ sz, const double* __restrict__ sm, long long no,
    const int* __restrict__ near, double G, const int* __restrict__ tpos,
    double* __restrict__ part) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x;
    const int h = (c + TPL - 1) / TPL;
    const int groups = kEvalLanes / h;
    const int g = lane / h;
    const int tl = lane - g * h;
    const bool active = g < groups;
    F tx[TPL], ty[TPL], tz[TPL];
    int tpl_pos[TPL];  // sorted positions of the lane's targets
    double near_a[TPL][3];
#pragma unroll
    for (int k = 0; k < TPL; ++k) {
        const int ti = min(tl + k * h, c - 1);  // a spare slot repeats the last target
        tpl_pos[k] = target_pos(tpos, base, ti);
        tx[k] = static_cast<F>(sx[tpl_pos[k]] - cx);
        ty[k] = static_cast<F>(sy[tpl_pos[k]] - cy);
        tz[k] = static_cast<F>(sz[tpl_pos[k]] - cz);
        near_a[k][0] = near_a[k][1] = near_a[k][2] = 0.0;
    }
    // source j of the current piece: its sorted index (-1 past the end)
    auto locate = [&](int j, int total, int kq) -> int {
        if (j >= total) return -1;
        int lo = 0, hi = kq - 1;  // last q with qpre[q] <= j
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (qpre[mid] <= j) lo = mid;
            else hi = mid - 1;
        }
        return qs[lo] + (j - qpre[lo]);
    };
    for (int q0 = item.q0; q0 < item.q1; q0 += kNearPiece) {
        const int kq = min(kNearPiece, item.q1 - q0);
        __syncthreads();
        int len = 0;
        if (lane < kq) {
            const int sl = near[no + q0 + lane];
            qs[lane] = nb[sl];
            len = ne[sl] - nb[sl];
        }
        int incl = len;
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(incl, o, 64);
            if (lane >= o) incl += v;
        }
        qpre[lane + 1] = incl;
        if (lane == 0) qpre[0] = 0;
        const int total = __shfl(incl, 63, 64);
        __syncthreads();
        // the target leaf among this piece's near leaves, if it is there:
        // the particle at sorted position p of the leaf is then source
        // qpre[qself] + (p - lb) of the piece, and the leaf's particles are
        // sources [self_lo, self_lo + lsize)
        const unsigned long long own = __ballot(lane < kq && qs[lane] == lb);
        const int self_lo = own ? qpre[__ffsll(static_cast<long long>(own)) - 1] : -1;
        int selfpos[TPL];
#pragma unroll
        for (int k = 0; k < TPL; ++k) selfpos[k] = own ? self_lo + (tpl_pos[k] - lb) : -1;
        // registers hold the next round's sources
        double nxv[kNearPer][4];
        int nid[kNearPer];
#pragma unroll
        for (int u = 0; u < kNearPer; ++u) {
            nid[u] = locate(lane + u * kEvalLanes, total, kq);
            const int sj = nid[u] < 0 ? base : nid[u];
            nxv[u][0] = sx[sj];
            nxv[u][1] = sy[sj];
            nxv[u][2] = sz[sj];
            nxv[u][3] = sm[sj];
        }
        for (int cb = 0; cb < total; cb += kNearStage) {
            __syncthreads();
#pragma unroll
            for (int u = 0; u < kNearPer; ++u) {
                NearSrc<F> t;
                t.x = static_cast<F>(nxv[u][0] - cx);
                t.y = static_cast<F>(nxv[u][1] - cy);
                t.z = static_cast<F>(nxv[u][2] - cz);
                t.gm = nid[u] < 0 ? F(0) : static_cast<F>(G * nxv[u][3]);
                tile[lane + u * kEvalLanes] = t;
            }
            __syncthreads();
            if (cb + kNearStage < total) {
#pragma unroll
                for (int u = 0; u < kNearPer; ++u) {
                    nid[u] = locate(cb + kNearStage + lane + u * kEvalLanes, total, kq);
                    const int sj = nid[u] < 0 ? base : nid[u];
                    nxv[u][0] = sx[sj];
                    nxv[u][1] = sy[sj];
                    nxv[u][2] = sz[sj];
                    nxv[u][3] = sm[sj];
                }
            }
            if (!active) continue;
            const int cnt = min(kNearStage, total - cb);
            F px[TPL], py[TPL], pz[TPL];
#pragma unroll
            for (int q = 0; q < TPL; ++q) px[q] = py[q] = pz[q] = 0;
            // only a round holding the target leaf's own sources can meet a
            // self pair (wave-uniform test); the others skip the check
            if (own && cb < self_lo + lsize && self_lo < cb + cnt)
                near_round<F, SAFE, TPL, true>(tile, g, cnt, groups, cb, eps2, tx, ty, tz, selfpos,
                                               px, py, pz);
            else
                near_round<F, SAFE, TPL, false>(tile, g, cnt, groups, cb, eps2, tx, ty, tz, selfpos,
                                                px, py, pz);
#pragma unroll
            for (int q = 0; q < TPL; ++q) {
                near_a[q][0] += static_cast<double>(px[q]);
                near_a[q][1] += static_cast<double>(py[q]);
                near_a[q][2] += static_cast<double>(pz[q]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < TPL; ++q) {
        __syncthreads();
        for (int d = 0; d < 3; ++d) red[d][lane] = active ? near_a[q][d] : 0.0;
        __syncthreads();
        const int ti = tl + q * h;
        if (g != 0 || ti >= c) continue;
        double* row = part + 3 * (static_cast<size_t>(item.slot) + (tpl_pos[q] - lb));
        for (int d = 0; d < 3; ++d) {
            double acc = red[d][tl];
            for (int j = 1; j < groups; ++j) acc += red[d][tl + j * h];
            row[d] = acc;
        }
    }
}

template <typename F, bool SAFE>
__global__ __launch_bounds__(kEvalLanes) void eval_near_kernel(
    const NearItem* __restrict__ items, const int* __restrict__ leaves,
    const double* __restrict__ geom, const int* __restrict__ nb, const int* __restrict__ ne,
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    const double* __restrict__ sm, const long long* __restrict__ near_off,
    const int* __restrict__ near, double G, double eps, const int* __restrict__ tpos,
    const int* __restrict__ toff, const int* __restrict__ tcnt, double* __restrict__ part) {
    __shared__ NearSrc<F> tile[kNearStage + kNearPad];
    __shared__ int qs[kNearPiece], qpre[kNearPiece + 1];
    __shared__ double red[3][kEvalLanes];
    const NearItem item = items[blockIdx.x];
    // the padding: zero mass at the leaf's centre (never rewritten; the
    // stage's own entries past a round's end carry zero mass too)
    tile[kNearStage + threadIdx.x] = NearSrc<F>{F(0), F(0), F(0), F(0)};
    const int leaf = leaves[item.leaf];
    const int lb = nb[leaf], lsize = ne[leaf] - lb;
    // targets: the leaf's particles, or its entries of the range's target list
    const int tb = tpos ? toff[item.leaf] : lb;
    const int te = tpos ? tb + tcnt[item.leaf] : lb + lsize;
    if (tb >= te) return;
    const double cx = geom[4 * leaf], cy = geom[4 * leaf + 1], cz = geom[4 * leaf + 2];
    const F eps2 = static_cast<F>(eps * eps);
    const long long no = near_off[item.leaf];
    for (int base = tb; base < te; base += kEvalLanes) {
        const int c = min(kEvalLanes, te - base);
#define GC_CHUNK(T)                                                                              \
    near_leaf_chunk<F, SAFE, T>(tile, qs, qpre, red, item, lb, lsize, base, c, cx, cy, cz, eps2,   \
                                nb, ne, sx, sy, sz, sm, no, near, G, tpos, part)
        // at most 2 targets per lane: a third costs more registers than the
        // slots it fills save here (measured)
        if (far_targets_per_lane(c) >= 2) GC_CHUNK(2);  // uniform per block
        else GC_CHUNK(1);
#undef GC_CHUNK
    }
}

// Near field in FP32 with the sources across the lanes.  A chunk of c <= 24
// targets is evaluated for NT = c rounded up to an even number of target
// slots, each with one FP32 sum per component in every lane's registers
// (a spare slot holds a padding target whose sums are never written; the
// network below adds slots in fours, the slots past NT holding zeros).
// The item's sources, 64 near leaves at a time laid end to end, pass 64 per
// step: lane k takes source k of the step, whose leaf it finds with two
// mbcnt instructions from a bitmap of the leaves' first sources (NearPiece),
// and adds its pair term to every slot, the target read from LDS (the same
// address for all lanes).  A source is its FP32 record (offset from its own
// leaf's centre, G m; near_records_kernel) plus the FP32 offset of its
// leaf's centre from the target leaf's; the next step's record is loaded
// while this step is evaluated.  Sources past a piece's end belong to a
// sentinel leaf whose records are the zero records after the last
// particle's, placed at the padding point.  After the item's last piece the
// slot sums are added over the wave by the far field's fixed network
// (row_quad_sum, then the four row sums in row order) into FP64 sums held
// by lane t for target t.
// The self pair (fmm_solver.cpp:295-310) adds exactly zero here: its offset
// is zero (the target leaf's centre offset is 0, so the source is its
// record, the target's own offset), and its factor G m r^3 is finite (the
// caller uses this kernel only when G times the largest mass times eps^-3
// is far below FP32's range) or zero (SAFE: a zero distance gives r = 0).
constexpr int kNearTargets = 24;

// Source records of the listed leaves' particles (sorted positions): offset
// from the leaf's centre and G m, in FP32.  One block per leaf; block 0 also
// writes the 2 kEvalLanes zero records at [n, n + 2 kEvalLanes) that stand
// for the sources past the end of a piece.
__global__ __launch_bounds__(kEvalLanes) void near_records_kernel(
    const int* __restrict__ leaves, const double* __restrict__ geom, const int* __restrict__ nb,
    const int* __restrict__ ne, const double* __restrict__ sx, const double* __restrict__ sy,
    const double* __restrict__ sz, const double* __restrict__ sm, double G, int n,
    float4* __restrict__ rec) {
    if (blockIdx.x == 0) {
        rec[n + threadIdx.x] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        rec[n + kEvalLanes + threadIdx.x] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    const int leaf = leaves[blockIdx.x];
    const double cx = geom[4 * leaf], cy = geom[4 * leaf + 1], cz = geom[4 * leaf + 2];
    for (int e = nb[leaf] + static_cast<int>(threadIdx.x); e < ne[leaf]; e += kEvalLanes)
        rec[e] = make_float4(static_cast<float>(sx[e] - cx), static_cast<float>(sy[e] - cy),
                             static_cast<float>(sz[e] - cz), static_cast<float>(G * sm[e]));
}

// A piece of up to kNearPiece near leaves in LDS, then a sentinel leaf.
// info: per leaf (first sorted index - first piece index, centre offset
// from the target leaf's centre as FP32 bits); the sentinel's first index
// maps piece source `total` to zero record n.  qpre: each leaf's first piece
// index (total for the sentinel).  Pieces of at most 64 x 64 sources also
// mark each leaf's first source, and the sentinel's, in a bitmap of 65
// words; word w holds the bitmap word shifted down by one bit (low half in
// .x, high half in .y) and, in .z, the marks before the word minus one plus
// the word's bit 0, so that source 64 w + l belongs to leaf
// .z + (marks of .x/.y below lane l).  Longer pieces search qpre.
struct NearPiece {
    int4 info[kNearPiece + 1];
    int4 word[kNearPiece + 1];
    unsigned long long starts[kNearPiece + 1];
    int qpre[kNearPiece + 1];
};

template <bool SAFE, bool UNI, int NT>
__device__ __forceinline__ void near_cells_chunk(
    const NearItem& item, double cx, double cy, double cz, float pad,
    const double* __restrict__ geom, const int* __restrict__ nb, const int* __restrict__ ne,
    const float4* __restrict__ srec, int nrec0, long long no, const int* __restrict__ near,
    const float4* tgt, NearPiece& pc, float (&red)[4][kRowQuad][4], double (&a64)[3]) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x, q = lane & 3, row = lane >> 4;
    constexpr int NTR = (NT + kRowQuad - 1) / kRowQuad * kRowQuad;  // network slots
    float acc[NTR][3];
#pragma unroll
    for (int t = 0; t < NTR; ++t) acc[t][0] = acc[t][1] = acc[t][2] = 0.0f;
    for (int q0 = item.q0; q0 < item.q1; q0 += kNearPiece) {
        const int kq = min(kNearPiece, item.q1 - q0);
        __syncthreads();  // the previous piece's readers are done
        int len = 0, first = 0;
        float ox = 0.0f, oy = 0.0f, oz = 0.0f;
        if (lane < kq) {
            const int sl = near[no + q0 + lane];
            first = nb[sl];
            len = ne[sl] - first;
            ox = static_cast<float>(geom[4 * sl] - cx);
            oy = static_cast<float>(geom[4 * sl + 1] - cy);
            oz = static_cast<float>(geom[4 * sl + 2] - cz);
        }
        int incl = len;
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(incl, o, 64);
            if (lane >= o) incl += v;
        }
        const int excl = incl - len;
        // uniform, and in scalar registers: the step loop below has no
        // divergent branch
        const int total = __builtin_amdgcn_readfirstlane(__shfl(incl, 63, 64));
        const bool bits = total <= kNearPiece * 64;
        if (lane < kq)
            pc.info[lane] = make_int4(first - excl, __float_as_int(ox), __float_as_int(oy), __float_as_int(oz));
        if (lane == 0) {
            // with one mass for all (UNI) the sentinel's zero records carry
            // no G m = 0: they sit 1e16 away, where r^3 underflows to 0
            const float sp = UNI ? 1e16f : pad;
            pc.info[kq] = make_int4(nrec0 - total, __float_as_int(sp), __float_as_int(sp), __float_as_int(sp));
            pc.qpre[0] = 0;
            pc.starts[kNearPiece] = 0ull;
        }
        if (lane < kq) pc.qpre[lane + 1] = incl;
        pc.starts[lane] = 0ull;
        __syncthreads();
        if (bits) {
            if (lane < kq) atomicOr(&pc.starts[excl >> 6], 1ull << (excl & 63));
            if (lane == 0) atomicOr(&pc.starts[total >> 6], 1ull << (total & 63));
        }
        __syncthreads();
        if (bits) {
            const unsigned long long sw = pc.starts[lane];
            const int cnt = __popcll(sw);
            int wi = cnt;
            for (int o = 1; o < 64; o <<= 1) {
                const int v = __shfl_up(wi, o, 64);
                if (lane >= o) wi += v;
            }
            pc.word[lane] = make_int4(static_cast<int>(sw >> 1), static_cast<int>(sw >> 33),
                                      wi - cnt - 1 + static_cast<int>(sw & 1ull), 0);
            if (lane == 63) {
                const unsigned long long s64 = pc.starts[kNearPiece];
                pc.word[kNearPiece] = make_int4(static_cast<int>(s64 >> 1), static_cast<int>(s64 >> 33),
                                                wi - 1 + static_cast<int>(s64 & 1ull), 0);
            }
        }
        __syncthreads();
        // leaf slot of piece source j = 64 w + lane (the sentinel's past the end)
        auto locate = [&](int w, int j) -> int {
            if (bits) {
                const int4 wd = pc.word[w];
                return static_cast<int>(__builtin_amdgcn_mbcnt_hi(
                    static_cast<unsigned>(wd.y),
                    __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(wd.x), static_cast<unsigned>(wd.z))));
            }
            int lo = 0, hi = kq;  // last leaf with qpre <= j
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (pc.qpre[mid] <= j) lo = mid;
                else hi = mid - 1;
            }
            return lo;
        };
        int4 ninf = pc.info[locate(0, lane)];
        float4 nrec = srec[lane + ninf.x];
        for (int cb = 0; cb < total; cb += kEvalLanes) {
            const float x = __int_as_float(ninf.y) + nrec.x;
            const float y = __int_as_float(ninf.z) + nrec.y;
            const float z = __int_as_float(ninf.w) + nrec.z;
            const float gm = nrec.w;
            // the next step's source, also after the last step (the
            // sentinel's: j - total < 2 kEvalLanes, and w <= kNearPiece)
            const int jn = cb + kEvalLanes + lane;
            ninf = pc.info[locate((cb >> 6) + 1, jn)];
            nrec = srec[jn + ninf.x];
            // the targets stay in LDS (read each step, not held in registers),
            // each read two targets ahead of its use (3,947 us against 4,065
            // us reading each at its use, Plummer N = 2^20)
            asm volatile("" ::: "memory");
            float4 Tn = tgt[0], Tn2 = tgt[1];
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const float4 T = Tn;
                Tn = Tn2;
                if (t + 2 < NT) Tn2 = tgt[t + 2];
                const float dx = x - T.x, dy = y - T.y, dz = z - T.z;
                // T.w: eps^2 (a VGPR operand, not an SGPR one)
                const float s2 = gc_fma(dz, dz, gc_fma(dy, dy, gc_fma(dx, dx, T.w)));
                float r;
                if (SAFE) {
                    r = s2 > 1e-30f ? gc_rsq(s2) : 0.0f;
                } else {
                    r = gc_rsq(s2);
                }
                const float kk = UNI ? (r * r) * r : (gm * r) * (r * r);
                acc[t][0] = gc_fma(dx, kk, acc[t][0]);
                acc[t][1] = gc_fma(dy, kk, acc[t][1]);
                acc[t][2] = gc_fma(dz, kk, acc[t][2]);
            }
        }
    }
#pragma unroll
    for (int t0 = 0; t0 < NTR; t0 += kRowQuad) {
        float r[3];
#pragma unroll
        for (int d = 0; d < 3; ++d)
            r[d] = row_quad_sum(acc[t0][d], acc[t0 + 1][d], acc[t0 + 2][d], acc[t0 + 3][d], lane);
        __syncthreads();
        if ((lane & 15) < kRowQuad)
            for (int d = 0; d < 3; ++d) red[row][q][d] = r[d];
        __syncthreads();
        if ((lane >> 2) == (t0 >> 2)) {
#pragma unroll
            for (int d = 0; d < 3; ++d)
                a64[d] += static_cast<double>(((red[0][q][d] + red[1][q][d]) + red[2][q][d]) +
                                              red[3][q][d]);
        }
    }
}

// UNI: every particle has the G m `gmu`; the pair terms leave it out and
// the target sums are multiplied by it in FP64 before they are written.
template <bool SAFE, bool UNI>
__global__ __launch_bounds__(kEvalLanes) __attribute__((amdgpu_waves_per_eu(4))) void eval_near_cells_kernel(
    const NearItem* __restrict__ items, const int* __restrict__ leaves,
    const double* __restrict__ geom, const int* __restrict__ nb, const int* __restrict__ ne,
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    const float4* __restrict__ srec, int nrec0, const long long* __restrict__ near_off,
    const int* __restrict__ near, double eps, float gmu, const int* __restrict__ tpos,
    const int* __restrict__ toff, const int* __restrict__ tcnt, double* __restrict__ part) {
    __shared__ float4 tgt[kNearTargets];
    __shared__ int tgp[kNearTargets];  // sorted positions of the chunk's targets
    __shared__ NearPiece pc;
    __shared__ float red[4][kRowQuad][4];
    const NearItem item = items[blockIdx.x];
    const int leaf = leaves[item.leaf];
    const int lb = nb[leaf], lsize = ne[leaf] - lb;
    // targets: the leaf's particles, or its entries of the range's target list
    const int tb = tpos ? toff[item.leaf] : lb;
    const int te = tpos ? tb + tcnt[item.leaf] : lb + lsize;
    if (tb >= te) return;
    const double cx = geom[4 * leaf], cy = geom[4 * leaf + 1], cz = geom[4 * leaf + 2];
    // padding sources and targets sit here: at least 1 + 3 half-widths from
    // the leaf's targets, so their terms stay finite
    const float pad = static_cast<float>(4.0 * geom[4 * leaf + 3] + 1.0);
    const float eps2 = static_cast<float>(eps * eps);
    const long long no = near_off[item.leaf];
    const int lane = threadIdx.x;
    for (int base = tb; base < te; base += kNearTargets) {
        const int c = min(kNearTargets, te - base);
        __syncthreads();
        if (lane < kNearTargets) {
            if (lane < c) {
                const int pos = target_pos(tpos, base, lane);
                tgt[lane] = make_float4(static_cast<float>(sx[pos] - cx), static_cast<float>(sy[pos] - cy),
                                        static_cast<float>(sz[pos] - cz), eps2);
                tgp[lane] = pos;
            } else {
                tgt[lane] = make_float4(-pad, -pad, -pad, eps2);
                tgp[lane] = -1;
            }
        }
        double a64[3] = {0.0, 0.0, 0.0};  // lane t: target t
#define GC_NC(NT)                                                                                  \
    near_cells_chunk<SAFE, UNI, NT>(item, cx, cy, cz, pad, geom, nb, ne, srec, nrec0, no, near, tgt, \
                               pc, red, a64)
        switch ((c + 1) / 2) {  // uniform per block
        case 1: GC_NC(2); break;
        case 2: GC_NC(4); break;
        case 3: GC_NC(6); break;
        case 4: GC_NC(8); break;
        case 5: GC_NC(10); break;
        case 6: GC_NC(12); break;
        case 7: GC_NC(14); break;
        case 8: GC_NC(16); break;
        case 9: GC_NC(18); break;
        case 10: GC_NC(20); break;
        case 11: GC_NC(22); break;
        default: GC_NC(24); break;
        }
#undef GC_NC
        __syncthreads();
        if (lane < c) {
            double* out = part + 3 * (static_cast<size_t>(item.slot) + (tgp[lane] - lb));
            const double g = UNI ? static_cast<double>(gmu) : 1.0;
            out[0] = a64[0] * g;
            out[1] = a64[1] * g;
            out[2] = a64[2] * g;
        }
    }
}

// a = far (already in ax) + the leaf's near item partials in item order;
// optional closing half kick.  One thread per target.
__global__ __launch_bounds__(kEvalLanes) void near_merge_kernel(
    const int* __restrict__ leaves, const int* __restrict__ nb, const int* __restrict__ ne,
    const int* __restrict__ sidx, const int* __restrict__ icount, const int* __restrict__ ioff,
    const NearItem* __restrict__ items, const double* __restrict__ part, int begin, int end,
    const int* __restrict__ tpos, const int* __restrict__ toff, const int* __restrict__ tcnt,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az,
    double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz, int close_kick,
    double half_dt) {
#pragma clang fp contract(off)
    const int li = blockIdx.x;
    const int leaf = leaves[li];
    const int lb = nb[leaf];
    const int tb = tpos ? toff[li] : lb;
    const int te = tpos ? tb + tcnt[li] : ne[leaf];
    const int k = icount[li], i0 = ioff[li];
    for (int e = tb + static_cast<int>(threadIdx.x); e < te; e += kEvalLanes) {
        const int t = tpos ? tpos[e] : e;  // sorted position
        double a[3] = {0.0, 0.0, 0.0};
        for (int j = 0; j < k; ++j) {
            const double* row = part + 3 * (static_cast<size_t>(items[i0 + j].slot) + (t - lb));
            a[0] += row[0];
            a[1] += row[1];
            a[2] += row[2];
        }
        const int me = sidx[t];
        if (me < begin || me >= end) continue;
        a[0] = ax[me] + a[0];
        a[1] = ay[me] + a[1];
        a[2] = az[me] + a[2];
        ax[me] = a[0];
        ay[me] = a[1];
        az[me] = a[2];
        if (close_kick) {
            vx[me] = vx[me] + a[0] * half_dt;
            vy[me] = vy[me] + a[1] * half_dt;
            vz[me] = vz[me] + a[2] * half_dt;
        }
    }
}

// Particle state in key order, for coalesced source and target reads.
__global__ __launch_bounds__(256) void gather_sorted_kernel(
    int n, const int* __restrict__ sidx, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ m, double* __restrict__ sx,
    double* __restrict__ sy, double* __restrict__ sz, double* __restrict__ sm) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const int i = sidx[j];
    sx[j] = x[i];
    sy[j] = y[i];
    sz[j] = z[i];
    sm[j] = m[i];
}


// zero accelerations of [begin, end) (particles in massless leaves keep 0,
// as in fmm_solver.cpp:31 + :210-226)
// ---- Barnes-Hut (cpp/fmm/quadtree.cpp) on the same tree -------------------
// The reference's tree solver builds the FMM's tree (same root cube and
// insertion rule), and walks it once per target particle
// (accumulate_from_node, quadtree.cpp:193-243): a massless node gives
// nothing; a leaf gives its particles' pair forces (self skipped); an
// internal node gives one M2P when the target lies outside its cube, at a
// non-zero distance from its centre of mass and width / distance < theta,
// and otherwise its children's sums.
//
// Here one wave walks for 64 consecutive targets in key order (spatially
// compact), one per lane.  A lane's sums depend only on the nodes it visits
// itself, in depth-first order, so grouping targets changes the work
// shared by a wave, not any target's walk.  The walk is depth first with a
// stack in LDS of (node, mask of lanes that still need it).  At an
// internal node each lane takes the reference's decision for
// its own target (in FP64, so decisions match it exactly); lanes that accept
// evaluate the M2P from the node's coefficient table (far_m2p.hpp, read by
// scalar loads since the node is uniform), the others carry the node's
// children.  A leaf gives the pair forces to the lanes still carrying it.
// Every target therefore sees exactly the reference's cells and leaves;
// only the order of the sums differs.
constexpr int kTreeStack = 256;  // >= 31 levels x 7 siblings + 8 children (depth 32)

// width / distance < theta on the reference's roundings (quadtree.cpp:
// 222-228): decided from width^2 against theta^2 d2 when they differ by
// more than any rounding can move them, else by the exact expression.
__device__ __forceinline__ bool tree_accepts(double dx, double dy, double dz, double width,
                                             double theta) {
#pragma clang fp contract(off)
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (!(d2 > 0.0)) return false;  // distance > 0
    const double lhs = width * width, rhs = theta * theta * d2;
    if (lhs < rhs * (1.0 - 1e-12)) return true;
    if (lhs > rhs * (1.0 + 1e-12)) return false;
    return width / __builtin_sqrt(d2) < theta;
}

// Eight waves per SIMD (at most 96 SGPRs; 112 allowed only 7): the grid's
// one wave per 64 particles then runs in whole rounds of 8,192 waves, and
// the latency-bound walks have one more wave to switch to (11.24 ms against
// 12.11 ms per evaluation at Plummer N = 2^20).
template <typename F, int ORDER, bool SAFE>
__global__ __launch_bounds__(kEvalLanes) __attribute__((amdgpu_waves_per_eu(8))) void tree_eval_kernel(
    int n, const double* __restrict__ geom,
    const double* __restrict__ mcom, const F* __restrict__ kcoef, const int* __restrict__ nb,
    const int* __restrict__ ne, const int* __restrict__ first_child,
    const int* __restrict__ nchild, const int* __restrict__ sidx, const double* __restrict__ sx, const double* __restrict__ sy,
    const double* __restrict__ sz, const double* __restrict__ sm, double G, double eps,
    double theta, int begin, int end, double* __restrict__ ax, double* __restrict__ ay,
    double* __restrict__ az, double* __restrict__ vx, double* __restrict__ vy,
    double* __restrict__ vz, int close_kick, double half_dt,
    unsigned long long* __restrict__ counters) {
#pragma clang fp contract(off)
    __shared__ int stack_node[kTreeStack];
    __shared__ unsigned long long stack_mask[kTreeStack];
    __shared__ NearSrc<F> leaf_src[kEvalLanes];
    // the wave's targets: 64 consecutive particles in key order
    const int tb = static_cast<int>(blockIdx.x) * kEvalLanes;
    if (tb >= n) return;
    const int te = min(n, tb + kEvalLanes);
    const int lane = threadIdx.x;
    const F eps2 = static_cast<F>(eps * eps);
    // offsets are taken from the centre of the targets' bounding box, so
    // FP32 differences stay small for nearby cells and sources
    double cx, cy, cz;
    {
        const int q = tb + min(lane, te - tb - 1);
        double lx = sx[q], ly = sy[q], lz = sz[q], hx = lx, hy = ly, hz = lz;
        for (int o = 32; o > 0; o >>= 1) {
            lx = fmin(lx, __shfl_xor(lx, o, 64));
            ly = fmin(ly, __shfl_xor(ly, o, 64));
            lz = fmin(lz, __shfl_xor(lz, o, 64));
            hx = fmax(hx, __shfl_xor(hx, o, 64));
            hy = fmax(hy, __shfl_xor(hy, o, 64));
            hz = fmax(hz, __shfl_xor(hz, o, 64));
        }
        cx = (lx + hx) * 0.5;
        cy = (ly + hy) * 0.5;
        cz = (lz + hz) * 0.5;
    }
    unsigned long long n_m2p = 0, n_p2p = 0;
    for (int base = tb; base < te; base += kEvalLanes) {
        const int c = min(kEvalLanes, te - base);
        const bool active = lane < c;
        const int me = base + (active ? lane : 0);  // sorted index of this lane's target
        const int id = sidx[me];
        // only targets in [begin, end) walk; a chunk without one is skipped
        // (the offsets' origin above stays that of the whole chunk, so a
        // target's sums do not depend on the range)
        const bool want = active && id >= begin && id < end;
        const unsigned long long walkers = __ballot(want);
        if (walkers == 0) continue;
        const double x = sx[me], y = sy[me], z = sz[me];
        const F tx = static_cast<F>(x - cx), ty = static_cast<F>(y - cy), tz = static_cast<F>(z - cz);
        double a[3] = {0.0, 0.0, 0.0};
        int sp = 0;
        if (lane == 0) {
            stack_node[0] = 0;
            stack_mask[0] = walkers;
        }
        __syncthreads();
        sp = 1;
        while (sp > 0) {
            --sp;
            const int node = __builtin_amdgcn_readfirstlane(stack_node[sp]);
            const unsigned long long mask = stack_mask[sp];
            __syncthreads();  // the slot may be overwritten by the pushes below
            const bool mine = (mask >> lane) & 1ull;
            const double mass = mcom[4 * node];
            if (!(mass > 0.0)) continue;
            const int nc = nchild[node];
            if (nc == 0) {
                // leaf: pair forces for the lanes still carrying it; the
                // leaf's sources are converted once (one lane each) into
                // LDS, 64 at a time, and read back by every lane
                const int sb = nb[node], se = ne[node];
                F px = 0, py = 0, pz = 0;
                for (int j0 = sb; j0 < se; j0 += kEvalLanes) {
                    const int cnt = min(kEvalLanes, se - j0);
                    if (lane < cnt) {
                        const int j = j0 + lane;
                        NearSrc<F> q;
                        q.x = static_cast<F>(sx[j] - cx);
                        q.y = static_cast<F>(sy[j] - cy);
                        q.z = static_cast<F>(sz[j] - cz);
                        q.gm = static_cast<F>(sm[j]);  // G is applied to the whole sum
                        leaf_src[lane] = q;
                    }
                    __syncthreads();
                    if (mine) {
                        for (int k = 0; k < cnt; ++k) {
                            const NearSrc<F> q = leaf_src[k];
                            const F dx = q.x - tx, dy = q.y - ty, dz = q.z - tz;
                            const F s2 = gc_fma(dz, dz, gc_fma(dy, dy, gc_fma(dx, dx, eps2)));
                            F r;
                            if (SAFE) {
                                r = s2 > F(1e-30) ? gc_rsq(s2) : F(0);
                            } else {
                                r = gc_rsq(s2);
                            }
                            r = j0 + k == me ? F(0) : r;
                            const F kk = (q.gm * r) * (r * r);
                            px = gc_fma(dx, kk, px);
                            py = gc_fma(dy, kk, py);
                            pz = gc_fma(dz, kk, pz);
                        }
                    }
                    __syncthreads();  // leaf_src is refilled next
                }
                a[0] += static_cast<double>(px);
                a[1] += static_cast<double>(py);
                a[2] += static_cast<double>(pz);
                // directed pairs, self excluded
                const unsigned long long carried = static_cast<unsigned long long>(__popcll(mask));
                const unsigned long long self =
                    static_cast<unsigned long long>(__popcll(mask & __ballot(me >= sb && me < se)));
                n_p2p += carried * (se - sb) - self;
                continue;
            }
            const double ccx = geom[4 * node], ccy = geom[4 * node + 1], ccz = geom[4 * node + 2];
            const double hw = geom[4 * node + 3];
            const double comx = mcom[4 * node + 1], comy = mcom[4 * node + 2],
                         comz = mcom[4 * node + 3];
            bool accept = false;
            if (mine) {
                const bool inside = __builtin_fabs(x - ccx) <= hw && __builtin_fabs(y - ccy) <= hw &&
                                    __builtin_fabs(z - ccz) <= hw;
                accept = !inside && tree_accepts(comx - x, comy - y, comz - z, 2.0 * hw, theta);
            }
            const unsigned long long amask = __ballot(accept);
            if (amask) {
                const F* K = kcoef + static_cast<size_t>(kFarCoef) * node;
                const F yx = static_cast<F>(comx - cx) - tx, yy = static_cast<F>(comy - cy) - ty,
                        yz = static_cast<F>(comz - cz) - tz;
                F fx = 0, fy = 0, fz = 0;
                if (accept) far_m2p<F, ORDER>(yx, yy, yz, eps2, K, fx, fy, fz);
                a[0] += static_cast<double>(fx);
                a[1] += static_cast<double>(fy);
                a[2] += static_cast<double>(fz);
                n_m2p += static_cast<unsigned long long>(__popcll(amask));
            }
            const unsigned long long rest = mask & ~amask;
            if (rest) {
                // children in reverse so that child 0 is walked first
                const int fc = first_child[node];
                if (lane < nc) {
                    stack_node[sp + nc - 1 - lane] = fc + lane;
                    stack_mask[sp + nc - 1 - lane] = rest;
                }
                sp += nc;
                __syncthreads();
            }
        }
        if (want) {
            const double gx = G * a[0], gy = G * a[1], gz = G * a[2];
            ax[id] = gx;
            ay[id] = gy;
            az[id] = gz;
            if (close_kick) {
                vx[id] = vx[id] + gx * half_dt;
                vy[id] = vy[id] + gy * half_dt;
                vz[id] = vz[id] + gz * half_dt;
            }
        }
        __syncthreads();
    }
    if (lane == 0) {
        atomicAdd(counters + 0, n_m2p);
        atomicAdd(counters + 1, n_p2p);
    }
}

// Counts the leaves holding particles (the tree's leaf statistic).
__global__ __launch_bounds__(256) void tree_leaf_count_kernel(int nnodes, const int* __restrict__ nb,
                                                              const int* __restrict__ ne,
                                                              const int* __restrict__ nchild,
                                                              unsigned long long* __restrict__ counters) {
    __shared__ int wave_leaves[4];
    const int k = blockIdx.x * 256 + threadIdx.x;
    const bool leaf = k < nnodes && nchild[k] == 0 && ne[k] > nb[k];
    const unsigned long long m = __ballot(leaf);
    if ((threadIdx.x & 63) == 0) wave_leaves[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {  // one atomic per block
        const int c = wave_leaves[0] + wave_leaves[1] + wave_leaves[2] + wave_leaves[3];
        if (c) atomicAdd(counters + 2, static_cast<unsigned long long>(c));
    }
}

__global__ __launch_bounds__(256) void zero_range_kernel(int begin, int end, double* __restrict__ ax,
                                                         double* __restrict__ ay,
                                                         double* __restrict__ az) {
    const int i = begin + blockIdx.x * 256 + threadIdx.x;
    if (i >= end) return;
    ax[i] = 0.0;
    ay[i] = 0.0;
    az[i] = 0.0;
}

// Targets of a range [begin, end): per leaf the count of its particles in
// the range, then (after a scan) their sorted positions in ascending order.
__global__ __launch_bounds__(256) void range_count_kernel(int nleaves, const int* __restrict__ leaves,
                                                          const int* __restrict__ nb,
                                                          const int* __restrict__ ne,
                                                          const int* __restrict__ sidx, int begin,
                                                          int end, int* __restrict__ tcnt) {
    const int li = blockIdx.x * 256 + threadIdx.x;
    if (li >= nleaves) return;
    const int l = leaves[li];
    int c = 0;
    for (int t = nb[l]; t < ne[l]; ++t) c += (sidx[t] >= begin && sidx[t] < end) ? 1 : 0;
    tcnt[li] = c;
}

__global__ __launch_bounds__(256) void range_fill_kernel(int nleaves, const int* __restrict__ leaves,
                                                         const int* __restrict__ nb,
                                                         const int* __restrict__ ne,
                                                         const int* __restrict__ sidx, int begin,
                                                         int end, const int* __restrict__ toff,
                                                         int* __restrict__ tpos) {
    const int li = blockIdx.x * 256 + threadIdx.x;
    if (li >= nleaves) return;
    const int l = leaves[li];
    int o = toff[li];
    for (int t = nb[l]; t < ne[l]; ++t)
        if (sidx[t] >= begin && sidx[t] < end) tpos[o++] = t;
}

__global__ __launch_bounds__(256) void stats_kernel(
    int nleaves, const int* __restrict__ leaves, const int* __restrict__ nb,
    const int* __restrict__ ne, const int* __restrict__ nfar, const int* __restrict__ nnear,
    const int* __restrict__ nsrc, unsigned long long* __restrict__ out) {
    const int li = blockIdx.x * 256 + threadIdx.x;
    unsigned long long v[4] = {0ull, 0ull, 0ull, 0ull};
    if (li < nleaves) {
        const long long cnt = ne[leaves[li]] - nb[leaves[li]];
        v[0] = static_cast<unsigned long long>(nfar[li]);
        v[1] = static_cast<unsigned long long>(nnear[li]);
        v[2] = static_cast<unsigned long long>(cnt * nfar[li]);
        v[3] = static_cast<unsigned long long>(cnt * nsrc[li] - cnt);
    }
    // one atomic per wave and counter instead of one per leaf
    for (int k = 0; k < 4; ++k) {
        for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
        if ((threadIdx.x & 63) == 0 && v[k]) atomicAdd(&out[k], v[k]);
    }
}

struct ToI64 {
    __host__ __device__ long long operator()(int v) const { return static_cast<long long>(v); }
};

// Partial rows a leaf's near items need: items x targets.
// The totals a build's host side sizes its buffers by, read in one copy
// with the statistics: list overflow flags, far and near list lengths,
// near work items and their partial rows.
__global__ void build_totals_kernel(int nleaves, const int* __restrict__ leaves,
                                    const int* __restrict__ nb, const int* __restrict__ ne,
                                    const int* __restrict__ nfar, const int* __restrict__ nnear,
                                    const long long* __restrict__ far_off,
                                    const long long* __restrict__ near_off,
                                    const int* __restrict__ icount, const int* __restrict__ ioff,
                                    const long long* __restrict__ row_off,
                                    const int* __restrict__ overflow,
                                    unsigned long long* __restrict__ out) {
    const int l = nleaves - 1;
    const int leaf = leaves[l];
    out[0] = static_cast<unsigned long long>(*overflow);
    out[1] = static_cast<unsigned long long>(far_off[l] + nfar[l]);
    out[2] = static_cast<unsigned long long>(near_off[l] + nnear[l]);
    out[3] = static_cast<unsigned long long>(ioff[l] + icount[l]);
    out[4] = static_cast<unsigned long long>(row_off[l] +
                                             static_cast<long long>(icount[l]) * (ne[leaf] - nb[leaf]));
}

struct AbsValue {
    __host__ __device__ double operator()(double v) const { return v < 0.0 ? -v : v; }
};

struct RowsPerLeaf {
    const int *leaves, *nb, *ne, *icount;
    __host__ __device__ long long operator()(int li) const {
        const int l = leaves[li];
        return static_cast<long long>(icount[li]) * (ne[l] - nb[l]);
    }
};

template <typename T>
hipError_t grow(T** p, size_t* cap, size_t need, size_t keep, hipStream_t st) {
    if (need <= *cap) return hipSuccess;
    size_t ncap = std::max(need, *cap * 2);
    T* q = nullptr;
    hipError_t e = hipMalloc(&q, sizeof(T) * ncap);
    if (e != hipSuccess) return e;
    if (*p && keep) {
        e = hipMemcpyAsync(q, *p, sizeof(T) * keep, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
    }
    (void)hipFree(*p);
    *p = q;
    *cap = ncap;
    return hipSuccess;
}

// Frees *p and allocates `count` elements; *cap is `count` on success and 0
// (with *p null) on failure, so a failed call is retried by the next one.
template <typename T>
hipError_t realloc_dev(T** p, size_t* cap, size_t count) {
    (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    hipError_t e = hipMalloc(p, sizeof(T) * std::max<size_t>(1, count));
    if (e != hipSuccess) {
        *p = nullptr;
        return e;
    }
    *cap = count;
    return hipSuccess;
}

}  // namespace

struct FmmEngine::Impl {
    // particles
    size_t pcap = 0;
    unsigned long long *keys = nullptr, *keys_sorted = nullptr;
    unsigned long long *keys_lo = nullptr, *keys_lo_sorted = nullptr;  // levels 22-32
    int *idx = nullptr, *idx_sorted = nullptr;
    double *sxs = nullptr, *sys = nullptr, *szs = nullptr, *sms = nullptr;  // key order
    void* sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    double* bbox_part = nullptr;
    // nodes: topology (geom, nb, ne, first_child, nchild) sized by the
    // split's bound, moments (mcom, mom, kc64, kc32) by the node count
    size_t ncap = 0, mcap = 0;
    double* geom = nullptr;  // 4 per node: cx, cy, cz, hw
    double* mcom = nullptr;  // 4 per node: mass, com
    double* mom = nullptr;   // 35 per node
    double* kc64 = nullptr;  // kFarCoef per node: far-field coefficient tables
    float* kc32 = nullptr;   // the same in FP32
    int *nb = nullptr, *ne = nullptr, *first_child = nullptr, *nchild = nullptr;
    // per-level scratch
    size_t lcap = 0;
    int *bounds = nullptr, *count = nullptr, *offset = nullptr;
    void* scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
    int* overflow = nullptr;
    int* lvl = nullptr;       // kMaxLevels + 4 level offsets (device)
    int* deepest = nullptr;   // deepest level holding more than leaf_capacity particles
    double* geom_root = nullptr;  // root cube (centre, half width)
    // leaves and lists
    size_t leafcap = 0;
    int *flag = nullptr, *flag_pos = nullptr, *leaves = nullptr;
    int *nfar = nullptr, *nnear = nullptr, *nsrc = nullptr;
    long long *far_off = nullptr, *near_off = nullptr;
    // list traversal frontiers: 2 * fcap nodes per persistent wave
    int* frontier = nullptr;
    size_t frontier_ints = 0;
    int fcap = 4096;
    size_t farcap = 0, nearcap = 0;
    int *far = nullptr, *near = nullptr;
    // one-pass list build: per-wave stash regions (0: count first), sized
    // from the per-wave totals of the last two-pass build
    long long wcap_far = 0, wcap_near = 0;
    size_t stash_far_cap = 0, stash_near_cap = 0, lpos_cap = 0;
    int *stash_far = nullptr, *stash_near = nullptr;
    long long *lpos_far = nullptr, *lpos_near = nullptr, *wtot = nullptr;
    // near-field work items and their partial sums
    int *icount = nullptr, *ioff = nullptr;
    long long* row_off = nullptr;
    // target range: per leaf its count and first entry, and the list of
    // sorted positions of the range's targets
    int *tcnt = nullptr, *toff = nullptr, *tpos = nullptr;
    size_t tleaf_cap = 0, tpos_cap = 0;
    size_t icap = 0, itemcap = 0, partcap = 0;
    NearItem* items = nullptr;
    double* npart = nullptr;
    int nitems = 0;
    unsigned long long* stats_dev = nullptr;
    double mass_abs_max = 0.0;  // largest |m| of the last build
    double mass_min = 0.0;      // smallest m of the last build
    float4* srec = nullptr;     // FP32 source records of the near field
    size_t srec_cap = 0;
    // Barnes-Hut work counts (M2P, pairs, leaves) of the last evaluation, in
    // pinned host memory; read once the evaluation has finished
    unsigned long long* tree_counts = nullptr;
    bool tree_counts_pending = false;
    // host-side results of the last build
    std::vector<int> level_begin;
    int nnodes = 0, nleaves = 0;
    gc_fmm_stats last{};
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    // the near field runs on a second stream (default priority) beside the
    // far field: fork after the build, join before the merge
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;

    ~Impl() {
        if (side) {
            (void)hipStreamSynchronize(side);
            (void)hipStreamDestroy(side);
        }
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        void* ptrs[] = {keys, keys_sorted, keys_lo, keys_lo_sorted, idx, idx_sorted, sxs, sys, szs, sms, sort_tmp, bbox_part, geom, mcom, mom,
                        kc64, kc32, nb, ne, first_child, nchild, bounds, count, offset,
                        scan_tmp, overflow, lvl, deepest, geom_root, flag, flag_pos, leaves, nfar, nnear, nsrc, far_off,
                        near_off, frontier, far, near, stash_far, stash_near, lpos_far, lpos_near, wtot, icount, ioff, row_off, items, npart, stats_dev,
                        tcnt, toff, tpos, srec};
        for (void* p : ptrs) (void)hipFree(p);
        if (tree_counts) (void)hipHostFree(tree_counts);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    }

    // Per-node leaf and list-count arrays for `nodes` nodes (contents not kept).
    hipError_t grow_leaf_arrays(size_t nodes) {
        if (nodes <= leafcap) return hipSuccess;
        void* ptrs[] = {flag, flag_pos, leaves, nfar, nnear, nsrc, far_off, near_off};
        for (void* q : ptrs) (void)hipFree(q);
        flag = flag_pos = leaves = nfar = nnear = nsrc = nullptr;
        far_off = near_off = nullptr;
        leafcap = 0;
        const size_t cap = nodes * 2;
        hipError_t e = hipMalloc(&flag, sizeof(int) * cap);
        if (e == hipSuccess) e = hipMalloc(&flag_pos, sizeof(int) * cap);
        if (e == hipSuccess) e = hipMalloc(&leaves, sizeof(int) * cap);
        if (e == hipSuccess) e = hipMalloc(&nfar, sizeof(int) * cap);
        if (e == hipSuccess) e = hipMalloc(&nnear, sizeof(int) * cap);
        if (e == hipSuccess) e = hipMalloc(&nsrc, sizeof(int) * cap);
        if (e == hipSuccess) e = hipMalloc(&far_off, sizeof(long long) * cap);
        if (e == hipSuccess) e = hipMalloc(&near_off, sizeof(long long) * cap);
        if (e == hipSuccess) leafcap = cap;
        return e;
    }

    // Topology arrays for `nodes` nodes (contents not kept: every build
    // writes them from the root down).
    hipError_t grow_topology(size_t nodes, hipStream_t st) {
        if (nodes <= ncap) return hipSuccess;
        size_t c;
        hipError_t e;
        c = 4 * ncap; if ((e = grow(&geom, &c, 4 * nodes, 0, st)) != hipSuccess) return e;
        c = ncap; if ((e = grow(&nb, &c, nodes, 0, st)) != hipSuccess) return e;
        c = ncap; if ((e = grow(&ne, &c, nodes, 0, st)) != hipSuccess) return e;
        c = ncap; if ((e = grow(&first_child, &c, nodes, 0, st)) != hipSuccess) return e;
        c = ncap; if ((e = grow(&nchild, &c, nodes, 0, st)) != hipSuccess) return e;
        ncap = std::max(nodes, 2 * ncap);  // grow() doubles the same way
        return hipSuccess;
    }

    // Mass, moment and coefficient arrays for `nodes` nodes (rewritten by
    // every upward pass).
    hipError_t grow_moments(size_t nodes, hipStream_t st) {
        if (nodes <= mcap) return hipSuccess;
        size_t c;
        hipError_t e;
        c = 4 * mcap; if ((e = grow(&mcom, &c, 4 * nodes, 0, st)) != hipSuccess) return e;
        c = kMom * mcap; if ((e = grow(&mom, &c, kMom * nodes, 0, st)) != hipSuccess) return e;
        c = kFarCoef * mcap; if ((e = grow(&kc64, &c, kFarCoef * nodes, 0, st)) != hipSuccess) return e;
        c = kFarCoef * mcap; if ((e = grow(&kc32, &c, kFarCoef * nodes, 0, st)) != hipSuccess) return e;
        mcap = std::max(nodes, 2 * mcap);
        return hipSuccess;
    }

    int build(const DeviceState& s, const gc_params& p, hipStream_t st, std::string* err);
    int launch_eval(const DeviceState& s, const gc_params& p, size_t begin, size_t end,
                    bool close_kick, double half_dt, hipStream_t st, std::string* err);
};

#define FMM_HIP(call, what)                                                                     \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            if (err) *err = std::string(what) + ": " + hipGetErrorString(e_);                   \
            return e_ == hipErrorOutOfMemory ? GC_ERR_OUT_OF_MEMORY : GC_ERR_HIP;               \
        }                                                                                       \
    } while (0)

int FmmEngine::Impl::build(const DeviceState& s, const gc_params& p, hipStream_t st,
                           std::string* err) {
    const int n = static_cast<int>(s.n);
    const int cap = std::max(1, p.leaf_capacity);
    if (p.max_depth > kMaxLevels) {
        if (err) *err = "max_depth above 32 (the reference's depth cap; keys hold 32 levels)";
        return GC_ERR_UNSUPPORTED;
    }
    const int max_depth = std::max(1, p.max_depth);
    const double theta = std::max(1.0e-6, p.theta);
    const int bblocks = std::min(1024, (n + 255) / 256);

    // particle buffers: all of them, or none (pcap stays 0 until the last
    // allocation succeeds, so a failed build allocates again next time)
    if (static_cast<size_t>(n) > pcap) {
        void** ptrs[] = {(void**)&keys, (void**)&keys_sorted, (void**)&keys_lo,
                         (void**)&keys_lo_sorted, (void**)&idx, (void**)&idx_sorted,
                         (void**)&sxs, (void**)&sys, (void**)&szs, (void**)&sms, &sort_tmp,
                         (void**)&bbox_part};
        for (void** q : ptrs) {
            (void)hipFree(*q);
            *q = nullptr;
        }
        pcap = 0;
        const size_t np = static_cast<size_t>(n);
        FMM_HIP(hipMalloc(&sxs, sizeof(double) * np), "allocate sorted state");
        FMM_HIP(hipMalloc(&sys, sizeof(double) * np), "allocate sorted state");
        FMM_HIP(hipMalloc(&szs, sizeof(double) * np), "allocate sorted state");
        FMM_HIP(hipMalloc(&sms, sizeof(double) * np), "allocate sorted state");
        FMM_HIP(hipMalloc(&keys, sizeof(unsigned long long) * np), "allocate keys");
        FMM_HIP(hipMalloc(&keys_sorted, sizeof(unsigned long long) * np), "allocate keys");
        FMM_HIP(hipMalloc(&keys_lo, sizeof(unsigned long long) * np), "allocate keys");
        FMM_HIP(hipMalloc(&keys_lo_sorted, sizeof(unsigned long long) * np), "allocate keys");
        FMM_HIP(hipMalloc(&idx, sizeof(int) * np), "allocate indices");
        FMM_HIP(hipMalloc(&idx_sorted, sizeof(int) * np), "allocate indices");
        FMM_HIP(hipMalloc(&bbox_part, sizeof(double) * 6 * 1024), "allocate bbox partials");
        sort_tmp_bytes = 0;
        size_t sort_bytes = 0, sort_bytes_lo = 0;
        FMM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, keys, keys_sorted, idx,
                                                   idx_sorted, n, 0, 3 * kLevels, st),
                "size radix sort");
        FMM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes_lo, keys_lo, keys_lo_sorted,
                                                   idx, idx_sorted, n, 0, 3 * kDeepLevels, st),
                "size radix sort");
        sort_bytes = std::max(sort_bytes, sort_bytes_lo);
        FMM_HIP(hipMalloc(&sort_tmp, sort_bytes), "allocate radix sort storage");
        sort_tmp_bytes = sort_bytes;
        if (!overflow) FMM_HIP(hipMalloc(&overflow, sizeof(int)), "allocate flag");
        if (!lvl) FMM_HIP(hipMalloc(&lvl, sizeof(int) * (kMaxLevels + 4)), "allocate level offsets");
        if (!deepest) FMM_HIP(hipMalloc(&deepest, sizeof(int)), "allocate depth");
        if (!geom_root) FMM_HIP(hipMalloc(&geom_root, sizeof(double) * 4), "allocate root");
        if (!stats_dev) FMM_HIP(hipMalloc(&stats_dev, sizeof(unsigned long long) * kStats), "allocate stats");
        if (!tree_counts)
            FMM_HIP(hipHostMalloc(&tree_counts, sizeof(unsigned long long) * kStats), "allocate counts");
        pcap = np;
    }
    // 1-3: root, keys, sort on the first key word
    const bool deep_keys = max_depth > kLevels;
    hipLaunchKernelGGL(bbox_partial_kernel, dim3(bblocks), dim3(256), 0, st, n, s.x, s.y, s.z,
                       bbox_part);
    hipLaunchKernelGGL(bbox_final_kernel, dim3(1), dim3(64), 0, st, bblocks, bbox_part,
                       p.softening, geom_root, lvl);
    hipLaunchKernelGGL(keys_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, s.x, s.y, s.z,
                       geom_root, keys, deep_keys ? keys_lo : nullptr, idx);
    FMM_HIP(hipcub::DeviceRadixSort::SortPairs(sort_tmp, sort_tmp_bytes, keys, keys_sorted, idx,
                                               idx_sorted, n, 0, 3 * kLevels, st),
            "radix sort keys");
    const unsigned dblocks = static_cast<unsigned>(std::min(1024, (n + 255) / 256));
    FMM_HIP(hipMemsetAsync(deepest, 0xFF, sizeof(int), st), "clear depth");
    hipLaunchKernelGGL(deepest_kernel, dim3(dblocks), dim3(256), 0, st, n, cap, keys_sorted,
                       nullptr, deepest);
    int D = -1;
    FMM_HIP(hipMemcpyAsync(&D, deepest, sizeof(int), hipMemcpyDeviceToHost, st), "read depth");
    FMM_HIP(hipStreamSynchronize(st), "tree depth");
    const unsigned pblocks = static_cast<unsigned>((n + 255) / 256);
    const bool deep = deep_keys && D >= kLevels;
    if (deep) {
        // more than leaf_capacity particles share all 21 levels of the first
        // word: order by both words (stable sort on the second, then on the
        // first), and take the depth from both
        FMM_HIP(hipcub::DeviceRadixSort::SortPairs(sort_tmp, sort_tmp_bytes, keys_lo, keys_lo_sorted,
                                                   idx, idx_sorted, n, 0, 3 * kDeepLevels, st),
                "radix sort deep keys");
        hipLaunchKernelGGL(permute_kernel<unsigned long long>, dim3(pblocks), dim3(256), 0, st, n,
                           idx_sorted, keys, keys_sorted);
        FMM_HIP(hipMemcpyAsync(idx, idx_sorted, sizeof(int) * n, hipMemcpyDeviceToDevice, st),
                "deep key order");
        FMM_HIP(hipcub::DeviceRadixSort::SortPairs(sort_tmp, sort_tmp_bytes, keys_sorted, keys, idx,
                                                   idx_sorted, n, 0, 3 * kLevels, st),
                "radix sort keys");
        std::swap(keys, keys_sorted);
        hipLaunchKernelGGL(permute_kernel<unsigned long long>, dim3(pblocks), dim3(256), 0, st, n,
                           idx_sorted, keys_lo, keys_lo_sorted);
        FMM_HIP(hipMemsetAsync(deepest, 0xFF, sizeof(int), st), "clear depth");
        hipLaunchKernelGGL(deepest_kernel, dim3(dblocks), dim3(256), 0, st, n, cap, keys_sorted,
                           keys_lo_sorted, deepest);
        FMM_HIP(hipMemcpyAsync(&D, deepest, sizeof(int), hipMemcpyDeviceToHost, st), "read depth");
        FMM_HIP(hipStreamSynchronize(st), "tree depth");
    }
    FMM_HIP(hipMemsetAsync(overflow, 0, sizeof(int), st), "clear flag");
    hipLaunchKernelGGL(gather_sorted_kernel, dim3(pblocks), dim3(256), 0, st, n, idx_sorted,
                       s.x, s.y, s.z, s.m, sxs, sys, szs, sms);
    {  // largest |m|, read with the statistics (bounds the near field's self term)
        hipcub::TransformInputIterator<double, AbsValue, const double*> am(sms, AbsValue{});
        double* mx = reinterpret_cast<double*>(stats_dev + kStatMassMax);
        size_t need = 0;
        FMM_HIP(hipcub::DeviceReduce::Max(nullptr, need, am, mx, n, st), "size mass reduction");
        if (need > scan_tmp_bytes)
            FMM_HIP(realloc_dev(reinterpret_cast<unsigned char**>(&scan_tmp), &scan_tmp_bytes, need),
                    "allocate scan storage");
        FMM_HIP(hipcub::DeviceReduce::Max(scan_tmp, scan_tmp_bytes, am, mx, n, st), "mass reduction");
        // smallest m: equal to the largest |m| when every particle has one
        // mass, which lets the near field leave out its per-pair G m
        double* mn = reinterpret_cast<double*>(stats_dev + kStatMassMin);
        FMM_HIP(hipcub::DeviceReduce::Min(nullptr, need, sms, mn, n, st), "size mass reduction");
        if (need > scan_tmp_bytes)
            FMM_HIP(realloc_dev(reinterpret_cast<unsigned char**>(&scan_tmp), &scan_tmp_bytes, need),
                    "allocate scan storage");
        FMM_HIP(hipcub::DeviceReduce::Min(scan_tmp, scan_tmp_bytes, sms, mn, n, st), "mass reduction");
    }

    // 4: levels.  Levels 0..lmax hold internal nodes; level lmax + 1 only
    // leaves (it is split too, to mark them).  A level's node count is at
    // most 8 x its parents' internal nodes, each holding > cap particles,
    // and at most n: the bound each level's launch covers.
    const int lmax = D < 0 ? -1 : std::min(D, max_depth - 1);
    const int nlev = lmax + 2;  // levels split: 0 .. lmax + 1
    std::vector<long long> bound(static_cast<size_t>(nlev) + 1, 0);
    bound[0] = 1;
    const long long max_internal = static_cast<long long>(n) / (static_cast<long long>(cap) + 1);
    long long total_bound = 1, level_bound = 1;
    for (int L = 0; L + 1 < nlev; ++L) {
        bound[L + 1] = std::min<long long>(n, 8 * std::min(max_internal, bound[L]));
        total_bound += bound[L + 1];
        level_bound = std::max(level_bound, bound[L + 1]);
    }
    FMM_HIP(grow_topology(static_cast<size_t>(total_bound), st), "allocate nodes");
    hipLaunchKernelGGL(root_kernel, dim3(1), dim3(1), 0, st, n, geom_root, geom, nb, ne);
    if (static_cast<size_t>(level_bound) > lcap) {
        void** ptrs[] = {(void**)&bounds, (void**)&count, (void**)&offset, &scan_tmp};
        for (void** q : ptrs) {
            (void)hipFree(*q);
            *q = nullptr;
        }
        lcap = 0;
        scan_tmp_bytes = 0;
        const size_t want = static_cast<size_t>(level_bound);
        FMM_HIP(hipMalloc(&bounds, sizeof(int) * 9 * want), "allocate level scratch");
        FMM_HIP(hipMalloc(&count, sizeof(int) * want), "allocate level scratch");
        FMM_HIP(hipMalloc(&offset, sizeof(int) * want), "allocate level scratch");
        size_t scan_bytes = 0;
        FMM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, count, offset,
                                                 static_cast<int>(want), st),
                "size scan");
        FMM_HIP(hipMalloc(&scan_tmp, scan_bytes), "allocate scan storage");
        scan_tmp_bytes = scan_bytes;
        lcap = want;
    }
    const unsigned long long* klo = deep ? keys_lo_sorted : nullptr;
    for (int L = 0; L < nlev; ++L) {
        const int bl = static_cast<int>(bound[L]);
        const unsigned blocks = static_cast<unsigned>((bl + 255) / 256);
        const unsigned sblocks = static_cast<unsigned>((static_cast<size_t>(bl) * kSplitLanes + 255) / 256);
        hipLaunchKernelGGL(split_count_kernel, dim3(sblocks), dim3(256), 0, st, L, bl, lvl, cap,
                           max_depth, keys_sorted, klo, nb, ne, bounds, count, overflow);
        FMM_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_tmp_bytes, count, offset, bl, st),
                "scan children");
        hipLaunchKernelGGL(split_emit_kernel, dim3(blocks), dim3(256), 0, st, L, bl, lvl,
                           static_cast<int>(std::min<size_t>(ncap, INT_MAX)), bounds, count,
                           offset, geom, nb, ne,
                           first_child, nchild, overflow);
    }
    std::vector<int> lv(static_cast<size_t>(nlev) + 2, 0);
    int ovf = 0;
    FMM_HIP(hipMemcpyAsync(lv.data(), lvl, sizeof(int) * lv.size(), hipMemcpyDeviceToHost, st),
            "read level offsets");
    FMM_HIP(hipMemcpyAsync(&ovf, overflow, sizeof(int), hipMemcpyDeviceToHost, st), "read flag");
    FMM_HIP(hipStreamSynchronize(st), "tree build");
    if (ovf) {
        if (err) *err = "octree split exceeded its key levels or node bound (internal error)";
        return GC_ERR_STATE;
    }
    level_begin.assign(lv.begin(), lv.end());  // level L: [level_begin[L], level_begin[L + 1])
    nnodes = level_begin.back();
    int depth = 0;  // deepest non-empty level
    for (int L = 0; L + 1 < static_cast<int>(level_begin.size()); ++L)
        if (level_begin[L + 1] > level_begin[L]) depth = L;
    FMM_HIP(grow_moments(static_cast<size_t>(nnodes), st), "allocate node moments");

    // 5: upward, deepest level first
    for (int level = depth; level >= 0; --level) {
        const int lb = level_begin[level], le = level_begin[level + 1];
        if (le <= lb) continue;
        hipLaunchKernelGGL(upward_kernel, dim3((kUpLanes * (le - lb) + 63) / 64), dim3(64), 0, st, lb, le,
                           nb, ne, first_child, nchild, geom, sxs, sys, szs, sms, mcom, mom);
    }
    hipLaunchKernelGGL(far_coef_kernel, dim3((nnodes + 63) / 64), dim3(64), 0, st, nnodes, mcom, mom,
                       kc64, kc32);
    // internal nodes, counted on the device (read with the other statistics)
    FMM_HIP(hipMemsetAsync(stats_dev + kStatInternal, 0, sizeof(unsigned long long), st),
            "clear counts");
    hipLaunchKernelGGL(internal_count_kernel, dim3((nnodes + 255) / 256), dim3(256), 0, st, nnodes,
                       nchild, stats_dev + kStatInternal);
    if (p.solver == GC_SOLVER_TREE) {
        // the Barnes-Hut walk needs no lists; its work counts and the node
        // count come with the evaluation (stats())
        last = gc_fmm_stats{};
        last.max_depth = depth;
        last.tree_nodes = static_cast<size_t>(nnodes);
        nleaves = nnodes;  // evaluation grid: every node, leaves holding particles act
        return GC_OK;
    }

    // leaves (mass > 0)
    FMM_HIP(grow_leaf_arrays(static_cast<size_t>(nnodes)), "allocate leaves");
    {
        size_t need = 0;
        FMM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need, flag, flag_pos, nnodes, st), "size scan");
        size_t need2 = 0;
        FMM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need2, nfar, far_off, nnodes, st), "size scan");
        need = std::max(need, need2);
        if (need > scan_tmp_bytes)
            FMM_HIP(realloc_dev(reinterpret_cast<unsigned char**>(&scan_tmp), &scan_tmp_bytes, need),
                    "allocate scan storage");
    }
    const unsigned nblk = static_cast<unsigned>((nnodes + 255) / 256);
    hipLaunchKernelGGL(leaf_flag_kernel, dim3(nblk), dim3(256), 0, st, nnodes, nchild, mcom, flag);
    FMM_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_tmp_bytes, flag, flag_pos, nnodes, st),
            "scan leaves");
    hipLaunchKernelGGL(leaf_collect_kernel, dim3(nblk), dim3(256), 0, st, nnodes, flag, flag_pos,
                       leaves);
    int lastf[2] = {0, 0};
    FMM_HIP(hipMemcpyAsync(&lastf[0], flag_pos + nnodes - 1, sizeof(int), hipMemcpyDeviceToHost, st), "leaf count");
    FMM_HIP(hipMemcpyAsync(&lastf[1], flag + nnodes - 1, sizeof(int), hipMemcpyDeviceToHost, st), "leaf count");
    FMM_HIP(hipStreamSynchronize(st), "leaf count");
    nleaves = lastf[0] + lastf[1];
    if (nleaves == 0) return GC_OK;

    // 6: lists; persistent waves, larger frontiers on overflow.  Once a
    // two-pass build (count, scan, write) has sized the per-wave stash
    // regions, the lists are counted and stashed in one traversal and moved
    // to their places after the scan; a wave whose lists outgrow its region
    // falls back to the writing traversal.
    const int lwaves = std::min(nleaves, kListWaves);
    const bool stash = wcap_far > 0 && wcap_near > 0;
    if (stash) {
        const size_t fneed = static_cast<size_t>(lwaves) * wcap_far;
        const size_t nneed = static_cast<size_t>(lwaves) * wcap_near;
        if (fneed > stash_far_cap)
            FMM_HIP(realloc_dev(&stash_far, &stash_far_cap, fneed), "allocate list stash");
        if (nneed > stash_near_cap)
            FMM_HIP(realloc_dev(&stash_near, &stash_near_cap, nneed), "allocate list stash");
    }
    if (static_cast<size_t>(nleaves) > lpos_cap) {
        size_t c2 = 0;
        lpos_cap = 0;  // set by the last allocation only
        FMM_HIP(realloc_dev(&lpos_near, &c2, static_cast<size_t>(nleaves) * 2), "allocate list stash");
        FMM_HIP(realloc_dev(&lpos_far, &lpos_cap, c2), "allocate list stash");
    }
    if (static_cast<size_t>(nleaves) > icap) {
        size_t c2 = 0, c3 = 0;
        icap = 0;  // set by the last allocation only
        FMM_HIP(realloc_dev(&ioff, &c2, static_cast<size_t>(nleaves) * 2), "allocate near items");
        FMM_HIP(realloc_dev(&row_off, &c3, c2), "allocate near items");
        FMM_HIP(realloc_dev(&icount, &icap, c3), "allocate near items");
    }
    // 64-bit offsets: widen the counts with a scan over long long
    auto widen = [&](const int* in, long long* out) -> hipError_t {
        size_t need = 0;
        hipcub::TransformInputIterator<long long, ToI64, const int*> it(in, ToI64());
        hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, need, it, out, nleaves, st);
        if (e != hipSuccess) return e;
        if (need > scan_tmp_bytes) {
            e = realloc_dev(reinterpret_cast<unsigned char**>(&scan_tmp), &scan_tmp_bytes, need);
            if (e != hipSuccess) return e;
        }
        return hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_tmp_bytes, it, out, nleaves, st);
    };
    // The walk counts (and stashes) the lists; everything that depends
    // only on the counts follows on the device (list offsets, near work
    // items, statistics, totals), and one copy reads the totals.  A frontier
    // overflow repeats the walk with larger frontiers.
    unsigned long long sv[kStatTotals + 5];
    for (;;) {
        const size_t need = static_cast<size_t>(lwaves) * 2 * fcap;
        if (need > frontier_ints)
            FMM_HIP(realloc_dev(&frontier, &frontier_ints, need), "allocate list frontiers");
        FMM_HIP(hipMemsetAsync(overflow, 0, sizeof(int), st), "clear flag");
        if (stash)
            hipLaunchKernelGGL((lists_kernel<2>), dim3(lwaves), dim3(64), 0, st, nleaves, leaves,
                               geom, mcom, first_child, nchild, nb, ne, theta, nfar, nnear, nsrc,
                               far_off, near_off, wcap_far, wcap_near, lpos_far, lpos_near,
                               stash_far, stash_near, frontier, fcap, overflow);
        else
            hipLaunchKernelGGL((lists_kernel<0>), dim3(lwaves), dim3(64), 0, st, nleaves, leaves,
                               geom, mcom, first_child, nchild, nb, ne, theta, nfar, nnear, nsrc,
                               far_off, near_off, 0ll, 0ll, lpos_far, lpos_near, far, near,
                               frontier, fcap, overflow);
        FMM_HIP(widen(nfar, far_off), "scan far lists");
        FMM_HIP(widen(nnear, near_off), "scan near lists");
        // near-field work items: leaf li gets max(1, ceil(nnear / kNearItem))
        // items, each with one partial row per target of the leaf
        hipLaunchKernelGGL(near_item_count_kernel, dim3((nleaves + 255) / 256), dim3(256), 0, st,
                           nleaves, nnear, icount);
        {
            size_t need1 = 0;
            FMM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need1, icount, ioff, nleaves, st), "size scan");
            RowsPerLeaf rows_fn{leaves, nb, ne, icount};
            hipcub::CountingInputIterator<int> ci(0);
            hipcub::TransformInputIterator<long long, RowsPerLeaf, hipcub::CountingInputIterator<int>> rit(ci, rows_fn);
            size_t need2 = 0;
            FMM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need2, rit, row_off, nleaves, st), "size scan");
            need1 = std::max(need1, need2);
            if (need1 > scan_tmp_bytes)
                FMM_HIP(realloc_dev(reinterpret_cast<unsigned char**>(&scan_tmp), &scan_tmp_bytes, need1),
                        "allocate scan storage");
            FMM_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_tmp_bytes, icount, ioff, nleaves, st),
                    "scan near items");
            FMM_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_tmp_bytes, rit, row_off, nleaves, st),
                    "scan near rows");
        }
        // statistics (FmmStats, fmm_solver.hpp:20-25, plus pair counts)
        FMM_HIP(hipMemsetAsync(stats_dev, 0, sizeof(unsigned long long) * 4, st), "clear stats");
        hipLaunchKernelGGL(stats_kernel, dim3((nleaves + 255) / 256), dim3(256), 0, st, nleaves, leaves,
                           nb, ne, nfar, nnear, nsrc, stats_dev);
        hipLaunchKernelGGL(build_totals_kernel, dim3(1), dim3(1), 0, st, nleaves, leaves, nb, ne, nfar,
                           nnear, far_off, near_off, icount, ioff, row_off, overflow,
                           stats_dev + kStatTotals);
        FMM_HIP(hipMemcpyAsync(sv, stats_dev, sizeof(sv), hipMemcpyDeviceToHost, st), "read totals");
        FMM_HIP(hipStreamSynchronize(st), "list counts");
        if (!(sv[kStatTotals] & 1)) break;
        if (fcap >= (1 << 24)) {
            if (err) *err = "interaction-list frontier beyond 16M nodes";
            return GC_ERR_UNSUPPORTED;
        }
        fcap *= 4;
    }
    const int lovf = static_cast<int>(sv[kStatTotals]);
    const long long total_far = static_cast<long long>(sv[kStatTotals + 1]);
    const long long total_near = static_cast<long long>(sv[kStatTotals + 2]);
    if (static_cast<size_t>(total_far) + 1 > farcap)
        FMM_HIP(realloc_dev(&far, &farcap,
                            static_cast<size_t>(total_far) + 1 + static_cast<size_t>(total_far) / 4),
                "allocate far lists");
    if (static_cast<size_t>(total_near) + 1 > nearcap)
        FMM_HIP(realloc_dev(&near, &nearcap,
                            static_cast<size_t>(total_near) + 1 + static_cast<size_t>(total_near) / 4),
                "allocate near lists");
    if (stash && !(lovf & 2)) {
        hipLaunchKernelGGL(list_copy_kernel, dim3(nleaves), dim3(256), 0, st, nleaves, lwaves, nfar,
                           nnear, far_off, near_off, lpos_far, lpos_near, wcap_far, wcap_near,
                           stash_far, stash_near, far, near);
    } else {
        hipLaunchKernelGGL((lists_kernel<1>), dim3(lwaves), dim3(64), 0, st, nleaves, leaves, geom,
                           mcom, first_child, nchild, nb, ne, theta, nfar, nnear, nsrc, far_off,
                           near_off, 0ll, 0ll, lpos_far, lpos_near, far, near, frontier, fcap,
                           overflow);
        // size the stash regions for the next builds from this build's
        // per-wave totals, while they stay within 8x the lists' size
        if (!wtot) FMM_HIP(hipMalloc(&wtot, sizeof(long long) * 2 * kListWaves), "allocate wave totals");
        hipLaunchKernelGGL(wave_totals_kernel, dim3((lwaves + 255) / 256), dim3(256), 0, st, nleaves,
                           lwaves, nfar, nnear, wtot);
        std::vector<long long> wt(2 * static_cast<size_t>(lwaves));
        FMM_HIP(hipMemcpyAsync(wt.data(), wtot, sizeof(long long) * wt.size(), hipMemcpyDeviceToHost, st),
                "wave totals");
        FMM_HIP(hipStreamSynchronize(st), "wave totals");
        long long mf = 0, mn = 0;
        for (int w = 0; w < lwaves; ++w) {
            mf = std::max(mf, wt[w]);
            mn = std::max(mn, wt[lwaves + w]);
        }
        wcap_far = mf + mf / 4 + 1024;
        wcap_near = mn + mn / 4 + 1024;
        if (static_cast<long long>(lwaves) * (wcap_far + wcap_near) > 8 * (total_far + total_near) + (1 << 20))
            wcap_far = wcap_near = 0;
    }

    nitems = static_cast<int>(sv[kStatTotals + 3]);
    const size_t rows = static_cast<size_t>(sv[kStatTotals + 4]);
    if (static_cast<size_t>(nitems) > itemcap)
        FMM_HIP(realloc_dev(&items, &itemcap,
                            static_cast<size_t>(nitems) + static_cast<size_t>(nitems) / 4),
                "allocate near items");
    if (3 * rows > partcap)
        FMM_HIP(realloc_dev(&npart, &partcap, 3 * rows + 3 * rows / 4), "allocate near partials");
    hipLaunchKernelGGL(near_item_fill_kernel, dim3((nleaves + 255) / 256), dim3(256), 0, st, nleaves,
                       leaves, nb, ne, nnear, icount, ioff, row_off, items);
    std::memcpy(&mass_abs_max, &sv[kStatMassMax], sizeof(double));
    std::memcpy(&mass_min, &sv[kStatMassMin], sizeof(double));
    last.node_count = 1 + 8 * static_cast<size_t>(sv[kStatInternal]);
    last.leaf_count = static_cast<size_t>(nleaves);
    last.far_interactions = sv[0];
    last.near_interactions = sv[1];
    last.m2p_evaluations = sv[2];
    last.p2p_pairs = sv[3];
    last.max_depth = depth;
    last.tree_nodes = static_cast<size_t>(nnodes);
    return GC_OK;
}

int FmmEngine::Impl::launch_eval(const DeviceState& s, const gc_params& p, size_t begin, size_t end,
                                 bool close_kick, double half_dt, hipStream_t st, std::string* err) {
    int order = std::min(4, std::max(0, p.expansion_order));
    if (order == 1 || order == 3) ++order;
    const bool mixed = p.precision == GC_PRECISION_MIXED;
    const bool safe = !(static_cast<float>(p.softening * p.softening) >= 1e-30f);
    const dim3 grid(static_cast<unsigned>(nleaves)), block(kEvalLanes);
    const int b = static_cast<int>(begin), e = static_cast<int>(end);
    const double G = p.gravitational_constant, eps = p.softening;
    if (p.solver == GC_SOLVER_TREE) {
        FMM_HIP(hipMemsetAsync(stats_dev, 0, sizeof(unsigned long long) * 4, st), "clear counts");
        hipLaunchKernelGGL(tree_leaf_count_kernel, dim3((nnodes + 255) / 256), dim3(256), 0, st, nnodes,
                           nb, ne, nchild, stats_dev);
        const double theta = p.theta;
        const int n = static_cast<int>(s.n);
#define GC_TREE(F, O, S, KC)                                                                     \
    hipLaunchKernelGGL((tree_eval_kernel<F, O, S>), dim3(static_cast<unsigned>((n + kEvalLanes - 1) / kEvalLanes)), \
                       block, 0, st, n, geom, mcom, KC, nb, ne, first_child, nchild, idx_sorted, sxs,    \
                       sys, szs, sms, G, eps, theta, b, e, s.ax, s.ay, s.az, s.vx, s.vy, s.vz,     \
                       close_kick ? 1 : 0, half_dt, stats_dev)
#define GC_TREE_O(F, S, KC)                                                                      \
    do {                                                                                         \
        if (order == 0) GC_TREE(F, 0, S, KC);                                                    \
        else if (order == 2) GC_TREE(F, 2, S, KC);                                               \
        else GC_TREE(F, 4, S, KC);                                                               \
    } while (0)
        if (n > 0) {
            if (mixed) {
                if (safe) GC_TREE_O(float, true, kc32);
                else GC_TREE_O(float, false, kc32);
            } else {
                if (safe) GC_TREE_O(double, true, kc64);
                else GC_TREE_O(double, false, kc64);
            }
        }
#undef GC_TREE_O
#undef GC_TREE
        FMM_HIP(hipGetLastError(), "launch tree evaluation");
        FMM_HIP(hipMemcpyAsync(tree_counts, stats_dev, sizeof(unsigned long long) * (kStatInternal + 1),
                               hipMemcpyDeviceToHost, st), "tree counts");
        tree_counts_pending = true;
        return GC_OK;
    }
    // a target range evaluates only its own targets (the tree and lists are
    // the whole set's, as in compute_targets, fmm_solver.cpp:24-46)
    const int* tp = nullptr;
    if (begin > 0 || end < s.n) {
        if (static_cast<size_t>(nleaves) > tleaf_cap) {
            size_t c2 = 0;
            tleaf_cap = 0;  // set by the last allocation only
            FMM_HIP(realloc_dev(&toff, &c2, static_cast<size_t>(nleaves) * 2), "allocate range lists");
            FMM_HIP(realloc_dev(&tcnt, &tleaf_cap, c2), "allocate range lists");
        }
        if (end - begin > tpos_cap)
            FMM_HIP(realloc_dev(&tpos, &tpos_cap, end - begin), "allocate range lists");
        size_t need = 0;
        FMM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need, tcnt, toff, nleaves, st), "size scan");
        if (need > scan_tmp_bytes)
            FMM_HIP(realloc_dev(reinterpret_cast<unsigned char**>(&scan_tmp), &scan_tmp_bytes, need),
                    "allocate scan storage");
        const unsigned lblk = static_cast<unsigned>((nleaves + 255) / 256);
        hipLaunchKernelGGL(range_count_kernel, dim3(lblk), dim3(256), 0, st, nleaves, leaves, nb, ne,
                           idx_sorted, b, e, tcnt);
        FMM_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_tmp_bytes, tcnt, toff, nleaves, st),
                "scan range targets");
        hipLaunchKernelGGL(range_fill_kernel, dim3(lblk), dim3(256), 0, st, nleaves, leaves, nb, ne,
                           idx_sorted, b, e, toff, tpos);
        tp = tpos;
    }
#define GC_FAR(F, O, KC)                                                                         \
    hipLaunchKernelGGL((eval_far_kernel<F, O>), grid, block, 0, st, leaves, geom, mcom, KC, nb, ne, \
                       idx_sorted, sxs, sys, szs, far_off, nfar, far, G, eps, b, e, tp, toff,     \
                       tcnt, s.ax, s.ay, s.az)
#define GC_NEAR(F, S)                                                                            \
    hipLaunchKernelGGL((eval_near_kernel<F, S>), dim3(static_cast<unsigned>(nitems)), block, 0, nst, \
                       items, leaves, geom, nb, ne, sxs, sys, szs, sms, near_off, near, G, eps,   \
                       tp, toff, tcnt, npart)
#define GC_FAR32(O)                                                                              \
    hipLaunchKernelGGL((eval_far_cells_kernel<O>), grid, block, 0, st, leaves, geom, mcom, kc32, nb, \
                       ne, idx_sorted, sxs, sys, szs, far_off, nfar, far, G, eps, b, e, tp, toff, \
                       tcnt, s.ax, s.ay, s.az)
#define GC_NEAR32(S, U)                                                                          \
    hipLaunchKernelGGL((eval_near_cells_kernel<S, U>), dim3(static_cast<unsigned>(nitems)), block, 0, nst, \
                       items, leaves, geom, nb, ne, sxs, sys, szs, srec, static_cast<int>(s.n),    \
                       near_off, near, eps, gmu, tp, toff, tcnt, npart)
    // far field on st, near field on the side stream: independent until the
    // merge (the far kernels write the accelerations, the near kernels their
    // own partial rows)
    // the particles' records, then 2 kEvalLanes zero records
    if (mixed && static_cast<size_t>(s.n) + 2 * kEvalLanes > srec_cap)
        FMM_HIP(realloc_dev(&srec, &srec_cap, static_cast<size_t>(s.n) + 2 * kEvalLanes),
                "allocate source records");
    if (!side) FMM_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking), "create FMM side stream");
    if (!fork) FMM_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming), "create FMM events");
    if (!join) FMM_HIP(hipEventCreateWithFlags(&join, hipEventDisableTiming), "create FMM events");
    FMM_HIP(hipEventRecord(fork, st), "fork near field");
    FMM_HIP(hipStreamWaitEvent(side, fork, 0), "fork near field");
    const hipStream_t nst = side;
    if (mixed) {
        // the sources-across-lanes kernel adds the self pair as an exact
        // zero, which needs its factor G m eps^-3 finite in FP32
        const double e2 = static_cast<double>(static_cast<float>(eps * eps));
        const double rs = e2 > 0.0 ? 1.0 / std::sqrt(e2) : HUGE_VAL;
        const bool self_zero = safe || (rs * rs <= 1e30 &&
                                        std::fabs(G) * mass_abs_max * std::max(rs, rs * rs * rs) <= 1e30);
        // one mass for every particle: the pair terms leave out G m (the
        // self pair's factor is then eps^-3, finite by rs^3 <= 1e30)
        const bool uni = !safe && mass_min == mass_abs_max && mass_min > 0.0 &&
                         rs * rs * rs <= 1e30 && std::isfinite(G * mass_min) && G * mass_min > 0.0 &&
                         static_cast<float>(G * mass_min) > 0.0f;
        const float gmu = uni ? static_cast<float>(G * mass_min) : 0.0f;
        if (self_zero) {
            hipLaunchKernelGGL(near_records_kernel, grid, block, 0, nst, leaves, geom, nb, ne, sxs, sys,
                               szs, sms, G, static_cast<int>(s.n), srec);
            if (safe) GC_NEAR32(true, false);
            else if (uni) GC_NEAR32(false, true);
            else GC_NEAR32(false, false);
        } else {
            GC_NEAR(float, false);
        }
        if (order == 0) GC_FAR32(0);
        else if (order == 2) GC_FAR32(2);
        else GC_FAR32(4);
    } else {
        if (safe) GC_NEAR(double, true);
        else GC_NEAR(double, false);
        if (order == 0) GC_FAR(double, 0, kc64);
        else if (order == 2) GC_FAR(double, 2, kc64);
        else GC_FAR(double, 4, kc64);
    }
    FMM_HIP(hipGetLastError(), "launch FMM far and near fields");
    FMM_HIP(hipEventRecord(join, nst), "join near field");
    FMM_HIP(hipStreamWaitEvent(st, join, 0), "join near field");
#undef GC_FAR
#undef GC_FAR32
#undef GC_NEAR32
#undef GC_NEAR
    hipLaunchKernelGGL(near_merge_kernel, grid, block, 0, st, leaves, nb, ne, idx_sorted, icount, ioff,
                       items, npart, b, e, tp, toff, tcnt, s.ax, s.ay, s.az, s.vx, s.vy, s.vz,
                       close_kick ? 1 : 0, half_dt);
    FMM_HIP(hipGetLastError(), "launch FMM evaluation");
    return GC_OK;
}

FmmEngine::FmmEngine() : impl_(new Impl()) {}
FmmEngine::~FmmEngine() { delete impl_; }

const gc_fmm_stats& FmmEngine::stats() {
    if (impl_->tree_counts_pending && impl_->ev[2] && hipEventSynchronize(impl_->ev[2]) == hipSuccess) {
        impl_->last.m2p_evaluations = impl_->tree_counts[0];
        impl_->last.p2p_pairs = impl_->tree_counts[1];
        impl_->last.leaf_count = impl_->tree_counts[2];
        impl_->last.node_count = 1 + 8 * static_cast<size_t>(impl_->tree_counts[kStatInternal]);
        impl_->tree_counts_pending = false;
    }
    return impl_->last;
}

int FmmEngine::compute(const DeviceState& s, const gc_params& p, size_t begin, size_t end,
                       bool close_kick, double half_dt, hipStream_t st, std::string* err) {
    end = std::min(end, s.n);
    if (s.n == 0 || begin >= end) return GC_OK;
    for (auto& e : impl_->ev)
        if (!e) FMM_HIP(hipEventCreate(&e), "create FMM events");
    FMM_HIP(hipEventRecord(impl_->ev[0], st), "record build start");
    int rc = impl_->build(s, p, st, err);
    if (rc != GC_OK) return rc;
    FMM_HIP(hipEventRecord(impl_->ev[1], st), "record build end");
    const unsigned zb = static_cast<unsigned>((end - begin + 255) / 256);
    hipLaunchKernelGGL(zero_range_kernel, dim3(zb), dim3(256), 0, st, static_cast<int>(begin),
                       static_cast<int>(end), s.ax, s.ay, s.az);
    if (impl_->nleaves == 0) return GC_OK;
    rc = impl_->launch_eval(s, p, begin, end, close_kick, half_dt, st, err);
    FMM_HIP(hipEventRecord(impl_->ev[2], st), "record eval end");
    return rc;
}

int FmmEngine::time_kernels(const DeviceState& s, const gc_params& p, int reps, hipStream_t st,
                            double* ms, double* interactions, std::string* err) {
    int rc = compute(s, p, 0, s.n, false, 0.0, st, err);
    if (rc != GC_OK) return rc;
    if (impl_->nleaves == 0) {
        *ms = 0.0;
        if (interactions) *interactions = 0.0;
        return GC_OK;
    }
    FMM_HIP(hipEventRecord(impl_->ev[3], st), "record");
    for (int r = 0; r < reps; ++r) {
        rc = impl_->launch_eval(s, p, 0, s.n, false, 0.0, st, err);
        if (rc != GC_OK) return rc;
    }
    FMM_HIP(hipEventRecord(impl_->ev[4], st), "record");
    FMM_HIP(hipEventSynchronize(impl_->ev[4]), "wait");
    float t = 0.f;
    FMM_HIP(hipEventElapsedTime(&t, impl_->ev[3], impl_->ev[4]), "elapsed");
    *ms = static_cast<double>(t) / reps;
    if (interactions) *interactions = static_cast<double>(stats().p2p_pairs);
    return GC_OK;
}

int FmmEngine::phase_times(float* build_ms, float* eval_ms) const {
    if (!impl_->ev[2]) return GC_ERR_STATE;
    if (hipEventSynchronize(impl_->ev[2]) != hipSuccess) return GC_ERR_HIP;
    if (hipEventElapsedTime(build_ms, impl_->ev[0], impl_->ev[1]) != hipSuccess) return GC_ERR_HIP;
    if (hipEventElapsedTime(eval_ms, impl_->ev[1], impl_->ev[2]) != hipSuccess) return GC_ERR_HIP;
    return GC_OK;
}

}  // namespace gc
