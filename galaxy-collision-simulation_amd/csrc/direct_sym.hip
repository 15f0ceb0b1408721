// direct_sym.hip — mixed-precision all-pairs forces, each unordered pair
// evaluated once (Newton's third law), as the reference's symmetric direct
// sum does (cpp/direct/direct_solver.cpp:37-57: one softened kernel per
// i < j, added to a_i and subtracted from a_j).
//
// Work is the upper triangle of source-tile pairs (I, J), J >= I, cut into
// items of a few consecutive J tiles of one row I.  A block of four waves
// holds the 256 targets of tile I, one per lane; the sources of tile J
// are staged in LDS (the mixed layout of direct_mixed.hip: FP32 offsets from
// the tile's FP64 centre and G*m).  A wave meets the 64 sources of a chunk
// in 64 steps; at step s lane l takes source (l + s) mod 64, so the 64 lanes
// see 64 distinct sources and every (target, source) pair of the chunk comes
// up once.  Per step and lane:
//   d = o_j - t_i, s2 = |d|^2 + eps^2, r = rsq(s2), r3 = r^3
//   a_i += (G m_j r3) d          (3 FMA into the lane's own target)
//   A_j -= (G m_i r3) d          (3 FMA into the source's running sum)
//   A <- A of lane l + 1         (3 DPP wave_rol:1 moves)
// The source sums ride along the wave with their sources, so one v_rsq_f32
// serves two interactions without any cross-lane reduction; after 64 steps
// lane l holds source l's sum.  The four waves' sums are added in wave
// order through LDS, and each (I, J) pair leaves one FP32 partial per
// target of tile J.  The diagonal tile J == I is an item of its own, each
// of its unordered pairs once at 4 targets per lane (the one-sided form,
// tile_pairs of mixed_tiles.hpp, at 1 or 2).  The i-side sums over all of an
// item's J tiles are the item's FP32 row partial.  The merge adds, per
// target, its row's item partials in J order and its column's pair partials
// in I order, in FP64: the result does not depend on scheduling.
#include "gc_internal.hpp"
#include "mixed_tiles.hpp"

#include <algorithm>
#include <vector>

namespace gc {
namespace {

// DPP wave_rol:1 (GFX9 encoding): lane l receives lane (l + 1) mod 64.
constexpr int kWaveRol1 = 0x134;

__device__ __forceinline__ float wave_rol1(float v) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), kWaveRol1, 0xF, 0xF, true));
}

__host__ __device__ inline size_t pair_slot(int I, int J) {  // I < J
    return static_cast<size_t>(J) * static_cast<size_t>(J - 1) / 2 + static_cast<size_t>(I);
}


// A tile's shared G m is used by the scaled equal-mass pair step only within
// [2^-40, 2^40] and with eps^2 / (G m) in [1e-20, 1e30], where the scaled
// squares stay finite wherever the unscaled terms are above FP32's normal
// range.
__device__ __forceinline__ bool sym_weight_in_range(float w0, float eps2) {
    return w0 >= 0x1p-40f && w0 <= 0x1p40f && eps2 >= 1e-20f * w0 && eps2 <= 1e30f * w0;
}

#ifndef GC_SYM_LDS_PAD
#define GC_SYM_LDS_PAD 1024
#endif
// Extra (unused) LDS per block of the 4-targets-per-lane kernel: with its
// 8,960 bytes of source tile and reduction space, 1 KB more holds it at 16
// blocks per CU of 160 KB, 4 waves per SIMD: 4,096 resident items on 256
// CUs, which the work list fills in whole rounds (its VGPRs alone would
// allow 5 waves, its LDS 18 blocks).
constexpr unsigned kSymLdsPad = GC_SYM_LDS_PAD;

constexpr int kChunk = 64;                // sources per rotation (one wave)
constexpr int kChunks = kTile / kChunk;   // 4

// One target (offsets t, G m_i) and one source q of a rotation step: the
// pair's force is added to the target's sum p and subtracted from the
// source's running sum a.  live == false drops the pair (both sides).
template <bool SAFE>
__device__ __forceinline__ void sym_pair(float4 q, float tx, float ty, float tz, float mi,
                                         float eps2, bool live, float& px, float& py, float& pz,
                                         float& ax, float& ay, float& az) {
    const float dx = q.x - tx, dy = q.y - ty, dz = q.z - tz;
    const float s2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, eps2)));
    float r;
    if (SAFE) {
        r = s2 > kTinySeparation2 ? __builtin_amdgcn_rsqf(s2) : 0.f;
    } else {
        r = __builtin_amdgcn_rsqf(s2);
    }
    r = live ? r : 0.f;
    const float r3 = (r * r) * r;
    const float kj = q.w * r3, ki = mi * r3;
    px = __builtin_fmaf(dx, kj, px);
    py = __builtin_fmaf(dy, kj, py);
    pz = __builtin_fmaf(dz, kj, pz);
    ax = __builtin_fmaf(-dx, ki, ax);
    ay = __builtin_fmaf(-dy, ki, ay);
    az = __builtin_fmaf(-dz, ki, az);
}

// Chunk C of the diagonal tile with 4 targets per lane, one wave: the
// lane's target groups k (targets k * 64 + lane) are the tile's 64-particle
// chunks.  Groups k < C meet every source of chunk C; group C meets the
// sources of its own chunk at forward distance 1..32 (distance 32 from
// lanes 0..31 only), so each unordered pair of the tile comes up exactly
// once and the tile costs half a tile pair.  After the 64 rotation steps
// lane l holds the source sum of particle C * 64 + l, the lane's own target
// of group C, and adds it to that target's sum.
template <int C, bool SAFE>
__device__ __forceinline__ void diag_chunk(const float4* row, int lane, float eps2,
                                           const float* tx, const float* ty, const float* tz,
                                           const float* mi, float* px, float* py, float* pz) {
    float ax = 0.f, ay = 0.f, az = 0.f;
    auto step = [&](int s, bool half, bool live) {
        const float4 q = row[s];
#pragma unroll
        for (int k = 0; k < C; ++k)
            sym_pair<SAFE>(q, tx[k], ty[k], tz[k], mi[k], eps2, true, px[k], py[k], pz[k], ax, ay,
                           az);
        if (half)
            sym_pair<SAFE>(q, tx[C], ty[C], tz[C], mi[C], eps2, live, px[C], py[C], pz[C], ax, ay,
                           az);
        ax = wave_rol1(ax);
        ay = wave_rol1(ay);
        az = wave_rol1(az);
    };
    step(0, false, false);
#pragma unroll 8
    for (int s = 1; s < 32; ++s) step(s, true, true);
    step(32, true, lane < 32);
#pragma unroll 8
    for (int s = 33; s < kChunk; ++s) step(s, false, false);
    px[C] += ax;
    py[C] += ay;
    pz[C] += az;
}

// IPT targets per lane: a block of kTile / IPT threads holds the kTile
// targets of tile I (thread t: targets t + k kTile / IPT).  The rotating
// source sums are shared by a lane's IPT targets, so the three DPP moves of
// a step serve 2 * IPT interactions.
template <int IPT, bool SAFE>
__global__ __launch_bounds__(kTile / IPT) __attribute__((amdgpu_waves_per_eu(4, 4))) void direct_sym_kernel(
    const SymItem* __restrict__ items, int nitems, int n, const int* __restrict__ tile_valid,
    int ntiles, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, const float4* __restrict__ jp,
    const double* __restrict__ centre, float eps2, const float* __restrict__ wt,
    float* __restrict__ ipart, float* __restrict__ jpart) {
    constexpr int kThreads = kTile / IPT;
    constexpr int kWaves = kThreads / 64;
    // Source tile J in LDS, each 64-source chunk stored twice in a row so
    // that lane l at step s reads entry l + s of its chunk without wrapping.
    __shared__ float4 src2[kChunks][2 * kChunk];
    // waves 1.. hand their lanes' chunk sums of -G m_i r3 d to wave 0
    __shared__ float red[kWaves > 1 ? kWaves - 1 : 1][3][kChunk];
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const SymItem item = items[it];
        const int I = item.I;
        // G m of every particle of tile I when they are all equal (0: not)
        const float wI = (!SAFE && IPT == 4 && wt != nullptr) ? wt[I] : 0.f;
        double xi[IPT], yi[IPT], zi[IPT];
        float mi[IPT];
        const int valid = tile_valid ? tile_valid[I] : n - I * kTile;
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            const int i = I * kTile + k * kThreads + t;
            const bool active = k * kThreads + t < valid;
            xi[k] = active ? x[i] : 0.0;
            yi[k] = active ? y[i] : 0.0;
            zi[k] = active ? z[i] : 0.0;
            mi[k] = jp[i].w;  // G m_i; zero past n
        }
        // the item's row partial: its targets' FP32 sums over all of its
        // source tiles, at the item's slot
        float qx[IPT], qy[IPT], qz[IPT];
#pragma unroll
        for (int k = 0; k < IPT; ++k) qx[k] = qy[k] = qz[k] = 0.f;
        auto store_row = [&]() {
            float* base = ipart + static_cast<size_t>(item.islot) * 3 * kTile;
#pragma unroll
            for (int k = 0; k < IPT; ++k) {
                const int e = k * kThreads + t;
                base[e] = qx[k];
                base[kTile + e] = qy[k];
                base[2 * kTile + e] = qz[k];
            }
        };
        if (item.jb == I) {
            // the diagonal tile, always an item of its own
            float tx[IPT], ty[IPT], tz[IPT];
            tile_offsets<IPT>(centre, ntiles, I, xi, yi, zi, tx, ty, tz);
            const float4* src = jp + static_cast<size_t>(I) * kTile;
            if constexpr (IPT == 4) {
                // each unordered pair once
                for (int e = t; e < kTile; e += kThreads) {
                    const float4 v = src[e];
                    src2[e / kChunk][e % kChunk] = v;
                    src2[e / kChunk][kChunk + e % kChunk] = v;
                }
                __syncthreads();
                // a chunk of padding (no particle at or past its first slot)
                // adds nothing: its sources have zero mass and its source
                // sums belong to no particle
                diag_chunk<0, SAFE>(&src2[0][lane], lane, eps2, tx, ty, tz, mi, qx, qy, qz);
                if (valid > kChunk)
                    diag_chunk<1, SAFE>(&src2[1][lane], lane, eps2, tx, ty, tz, mi, qx, qy, qz);
                if (valid > 2 * kChunk)
                    diag_chunk<2, SAFE>(&src2[2][lane], lane, eps2, tx, ty, tz, mi, qx, qy, qz);
                if (valid > 3 * kChunk)
                    diag_chunk<3, SAFE>(&src2[3][lane], lane, eps2, tx, ty, tz, mi, qx, qy, qz);
            } else {
                // one-sided: every target over the tile's sources
                tile_pairs<IPT, SAFE>(src, eps2, tx, ty, tz, qx, qy, qz);
            }
            store_row();
            __syncthreads();  // LDS reuse by the next item
            continue;
        }
        // the next source tile of the item is loaded into registers while
        // this one is evaluated
        constexpr int kPer = kTile / kThreads;
        float4 pre[kPer];
        for (int J = item.jb; J < item.je; ++J) {
            float tx[IPT], ty[IPT], tz[IPT];
            float* px = qx;
            float* py = qy;
            float* pz = qz;
            // Equal-mass tiles I and J (w = G m): offsets scaled by
            // w_J^-1/2 and eps^2 by 1 / w_J give d' r'^3 = w_J d r^-3, the
            // target's term, with no mass product; the source sums
            // sum d' r'^3 are scaled by -w_I / w_J once per chunk.
            // eps'^2 rides in the fourth slot of each staged source.
            const float wJ = wI != 0.f ? wt[J] : 0.f;
            const bool uni = wJ != 0.f;
            double sd = 1.0;
            float sJ = 1.f, eps2u = eps2, fJ = 0.f;
            if (uni) {
                sd = 1.0 / sqrt(static_cast<double>(wJ));
                sJ = static_cast<float>(sd);
                eps2u = static_cast<float>(static_cast<double>(eps2) / static_cast<double>(wJ));
                fJ = static_cast<float>(-static_cast<double>(wI) / static_cast<double>(wJ));
                const double cx = centre[J], cy = centre[ntiles + J], cz = centre[2 * ntiles + J];
#pragma unroll
                for (int k = 0; k < IPT; ++k) {
                    tx[k] = static_cast<float>((xi[k] - cx) * sd);
                    ty[k] = static_cast<float>((yi[k] - cy) * sd);
                    tz[k] = static_cast<float>((zi[k] - cz) * sd);
                }
            } else {
                tile_offsets<IPT>(centre, ntiles, J, xi, yi, zi, tx, ty, tz);
            }
            const float4* src = jp + static_cast<size_t>(J) * kTile;
            // source chunks holding particles of tile J (the rest is padding:
            // zero-mass sources whose sums belong to no particle, which the
            // merge never reads)
            const int jvalid = tile_valid ? tile_valid[J] : n - J * kTile;
            {
#pragma unroll
                for (int u = 0; u < kPer; ++u) {
                    const int e = t + u * kThreads;
                    float4 v = J > item.jb ? pre[u] : src[e];
                    if (uni) v = make_float4(v.x * sJ, v.y * sJ, v.z * sJ, eps2u);
                    src2[e / kChunk][e % kChunk] = v;
                    src2[e / kChunk][kChunk + e % kChunk] = v;
                }
                __syncthreads();
                if (J + 1 < item.je) {
                    const float4* nsrc = src + kTile;
#pragma unroll
                    for (int u = 0; u < kPer; ++u) pre[u] = nsrc[t + u * kThreads];
                }
#pragma unroll
                for (int c = 0; c < kChunks; ++c) {
                    if (c * kChunk >= jvalid) break;
                    // lane l holds the running sum of source c*64 + (l + s) % 64
                    // at step s; wave_rol1 after each step keeps that true.
                    float ax = 0.f, ay = 0.f, az = 0.f;
                    const float4* row = &src2[c][lane];
                    if (uni) {
#pragma unroll 8
                        for (int s = 0; s < kChunk; ++s) {
                            const float4 q = row[s];
#pragma unroll
                            for (int k = 0; k < IPT; ++k) {
                                const float dx = q.x - tx[k], dy = q.y - ty[k], dz = q.z - tz[k];
                                const float s2 = __builtin_fmaf(
                                    dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, q.w)));
                                const float r = __builtin_amdgcn_rsqf(s2);
                                const float r3 = (r * r) * r;
                                px[k] = __builtin_fmaf(dx, r3, px[k]);
                                py[k] = __builtin_fmaf(dy, r3, py[k]);
                                pz[k] = __builtin_fmaf(dz, r3, pz[k]);
                                ax = __builtin_fmaf(dx, r3, ax);
                                ay = __builtin_fmaf(dy, r3, ay);
                                az = __builtin_fmaf(dz, r3, az);
                            }
                            ax = wave_rol1(ax);
                            ay = wave_rol1(ay);
                            az = wave_rol1(az);
                        }
                        ax *= fJ;
                        ay *= fJ;
                        az *= fJ;
                    } else {
#pragma unroll 8
                    for (int s = 0; s < kChunk; ++s) {
                        const float4 q = row[s];
#pragma unroll
                        for (int k = 0; k < IPT; ++k) {
                            const float dx = q.x - tx[k], dy = q.y - ty[k], dz = q.z - tz[k];
                            const float s2 = __builtin_fmaf(
                                dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, eps2)));
                            float r;
                            if (SAFE) {
                                r = s2 > kTinySeparation2 ? __builtin_amdgcn_rsqf(s2) : 0.f;
                            } else {
                                r = __builtin_amdgcn_rsqf(s2);
                            }
                            const float r3 = (r * r) * r;
                            const float kj = q.w * r3, ki = mi[k] * r3;
                            px[k] = __builtin_fmaf(dx, kj, px[k]);
                            py[k] = __builtin_fmaf(dy, kj, py[k]);
                            pz[k] = __builtin_fmaf(dz, kj, pz[k]);
                            ax = __builtin_fmaf(-dx, ki, ax);
                            ay = __builtin_fmaf(-dy, ki, ay);
                            az = __builtin_fmaf(-dz, ki, az);
                        }
                        ax = wave_rol1(ax);
                        ay = wave_rol1(ay);
                        az = wave_rol1(az);
                    }
                    }
                    // after 64 steps lane l holds source c*64 + l; wave 0 adds
                    // the other waves' sums in wave order and stores the chunk
                    float* out = jpart + (item.pbase + pair_slot(I - item.t0, J - item.t0)) * 3 * kTile + c * kChunk + lane;
                    if (kWaves > 1) {
                        if (wave > 0) {
                            red[wave - 1][0][lane] = ax;
                            red[wave - 1][1][lane] = ay;
                            red[wave - 1][2][lane] = az;
                        }
                        __syncthreads();
                        if (wave == 0) {
#pragma unroll
                            for (int w = 0; w < kWaves - 1; ++w) {
                                ax += red[w][0][lane];
                                ay += red[w][1][lane];
                                az += red[w][2][lane];
                            }
                        }
                        __syncthreads();
                    }
                    if (wave == 0) {
                        out[0] = ax;
                        out[kTile] = ay;
                        out[2 * kTile] = az;
                    }
                }
            }
        }
        store_row();
        __syncthreads();  // LDS reuse by the next item
    }
}

// a_i = sum over row T's items of their row partials (J order) + sum over
// I < T of the (I, T) column partials (I order), in FP64; optional closing
// half kick.  A block of 16 groups of 64 lanes merges tile T, 4 targets per
// lane (float4 loads); group g adds the g-th of 16 consecutive ranges of
// that sequence of partials, and the 16 group sums are added in group
// order, so the result is deterministic.  Every partial is 3 KB.  Blocks
// run the longest columns (largest T) first.
constexpr int kMergeGroups = 16;
constexpr int kMergeLanes = kTile / 4;  // 64

// Row and column partials of target tile T summed per group into part.
__device__ __forceinline__ void merge_sym_groups(
    int T, const int* __restrict__ tile_t0, const int* __restrict__ row_start,
    const long long* __restrict__ tile_pbase, const float* __restrict__ ipart,
    const float* __restrict__ jpart, double (*part)[3][kTile]) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x % kMergeLanes;
    const int g = threadIdx.x / kMergeLanes;
    double a[3][4];
#pragma unroll
    for (int c = 0; c < 3; ++c) a[c][0] = a[c][1] = a[c][2] = a[c][3] = 0.0;
    const int r = T - tile_t0[T];
    const long long pbase = tile_pbase[T];
    const int s0 = row_start[T];
    const int nrow = row_start[T + 1] - s0;  // row T's items, in J order
    const int total = nrow + r;
    const int q0 = g * total / kMergeGroups, q1 = (g + 1) * total / kMergeGroups;
    for (int q = q0; q < q1; ++q) {
        const float* b;
        if (q < nrow) {
            b = ipart + static_cast<size_t>(s0 + q) * 3 * kTile;
        } else {
            b = jpart + (pbase + pair_slot(q - nrow, r)) * 3 * kTile;
        }
        b += 4 * lane;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float4 v = *reinterpret_cast<const float4*>(b + c * kTile);
            a[c][0] += static_cast<double>(v.x);
            a[c][1] += static_cast<double>(v.y);
            a[c][2] += static_cast<double>(v.z);
            a[c][3] += static_cast<double>(v.w);
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < 4; ++k) part[g][c][4 * lane + k] = a[c][k];
}

__global__ __launch_bounds__(kMergeLanes * kMergeGroups) void merge_sym_kernel(
    int n, int ntiles, const int* __restrict__ tile_t0, const int* __restrict__ row_start,
    const long long* __restrict__ tile_pbase, const int* __restrict__ tile_valid,
    const float* __restrict__ ipart, const float* __restrict__ jpart, double* __restrict__ ax,
    double* __restrict__ ay, double* __restrict__ az, double* __restrict__ vx,
    double* __restrict__ vy, double* __restrict__ vz, int close_kick, double half_dt) {
#pragma clang fp contract(off)
    __shared__ double part[kMergeGroups][3][kTile];
    const int T = ntiles - 1 - static_cast<int>(blockIdx.x);
    merge_sym_groups(T, tile_t0, row_start, tile_pbase, ipart, jpart, part);
    __syncthreads();
    // thread (c, e) of the first 3 * kTile adds target e's component c
    const int t = threadIdx.x;
    if (t >= 3 * kTile) return;
    const int c = t / kTile, e = t % kTile;
    double sum = 0.0;
#pragma unroll
    for (int k = 0; k < kMergeGroups; ++k) sum += part[k][c][e];
    const int valid = tile_valid ? tile_valid[T] : n - T * kTile;
    if (e >= valid) return;
    const int i = T * kTile + e;
    double* acc = c == 0 ? ax : (c == 1 ? ay : az);
    double* vel = c == 0 ? vx : (c == 1 ? vy : vz);
    acc[i] = sum;
    if (close_kick) vel[i] = vel[i] + sum * half_dt;
}

// The merge of one KDK step fused with the next step's opening: closing
// kick, then the next opening kick and drift (open_step_kernel's
// arithmetic), then tile T's source entries for the next force evaluation
// (pack_mixed_kernel's: bounding-box centre, FP32 offsets, G m).  Every
// operation and rounding is the separate kernels', so the trajectory is
// unchanged; the state and the tiles are touched once instead of three
// times.
__global__ __launch_bounds__(kMergeLanes * kMergeGroups) void merge_sym_open_kernel(
    int n, int ntiles, const int* __restrict__ tile_t0, const int* __restrict__ row_start,
    const long long* __restrict__ tile_pbase, const float* __restrict__ ipart,
    const float* __restrict__ jpart, double* __restrict__ x, double* __restrict__ y,
    double* __restrict__ z, double* __restrict__ vx, double* __restrict__ vy,
    double* __restrict__ vz, double* __restrict__ ax, double* __restrict__ ay,
    double* __restrict__ az, const double* __restrict__ m, double half_dt, double dt, double G,
    float eps2, float4* __restrict__ jp, double* __restrict__ centre, float* __restrict__ wt) {
#pragma clang fp contract(off)
    __shared__ double part[kMergeGroups][3][kTile];
    __shared__ double box[2][3][4];   // per component: wave minima, maxima
    __shared__ float offs[3][kTile];
    const int T = ntiles - 1 - static_cast<int>(blockIdx.x);
    merge_sym_groups(T, tile_t0, row_start, tile_pbase, ipart, jpart, part);
    __syncthreads();
    const int t = threadIdx.x;
    const bool active = t < 3 * kTile;  // thread (c, e): component c of target e
    const int c = active ? t / kTile : 0, e = t % kTile;
    const int valid = n - T * kTile;
    const bool own = active && e < valid;
    const int i = T * kTile + e;
    double pos = 0.0;
    if (own) {
        double sum = 0.0;
#pragma unroll
        for (int k = 0; k < kMergeGroups; ++k) sum += part[k][c][e];
        double* acc = c == 0 ? ax : (c == 1 ? ay : az);
        double* vel = c == 0 ? vx : (c == 1 ? vy : vz);
        double* xc = c == 0 ? x : (c == 1 ? y : z);
        acc[i] = sum;
        const double closed = vel[i] + sum * half_dt;   // closing kick
        const double opened = closed + sum * half_dt;   // next opening kick
        vel[i] = opened;
        pos = xc[i] + opened * dt;                       // drift
        xc[i] = pos;
    }
    // tile bounding box per component (4 waves per component)
    double lo = own ? pos : INFINITY, hi = own ? pos : -INFINITY;
    lo = wave_min(lo);
    hi = wave_max(hi);
    const int w = (t >> 6) & 3;
    if (active && (t & 63) == 0) {
        box[0][c][w] = lo;
        box[1][c][w] = hi;
    }
    __syncthreads();
    if (active) {
        double l = box[0][c][0], h = box[1][c][0];
#pragma unroll
        for (int k = 1; k < 4; ++k) {
            l = fmin(l, box[0][c][k]);
            h = fmax(h, box[1][c][k]);
        }
        const double mid = (l + h) * 0.5;
        offs[c][e] = own ? static_cast<float>(pos - mid) : 0.0f;
        if (e == 0) centre[c * ntiles + T] = mid;
    }
    __syncthreads();
    // the tile's weight for the next evaluation (sym_tile_weights_kernel's)
    const float w0 = static_cast<float>(G * m[T * kTile]);
    bool same = true;
    if (t < kTile) {
        float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
        if (t < valid) o = make_float4(offs[0][t], offs[1][t], offs[2][t],
                                       static_cast<float>(G * m[i]));
        jp[i] = o;
        same = t < valid && o.w == w0;
    }
    const bool all_same = __syncthreads_and(same) != 0;
    if (wt != nullptr && t == 0) wt[T] = all_same && sym_weight_in_range(w0, eps2) ? w0 : 0.f;
}

// ---- force VJP, each unordered pair once ---------------------------------
// With w = G m_i lambda_j - G m_j lambda_i and J(d) = r3 I - 3 r5 d d^T
// (d = x_j - x_i, even in d), the pair gives xbar_i += J w and
// xbar_j -= J w, and the mass terms mbar_i -= r3 lambda_j . d,
// mbar_j += r3 lambda_i . d (G applied at the merge).  The schedule is the
// force kernel's: items of row I, sources of tile J rotating through the
// wave with their four running sums; the diagonal tile runs the one-sided
// form with the self pair dropped by index.
template <int IPT, bool SAFE>
__global__ __launch_bounds__(kTile / IPT) void vjp_sym_kernel(
    const SymItem* __restrict__ items, int nitems, int n, int ntiles,
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const float4* __restrict__ jp, const float4* __restrict__ lp,
    const double* __restrict__ centre, float eps2, float* __restrict__ ipart,
    float* __restrict__ jpart) {
    constexpr int kThreads = kTile / IPT;
    constexpr int kWaves = kThreads / 64;
    __shared__ float4 src2[kChunks][2 * kChunk];
    __shared__ float4 lam2[kChunks][2 * kChunk];
    __shared__ float red[kWaves > 1 ? kWaves - 1 : 1][4][kChunk];
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const SymItem item = items[it];
        const int I = item.I;
        double xi[IPT], yi[IPT], zi[IPT];
        float mi[IPT], li0[IPT], li1[IPT], li2[IPT];
        float h[4][IPT], l[4][IPT];  // (hi, lo) of xbar x, y, z and mbar
        const int valid = n - I * kTile;
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            const int i = I * kTile + k * kThreads + t;
            const bool active = k * kThreads + t < valid;
            xi[k] = active ? x[i] : 0.0;
            yi[k] = active ? y[i] : 0.0;
            zi[k] = active ? z[i] : 0.0;
            mi[k] = jp[i].w;  // G m_i; zero past n
            const float4 ml = lp[i];  // zero past n
            li0[k] = ml.x;
            li1[k] = ml.y;
            li2[k] = ml.z;
#pragma unroll
            for (int c = 0; c < 4; ++c) h[c][k] = l[c][k] = 0.f;
        }
        for (int J = item.jb; J < item.je; ++J) {
            float tx[IPT], ty[IPT], tz[IPT], p[4][IPT];
            tile_offsets<IPT>(centre, ntiles, J, xi, yi, zi, tx, ty, tz);
#pragma unroll
            for (int k = 0; k < IPT; ++k) p[0][k] = p[1][k] = p[2][k] = p[3][k] = 0.f;
            const float4* src = jp + static_cast<size_t>(J) * kTile;
            const float4* lsrc = lp + static_cast<size_t>(J) * kTile;
            if (J == I) {
                // one-sided over the tile's own sources, self pair excluded
#pragma unroll 4
                for (int jj = 0; jj < kTile; ++jj) {
                    const float4 q = src[jj];
                    const float4 lj = lsrc[jj];
#pragma unroll
                    for (int k = 0; k < IPT; ++k) {
                        const float dx = q.x - tx[k], dy = q.y - ty[k], dz = q.z - tz[k];
                        const float s2 = __builtin_fmaf(
                            dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, eps2)));
                        float r;
                        if (SAFE) {
                            r = s2 > kTinySeparation2 ? __builtin_amdgcn_rsqf(s2) : 0.f;
                        } else {
                            r = __builtin_amdgcn_rsqf(s2);
                        }
                        r = jj == k * kThreads + t ? 0.f : r;
                        const float r2 = r * r, r3 = r2 * r;
                        const float w0 = __builtin_fmaf(mi[k], lj.x, -q.w * li0[k]);
                        const float w1 = __builtin_fmaf(mi[k], lj.y, -q.w * li1[k]);
                        const float w2 = __builtin_fmaf(mi[k], lj.z, -q.w * li2[k]);
                        const float dw = __builtin_fmaf(dz, w2, __builtin_fmaf(dy, w1, dx * w0));
                        const float cc = 3.f * r3 * r2 * dw;
                        p[0][k] = __builtin_fmaf(r3, w0, __builtin_fmaf(-cc, dx, p[0][k]));
                        p[1][k] = __builtin_fmaf(r3, w1, __builtin_fmaf(-cc, dy, p[1][k]));
                        p[2][k] = __builtin_fmaf(r3, w2, __builtin_fmaf(-cc, dz, p[2][k]));
                        const float ldj = __builtin_fmaf(lj.z, dz, __builtin_fmaf(lj.y, dy, lj.x * dx));
                        p[3][k] = __builtin_fmaf(-r3, ldj, p[3][k]);
                    }
                }
            } else {
                for (int e = t; e < kTile; e += kThreads) {
                    const float4 v = src[e], w = lsrc[e];
                    src2[e / kChunk][e % kChunk] = v;
                    src2[e / kChunk][kChunk + e % kChunk] = v;
                    lam2[e / kChunk][e % kChunk] = w;
                    lam2[e / kChunk][kChunk + e % kChunk] = w;
                }
                __syncthreads();
                for (int c = 0; c < kChunks; ++c) {
                    // lane l holds the running sums of source c*64 + (l + s) % 64
                    float b0 = 0.f, b1 = 0.f, b2 = 0.f, bm = 0.f;
                    const float4* row = &src2[c][lane];
                    const float4* lrow = &lam2[c][lane];
#pragma unroll 4
                    for (int s = 0; s < kChunk; ++s) {
                        const float4 q = row[s];
                        const float4 lj = lrow[s];
#pragma unroll
                        for (int k = 0; k < IPT; ++k) {
                            const float dx = q.x - tx[k], dy = q.y - ty[k], dz = q.z - tz[k];
                            const float s2 = __builtin_fmaf(
                                dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, eps2)));
                            float r;
                            if (SAFE) {
                                r = s2 > kTinySeparation2 ? __builtin_amdgcn_rsqf(s2) : 0.f;
                            } else {
                                r = __builtin_amdgcn_rsqf(s2);
                            }
                            const float r2 = r * r, r3 = r2 * r;
                            const float w0 = __builtin_fmaf(mi[k], lj.x, -q.w * li0[k]);
                            const float w1 = __builtin_fmaf(mi[k], lj.y, -q.w * li1[k]);
                            const float w2 = __builtin_fmaf(mi[k], lj.z, -q.w * li2[k]);
                            const float dw = __builtin_fmaf(dz, w2, __builtin_fmaf(dy, w1, dx * w0));
                            const float cc = 3.f * r3 * r2 * dw;
                            p[0][k] = __builtin_fmaf(r3, w0, __builtin_fmaf(-cc, dx, p[0][k]));
                            p[1][k] = __builtin_fmaf(r3, w1, __builtin_fmaf(-cc, dy, p[1][k]));
                            p[2][k] = __builtin_fmaf(r3, w2, __builtin_fmaf(-cc, dz, p[2][k]));
                            b0 = __builtin_fmaf(-r3, w0, __builtin_fmaf(cc, dx, b0));
                            b1 = __builtin_fmaf(-r3, w1, __builtin_fmaf(cc, dy, b1));
                            b2 = __builtin_fmaf(-r3, w2, __builtin_fmaf(cc, dz, b2));
                            const float ldj = __builtin_fmaf(lj.z, dz, __builtin_fmaf(lj.y, dy, lj.x * dx));
                            const float ldi = __builtin_fmaf(li2[k], dz, __builtin_fmaf(li1[k], dy, li0[k] * dx));
                            p[3][k] = __builtin_fmaf(-r3, ldj, p[3][k]);
                            bm = __builtin_fmaf(r3, ldi, bm);
                        }
                        b0 = wave_rol1(b0);
                        b1 = wave_rol1(b1);
                        b2 = wave_rol1(b2);
                        bm = wave_rol1(bm);
                    }
                    float* out = jpart + (item.pbase + pair_slot(I - item.t0, J - item.t0)) * 4 * kTile + c * kChunk + lane;
                    if (kWaves > 1) {
                        if (wave > 0) {
                            red[wave - 1][0][lane] = b0;
                            red[wave - 1][1][lane] = b1;
                            red[wave - 1][2][lane] = b2;
                            red[wave - 1][3][lane] = bm;
                        }
                        __syncthreads();
                        if (wave == 0) {
#pragma unroll
                            for (int w = 0; w < kWaves - 1; ++w) {
                                b0 += red[w][0][lane];
                                b1 += red[w][1][lane];
                                b2 += red[w][2][lane];
                                bm += red[w][3][lane];
                            }
                        }
                        __syncthreads();
                    }
                    if (wave == 0) {
                        out[0] = b0;
                        out[kTile] = b1;
                        out[2 * kTile] = b2;
                        out[3 * kTile] = bm;
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < IPT; ++k)
#pragma unroll
                for (int c = 0; c < 4; ++c) two_sum_acc(h[c][k], l[c][k], p[c][k]);
        }
        float* base = ipart + static_cast<size_t>(item.islot) * 8 * kTile;
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            const int e = k * kThreads + t;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                base[c * kTile + e] = h[c][k];
                base[(4 + c) * kTile + e] = l[c][k];
            }
        }
        __syncthreads();  // LDS reuse by the next item
    }
}

// xbar_T, mbar_T = the row's item partials (hi + lo, J order) + the
// column's pair partials (I order), in FP64, by 8 groups of 64 lanes (4
// targets per lane) whose sums are added in group order.
constexpr int kVjpMergeGroups = 8;

__global__ __launch_bounds__(kMergeLanes * kVjpMergeGroups) void vjp_sym_merge_kernel(
    int n, int ntiles, const int* __restrict__ row_start, const int* __restrict__ tile_t0,
    const long long* __restrict__ tile_pbase, const float* __restrict__ ipart,
    const float* __restrict__ jpart, double G, double* __restrict__ xbar,
    double* __restrict__ mbar, int ostride, int accumulate) {
#pragma clang fp contract(off)
    __shared__ double part[kVjpMergeGroups][4][kTile];
    const int T = ntiles - 1 - static_cast<int>(blockIdx.x);
    const int lane = threadIdx.x % kMergeLanes;
    const int g = threadIdx.x / kMergeLanes;
    double a[4][4];
#pragma unroll
    for (int c = 0; c < 4; ++c) a[c][0] = a[c][1] = a[c][2] = a[c][3] = 0.0;
    const int r0 = row_start[T], nr = row_start[T + 1] - r0;
    const int s1 = r0 + (g + 1) * nr / kVjpMergeGroups;
    for (int s = r0 + g * nr / kVjpMergeGroups; s < s1; ++s) {
        const float* b = ipart + static_cast<size_t>(s) * 8 * kTile + 4 * lane;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float4 hi = *reinterpret_cast<const float4*>(b + c * kTile);
            const float4 lo = *reinterpret_cast<const float4*>(b + (c + 4) * kTile);
            a[c][0] += static_cast<double>(hi.x) + static_cast<double>(lo.x);
            a[c][1] += static_cast<double>(hi.y) + static_cast<double>(lo.y);
            a[c][2] += static_cast<double>(hi.z) + static_cast<double>(lo.z);
            a[c][3] += static_cast<double>(hi.w) + static_cast<double>(lo.w);
        }
    }
    const int t0 = tile_t0[T], cols = T - t0;
    const float* col = jpart + (tile_pbase[T] + pair_slot(0, cols)) * 4 * kTile + 4 * lane;
    const int i1 = (g + 1) * cols / kVjpMergeGroups;
    for (int I = g * cols / kVjpMergeGroups; I < i1; ++I) {
        const float* b = col + static_cast<size_t>(I) * 4 * kTile;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float4 v = *reinterpret_cast<const float4*>(b + c * kTile);
            a[c][0] += static_cast<double>(v.x);
            a[c][1] += static_cast<double>(v.y);
            a[c][2] += static_cast<double>(v.z);
            a[c][3] += static_cast<double>(v.w);
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int k = 0; k < 4; ++k) part[g][c][4 * lane + k] = a[c][k];
    __syncthreads();
    const int valid = n - T * kTile;
    for (int q = threadIdx.x; q < 4 * kTile; q += kMergeLanes * kVjpMergeGroups) {
        const int c = q / kTile, e = q % kTile;
        if (e >= valid) continue;
        double sum = 0.0;
#pragma unroll
        for (int k = 0; k < kVjpMergeGroups; ++k) sum += part[k][c][e];
        const int i = T * kTile + e;
        double* out = c < 3 ? xbar + static_cast<size_t>(c) * ostride + i : mbar + i;
        const double v = c < 3 ? sum : G * sum;
        *out = accumulate ? *out + v : v;
    }
}

// wt[T] = the G m of tile T's particles when the tile is full and they all
// share it (sym_weight_in_range), else 0.  One wave per tile.
__global__ __launch_bounds__(256) void sym_tile_weights_kernel(const float4* __restrict__ jp,
                                                               const int* __restrict__ tile_valid,
                                                               int n, int ntiles, float eps2,
                                                               float* __restrict__ wt) {
    const int T = static_cast<int>(blockIdx.x) * 4 + static_cast<int>(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (T >= ntiles) return;
    const int valid = tile_valid ? tile_valid[T] : n - T * kTile;
    const float4* src = jp + static_cast<size_t>(T) * kTile;
    const float w0 = src[0].w;
    bool same = valid >= kTile;
#pragma unroll
    for (int k = 0; k < kTile / 64; ++k) same = same && src[k * 64 + lane].w == w0;
    const bool all_same = __all(same);
    if (lane == 0) wt[T] = all_same && sym_weight_in_range(w0, eps2) ? w0 : 0.f;
}

}  // namespace

int sym_prepare_runs(SymPlan& plan, const int* run_t0, const int* run_nt, int nruns, int len,
                     hipStream_t st, std::string* err) {
    if (plan.items) (void)hipStreamSynchronize(st);
    sym_release(plan);
    SymHostPlan h;
    if (!build_sym_plan(run_t0, run_nt, nruns, len, kSymMaxTiles, h, err)) return 1;
    plan.nitems = static_cast<int>(h.items.size());
    plan.ipart_floats = std::max<size_t>(1, h.items.size()) * 3 * kTile;
    plan.jpart_floats = std::max<size_t>(1, static_cast<size_t>(h.pairs)) * 3 * kTile;
    hipError_t e = hipMalloc(&plan.items, sizeof(SymItem) * h.items.size());
    if (e == hipSuccess) e = hipMalloc(&plan.row_start, sizeof(int) * h.row_start.size());
    if (e == hipSuccess) e = hipMalloc(&plan.tile_t0, sizeof(int) * h.tile_t0.size());
    if (e == hipSuccess) e = hipMalloc(&plan.tile_pbase, sizeof(long long) * h.tile_pbase.size());
    if (e == hipSuccess) e = hipMalloc(&plan.ipart, sizeof(float) * plan.ipart_floats);
    if (e == hipSuccess) e = hipMalloc(&plan.jpart, sizeof(float) * plan.jpart_floats);
    if (e == hipSuccess) e = hipMalloc(&plan.wt, sizeof(float) * std::max(1, h.ntiles));
    if (e == hipSuccess)
        e = hipMemcpyAsync(plan.items, h.items.data(), sizeof(SymItem) * h.items.size(),
                           hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(plan.row_start, h.row_start.data(), sizeof(int) * h.row_start.size(),
                           hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(plan.tile_t0, h.tile_t0.data(), sizeof(int) * h.tile_t0.size(),
                           hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(plan.tile_pbase, h.tile_pbase.data(),
                           sizeof(long long) * h.tile_pbase.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);  // host vectors end here
    if (e != hipSuccess) {
        *err = std::string("symmetric form buffers: ") + hipGetErrorString(e);
        sym_release(plan);
        return e == hipErrorOutOfMemory ? 2 : 3;
    }
    plan.ntiles = h.ntiles;
    return 0;
}

int sym_prepare(SymPlan& plan, int ntiles, int len, hipStream_t st, std::string* err) {
    if (ntiles == plan.ntiles && len == plan.len && plan.items) return 0;
    if (ntiles < 1 || ntiles > kSymMaxTiles) {
        *err = "symmetric form: tile count out of range";
        return 1;
    }
    const int t0 = 0;
    const int rc = sym_prepare_runs(plan, &t0, &ntiles, 1, len, st, err);
    if (rc == 0) plan.len = len;
    return rc;
}

void sym_release(SymPlan& plan) {
    (void)hipFree(plan.items);
    (void)hipFree(plan.row_start);
    (void)hipFree(plan.tile_t0);
    (void)hipFree(plan.tile_pbase);
    (void)hipFree(plan.ipart);
    (void)hipFree(plan.jpart);
    (void)hipFree(plan.wt);
    plan = SymPlan{};
}

template <int IPT>
void launch_sym_ipt(bool safe, const DeviceState& s, const float4* jp, const double* centre,
                    const int* tile_valid, int ntiles, const SymPlan& plan, float eps2, int blocks,
                    bool weights_ready, hipStream_t st) {
    const int n = static_cast<int>(s.n);
    const dim3 grid(static_cast<unsigned>(blocks > 0 ? std::min(blocks, plan.nitems) : plan.nitems));
    const unsigned lds = IPT == 4 ? kSymLdsPad : 0;
    if (safe) {
        hipLaunchKernelGGL((direct_sym_kernel<IPT, true>), grid, dim3(kTile / IPT), lds, st,
                           plan.items, plan.nitems, n, tile_valid, ntiles, s.x, s.y, s.z, jp, centre,
                           eps2, nullptr, plan.ipart, plan.jpart);
    } else {
        const float* wt = nullptr;
        if (IPT == 4 && plan.wt != nullptr && ntiles <= plan.ntiles) {
            if (!weights_ready)
                hipLaunchKernelGGL(sym_tile_weights_kernel, dim3((ntiles + 3) / 4), dim3(256), 0, st,
                                   jp, tile_valid, n, ntiles, eps2, plan.wt);
            wt = plan.wt;
        }
        hipLaunchKernelGGL((direct_sym_kernel<IPT, false>), grid, dim3(kTile / IPT), lds, st,
                           plan.items, plan.nitems, n, tile_valid, ntiles, s.x, s.y, s.z, jp, centre,
                           eps2, wt, plan.ipart, plan.jpart);
    }
}

void launch_direct_sym_tiles(const DeviceState& s, const float4* jp, const double* centre,
                             const int* tile_valid, int ntiles, const SymPlan& plan, double eps,
                             int ipt, int blocks, hipStream_t st, bool weights_ready) {
    const float eps2 = static_cast<float>(eps * eps);
    const bool safe = !(eps2 >= kTinySeparation2);
    const bool wr = weights_ready;
    switch (ipt) {
    case 4: launch_sym_ipt<4>(safe, s, jp, centre, tile_valid, ntiles, plan, eps2, blocks, wr, st); break;
    case 2: launch_sym_ipt<2>(safe, s, jp, centre, tile_valid, ntiles, plan, eps2, blocks, wr, st); break;
    default: launch_sym_ipt<1>(safe, s, jp, centre, tile_valid, ntiles, plan, eps2, blocks, wr, st); break;
    }
}

void launch_direct_sym(const DeviceState& s, const MixedLayout& L, const SymPlan& plan, double eps,
                       int ipt, hipStream_t st, bool weights_ready) {
    launch_direct_sym_tiles(s, L.jp, L.centre, nullptr, static_cast<int>(L.ntiles), plan, eps, ipt,
                            0, st, weights_ready);
}

void launch_sym_weights(const DeviceState& s, const MixedLayout& L, const SymPlan& plan, double eps,
                        hipStream_t st) {
    const int ntiles = static_cast<int>(L.ntiles);
    if (plan.wt == nullptr || ntiles > plan.ntiles || ntiles < 1) return;
    hipLaunchKernelGGL(sym_tile_weights_kernel, dim3((ntiles + 3) / 4), dim3(256), 0, st, L.jp,
                       nullptr, static_cast<int>(s.n), ntiles, static_cast<float>(eps * eps),
                       plan.wt);
}

void launch_merge_sym(const DeviceState& s, const SymPlan& plan, const int* tile_valid,
                      bool close_kick, double half_dt, hipStream_t st) {
    hipLaunchKernelGGL(merge_sym_kernel, dim3(plan.ntiles), dim3(kMergeLanes * kMergeGroups), 0, st,
                       static_cast<int>(s.n), plan.ntiles, plan.tile_t0, plan.row_start,
                       plan.tile_pbase, tile_valid, plan.ipart, plan.jpart, s.ax, s.ay, s.az, s.vx,
                       s.vy, s.vz, close_kick ? 1 : 0, half_dt);
}

void launch_vjp_sym(const DeviceState& s, const MixedLayout& L, const SymPlan& plan,
                    const float4* lpack, float* vipart, float* vjpart, double G, double eps,
                    int ipt, double* xbar, double* mbar, int ostride, bool accumulate,
                    hipStream_t st) {
    if (s.n == 0 || plan.nitems == 0) return;
    const float eps2 = static_cast<float>(eps * eps);
    const bool safe = !(eps2 >= kTinySeparation2);
    const int n = static_cast<int>(s.n), nt = plan.ntiles;
#define GC_VJP_SYM(IPT)                                                                          \
    do {                                                                                         \
        if (safe)                                                                                \
            hipLaunchKernelGGL((vjp_sym_kernel<IPT, true>), dim3(plan.nitems), dim3(kTile / IPT), \
                               0, st, plan.items, plan.nitems, n, nt, s.x, s.y, s.z, L.jp, lpack, \
                               L.centre, eps2, vipart, vjpart);                                  \
        else                                                                                     \
            hipLaunchKernelGGL((vjp_sym_kernel<IPT, false>), dim3(plan.nitems),                   \
                               dim3(kTile / IPT), 0, st, plan.items, plan.nitems, n, nt, s.x,    \
                               s.y, s.z, L.jp, lpack, L.centre, eps2, vipart, vjpart);           \
    } while (0)
    if (ipt == 4) GC_VJP_SYM(4);
    else if (ipt == 1) GC_VJP_SYM(1);
    else GC_VJP_SYM(2);
#undef GC_VJP_SYM
    hipLaunchKernelGGL(vjp_sym_merge_kernel, dim3(nt), dim3(kMergeLanes * kVjpMergeGroups), 0, st,
                       n, nt, plan.row_start, plan.tile_t0, plan.tile_pbase, vipart, vjpart, G,
                       xbar, mbar, ostride, accumulate ? 1 : 0);
}

void launch_merge_sym_open(const DeviceState& s, const SymPlan& plan, MixedLayout& L, double G,
                           double eps, double half_dt, double dt, hipStream_t st) {
    hipLaunchKernelGGL(merge_sym_open_kernel, dim3(plan.ntiles), dim3(kMergeLanes * kMergeGroups),
                       0, st, static_cast<int>(s.n), plan.ntiles, plan.tile_t0, plan.row_start,
                       plan.tile_pbase, plan.ipart, plan.jpart, s.x, s.y, s.z, s.vx, s.vy, s.vz,
                       s.ax, s.ay, s.az, s.m, half_dt, dt, G, static_cast<float>(eps * eps), L.jp,
                       L.centre, plan.wt);
}

}  // namespace gc
