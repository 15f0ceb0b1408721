// mixed_tiles.hpp — the mixed-precision source-tile sweep shared by the
// single-run kernel (direct_mixed.hip) and the batched persistent kernel
// (batch.hip), so a run gives bit-identical partial sums in both.
#pragma once

#include "gc_internal.hpp"

namespace gc {

constexpr float kTinySeparation2 = 1e-30f;

// Wavefront minimum / maximum (butterfly over the 64 lanes).
template <typename T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ void two_sum_acc(float& hi, float& lo, float v) {
    const float s = hi + v;
    const float bb = s - hi;
    const float err = (hi - (s - bb)) + (v - bb);
    hi = s;
    lo += err;
}

// One source s (offset from the tile centre, G*m) on IPT targets.
template <int IPT, bool SAFE>
__device__ __forceinline__ void pair_step(float4 s, float eps2, const float* tx, const float* ty,
                                          const float* tz, float* px, float* py, float* pz) {
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const float dx = s.x - tx[k];
        const float dy = s.y - ty[k];
        const float dz = s.z - tz[k];
        const float s2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, eps2)));
        float r;
        if (SAFE) {
            r = s2 > kTinySeparation2 ? __builtin_amdgcn_rsqf(s2) : 0.f;
        } else {
            r = __builtin_amdgcn_rsqf(s2);
        }
        const float kk = (s.w * r) * (r * r);
        px[k] = __builtin_fmaf(dx, kk, px[k]);
        py[k] = __builtin_fmaf(dy, kk, py[k]);
        pz[k] = __builtin_fmaf(dz, kk, pz[k]);
    }
}

// FP32 pair sums of IPT targets (offsets t from the tile centre) over one
// LDS source tile, in ascending source order.
template <int IPT, bool SAFE>
__device__ __forceinline__ void tile_pairs(const float4* src, float eps2, const float* tx,
                                           const float* ty, const float* tz, float* px,
                                           float* py, float* pz) {
#pragma unroll 8
    for (int jj = 0; jj < kTile; ++jj) pair_step<IPT, SAFE>(src[jj], eps2, tx, ty, tz, px, py, pz);
}

// Target offsets from source tile tl's FP64 centre, rounded once to FP32.
template <int IPT>
__device__ __forceinline__ void tile_offsets(const double* __restrict__ centre, int cstride,
                                             int tl, const double* xi, const double* yi,
                                             const double* zi, float* tx, float* ty, float* tz) {
    const double cx = centre[tl], cy = centre[cstride + tl], cz = centre[2 * cstride + tl];
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        tx[k] = static_cast<float>(xi[k] - cx);
        ty[k] = static_cast<float>(yi[k] - cy);
        tz[k] = static_cast<float>(zi[k] - cz);
    }
}

template <int IPT>
__device__ __forceinline__ void fold_tile(const float* px, const float* py, const float* pz,
                                          float* hx, float* hy, float* hz, float* lx, float* ly,
                                          float* lz) {
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        two_sum_acc(hx[k], lx[k], px[k]);
        two_sum_acc(hy[k], ly[k], py[k]);
        two_sum_acc(hz[k], lz[k], pz[k]);
    }
}

// Sweeps source tiles [t0, t1) for IPT targets per lane.  Each tile is
// staged in `tile` (LDS, kTile float4) with a register prefetch of the next;
// target offsets from the tile's FP64 centre are formed once per tile; the
// tile's FP32 pair sum is folded into (h, l) by two-sum.  The whole block
// must call it (it synchronises).
template <int IPT, bool SAFE>
__device__ __forceinline__ void tile_sweep(float4* tile, int t0, int t1,
                                           const float4* __restrict__ jp,
                                           const double* __restrict__ centre, int cstride,
                                           float eps2, const double* xi, const double* yi,
                                           const double* zi, float* hx, float* hy, float* hz,
                                           float* lx, float* ly, float* lz) {
    const int t = threadIdx.x;
    float4 next = t0 < t1 ? jp[static_cast<size_t>(t0) * kTile + t] : make_float4(0.f, 0.f, 0.f, 0.f);
    for (int tl = t0; tl < t1; ++tl) {
        __syncthreads();
        tile[t] = next;
        __syncthreads();
        if (tl + 1 < t1) next = jp[static_cast<size_t>(tl + 1) * kTile + t];
        float tx[IPT], ty[IPT], tz[IPT], px[IPT], py[IPT], pz[IPT];
        tile_offsets<IPT>(centre, cstride, tl, xi, yi, zi, tx, ty, tz);
#pragma unroll
        for (int k = 0; k < IPT; ++k) px[k] = py[k] = pz[k] = 0.f;
        tile_pairs<IPT, SAFE>(tile, eps2, tx, ty, tz, px, py, pz);
        fold_tile<IPT>(px, py, pz, hx, hy, hz, lx, ly, lz);
    }
}

// The same sweep with sources read straight from global memory at
// wave-uniform addresses: they arrive by scalar loads into SGPRs and enter
// the VALU ops as their scalar operand, so waves need no LDS and no barrier
// and run independently.  Loads run one batch of kStreamBatch sources ahead
// of the arithmetic (across tile ends too), so `jp` must hold kStreamBatch
// readable entries past the last tile.  Same operations in the same order
// as tile_sweep.
template <int IPT, bool SAFE>
__device__ __forceinline__ void tile_stream(int t0, int t1, const float4* __restrict__ jp,
                                            const double* __restrict__ centre, int cstride,
                                            float eps2, const double* xi, const double* yi,
                                            const double* zi, float* hx, float* hy, float* hz,
                                            float* lx, float* ly, float* lz) {
    constexpr int B = kStreamBatch;
    const float4* p = jp + static_cast<size_t>(t0) * kTile;
    float4 cur[B];
#pragma unroll
    for (int q = 0; q < B; ++q) cur[q] = p[q];
    for (int tl = t0; tl < t1; ++tl) {
        float tx[IPT], ty[IPT], tz[IPT], px[IPT], py[IPT], pz[IPT];
        tile_offsets<IPT>(centre, cstride, tl, xi, yi, zi, tx, ty, tz);
#pragma unroll
        for (int k = 0; k < IPT; ++k) px[k] = py[k] = pz[k] = 0.f;
        for (int jj = 0; jj < kTile; jj += B) {
            p += B;
            float4 nxt[B];
#pragma unroll
            for (int q = 0; q < B; ++q) nxt[q] = p[q];
#pragma unroll
            for (int q = 0; q < B; ++q) pair_step<IPT, SAFE>(cur[q], eps2, tx, ty, tz, px, py, pz);
#pragma unroll
            for (int q = 0; q < B; ++q) cur[q] = nxt[q];
        }
        fold_tile<IPT>(px, py, pz, hx, hy, hz, lx, ly, lz);
    }
}

}  // namespace gc
