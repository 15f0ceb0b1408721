// mixed_tiles.hpp — the mixed-precision source-tile sweep shared by the
// single-run kernel (direct_mixed.hip) and the batched persistent kernel
// (batch.hip), so a run gives bit-identical partial sums in both.
#pragma once

#include "gc_internal.hpp"

namespace gc {

constexpr float kTinySeparation2 = 1e-30f;

__device__ __forceinline__ void two_sum_acc(float& hi, float& lo, float v) {
    const float s = hi + v;
    const float bb = s - hi;
    const float err = (hi - (s - bb)) + (v - bb);
    hi = s;
    lo += err;
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
        const double cx = centre[tl], cy = centre[cstride + tl], cz = centre[2 * cstride + tl];
        float tx[IPT], ty[IPT], tz[IPT], px[IPT], py[IPT], pz[IPT];
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            tx[k] = static_cast<float>(xi[k] - cx);
            ty[k] = static_cast<float>(yi[k] - cy);
            tz[k] = static_cast<float>(zi[k] - cz);
            px[k] = py[k] = pz[k] = 0.f;
        }
#pragma unroll 8
        for (int jj = 0; jj < kTile; ++jj) {
            const float4 s = tile[jj];
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
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            two_sum_acc(hx[k], lx[k], px[k]);
            two_sum_acc(hy[k], ly[k], py[k]);
            two_sum_acc(hz[k], lz[k], pz[k]);
        }
    }
}

}  // namespace gc
