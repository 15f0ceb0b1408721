// direct_mixed.hip — mixed-precision softened all-pairs forces on gfx950.
//
// Layout.  Sources are grouped in tiles of kTile consecutive particles.  Each
// tile has an FP64 centre c (midpoint of its bounding box) and each particle
// is stored as the FP32 offset o = x - c plus its FP32 G*m (16 B, one
// float4).  For a target at FP64 position x_i the kernel forms, once per
// source tile, t = float(x_i - c) in FP64-then-round, so the pair separation
// o_j - t is an FP32 difference of two quantities bounded by the tile extent
// and the target's distance to the tile, not by |x| itself.
//
// Arithmetic per directed pair (FP32, 12 VALU + 1 v_rsq_f32):
//   d = o_j - t;  s2 = fma(dz,dz, fma(dy,dy, fma(dx,dx, eps^2)))
//   r = rsq(s2);  k = (G m_j * r) * (r * r);  p += d * k   (3 FMA)
// The pair sum over one tile is an FP32 partial; after the tile it is folded
// into an (hi, lo) FP32 accumulator by Knuth's two-sum, so compensation costs
// 18 VALU per kTile pairs.  Sources are split into `jsplit` contiguous tile
// ranges (grid.y) so that a launch holds enough waves to fill 256 CUs; the
// merge kernel adds the per-split (hi + lo) in FP64 in split order, so the
// result is deterministic.
//
// The j == i pair needs no test: both operands are the same rounded offset,
// d = 0 and the pair contributes 0 when eps^2 is a normal FP32 number.  With
// eps^2 below kTinySeparation2 (eps == 0 or denormal-small) the SAFE variant
// skips pairs with s2 <= kTinySeparation2, the mixed-precision form of the
// reference's s2 == 0 skip (cpp/direct/direct_solver.cpp:17-19); without it
// r^3 would overflow and 0 * inf would poison the sum.
#include "gc_internal.hpp"
#include "mixed_tiles.hpp"

#include <algorithm>
#include <cmath>

namespace gc {
namespace {


// One block per source tile: FP64 bounding-box centre, then FP32 offsets.
__global__ __launch_bounds__(kTile) void pack_mixed_kernel(
    int n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ m, double G, int ntiles,
    float4* __restrict__ jp, double* __restrict__ centre) {
    __shared__ double red[6][kTile / 64];
    const int t = threadIdx.x;
    const int i = blockIdx.x * kTile + t;
    const bool valid = i < n;
    const double px = valid ? x[i] : 0.0, py = valid ? y[i] : 0.0, pz = valid ? z[i] : 0.0;
    double lo0 = valid ? px : INFINITY, lo1 = valid ? py : INFINITY, lo2 = valid ? pz : INFINITY;
    double hi0 = valid ? px : -INFINITY, hi1 = valid ? py : -INFINITY,
           hi2 = valid ? pz : -INFINITY;
    lo0 = wave_min(lo0);
    lo1 = wave_min(lo1);
    lo2 = wave_min(lo2);
    hi0 = wave_max(hi0);
    hi1 = wave_max(hi1);
    hi2 = wave_max(hi2);
    const int w = t >> 6;
    if ((t & 63) == 0) {
        red[0][w] = lo0;
        red[1][w] = lo1;
        red[2][w] = lo2;
        red[3][w] = hi0;
        red[4][w] = hi1;
        red[5][w] = hi2;
    }
    __syncthreads();
    double c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double lo = red[d][0], hi = red[3 + d][0];
#pragma unroll
        for (int k = 1; k < kTile / 64; ++k) {
            lo = fmin(lo, red[d][k]);
            hi = fmax(hi, red[3 + d][k]);
        }
        c[d] = (lo + hi) * 0.5;
    }
    float4 o;
    if (valid) {
        o.x = static_cast<float>(px - c[0]);
        o.y = static_cast<float>(py - c[1]);
        o.z = static_cast<float>(pz - c[2]);
        o.w = static_cast<float>(G * m[i]);
    } else {
        o = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    jp[i] = o;
    if (t == 0) {
        centre[blockIdx.x] = c[0];
        centre[ntiles + blockIdx.x] = c[1];
        centre[2 * ntiles + blockIdx.x] = c[2];
    }
}

template <int IPT, bool SAFE, bool STREAM>
__global__ __launch_bounds__(kTile) void direct_mixed_kernel(
    int begin, int end, int ntiles, int tiles_per_split, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, const float4* __restrict__ jp,
    const double* __restrict__ centre, float eps2, float* __restrict__ part, int cap) {
    __shared__ float4 tile[kTile];
    const int t = threadIdx.x;
    const int split = blockIdx.y;
    const int t0 = split * tiles_per_split;
    const int t1 = min(t0 + tiles_per_split, ntiles);

    int idx[IPT];
    double xi[IPT], yi[IPT], zi[IPT];
    float hx[IPT], hy[IPT], hz[IPT], lx[IPT], ly[IPT], lz[IPT];
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        idx[k] = begin + (blockIdx.x * IPT + k) * kTile + t;
        const bool a = idx[k] < end;
        xi[k] = a ? x[idx[k]] : 0.0;
        yi[k] = a ? y[idx[k]] : 0.0;
        zi[k] = a ? z[idx[k]] : 0.0;
        hx[k] = hy[k] = hz[k] = lx[k] = ly[k] = lz[k] = 0.f;
    }

    if (STREAM) {
        tile_stream<IPT, SAFE>(t0, t1, jp, centre, ntiles, eps2, xi, yi, zi, hx, hy, hz, lx, ly, lz);
    } else {
        tile_sweep<IPT, SAFE>(tile, t0, t1, jp, centre, ntiles, eps2, xi, yi, zi, hx, hy, hz, lx,
                              ly, lz);
    }
    float* base = part + static_cast<size_t>(split) * 6 * cap;
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        if (idx[k] >= end) continue;
        base[idx[k]] = hx[k];
        base[cap + idx[k]] = hy[k];
        base[2 * static_cast<size_t>(cap) + idx[k]] = hz[k];
        base[3 * static_cast<size_t>(cap) + idx[k]] = lx[k];
        base[4 * static_cast<size_t>(cap) + idx[k]] = ly[k];
        base[5 * static_cast<size_t>(cap) + idx[k]] = lz[k];
    }
}

__global__ __launch_bounds__(256) void merge_mixed_kernel(
    int begin, int end, int jsplit, int cap, const float* __restrict__ part,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az,
    double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz, int close_kick,
    double half_dt) {
#pragma clang fp contract(off)
    const int i = begin + blockIdx.x * 256 + threadIdx.x;
    if (i >= end) return;
    double a[3] = {0.0, 0.0, 0.0};
    for (int s = 0; s < jsplit; ++s) {
        const float* b = part + static_cast<size_t>(s) * 6 * cap;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            a[c] += static_cast<double>(b[static_cast<size_t>(c) * cap + i]) +
                    static_cast<double>(b[static_cast<size_t>(c + 3) * cap + i]);
    }
    ax[i] = a[0];
    ay[i] = a[1];
    az[i] = a[2];
    if (close_kick) {
        vx[i] = vx[i] + a[0] * half_dt;
        vy[i] = vy[i] + a[1] * half_dt;
        vz[i] = vz[i] + a[2] * half_dt;
    }
}

template <int IPT, bool STREAM>
void launch_ipt(bool safe, dim3 grid, hipStream_t st, int begin, int end, int ntiles, int tps,
                const DeviceState& s, const MixedLayout& L, float eps2, int cap) {
    if (safe) {
        hipLaunchKernelGGL((direct_mixed_kernel<IPT, true, STREAM>), grid, dim3(kTile), 0, st,
                           begin, end, ntiles, tps, s.x, s.y, s.z, L.jp, L.centre, eps2, L.part,
                           cap);
    } else {
        hipLaunchKernelGGL((direct_mixed_kernel<IPT, false, STREAM>), grid, dim3(kTile), 0, st,
                           begin, end, ntiles, tps, s.x, s.y, s.z, L.jp, L.centre, eps2, L.part,
                           cap);
    }
}

}  // namespace

void launch_pack_mixed(const DeviceState& s, MixedLayout& L, double G, hipStream_t st) {
    hipLaunchKernelGGL(pack_mixed_kernel, dim3(static_cast<unsigned>(L.ntiles)), dim3(kTile), 0,
                       st, static_cast<int>(s.n), s.x, s.y, s.z, s.m, G,
                       static_cast<int>(L.ntiles), L.jp, L.centre);
}

// jsplit here is the effective split count (ceil(ntiles / tiles_per_split)).
void launch_direct_mixed(const DeviceState& s, MixedLayout& L, double eps, int ipt, int jsplit,
                         bool stream, size_t begin, size_t end, hipStream_t st) {
    if (end <= begin) return;
    const int ntiles = static_cast<int>(L.ntiles);
    const int tps = (ntiles + jsplit - 1) / jsplit;
    const float eps2 = static_cast<float>(eps * eps);
    const bool safe = !(eps2 >= kTinySeparation2);
    const size_t span = static_cast<size_t>(kTile) * ipt;
    dim3 grid(static_cast<unsigned>((end - begin + span - 1) / span), static_cast<unsigned>(jsplit));
    const int b = static_cast<int>(begin), e = static_cast<int>(end), cap = static_cast<int>(s.cap);
    if (stream) {
        switch (ipt) {
        case 4: launch_ipt<4, true>(safe, grid, st, b, e, ntiles, tps, s, L, eps2, cap); break;
        case 2: launch_ipt<2, true>(safe, grid, st, b, e, ntiles, tps, s, L, eps2, cap); break;
        default: launch_ipt<1, true>(safe, grid, st, b, e, ntiles, tps, s, L, eps2, cap); break;
        }
    } else {
        switch (ipt) {
        case 4: launch_ipt<4, false>(safe, grid, st, b, e, ntiles, tps, s, L, eps2, cap); break;
        case 2: launch_ipt<2, false>(safe, grid, st, b, e, ntiles, tps, s, L, eps2, cap); break;
        default: launch_ipt<1, false>(safe, grid, st, b, e, ntiles, tps, s, L, eps2, cap); break;
        }
    }
}

void launch_merge_mixed(const DeviceState& s, const MixedLayout& L, int jsplit, size_t begin,
                        size_t end, bool close_kick, double half_dt, hipStream_t st) {
    if (end <= begin) return;
    const unsigned blocks = static_cast<unsigned>((end - begin + 255) / 256);
    hipLaunchKernelGGL(merge_mixed_kernel, dim3(blocks), dim3(256), 0, st, static_cast<int>(begin),
                       static_cast<int>(end), jsplit, static_cast<int>(s.cap), L.part, s.ax, s.ay,
                       s.az, s.vx, s.vy, s.vz, close_kick ? 1 : 0, half_dt);
}

}  // namespace gc
