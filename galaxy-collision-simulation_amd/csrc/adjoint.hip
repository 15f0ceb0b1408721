// adjoint.hip — vector-Jacobian product of the softened all-pairs force.
//
// With a_k = G sum_j m_j d_kj f(s_kj), d_kj = x_j - x_k, s = |d|^2 + eps^2,
// f = s^(-3/2), and J(d) = s^(-3/2) I - 3 s^(-5/2) d d^T (the derivative of
// d f(s), even in d), a cotangent lambda on the accelerations pulls back to
//
//   xbar_k = G sum_{j != k} J(d_kj) (m_k lambda_j - m_j lambda_k)
//   mbar_k = -G sum_{j != k} f(s_kj) lambda_j . d_kj
//
// Both are one all-pairs sweep per target with 7 source values
// (position, mass, lambda).  The pair term is odd under i <-> j, so the
// mixed VJP of a full-range state normally runs the pair-once form in
// direct_sym.hip (vjp_sym_kernel); the per-target form here serves target
// ranges, the per-target pair mode and states past the pair-once limit.  FP64 form: one target per lane, ascending j,
// sources staged in LDS.  Mixed form: sources as FP32 offsets from FP64 tile
// centres (the force kernel's packed tiles) plus FP32 lambda, FP32 pair
// terms, per-tile FP32 partials folded into FP64, source ranges split over
// grid.y and merged in split order (deterministic).
#include "gc_internal.hpp"

#include <algorithm>

namespace gc {
namespace {

struct alignas(64) SrcV64 {
    double x, y, z, m, lx, ly, lz, pad;
};

__global__ __launch_bounds__(kTile) void vjp_fp64_kernel(
    int n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ m, const double* __restrict__ lam,
    int lstride, double G, double eps2, double* __restrict__ xbar, double* __restrict__ mbar,
    int ostride, int accumulate) {
    __shared__ SrcV64 tile[kTile];
    const int t = threadIdx.x;
    const int i = blockIdx.x * kTile + t;
    const bool active = i < n;
    const double xi = active ? x[i] : 0.0, yi = active ? y[i] : 0.0, zi = active ? z[i] : 0.0;
    const double mi = active ? m[i] : 0.0;
    const double li0 = active ? lam[i] : 0.0, li1 = active ? lam[lstride + i] : 0.0,
                 li2 = active ? lam[2 * lstride + i] : 0.0;
    double g0 = 0.0, g1 = 0.0, g2 = 0.0, gm = 0.0;
    for (int j0 = 0; j0 < n; j0 += kTile) {
        __syncthreads();
        const int jl = j0 + t;
        if (jl < n) {
            SrcV64 s;
            s.x = x[jl];
            s.y = y[jl];
            s.z = z[jl];
            s.m = m[jl];
            s.lx = lam[jl];
            s.ly = lam[lstride + jl];
            s.lz = lam[2 * lstride + jl];
            s.pad = 0.0;
            tile[t] = s;
        }
        __syncthreads();
        const int cnt = min(kTile, n - j0);
        for (int jj = 0; jj < cnt; ++jj) {
            const SrcV64 s = tile[jj];
            if (j0 + jj == i) continue;
            const double dx = s.x - xi, dy = s.y - yi, dz = s.z - zi;
            const double s2 = dx * dx + dy * dy + dz * dz + eps2;
            if (s2 == 0.0) continue;
            const double r = 1.0 / __builtin_sqrt(s2);
            const double r3 = r * r * r;
            const double r5 = r3 * r * r;
            const double w0 = mi * s.lx - s.m * li0;
            const double w1 = mi * s.ly - s.m * li1;
            const double w2 = mi * s.lz - s.m * li2;
            const double dw = dx * w0 + dy * w1 + dz * w2;
            const double c = 3.0 * r5 * dw;
            g0 += r3 * w0 - c * dx;
            g1 += r3 * w1 - c * dy;
            g2 += r3 * w2 - c * dz;
            gm -= r3 * (s.lx * dx + s.ly * dy + s.lz * dz);
        }
    }
    if (!active) return;
    if (accumulate) {
        xbar[i] += G * g0;
        xbar[ostride + i] += G * g1;
        xbar[2 * ostride + i] += G * g2;
        mbar[i] += G * gm;
    } else {
        xbar[i] = G * g0;
        xbar[ostride + i] = G * g1;
        xbar[2 * ostride + i] = G * g2;
        mbar[i] = G * gm;
    }
}

// lambda (FP64, SoA with stride) -> FP32 float4 per particle, zero padding
__global__ __launch_bounds__(256) void pack_lambda_kernel(int n, int padded,
                                                          const double* __restrict__ lam,
                                                          int lstride, float4* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= padded) return;
    out[i] = i < n ? make_float4(static_cast<float>(lam[i]), static_cast<float>(lam[lstride + i]),
                                 static_cast<float>(lam[2 * lstride + i]), 0.f)
                   : make_float4(0.f, 0.f, 0.f, 0.f);
}

// Mixed VJP.  jp = (x - c, y - c, z - c, G m) per particle (pack_mixed_kernel),
// so w = Gm_k lambda_j - Gm_j lambda_k already carries G; the mass gradient
// is multiplied by G at the merge.
template <bool SAFE>
__global__ __launch_bounds__(kTile) void vjp_mixed_kernel(
    int n, int ntiles, int tiles_per_split, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, const float4* __restrict__ jp,
    const float4* __restrict__ lp, const double* __restrict__ centre, float eps2,
    double* __restrict__ part, int cap) {
    __shared__ float4 tj[kTile];
    __shared__ float4 tl[kTile];
    const int t = threadIdx.x;
    const int split = blockIdx.y;
    const int t0 = split * tiles_per_split;
    const int t1 = min(t0 + tiles_per_split, ntiles);
    const int i = blockIdx.x * kTile + t;
    const bool active = i < n;
    const double xi = active ? x[i] : 0.0, yi = active ? y[i] : 0.0, zi = active ? z[i] : 0.0;
    const float4 me = active ? jp[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 ml = active ? lp[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    double g0 = 0.0, g1 = 0.0, g2 = 0.0, gm = 0.0;
    for (int tile = t0; tile < t1; ++tile) {
        __syncthreads();
        tj[t] = jp[static_cast<size_t>(tile) * kTile + t];
        tl[t] = lp[static_cast<size_t>(tile) * kTile + t];
        __syncthreads();
        const double cx = centre[tile], cy = centre[ntiles + tile], cz = centre[2 * ntiles + tile];
        const float tx = static_cast<float>(xi - cx), ty = static_cast<float>(yi - cy),
                    tz = static_cast<float>(zi - cz);
        float p0 = 0.f, p1 = 0.f, p2 = 0.f, pm = 0.f;
#pragma unroll 4
        for (int jj = 0; jj < kTile; ++jj) {
            const float4 s = tj[jj];
            const float4 l = tl[jj];
            const float dx = s.x - tx, dy = s.y - ty, dz = s.z - tz;
            const float s2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, eps2)));
            float r;
            if (SAFE) {
                r = s2 > 1e-30f ? __builtin_amdgcn_rsqf(s2) : 0.f;
            } else {
                r = __builtin_amdgcn_rsqf(s2);
            }
            const float r2 = r * r;
            const float r3 = r2 * r;
            const float w0 = __builtin_fmaf(me.w, l.x, -s.w * ml.x);
            const float w1 = __builtin_fmaf(me.w, l.y, -s.w * ml.y);
            const float w2 = __builtin_fmaf(me.w, l.z, -s.w * ml.z);
            const float dw = __builtin_fmaf(dz, w2, __builtin_fmaf(dy, w1, dx * w0));
            const float c = 3.f * r3 * r2 * dw;
            p0 = __builtin_fmaf(r3, w0, __builtin_fmaf(-c, dx, p0));
            p1 = __builtin_fmaf(r3, w1, __builtin_fmaf(-c, dy, p1));
            p2 = __builtin_fmaf(r3, w2, __builtin_fmaf(-c, dz, p2));
            pm = __builtin_fmaf(-r3, __builtin_fmaf(l.z, dz, __builtin_fmaf(l.y, dy, l.x * dx)), pm);
        }
        g0 += static_cast<double>(p0);
        g1 += static_cast<double>(p1);
        g2 += static_cast<double>(p2);
        gm += static_cast<double>(pm);
    }
    if (!active) return;
    double* b = part + static_cast<size_t>(split) * 4 * cap;
    b[i] = g0;
    b[cap + i] = g1;
    b[2 * static_cast<size_t>(cap) + i] = g2;
    b[3 * static_cast<size_t>(cap) + i] = gm;
}

__global__ __launch_bounds__(256) void vjp_merge_kernel(int n, int jsplit, int cap,
                                                        const double* __restrict__ part, double G,
                                                        double* __restrict__ xbar,
                                                        double* __restrict__ mbar, int ostride,
                                                        int accumulate) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double g[4] = {0.0, 0.0, 0.0, 0.0};
    for (int s = 0; s < jsplit; ++s)
        for (int c = 0; c < 4; ++c) g[c] += part[(static_cast<size_t>(s) * 4 + c) * cap + i];
    const double gm = G * g[3];
    if (accumulate) {
        xbar[i] += g[0];
        xbar[ostride + i] += g[1];
        xbar[2 * ostride + i] += g[2];
        mbar[i] += gm;
    } else {
        xbar[i] = g[0];
        xbar[ostride + i] = g[1];
        xbar[2 * ostride + i] = g[2];
        mbar[i] = gm;
    }
}

}  // namespace

void launch_vjp_fp64(const DeviceState& s, double G, double eps, const double* lam, int lstride,
                     double* xbar, double* mbar, int ostride, bool accumulate, hipStream_t st) {
    if (s.n == 0) return;
    const unsigned blocks = static_cast<unsigned>((s.n + kTile - 1) / kTile);
    hipLaunchKernelGGL(vjp_fp64_kernel, dim3(blocks), dim3(kTile), 0, st, static_cast<int>(s.n),
                       s.x, s.y, s.z, s.m, lam, lstride, G, eps * eps, xbar, mbar, ostride,
                       accumulate ? 1 : 0);
}

void launch_pack_lambda(size_t n, size_t padded, const double* lam, int lstride, float4* out,
                        hipStream_t st) {
    if (padded == 0) return;
    hipLaunchKernelGGL(pack_lambda_kernel, dim3(static_cast<unsigned>((padded + 255) / 256)),
                       dim3(256), 0, st, static_cast<int>(n), static_cast<int>(padded), lam,
                       lstride, out);
}

void launch_vjp_mixed(const DeviceState& s, MixedLayout& L, float4* lpack, double* part,
                      int jsplit, double G, double eps, const double* lam, int lstride,
                      double* xbar, double* mbar, int ostride, bool accumulate, hipStream_t st) {
    if (s.n == 0) return;
    const int n = static_cast<int>(s.n);
    const int ntiles = static_cast<int>(L.ntiles);
    const int padded = ntiles * kTile;
    launch_pack_mixed(s, L, G, st);
    hipLaunchKernelGGL(pack_lambda_kernel, dim3((padded + 255) / 256), dim3(256), 0, st, n, padded,
                       lam, lstride, lpack);
    const int tps = (ntiles + jsplit - 1) / jsplit;
    const int js = (ntiles + tps - 1) / tps;
    const float eps2 = static_cast<float>(eps * eps);
    const dim3 grid(static_cast<unsigned>(ntiles), static_cast<unsigned>(js));
    if (!(eps2 >= 1e-30f)) {
        hipLaunchKernelGGL(vjp_mixed_kernel<true>, grid, dim3(kTile), 0, st, n, ntiles, tps, s.x,
                           s.y, s.z, L.jp, lpack, L.centre, eps2, part, static_cast<int>(s.cap));
    } else {
        hipLaunchKernelGGL(vjp_mixed_kernel<false>, grid, dim3(kTile), 0, st, n, ntiles, tps, s.x,
                           s.y, s.z, L.jp, lpack, L.centre, eps2, part, static_cast<int>(s.cap));
    }
    hipLaunchKernelGGL(vjp_merge_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, js,
                       static_cast<int>(s.cap), part, G, xbar, mbar, ostride, accumulate ? 1 : 0);
}

int vjp_mixed_splits(size_t n) {
    const size_t ntiles = (n + kTile - 1) / kTile;
    size_t js = (2048 + ntiles - 1) / std::max<size_t>(ntiles, 1);
    js = std::max<size_t>(1, std::min<size_t>({js, 64, ntiles}));
    const size_t tps = (ntiles + js - 1) / js;
    return static_cast<int>((ntiles + tps - 1) / tps);
}

}  // namespace gc
