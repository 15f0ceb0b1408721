// direct_fp64.hip — FP64 softened all-pairs forces on gfx950 ("the FP64 GPU
// path").
//
// One target per lane, sources streamed through LDS in kTile-particle tiles
// read with coalesced SoA loads.  At N = 65,536 that is one wave per SIMD,
// so the instruction-level parallelism comes from evaluating 8 sources side
// by side before adding their terms in order.  Each lane sums its sources in ascending j
// with the arithmetic of compute_direct_accelerations_for_targets
// (cpp/direct/direct_solver.cpp:8-23,59-83): r = x_j - x_i,
// s2 = ((rx*rx + ry*ry) + rz*rz) + eps^2, skip j == i and s2 == 0,
// inv_r = 1/sqrt(s2) (IEEE), k = (G*m_j) * ((inv_r*inv_r)*inv_r),
// a += r*k, with FMA contraction off — so the result is bit-identical to the
// reference's FP64 targets form.  The optional epilogue applies the closing
// half kick v += a*dt/2 (cpp/core/integrator.cpp:7-11) while a is still in
// registers.
#include "gc_internal.hpp"

#include <algorithm>

namespace gc {
namespace {

// Sources per step of a lane's sweep (kTile is a multiple).
constexpr int kUnroll = 8;
static_assert(kTile % kUnroll == 0, "tile holds whole steps");

__global__ __launch_bounds__(kTile) void direct_fp64_kernel(
    int n, int begin, int end, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ m, double G, double eps2,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az,
    double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz, int close_kick,
    double half_dt) {
#pragma clang fp contract(off)
    __shared__ Src64 tile[kTile];
    const int t = threadIdx.x;
    const int i = begin + blockIdx.x * kTile + t;
    const bool active = i < end;
    const double xi = active ? x[i] : 0.0;
    const double yi = active ? y[i] : 0.0;
    const double zi = active ? z[i] : 0.0;
    double sx = 0.0, sy = 0.0, sz = 0.0;

    for (int j0 = 0; j0 < n; j0 += kTile) {
        __syncthreads();
        const int jl = j0 + t;
        if (jl < n) {
            Src64 s;
            s.x = x[jl];
            s.y = y[jl];
            s.z = z[jl];
            s.gm = G * m[jl];
            tile[t] = s;
        }
        __syncthreads();
        const int cnt = min(kTile, n - j0);
        // kUnroll independent sources per step: their terms are computed side
        // by side and then added in ascending j.  A skipped source (j == i,
        // s2 == 0, or past the tile's end) contributes an exact +0.0, which
        // leaves the sum's bits unchanged: the sum starts at +0.0 and so is
        // never -0.0.
        for (int jj = 0; jj < cnt; jj += kUnroll) {
            double tx[kUnroll], ty[kUnroll], tz[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                const Src64 s = tile[jj + u];
                const double rx = s.x - xi;
                const double ry = s.y - yi;
                const double rz = s.z - zi;
                const double s2 = (rx * rx + ry * ry + rz * rz) + eps2;
                const double inv_r = 1.0 / __builtin_sqrt(s2);
                const double inv_r3 = inv_r * inv_r * inv_r;
                const double k = s.gm * inv_r3;
                const bool skip = jj + u >= cnt || j0 + jj + u == i || s2 == 0.0;
                tx[u] = skip ? 0.0 : rx * k;
                ty[u] = skip ? 0.0 : ry * k;
                tz[u] = skip ? 0.0 : rz * k;
            }
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                sx += tx[u];
                sy += ty[u];
                sz += tz[u];
            }
        }
    }
    if (!active) return;
    ax[i] = sx;
    ay[i] = sy;
    az[i] = sz;
    if (close_kick) {
        vx[i] = vx[i] + sx * half_dt;
        vy[i] = vy[i] + sy * half_dt;
        vz[i] = vz[i] + sz * half_dt;
    }
}

// ---- throughput form (precision GC_PRECISION_FP64_FAST) -----------------
// The same pair force in FP64 without the bit-exact constraints: IPT targets
// per lane, sources split into ranges (grid.y) so the launch fills every
// SIMD, sources read by wave-uniform scalar loads (SGPR operands of the FP64
// VALU ops, no LDS, no barriers), and 1/sqrt(s2) from v_rsq_f64 refined by
// one cubic-convergent correction y <- y (1 + h/2 + 3h^2/8), h = 1 - s2 y^2.
// Per pair: 3 sub, 3 FMA for s2, rsq + 5 for the correction, 2 for r^-3, 1 for
// G m r^-3, 3 FMA into the sum.  Each split's sums are FP64 partials added
// in split order by merge_fp64_kernel, so results are deterministic; they
// differ from the ascending-j form by rounding only.  The source array holds
// zero-mass padding past n (a multiple of kUnroll, plus kUnroll read-ahead
// entries), so a split covers whole batches of kUnroll sources.
__global__ __launch_bounds__(256) void pack_fp64_kernel(int n, int padded,
                                                        const double* __restrict__ x,
                                                        const double* __restrict__ y,
                                                        const double* __restrict__ z,
                                                        const double* __restrict__ m, double G,
                                                        Src64* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= padded) return;
    Src64 s{0.0, 0.0, 0.0, 0.0};
    if (i < n) {
        s.x = x[i];
        s.y = y[i];
        s.z = z[i];
        s.gm = G * m[i];
    }
    out[i] = s;
}

template <int IPT, bool SAFE>
__global__ __launch_bounds__(256) void direct_fp64_fast_kernel(
    int begin, int end, int per_split, int nsrc, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, const Src64* __restrict__ jp,
    double eps2, double* __restrict__ part, int cap) {
    const int t = threadIdx.x;
    const int j0 = blockIdx.y * per_split;
    const int j1 = min(j0 + per_split, nsrc);  // nsrc: padded to kUnroll
    int idx[IPT];
    double xi[IPT], yi[IPT], zi[IPT], sx[IPT], sy[IPT], sz[IPT];
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        idx[k] = begin + (blockIdx.x * IPT + k) * 256 + t;
        const bool a = idx[k] < end;
        xi[k] = a ? x[idx[k]] : 0.0;
        yi[k] = a ? y[idx[k]] : 0.0;
        zi[k] = a ? z[idx[k]] : 0.0;
        sx[k] = sy[k] = sz[k] = 0.0;
    }
    const Src64* p = jp + j0;
    Src64 cur[kUnroll];
#pragma unroll
    for (int q = 0; q < kUnroll; ++q) cur[q] = p[q];
    for (int j = j0; j < j1; j += kUnroll) {
        p += kUnroll;
        Src64 nxt[kUnroll];
#pragma unroll
        for (int q = 0; q < kUnroll; ++q) nxt[q] = p[q];
#pragma unroll
        for (int q = 0; q < kUnroll; ++q) {
#pragma unroll
            for (int k = 0; k < IPT; ++k) {
                const double rx = cur[q].x - xi[k];
                const double ry = cur[q].y - yi[k];
                const double rz = cur[q].z - zi[k];
                const double s2 = __builtin_fma(rz, rz, __builtin_fma(ry, ry, __builtin_fma(rx, rx, eps2)));
                double r = __builtin_amdgcn_rsq(s2);
                const double hr = s2 * r;
                const double h = __builtin_fma(-hr, r, 1.0);
                const double g = __builtin_fma(0.375, h, 0.5) * h;
                r = __builtin_fma(r, g, r);
                double kk = cur[q].gm * ((r * r) * r);
                if (SAFE) kk = s2 > 0.0 ? kk : 0.0;  // the reference skips s2 == 0
                sx[k] = __builtin_fma(rx, kk, sx[k]);
                sy[k] = __builtin_fma(ry, kk, sy[k]);
                sz[k] = __builtin_fma(rz, kk, sz[k]);
            }
        }
#pragma unroll
        for (int q = 0; q < kUnroll; ++q) cur[q] = nxt[q];
    }
    double* base = part + static_cast<size_t>(blockIdx.y) * 3 * cap;
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        if (idx[k] >= end) continue;
        base[idx[k]] = sx[k];
        base[cap + idx[k]] = sy[k];
        base[2 * static_cast<size_t>(cap) + idx[k]] = sz[k];
    }
}

__global__ __launch_bounds__(256) void merge_fp64_kernel(
    int begin, int end, int jsplit, int cap, const double* __restrict__ part,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az,
    double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz, int close_kick,
    double half_dt) {
#pragma clang fp contract(off)
    const int i = begin + blockIdx.x * 256 + threadIdx.x;
    if (i >= end) return;
    double a[3] = {0.0, 0.0, 0.0};
    for (int s = 0; s < jsplit; ++s) {
        const double* b = part + static_cast<size_t>(s) * 3 * cap;
#pragma unroll
        for (int c = 0; c < 3; ++c) a[c] += b[static_cast<size_t>(c) * cap + i];
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

template <int IPT>
void launch_fast_ipt(bool safe, dim3 grid, hipStream_t st, int begin, int end, int per_split,
                     int nsrc, const DeviceState& s, const Src64* jp, double eps2, double* part,
                     int cap) {
    if (safe)
        hipLaunchKernelGGL((direct_fp64_fast_kernel<IPT, true>), grid, dim3(256), 0, st, begin, end,
                           per_split, nsrc, s.x, s.y, s.z, jp, eps2, part, cap);
    else
        hipLaunchKernelGGL((direct_fp64_fast_kernel<IPT, false>), grid, dim3(256), 0, st, begin,
                           end, per_split, nsrc, s.x, s.y, s.z, jp, eps2, part, cap);
}

}  // namespace

size_t fp64_fast_sources(size_t n) { return round_up(std::max<size_t>(n, 1), kUnroll) + kUnroll; }

void launch_pack_fp64(const DeviceState& s, Src64* jp, double G, hipStream_t st) {
    const int padded = static_cast<int>(fp64_fast_sources(s.n));
    hipLaunchKernelGGL(pack_fp64_kernel, dim3((padded + 255) / 256), dim3(256), 0, st,
                       static_cast<int>(s.n), padded, s.x, s.y, s.z, s.m, G, jp);
}

void fp64_fast_shape(size_t n, size_t ntargets, int ipt, int jsplit, int* ipt_out,
                     int* jsplit_out, int* per_split_out) {
    if (ipt != 1 && ipt != 2 && ipt != 4) ipt = 2;
    const size_t span = static_cast<size_t>(256) * ipt;
    const size_t iblocks = std::max<size_t>(1, (ntargets + span - 1) / span);
    const size_t nsrc = round_up(std::max<size_t>(n, 1), kUnroll);
    // automatic: about 8,192 blocks (measured at N = 65,536, 2 targets per
    // lane: 64 splits 2.583 ms, 128 splits 2.606 ms, 16 splits 2.643 ms)
    size_t js = jsplit > 0 ? static_cast<size_t>(jsplit)
                           : std::min<size_t>(256, (8192 + iblocks - 1) / iblocks);
    js = std::max<size_t>(1, std::min(js, nsrc / kUnroll));
    const size_t per = round_up((nsrc + js - 1) / js, kUnroll);
    *ipt_out = ipt;
    *per_split_out = static_cast<int>(per);
    *jsplit_out = static_cast<int>((nsrc + per - 1) / per);
}

void launch_direct_fp64_fast(const DeviceState& s, const Src64* jp, double* part, double eps,
                             int ipt, int jsplit, int per_split, size_t begin, size_t end,
                             hipStream_t st) {
    if (end <= begin) return;
    const double eps2 = eps * eps;
    const bool safe = !(eps2 > 1e-200);
    const size_t span = static_cast<size_t>(256) * ipt;
    const dim3 grid(static_cast<unsigned>((end - begin + span - 1) / span),
                    static_cast<unsigned>(jsplit));
    const int b = static_cast<int>(begin), e = static_cast<int>(end), cap = static_cast<int>(s.cap);
    const int nsrc = static_cast<int>(round_up(std::max<size_t>(s.n, 1), kUnroll));
    switch (ipt) {
    case 4: launch_fast_ipt<4>(safe, grid, st, b, e, per_split, nsrc, s, jp, eps2, part, cap); break;
    case 1: launch_fast_ipt<1>(safe, grid, st, b, e, per_split, nsrc, s, jp, eps2, part, cap); break;
    default: launch_fast_ipt<2>(safe, grid, st, b, e, per_split, nsrc, s, jp, eps2, part, cap); break;
    }
}

void launch_merge_fp64(const DeviceState& s, const double* part, int jsplit, size_t begin,
                       size_t end, bool close_kick, double half_dt, hipStream_t st) {
    if (end <= begin) return;
    const unsigned blocks = static_cast<unsigned>((end - begin + 255) / 256);
    hipLaunchKernelGGL(merge_fp64_kernel, dim3(blocks), dim3(256), 0, st, static_cast<int>(begin),
                       static_cast<int>(end), jsplit, static_cast<int>(s.cap), part, s.ax, s.ay,
                       s.az, s.vx, s.vy, s.vz, close_kick ? 1 : 0, half_dt);
}

void launch_direct_fp64(const DeviceState& s, double G, double eps, size_t begin, size_t end,
                        bool close_kick, double half_dt, hipStream_t st) {
    if (end <= begin) return;
    const int blocks = static_cast<int>((end - begin + kTile - 1) / kTile);
    hipLaunchKernelGGL(direct_fp64_kernel, dim3(blocks), dim3(kTile), 0, st,
                       static_cast<int>(s.n), static_cast<int>(begin), static_cast<int>(end), s.x,
                       s.y, s.z, s.m, G, eps * eps, s.ax, s.ay, s.az, s.vx, s.vy, s.vz,
                       close_kick ? 1 : 0, half_dt);
}

}  // namespace gc
