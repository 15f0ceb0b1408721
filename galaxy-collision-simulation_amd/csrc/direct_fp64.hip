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

namespace gc {
namespace {

struct alignas(32) Src64 {
    double x, y, z, gm;
};

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

}  // namespace

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
