// diagnostics.hip — FP64 conservation diagnostics of the resident state
// (cpp/core/diagnostics.cpp:7-37) on the device.
//
//   linear pass  : total mass, KE = sum m|v|^2/2, P = sum m v, sum m x,
//                  L = sum m (x cross v); grid-stride lanes, wavefront
//                  (DPP/shuffle) reductions, one partial row per block.
//   potential    : PE = -sum_{i<j} G m_i m_j / sqrt(|x_j - x_i|^2 + eps^2),
//                  grid (target tile I, source chunk c): one target per
//                  lane over the sources j > i of kPotChunk tiles from
//                  tile I on, staged in LDS, sum_j m_j / d_ij per lane
//                  times G m_i, one partial per block (zero for chunks left
//                  of I); a row pass folds each target tile's chunks in
//                  order;
//   final        : one block folds the row sums and the linear partials in
//                  fixed order with compensation and writes the
//                  gc_diagnostics record.
// Deterministic (no atomics).  The reference forms each pair as
// (G*m_i)*m_j / sqrt(((dx*dx + dy*dy) + dz*dz) + eps^2); the potential here
// takes 1/d from a Newton-refined reciprocal square root (within 2 ulp) and
// applies G m_i once per target, so it agrees with the reference to
// rounding, like the summation order.  The linear sums follow the
// reference's per-particle arithmetic.
#include "gc_internal.hpp"

#include <algorithm>

namespace gc {
namespace {

constexpr int kLin = 11;  // M, KE, Px, Py, Pz, Cx, Cy, Cz, Lx, Ly, Lz
constexpr int kLinBlocksMax = 1024;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Neumaier compensated add of v into (s, c).
__device__ __forceinline__ void neumaier(double& s, double& c, double v) {
    const double t = s + v;
    if (fabs(s) >= fabs(v)) {
        c += (s - t) + v;
    } else {
        c += (v - t) + s;
    }
    s = t;
}

__global__ __launch_bounds__(256) void diag_linear_kernel(
    int n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ vx, const double* __restrict__ vy,
    const double* __restrict__ vz, const double* __restrict__ m, double* __restrict__ partial) {
#pragma clang fp contract(off)
    __shared__ double red[kLin][4];
    double acc[kLin];
#pragma unroll
    for (int k = 0; k < kLin; ++k) acc[k] = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const double mi = m[i], px = x[i], py = y[i], pz = z[i];
        const double ux = vx[i], uy = vy[i], uz = vz[i];
        acc[0] += mi;
        acc[1] += 0.5 * mi * (ux * ux + uy * uy + uz * uz);
        acc[2] += ux * mi;
        acc[3] += uy * mi;
        acc[4] += uz * mi;
        acc[5] += px * mi;
        acc[6] += py * mi;
        acc[7] += pz * mi;
        acc[8] += (py * uz - pz * uy) * mi;
        acc[9] += (pz * ux - px * uz) * mi;
        acc[10] += (px * uy - py * ux) * mi;
    }
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kLin; ++k) {
        const double v = wave_sum(acc[k]);
        if ((threadIdx.x & 63) == 0) red[k][w] = v;
    }
    __syncthreads();
    if (threadIdx.x < kLin) {
        const int k = threadIdx.x;
        partial[blockIdx.x * kLin + k] = ((red[k][0] + red[k][1]) + red[k][2]) + red[k][3];
    }
}

struct alignas(32) PotSrc {
    double x, y, z, m;
};

constexpr int kPotChunk = 4;  // source tiles per block

// 1/sqrt(x): the hardware estimate (about 23 correct bits) refined by two
// Newton steps y <- y + y (1/2 - e/2), e = (x y) y, within 2 ulp.  Forming
// (x y) y keeps every intermediate near sqrt(x) or 1, so normal and
// subnormal x both stay in range.  Returns 0 unless x is a positive finite
// number (a coincident pair at eps = 0 is skipped, as in the reference; an
// overflowed distance contributes 0, as m / inf does there).
__device__ __forceinline__ double inv_dist(double x) {
    double y = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double e = (x * y) * y;
        y = __builtin_fma(y, __builtin_fma(-0.5, e, 0.5), y);
    }
    // class bits 7 and 8: positive subnormal, positive normal
    return __builtin_amdgcn_class(x, (1 << 7) | (1 << 8)) ? y : 0.0;
}

// Sources of one LDS tile against this lane's target: sum of m_j / d_ij;
// DIAG: the target's own tile, only j > i counts.
template <bool DIAG>
__device__ __forceinline__ void potential_tile(const PotSrc* tile, int cnt, int jl_min, double xi,
                                               double yi, double zi, double eps2, double& s0,
                                               double& s1) {
    int jj = 0;
    for (; jj + 2 <= cnt; jj += 2) {
        const PotSrc a = tile[jj], b = tile[jj + 1];
        double dx = a.x - xi, dy = a.y - yi, dz = a.z - zi;
        const double ra = inv_dist(__builtin_fma(dz, dz, __builtin_fma(dy, dy, __builtin_fma(dx, dx, eps2))));
        dx = b.x - xi;
        dy = b.y - yi;
        dz = b.z - zi;
        const double rb = inv_dist(__builtin_fma(dz, dz, __builtin_fma(dy, dy, __builtin_fma(dx, dx, eps2))));
        s0 = __builtin_fma(DIAG && jj <= jl_min ? 0.0 : a.m, ra, s0);
        s1 = __builtin_fma(DIAG && jj + 1 <= jl_min ? 0.0 : b.m, rb, s1);
    }
    if (jj < cnt) {
        const PotSrc a = tile[jj];
        const double dx = a.x - xi, dy = a.y - yi, dz = a.z - zi;
        const double ra = inv_dist(__builtin_fma(dz, dz, __builtin_fma(dy, dy, __builtin_fma(dx, dx, eps2))));
        s0 = __builtin_fma(DIAG && jj <= jl_min ? 0.0 : a.m, ra, s0);
    }
}

// Grid (target tile I, source chunk c): lane i sums m_j / d_ij over the
// sources j > i of chunk c's tiles (from tile I on) in two interleaved
// FP64 sums (all terms positive: each sum's relative error is at most
// 512 ulp, typically ~20); the block's partial is G m_i times that,
// summed over its lanes.
__global__ __launch_bounds__(256) void diag_potential_kernel(
    int n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ m, double G, double eps2,
    double* __restrict__ partial) {
    __shared__ PotSrc tile[256];
    __shared__ double red[4];
    const int t = threadIdx.x;
    const int I = blockIdx.x, c = blockIdx.y;
    const int i = I * 256 + t;
    const int jt0 = max(I, c * kPotChunk), jt1 = min((c + 1) * kPotChunk, static_cast<int>(gridDim.x));
    double* out = partial + static_cast<size_t>(I) * gridDim.y + c;
    if (jt0 >= jt1) {  // chunk left of the target tile: no pairs
        if (t == 0) *out = 0.0;
        return;
    }
    const bool active = i < n;
    const double xi = active ? x[i] : 0.0, yi = active ? y[i] : 0.0, zi = active ? z[i] : 0.0;
    double s0 = 0.0, s1 = 0.0;
    for (int jt = jt0; jt < jt1; ++jt) {
        const int j0 = jt * 256;
        __syncthreads();
        if (j0 + t < n) {
            PotSrc q;
            q.x = x[j0 + t];
            q.y = y[j0 + t];
            q.z = z[j0 + t];
            q.m = m[j0 + t];
            tile[t] = q;
        }
        __syncthreads();
        const int cnt = min(256, n - j0);
        if (jt == I)
            potential_tile<true>(tile, cnt, t, xi, yi, zi, eps2, s0, s1);
        else
            potential_tile<false>(tile, cnt, -1, xi, yi, zi, eps2, s0, s1);
    }
    double v = active ? (G * m[i]) * (s0 + s1) : 0.0;
    v = wave_sum(v);
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    if (t == 0) *out = ((red[0] + red[1]) + red[2]) + red[3];
}

// One thread per target tile: its chunk partials in chunk order.
__global__ __launch_bounds__(256) void diag_rows_kernel(int tiles, int chunks,
                                                        const double* __restrict__ partial,
                                                        double* __restrict__ rows) {
#pragma clang fp contract(off)
    const int I = blockIdx.x * 256 + threadIdx.x;
    if (I >= tiles) return;
    double s = 0.0, c = 0.0;
    for (int k = 0; k < chunks; ++k) neumaier(s, c, partial[static_cast<size_t>(I) * chunks + k]);
    rows[I] = s + c;
}

// Writes the gc_diagnostics field order: KE, PE, E, M, P[3], COM[3], L[3].
// Wave k folds quantity k (k < kLin: the linear partials; k == kLin: the
// potential rows): lane l takes entries l, l + 64, ... in order, then the
// lanes' compensated pairs meet in a fixed butterfly (each step symmetric
// in its two operands, so every lane ends with the same, fixed result).
__global__ __launch_bounds__(64 * (kLin + 1)) void diag_final_kernel(
    int nlin, const double* __restrict__ lin, int npot, const double* __restrict__ pot,
    double* __restrict__ out) {
#pragma clang fp contract(off)
    __shared__ double tot[kLin + 1];
    const int k = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double s = 0.0, c = 0.0;
    if (k < kLin) {
        for (int b = lane; b < nlin; b += 64) neumaier(s, c, lin[b * kLin + k]);
    } else {
        for (int b = lane; b < npot; b += 64) neumaier(s, c, pot[b]);
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double s2 = __shfl_xor(s, o, 64), c2 = __shfl_xor(c, o, 64);
        c += c2;
        neumaier(s, c, s2);
    }
    if (lane == 0) tot[k] = k < kLin ? s + c : -(s + c);
    __syncthreads();
    if (threadIdx.x == 0) {
        const double M = tot[0];
        out[0] = tot[1];
        out[1] = tot[kLin];
        out[2] = tot[1] + tot[kLin];
        out[3] = M;
        out[4] = tot[2];
        out[5] = tot[3];
        out[6] = tot[4];
        out[7] = M > 0.0 ? tot[5] / M : tot[5];
        out[8] = M > 0.0 ? tot[6] / M : tot[6];
        out[9] = M > 0.0 ? tot[7] / M : tot[7];
        out[10] = tot[8];
        out[11] = tot[9];
        out[12] = tot[10];
    }
}

int linear_blocks(size_t n) {
    return static_cast<int>(std::max<size_t>(1, std::min<size_t>((n + 255) / 256, kLinBlocksMax)));
}

int potential_tiles(size_t n) { return static_cast<int>(std::max<size_t>(1, (n + 255) / 256)); }
int potential_chunks(size_t n) { return (potential_tiles(n) + kPotChunk - 1) / kPotChunk; }

}  // namespace

size_t diagnostics_scratch_doubles(size_t n) {
    const size_t T = static_cast<size_t>(potential_tiles(n));
    return static_cast<size_t>(linear_blocks(n)) * kLin + T * potential_chunks(n) + T;
}

void launch_diagnostics(const DeviceState& s, double G, double eps, double* scratch,
                        size_t scratch_doubles, double* out_dev, hipStream_t st) {
    const int nl = linear_blocks(s.n), T = potential_tiles(s.n), C = potential_chunks(s.n);
    if (scratch_doubles < diagnostics_scratch_doubles(s.n)) return;
    double* lin = scratch;
    double* pot = scratch + static_cast<size_t>(nl) * kLin;
    double* rows = pot + static_cast<size_t>(T) * C;
    const int n = static_cast<int>(s.n);
    hipLaunchKernelGGL(diag_linear_kernel, dim3(nl), dim3(256), 0, st, n, s.x, s.y, s.z, s.vx,
                       s.vy, s.vz, s.m, lin);
    hipLaunchKernelGGL(diag_potential_kernel, dim3(T, C), dim3(256), 0, st, n, s.x, s.y, s.z, s.m,
                       G, eps * eps, pot);
    hipLaunchKernelGGL(diag_rows_kernel, dim3((T + 255) / 256), dim3(256), 0, st, T, C, pot, rows);
    hipLaunchKernelGGL(diag_final_kernel, dim3(1), dim3(64 * (kLin + 1)), 0, st, nl, lin, T, rows,
                       out_dev);
}

}  // namespace gc
