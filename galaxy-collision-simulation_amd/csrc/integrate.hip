// integrate.hip — kick/drift on the resident FP64 SoA state and the
// interleaved <-> SoA transposes of the host boundary.
//
// kick: v += a*dt, drift: x += v*dt (cpp/core/integrator.cpp:7-17), FMA
// contraction off so each update is the reference's multiply-then-add.  The
// opening half of a KDK step (kick dt/2 then drift dt) is one fused pass:
// 72 B read + 48 B written per particle, HBM-bound.
#include "gc_internal.hpp"

namespace gc {
namespace {

__global__ __launch_bounds__(256) void open_step_kernel(
    int n, double* __restrict__ x, double* __restrict__ y, double* __restrict__ z,
    double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
    const double* __restrict__ ax, const double* __restrict__ ay, const double* __restrict__ az,
    double half_dt, double dt) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double ux = vx[i] + ax[i] * half_dt;
    const double uy = vy[i] + ay[i] * half_dt;
    const double uz = vz[i] + az[i] * half_dt;
    vx[i] = ux;
    vy[i] = uy;
    vz[i] = uz;
    x[i] = x[i] + ux * dt;
    y[i] = y[i] + uy * dt;
    z[i] = z[i] + uz * dt;
}

__global__ __launch_bounds__(256) void kick_kernel(
    int n, double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
    const double* __restrict__ ax, const double* __restrict__ ay, const double* __restrict__ az,
    double dt) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    vx[i] = vx[i] + ax[i] * dt;
    vy[i] = vy[i] + ay[i] * dt;
    vz[i] = vz[i] + az[i] * dt;
}

__global__ __launch_bounds__(256) void scatter_xyz_kernel(size_t n, const double* __restrict__ in,
                                                          double* __restrict__ x,
                                                          double* __restrict__ y,
                                                          double* __restrict__ z) {
    const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    x[i] = in[3 * i];
    y[i] = in[3 * i + 1];
    z[i] = in[3 * i + 2];
}

__global__ __launch_bounds__(256) void gather_xyz_kernel(size_t n, const double* __restrict__ x,
                                                         const double* __restrict__ y,
                                                         const double* __restrict__ z,
                                                         double* __restrict__ out) {
    const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    out[3 * i] = x[i];
    out[3 * i + 1] = y[i];
    out[3 * i + 2] = z[i];
}

inline unsigned blocks_for(size_t n) { return static_cast<unsigned>((n + 255) / 256); }

// Plain streaming copy (16 B per lane), the reference kernel for scaling
// the FETCH_SIZE / WRITE_SIZE counters.
__global__ __launch_bounds__(256) void calibration_copy_kernel(size_t n, const float4* __restrict__ in,
                                                               float4* __restrict__ out) {
    for (size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * 256)
        out[i] = in[i];
}

}  // namespace

hipError_t run_calibration_copy(int device, size_t bytes, int reps, double* ms) {
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return e;
    float4 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const size_t n = bytes / sizeof(float4);
    if ((e = hipMalloc(&a, bytes)) == hipSuccess && (e = hipMalloc(&b, bytes)) == hipSuccess &&
        (e = hipMemset(a, 0, bytes)) == hipSuccess && (e = hipEventCreate(&e0)) == hipSuccess &&
        (e = hipEventCreate(&e1)) == hipSuccess) {
        const unsigned grid = 4096;
        hipLaunchKernelGGL(calibration_copy_kernel, dim3(grid), dim3(256), 0, nullptr, n, a, b);
        (void)hipEventRecord(e0, nullptr);
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(calibration_copy_kernel, dim3(grid), dim3(256), 0, nullptr, n, a, b);
        (void)hipEventRecord(e1, nullptr);
        e = hipEventSynchronize(e1);
        float t = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        if (e == hipSuccess) *ms = static_cast<double>(t) / std::max(1, reps);
        if (e == hipSuccess) e = hipGetLastError();
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return e;
}

void launch_open_step(const DeviceState& s, double half_dt, double dt, hipStream_t st) {
    if (s.n == 0) return;
    hipLaunchKernelGGL(open_step_kernel, dim3(blocks_for(s.n)), dim3(256), 0, st,
                       static_cast<int>(s.n), s.x, s.y, s.z, s.vx, s.vy, s.vz, s.ax, s.ay, s.az,
                       half_dt, dt);
}

void launch_kick(const DeviceState& s, double dt, hipStream_t st) {
    if (s.n == 0) return;
    hipLaunchKernelGGL(kick_kernel, dim3(blocks_for(s.n)), dim3(256), 0, st,
                       static_cast<int>(s.n), s.vx, s.vy, s.vz, s.ax, s.ay, s.az, dt);
}

void launch_scatter_xyz(const double* interleaved, double* x, double* y, double* z, size_t n,
                        hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(scatter_xyz_kernel, dim3(blocks_for(n)), dim3(256), 0, st, n, interleaved,
                       x, y, z);
}

void launch_gather_xyz(const double* x, const double* y, const double* z, double* interleaved,
                       size_t n, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(gather_xyz_kernel, dim3(blocks_for(n)), dim3(256), 0, st, n, x, y, z,
                       interleaved);
}

}  // namespace gc
