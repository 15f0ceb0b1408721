// ic.hip — initial conditions generated in HBM: one thread per particle
// draws from the counter-based Philox4x32-10 stream of (seed, stream,
// particle) and applies the model of its segment (ic_models.hpp), writing
// the resident SoA state directly (accelerations zeroed).  No particle
// depends on another, so any N fills the chip in one launch, and nothing
// crosses PCIe.
#include "gc_internal.hpp"
#include "ic_models.hpp"

namespace gc {
namespace {

__device__ __forceinline__ int find_segment(const IcSegment* seg, int nseg, long long t) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {  // last segment with begin <= t
        const int mid = (lo + hi + 1) >> 1;
        if (seg[mid].begin <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void ic_kernel(const IcSegment* __restrict__ seg, int nseg,
                                                 long long total, double* x, double* y, double* z,
                                                 double* vx, double* vy, double* vz, double* ax,
                                                 double* ay, double* az, double* m) {
    const long long t = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    if (t >= total) return;
    const IcSegment& s = seg[find_segment(seg, nseg, t)];
    const unsigned long long i = static_cast<unsigned long long>(t - s.begin);
    double b0[2], b1[2], b2[2];
    ic::philox_uniform2(s.seed, i, s.stream, 0, b0);
    ic::philox_uniform2(s.seed, i, s.stream, 1, b1);
    ic::philox_uniform2(s.seed, i, s.stream, 2, b2);
    double p[3], v[3], mass;
    if (s.model == kIcRefDisk || s.model == kIcTwoDisk) {
        double n0, n1, n2, n3;
        ic::box_muller(b1[0], b1[1], &n0, &n1);
        ic::box_muller(b2[0], b2[1], &n2, &n3);
        if (s.model == kIcRefDisk)
            ic::ref_disk_particle(s.p, b0[0], b0[1], n0, n1, n2, n3, p, v, &mass);
        else
            ic::two_disk_particle(s.p, i, b0[0], b0[1], n0, n1, n2, n3, p, v, &mass);
    } else {  // Plummer
        double r;
        ic::plummer_position(s.p, b0[0], -1.0 + 2.0 * b0[1], 2.0 * ic::kPi * b1[0], &r, p);
        double q = 0.0;
        for (unsigned a = 0; a < kIcPlummerTries; ++a) {
            double w[2];
            ic::philox_uniform2(s.seed, i, s.stream, 3 + a, w);
            q = w[0];
            if (ic::plummer_accept(q, 0.1 * w[1])) break;
        }
        ic::plummer_velocity(s.p, r, q, -1.0 + 2.0 * b1[1], 2.0 * ic::kPi * b2[0], v, &mass);
    }
    const long long d = s.dst + static_cast<long long>(i);
    x[d] = p[0];
    y[d] = p[1];
    z[d] = p[2];
    vx[d] = v[0];
    vy[d] = v[1];
    vz[d] = v[2];
    ax[d] = 0.0;
    ay[d] = 0.0;
    az[d] = 0.0;
    m[d] = mass;
}

}  // namespace

void launch_ic(const IcSegment* segs, int nseg, long long total, const DeviceState& s,
               hipStream_t st) {
    if (total <= 0 || nseg <= 0) return;
    const long long blocks = (total + 255) / 256;
    hipLaunchKernelGGL(ic_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, segs,
                       nseg, total, s.x, s.y, s.z, s.vx, s.vy, s.vz, s.ax, s.ay, s.az, s.m);
}

}  // namespace gc
