// noalloc.hip — mixed-precision direct accelerations for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).
//
// gc_direct_accelerations_noalloc runs the kernels of the per-target mixed
// path (pack_mixed_kernel, direct_mixed_kernel with scalar-load sources,
// merge_mixed_kernel) on caller-owned device arrays.  The buffers those
// kernels need beyond the inputs and outputs (source tiles, tile centres,
// per-split partial sums) live in one file-scope __device__ array, so the
// call performs no hipMalloc, hipFree or other allocating routine.  Each
// device holds its own copy of the array for the life of the process; its
// address on a device is read once with hipGetSymbolAddress (with that
// device current) and kept per device.  Every call on one device shares that
// device's array: calls that overlap on one device (different streams or
// threads) race on it.  Nothing in the array needs clearing between calls:
// every entry a kernel reads was written earlier in the same call (the
// source read-ahead reads past the last tile are loaded and discarded).
#include "gc_internal.hpp"
#include "abi_guard.hpp"

#include <atomic>
#include <cstring>
#include <mutex>
#include <string>

namespace gc {
namespace {

// Scratch layout (bytes): source tiles (float4, max particles + read-ahead),
// tile centres (3 x max tiles FP64), per-split partials (FP32).
constexpr size_t kNoAllocMaxParticles = GC_NOALLOC_MAX_PARTICLES;
constexpr size_t kNoAllocMaxTiles = kNoAllocMaxParticles / kTile;
constexpr size_t kNoAllocPartFloats = GC_NOALLOC_PARTIAL_FLOATS;
constexpr size_t kJpBytes = round_up((kNoAllocMaxParticles + kStreamBatch) * 16, 256);
constexpr size_t kCentreBytes = round_up(3 * kNoAllocMaxTiles * sizeof(double), 256);
constexpr size_t kPartBytes = kNoAllocPartFloats * sizeof(float);
constexpr size_t kScratchBytes = kJpBytes + kCentreBytes + kPartBytes;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::atomic<unsigned char*> g_scratch_ptr[kMaxDevices];
std::mutex g_scratch_mutex;

// Error path only (builds the message for gc_last_error(NULL)).
gc_status nfail(gc_status code, const char* msg, hipError_t e = hipSuccess) {
    std::string m = msg;
    if (e != hipSuccess) m += std::string(": ") + hipGetErrorString(e);
    return set_error(m, code);
}

// The array's address on `device`, looked up once per device with the device
// current.  The caller has made `device` current.
gc_status scratch_on(int device, unsigned char** out) {
    unsigned char* p = g_scratch_ptr[device].load(std::memory_order_acquire);
    if (p) {
        *out = p;
        return GC_OK;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    p = g_scratch_ptr[device].load(std::memory_order_relaxed);
    if (!p) {
        hipDeviceProp_t prop;
        hipError_t e = hipGetDeviceProperties(&prop, device);
        if (e != hipSuccess) return nfail(GC_ERR_HIP, "direct_accelerations_noalloc: device properties", e);
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            return nfail(GC_ERR_NO_DEVICE, "direct_accelerations_noalloc: device is not gfx950 (MI355X)");
        void* addr = nullptr;
        e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !addr)
            return nfail(GC_ERR_HIP, "direct_accelerations_noalloc: scratch address", e);
        p = static_cast<unsigned char*>(addr);
        g_scratch_ptr[device].store(p, std::memory_order_release);
    }
    *out = p;
    return GC_OK;
}

// Source splits of the per-target kernel: the context path's automatic
// count, capped so that the partials fit the scratch.
int noalloc_splits(size_t n) {
    const size_t cap = round_up(n, kTile);
    const size_t ntiles = cap / kTile;
    const size_t iblocks = ntiles;  // one target per lane: kTile targets per block
    size_t js = std::min<size_t>(64, (16384 + iblocks - 1) / iblocks);
    js = std::min(js, kNoAllocPartFloats / (6 * cap));
    js = std::max<size_t>(1, std::min(js, ntiles));
    const size_t tps = (ntiles + js - 1) / js;
    return static_cast<int>((ntiles + tps - 1) / tps);
}

}  // namespace
}  // namespace gc

extern "C" gc_status gc_direct_accelerations_noalloc(const gc_params* p, size_t n,
                                                     const double* pos, const double* mass,
                                                     double* acc, void* stream) {
    using namespace gc;
    GC_ABI_BEGIN
    if (!p) return nfail(GC_ERR_INVALID_ARGUMENT, "direct_accelerations_noalloc: null params");
    if (p->solver != GC_SOLVER_DIRECT || p->precision != GC_PRECISION_MIXED)
        return nfail(GC_ERR_INVALID_ARGUMENT,
                     "direct_accelerations_noalloc: needs solver direct, precision mixed");
    if (!(p->softening >= 0.0))
        return nfail(GC_ERR_INVALID_ARGUMENT, "direct_accelerations_noalloc: softening must be non-negative");
    if (n > kNoAllocMaxParticles)
        return nfail(GC_ERR_INVALID_ARGUMENT,
                     "direct_accelerations_noalloc: more than GC_NOALLOC_MAX_PARTICLES particles");
    if (n == 0) return GC_OK;
    if (!pos || !mass || !acc)
        return nfail(GC_ERR_INVALID_ARGUMENT, "direct_accelerations_noalloc: null array");
    hipStream_t st = static_cast<hipStream_t>(stream);
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return nfail(GC_ERR_NO_DEVICE, "direct_accelerations_noalloc: no current device", e);
    int device = cur;
    if (st) {
        e = hipStreamGetDevice(st, &device);
        if (e != hipSuccess) return nfail(GC_ERR_HIP, "direct_accelerations_noalloc: stream device", e);
    }
    if (device < 0 || device >= kMaxDevices)
        return nfail(GC_ERR_UNSUPPORTED, "direct_accelerations_noalloc: device ordinal beyond 63");
    if (device != cur && (e = hipSetDevice(device)) != hipSuccess)
        return nfail(GC_ERR_HIP, "direct_accelerations_noalloc: select device", e);
    unsigned char* base = nullptr;
    gc_status status = scratch_on(device, &base);
    if (status == GC_OK) {
        DeviceState s;
        s.n = n;
        s.cap = round_up(n, kTile);
        s.x = const_cast<double*>(pos);
        s.y = const_cast<double*>(pos) + n;
        s.z = const_cast<double*>(pos) + 2 * n;
        s.m = const_cast<double*>(mass);
        s.ax = acc;
        s.ay = acc + n;
        s.az = acc + 2 * n;
        MixedLayout L;
        L.ntiles = s.cap / kTile;
        L.jp = reinterpret_cast<float4*>(base);
        L.centre = reinterpret_cast<double*>(base + kJpBytes);
        L.part = reinterpret_cast<float*>(base + kJpBytes + kCentreBytes);
        L.part_floats = kNoAllocPartFloats;
        const int js = noalloc_splits(n);
        launch_pack_mixed(s, L, p->gravitational_constant, st);
        launch_direct_mixed(s, L, p->softening, 1, js, true, 0, n, st);
        launch_merge_mixed(s, L, js, 0, n, false, 0.0, st);
        e = hipGetLastError();
        if (e != hipSuccess) status = nfail(GC_ERR_HIP, "direct_accelerations_noalloc: launch", e);
    }
    if (device != cur) (void)hipSetDevice(cur);
    return status;
    GC_ABI_END("gc_direct_accelerations_noalloc")
}
