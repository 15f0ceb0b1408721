// capi.hip — the gcsim C ABI (include/gcsim.h): contexts owning
// device-resident SoA particle state, the hip-direct force paths (fp64 and
// mixed), KDK stepping and diagnostics.  No exception crosses the ABI; every
// HIP error becomes GC_ERR_HIP with a message from gc_last_error.
#include "gc_internal.hpp"
#include "abi_guard.hpp"
#include "gc_fmm.hpp"
#include "ic_models.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

using namespace gc;

namespace {

// Everything a captured chunk of gc_step bakes into its kernel arguments:
// the graph is replayed only while all of it is unchanged.
struct StepKey {
    double dt = 0.0, G = 0.0, eps = 0.0;
    int solver = -1, precision = -1;
    int ipt = 0, jsplit = 0, sources = 0, pairs = 0, items = 0, steps = 0;
    size_t n = 0, cap = 0, ntiles = 0;
    const void* ptr[9] = {};
    int nitems = 0;
    const void* row_start = nullptr;

    bool operator==(const StepKey& o) const {
        if (dt != o.dt || G != o.G || eps != o.eps || solver != o.solver ||
            precision != o.precision || ipt != o.ipt || jsplit != o.jsplit ||
            sources != o.sources || pairs != o.pairs || items != o.items || steps != o.steps ||
            n != o.n || cap != o.cap || ntiles != o.ntiles)
            return false;
        for (int k = 0; k < 9; ++k)
            if (ptr[k] != o.ptr[k]) return false;
        return nitems == o.nitems && row_start == o.row_start;
    }
};

}  // namespace

struct gc_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    gc_params params{};
    DeviceState s;
    MixedLayout mixed;
    LaunchShape shape;
    SymPlan sym;
    double* stage = nullptr;  // 3 * cap doubles: interleaved host transfers
    double* diag_scratch = nullptr;
    size_t diag_scratch_doubles = 0;
    double* diag_out = nullptr;  // 13 doubles
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    gc::FmmEngine* fmm = nullptr;
    float4* vjp_lpack = nullptr;  // FP32 cotangents, cap entries
    double* vjp_part = nullptr;   // per-split VJP partials
    size_t vjp_lpack_n = 0, vjp_part_n = 0;
    float* vjp_sym_part = nullptr;  // pair-once VJP: item partials, then pair partials
    size_t vjp_sym_part_n = 0;
    Src64* jp64 = nullptr;       // FP64 throughput form: packed sources
    double* part64 = nullptr;    // ... and its per-split partial sums
    size_t jp64_n = 0, part64_n = 0;
    // gc_step: chunks of graph_steps steps replayed from a captured hipGraph
    // (graph_steps <= 1: every kernel launched eagerly).  Chunks are captured
    // on a stream of their own (created with the context, which keeps its
    // creation out of the first stepping call), which holds no other work,
    // so a capture never meets work or events that other threads wait on in
    // `stream`.
    int graph_steps = 10;
    hipGraphExec_t step_graph = nullptr;
    hipStream_t capture_stream = nullptr;
    StepKey step_key{};
    std::string err;
};

namespace {

thread_local std::string g_thread_err;

gc_status fail(gc_ctx* c, gc_status code, const std::string& msg) {
    if (c) c->err = msg;
    g_thread_err = msg;
    return code;
}

#define GC_HIP(ctx, call, what)                                                                  \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) {                                                                  \
            return fail(ctx, e_ == hipErrorOutOfMemory ? GC_ERR_OUT_OF_MEMORY : GC_ERR_HIP,      \
                        std::string(what) + ": " + hipGetErrorString(e_));                      \
        }                                                                                        \
    } while (0)

bool is_gfx950(int device) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return false;
    return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
}

gc_status check_params(gc_ctx* c, const gc_params* p) {
    if (!p) return fail(c, GC_ERR_INVALID_ARGUMENT, "params: null");
    if (!(p->softening >= 0.0)) return fail(c, GC_ERR_INVALID_ARGUMENT, "softening must be non-negative");
    if (p->solver != GC_SOLVER_DIRECT && p->solver != GC_SOLVER_FMM && p->solver != GC_SOLVER_TREE)
        return fail(c, GC_ERR_INVALID_ARGUMENT, "unknown solver kind");
    if (p->precision != GC_PRECISION_FP64 && p->precision != GC_PRECISION_MIXED &&
        p->precision != GC_PRECISION_FP64_FAST)
        return fail(c, GC_ERR_INVALID_ARGUMENT, "unknown precision");
    return GC_OK;
}

void free_state(gc_ctx* c) {
    sym_release(c->sym);
    (void)hipFree(c->s.base);
    (void)hipFree(c->mixed.jp);
    (void)hipFree(c->mixed.centre);
    (void)hipFree(c->mixed.part);
    (void)hipFree(c->stage);
    (void)hipFree(c->diag_scratch);
    c->s = DeviceState{};
    c->mixed = MixedLayout{};
    c->stage = nullptr;
    c->diag_scratch = nullptr;
    c->diag_scratch_doubles = 0;
}

gc_status ensure_capacity(gc_ctx* c, size_t n) {
    const size_t cap = std::max<size_t>(round_up(std::max<size_t>(n, 1), kTile), kTile);
    if (cap <= c->s.cap) {
        c->s.n = n;
        c->mixed.ntiles = (n + kTile - 1) / kTile;
        return GC_OK;
    }
    GC_HIP(c, hipStreamSynchronize(c->stream), "synchronize before reallocation");
    free_state(c);
    // every buffer, or none: a failure frees what was allocated and leaves
    // cap at 0, so the next call allocates again
    const size_t ntiles_cap = cap / kTile;
    const size_t diag_doubles = diagnostics_scratch_doubles(cap);
    const char* what = "allocate particle state";
    hipError_t e = hipMalloc(&c->s.base, sizeof(double) * 10 * cap);
    if (e == hipSuccess) {
        what = "allocate mixed source tiles";
        e = hipMalloc(&c->mixed.jp, sizeof(float4) * (cap + kStreamBatch));
    }
    if (e == hipSuccess) {
        what = "allocate tile centres";
        e = hipMalloc(&c->mixed.centre, sizeof(double) * 3 * ntiles_cap);
    }
    if (e == hipSuccess) {
        what = "allocate staging buffer";
        e = hipMalloc(&c->stage, sizeof(double) * 3 * cap);
    }
    if (e == hipSuccess) {
        what = "allocate diagnostics scratch";
        e = hipMalloc(&c->diag_scratch, sizeof(double) * diag_doubles);
    }
    if (e == hipSuccess) {
        what = "clear state";
        e = hipMemsetAsync(c->s.base, 0, sizeof(double) * 10 * cap, c->stream);
    }
    if (e != hipSuccess) {
        free_state(c);
        return fail(c, e == hipErrorOutOfMemory ? GC_ERR_OUT_OF_MEMORY : GC_ERR_HIP,
                    std::string(what) + ": " + hipGetErrorString(e));
    }
    double* b = c->s.base;
    c->s.x = b;
    c->s.y = b + cap;
    c->s.z = b + 2 * cap;
    c->s.vx = b + 3 * cap;
    c->s.vy = b + 4 * cap;
    c->s.vz = b + 5 * cap;
    c->s.ax = b + 6 * cap;
    c->s.ay = b + 7 * cap;
    c->s.az = b + 8 * cap;
    c->s.m = b + 9 * cap;
    c->diag_scratch_doubles = diag_doubles;
    c->s.cap = cap;
    c->s.n = n;
    c->mixed.ntiles = (n + kTile - 1) / kTile;
    return GC_OK;
}

// Source delivery of the mixed kernel: per-wave scalar loads (default) or
// LDS tiles shared by the block's four waves.  Measured at N = 65,536 with
// 64 splits: 1.0855 ms per launch with scalar loads, 1.0932 ms with LDS
// tiles (whole KDK step 1.118 ms against 1.122 ms).
bool scalar_sources(const gc_ctx* c) { return c->shape.sources != GC_SOURCES_LDS; }

// Mixed-kernel launch shape: ~16384 blocks of 4 waves (64 waves per SIMD
// over the launch, 8 resident), split over source tile ranges, so the last
// wave of blocks is short.  Measured at N = 65,536 (scalar loads): 1 target
// per lane with 64 splits runs 1.0855 ms, 32 splits 1.1107 ms, 16 splits
// 1.136 ms, 8 splits 1.195 ms; 128 splits 1.0785 ms but its partial sums
// cost more in the merge than they save (step 1.137 ms against 1.118 ms);
// 2 targets per lane are slower at every split count.
void mixed_shape(const gc_ctx* c, size_t ntargets, int* ipt_out, int* jsplit_out) {
    int ipt = c->shape.ipt;
    if (ipt != 1 && ipt != 2 && ipt != 4) ipt = 1;
    const size_t span = static_cast<size_t>(kTile) * ipt;
    const size_t iblocks = std::max<size_t>(1, (ntargets + span - 1) / span);
    const int ntiles = static_cast<int>(std::max<size_t>(1, c->mixed.ntiles));
    int js = c->shape.jsplit;
    if (js <= 0) js = static_cast<int>(std::min<size_t>(64, (16384 + iblocks - 1) / iblocks));
    js = std::max(1, std::min(js, ntiles));
    const int tps = (ntiles + js - 1) / js;
    *ipt_out = ipt;
    *jsplit_out = (ntiles + tps - 1) / tps;
}

// Each unordered pair once (direct_sym.hip) for full-range forces of up to
// kSymMaxTiles tiles (the default); target ranges run the per-target
// kernel.  Measured at N = 65,536: 1.003 ms per launch (2 targets per lane)
// against 1.090 ms for the per-target kernel; KDK step 1.088 ms against
// 1.124-1.155 ms.
bool symmetric_pairs(const gc_ctx* c, size_t begin, size_t end) {
    if (c->shape.pairs == GC_PAIRS_TARGETS) return false;
    return begin == 0 && end == c->s.n && c->mixed.ntiles >= 1 &&
           c->mixed.ntiles <= static_cast<size_t>(kSymMaxTiles);
}

// Targets per lane of the symmetric kernel unless set: the largest of 4, 2,
// 1 that still launches >= 8192 waves (8 per SIMD).  Measured at N = 65,536
// (16,512 items of 2 source tiles): 0.853 ms per launch with 4 per lane,
// 0.952 ms with 2; items of 4 tiles took 0.886 ms with 4.
int sym_ipt(const gc_ctx* c) {
    const int ipt = c->shape.ipt;
    if (ipt == 1 || ipt == 2 || ipt == 4) return ipt;
    const long waves4 = c->sym.nitems;  // one wave per item at 4 per lane
    if (waves4 >= 8192) return 4;
    if (2 * waves4 >= 8192) return 2;
    return 1;
}

// Targets per lane of the pair-once VJP unless set (gc_set_launch_shape).
constexpr int kVjpSymIpt = 2;

gc_status ensure_sym(gc_ctx* c) {
    std::string msg;
    const int rc = sym_prepare(c->sym, static_cast<int>(c->mixed.ntiles), c->shape.items,
                               c->stream, &msg);
    if (rc == 0) return GC_OK;
    return fail(c, rc == 2 ? GC_ERR_OUT_OF_MEMORY : GC_ERR_HIP, msg);
}

gc_status ensure_partials(gc_ctx* c, int jsplit) {
    const size_t need = static_cast<size_t>(jsplit) * 6 * c->s.cap;
    if (need <= c->mixed.part_floats) return GC_OK;
    GC_HIP(c, hipStreamSynchronize(c->stream), "synchronize before partials reallocation");
    (void)hipFree(c->mixed.part);
    c->mixed.part = nullptr;
    c->mixed.part_floats = 0;
    GC_HIP(c, hipMalloc(&c->mixed.part, sizeof(float) * need), "allocate mixed partial sums");
    c->mixed.part_floats = need;
    return GC_OK;
}

gc_status launch_status(gc_ctx* c, const char* what) {
    GC_HIP(c, hipGetLastError(), what);
    return GC_OK;
}

// Buffers of the FP64 throughput form for the resident particles and
// `jsplit` source splits.
gc_status ensure_fp64_fast(gc_ctx* c, int jsplit) {
    const size_t nsrc = fp64_fast_sources(c->s.n);
    const size_t npart = static_cast<size_t>(jsplit) * 3 * c->s.cap;
    if (nsrc <= c->jp64_n && npart <= c->part64_n) return GC_OK;
    GC_HIP(c, hipStreamSynchronize(c->stream), "synchronize before FP64 buffers");
    if (nsrc > c->jp64_n) {
        (void)hipFree(c->jp64);
        c->jp64 = nullptr;
        c->jp64_n = 0;
        GC_HIP(c, hipMalloc(&c->jp64, sizeof(Src64) * nsrc), "allocate FP64 sources");
        c->jp64_n = nsrc;
    }
    if (npart > c->part64_n) {
        (void)hipFree(c->part64);
        c->part64 = nullptr;
        c->part64_n = 0;
        GC_HIP(c, hipMalloc(&c->part64, sizeof(double) * npart), "allocate FP64 partial sums");
        c->part64_n = npart;
    }
    return GC_OK;
}

// a(x) for targets [begin, end); optionally the closing half kick.
gc_status force(gc_ctx* c, size_t begin, size_t end, bool close_kick, double half_dt) {
    const gc_params& p = c->params;
    if (p.solver == GC_SOLVER_FMM || p.solver == GC_SOLVER_TREE) {
        if (!c->fmm) c->fmm = new (std::nothrow) gc::FmmEngine();
        if (!c->fmm) return fail(c, GC_ERR_OUT_OF_MEMORY, "allocate FMM engine");
        std::string msg;
        const int rc = c->fmm->compute(c->s, p, begin, end, close_kick, half_dt, c->stream, &msg);
        if (rc != GC_OK)
            return fail(c, static_cast<gc_status>(rc),
                        (p.solver == GC_SOLVER_TREE ? "hip-tree: " : "hip-fmm: ") + msg);
        return GC_OK;
    }
    if (p.precision == GC_PRECISION_FP64) {
        launch_direct_fp64(c->s, p.gravitational_constant, p.softening, begin, end, close_kick,
                           half_dt, c->stream);
        return launch_status(c, "launch FP64 direct kernel");
    }
    if (p.precision == GC_PRECISION_FP64_FAST) {
        int ipt = 2, js = 1, per = 8;
        fp64_fast_shape(c->s.n, end - begin, c->shape.ipt, c->shape.jsplit, &ipt, &js, &per);
        gc_status st = ensure_fp64_fast(c, js);
        if (st != GC_OK) return st;
        launch_pack_fp64(c->s, c->jp64, p.gravitational_constant, c->stream);
        launch_direct_fp64_fast(c->s, c->jp64, c->part64, p.softening, ipt, js, per, begin, end,
                                c->stream);
        launch_merge_fp64(c->s, c->part64, js, begin, end, close_kick, half_dt, c->stream);
        return launch_status(c, "launch FP64 throughput kernels");
    }
    if (symmetric_pairs(c, begin, end)) {
        gc_status st = ensure_sym(c);
        if (st != GC_OK) return st;
        launch_pack_mixed(c->s, c->mixed, p.gravitational_constant, c->stream);
        launch_direct_sym(c->s, c->mixed, c->sym, p.softening, sym_ipt(c), c->stream);
        launch_merge_sym(c->s, c->sym, nullptr, close_kick, half_dt, c->stream);
        return launch_status(c, "launch symmetric mixed kernels");
    }
    int ipt = 1, js = 1;
    mixed_shape(c, end - begin, &ipt, &js);
    gc_status st = ensure_partials(c, js);
    if (st != GC_OK) return st;
    launch_pack_mixed(c->s, c->mixed, p.gravitational_constant, c->stream);
    launch_direct_mixed(c->s, c->mixed, p.softening, ipt, js, scalar_sources(c), begin, end,
                        c->stream);
    launch_merge_mixed(c->s, c->mixed, js, begin, end, close_kick, half_dt, c->stream);
    return launch_status(c, "launch mixed direct kernels");
}

// Buffers of the full-range direct force, allocated before a capture:
// force() then finds them in place and only launches kernels.
gc_status prepare_direct(gc_ctx* c) {
    const gc_params& p = c->params;
    if (p.precision == GC_PRECISION_FP64) return GC_OK;
    if (p.precision == GC_PRECISION_FP64_FAST) {
        int ipt = 2, js = 1, per = 8;
        fp64_fast_shape(c->s.n, c->s.n, c->shape.ipt, c->shape.jsplit, &ipt, &js, &per);
        return ensure_fp64_fast(c, js);
    }
    if (symmetric_pairs(c, 0, c->s.n)) return ensure_sym(c);
    int ipt = 1, js = 1;
    mixed_shape(c, c->s.n, &ipt, &js);
    return ensure_partials(c, js);
}

// k KDK steps, each its own opening kick + drift and force + closing kick.
gc_status launch_plain_steps(gc_ctx* c, double dt, int k) {
    for (int i = 0; i < k; ++i) {
        launch_open_step(c->s, 0.5 * dt, dt, c->stream);
        gc_status st = force(c, 0, c->s.n, true, 0.5 * dt);
        if (st != GC_OK) return st;
    }
    return GC_OK;
}

// k middle steps of the fused mixed symmetric sequence: the force, then the
// merge that closes the step, opens the next one and packs its sources.
void launch_sym_middle(gc_ctx* c, double dt, int k) {
    const int ipt = sym_ipt(c);
    for (int i = 0; i < k; ++i) {
        // the tile weights come from the pack before the first middle step
        // and from each merge after it
        launch_direct_sym(c->s, c->mixed, c->sym, c->params.softening, ipt, c->stream, true);
        launch_merge_sym_open(c->s, c->sym, c->mixed, c->params.gravitational_constant,
                              c->params.softening, 0.5 * dt, dt, c->stream);
    }
}

StepKey make_step_key(const gc_ctx* c, double dt) {
    StepKey k;
    k.dt = dt;
    k.G = c->params.gravitational_constant;
    k.eps = c->params.softening;
    k.solver = c->params.solver;
    k.precision = c->params.precision;
    k.ipt = c->shape.ipt;
    k.jsplit = c->shape.jsplit;
    k.sources = c->shape.sources;
    k.pairs = c->shape.pairs;
    k.items = c->shape.items;
    k.steps = c->graph_steps;
    k.n = c->s.n;
    k.cap = c->s.cap;
    k.ntiles = c->mixed.ntiles;
    const void* ptr[] = {c->s.base,    c->mixed.jp, c->mixed.centre, c->mixed.part,
                         c->sym.items, c->sym.ipart, c->sym.jpart, c->jp64, c->part64};
    static_assert(sizeof(ptr) == sizeof(k.ptr), "step key pointers");
    std::memcpy(k.ptr, ptr, sizeof(ptr));
    k.nitems = c->sym.nitems;
    k.row_start = c->sym.row_start;
    return k;
}

void drop_step_graph(gc_ctx* c) {
    if (c->step_graph) (void)hipGraphExecDestroy(c->step_graph);
    c->step_graph = nullptr;
    c->step_key = StepKey{};
}

// Chunks are replayed for the direct solver on a stream that is not itself
// being captured (a caller's own capture records the eager launches
// instead).
bool use_step_graph(gc_ctx* c) {
    if (c->graph_steps <= 1 || c->params.solver != GC_SOLVER_DIRECT) return false;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(c->stream, &cs) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return cs == hipStreamCaptureStatusNone;
}

// Makes c->step_graph the graph of one chunk of c->graph_steps steps,
// capturing it unless the graph of an identical chunk is already there:
// body() launches the chunk on c->stream, which points at the capture
// stream meanwhile (kernels only: no allocation, copy or memset).  Nothing
// runs on the device here.
template <class Body>
gc_status ensure_step_graph(gc_ctx* c, double dt, Body body) {
    const StepKey key = make_step_key(c, dt);
    if (c->step_graph && c->step_key == key) return GC_OK;
    drop_step_graph(c);
    if (!c->capture_stream)
        GC_HIP(c, hipStreamCreateWithFlags(&c->capture_stream, hipStreamNonBlocking),
               "create step capture stream");
    hipGraph_t g = nullptr;
    GC_HIP(c, hipStreamBeginCapture(c->capture_stream, hipStreamCaptureModeThreadLocal),
           "begin step capture");
    hipStream_t const user = c->stream;
    c->stream = c->capture_stream;
    gc_status st = body();
    c->stream = user;
    const hipError_t le = hipGetLastError();
    const hipError_t ee = hipStreamEndCapture(c->capture_stream, &g);
    if (st == GC_OK && le != hipSuccess)
        st = fail(c, GC_ERR_HIP, std::string("capture step kernels: ") + hipGetErrorString(le));
    if (st == GC_OK && ee != hipSuccess)
        st = fail(c, GC_ERR_HIP, std::string("end step capture: ") + hipGetErrorString(ee));
    if (st != GC_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    const hipError_t ie = hipGraphInstantiate(&c->step_graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
        c->step_graph = nullptr;
        return fail(c, ie == hipErrorOutOfMemory ? GC_ERR_OUT_OF_MEMORY : GC_ERR_HIP,
                    std::string("instantiate step graph: ") + hipGetErrorString(ie));
    }
    c->step_key = key;
    return GC_OK;
}

// Launches `chunks` replays of the chunk graph on the context's stream and
// returns true, or returns false when the chunk cannot be captured or
// instantiated: then nothing has been launched, the context's error text is
// left as it was, and the caller launches those steps eagerly (a step that
// has begun is never left half done for the sake of a graph).  A failed
// launch of the graph is an error of the step.
template <class Body>
bool replay_chunks(gc_ctx* c, double dt, int chunks, Body body, gc_status* st) {
    const std::string saved = c->err, saved_thread = g_thread_err;
    if (ensure_step_graph(c, dt, body) != GC_OK) {
        (void)hipGetLastError();
        drop_step_graph(c);
        c->err = saved;
        g_thread_err = saved_thread;
        *st = GC_OK;
        return false;
    }
    *st = GC_OK;
    for (int i = 0; i < chunks; ++i) {
        const hipError_t e = hipGraphLaunch(c->step_graph, c->stream);
        if (e != hipSuccess) {
            *st = fail(c, e == hipErrorOutOfMemory ? GC_ERR_OUT_OF_MEMORY : GC_ERR_HIP,
                       std::string("launch step graph: ") + hipGetErrorString(e));
            break;
        }
    }
    return true;
}

gc_status need_particles(gc_ctx* c) {
    if (!c) return fail(nullptr, GC_ERR_INVALID_ARGUMENT, "null context");
    GC_HIP(c, hipSetDevice(c->device), "select device");
    return GC_OK;
}

}  // namespace

namespace {

constexpr size_t kMaxIcParticles = static_cast<size_t>(1) << 30;

// Replaces the resident state with n particles drawn on the device from
// segs (begin/count cover [0, n), dst = begin).
gc_status generate_state(gc_ctx* c, std::vector<IcSegment>& segs, size_t n, const char* what) {
    gc_status st = need_particles(c);
    if (st != GC_OK) return st;
    if (n > kMaxIcParticles)
        return fail(c, GC_ERR_INVALID_ARGUMENT, std::string(what) + ": particle count exceeds 2^30");
    st = ensure_capacity(c, n);
    if (st != GC_OK || n == 0) return st;
    IcSegment* d = nullptr;
    GC_HIP(c, hipMalloc(&d, sizeof(IcSegment) * segs.size()), what);
    hipError_t e = hipMemcpyAsync(d, segs.data(), sizeof(IcSegment) * segs.size(),
                                  hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        launch_ic(d, static_cast<int>(segs.size()), static_cast<long long>(n), c->s, c->stream);
        e = hipGetLastError();
    }
    const hipError_t e2 = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e == hipSuccess) e = e2;
    if (e != hipSuccess) return fail(c, GC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return GC_OK;
}

}  // namespace

namespace gc {

gc_status set_error(const std::string& msg, gc_status code) { return fail(nullptr, code, msg); }

// One stream-ordered copy of the resident state block (x y z vx vy vz ax ay
// az m, each *cap doubles) into dst; ev, if given, is recorded after it.
// dst == nullptr only reports n and cap.
gc_status stage_state(gc_ctx* c, double* dst, size_t dst_doubles, size_t* n, size_t* cap,
                      hipEvent_t ev) {
    gc_status st = need_particles(c);
    if (st != GC_OK) return st;
    *n = c->s.n;
    *cap = c->s.cap;
    if (!dst) return GC_OK;
    if (dst_doubles < 10 * c->s.cap) return fail(c, GC_ERR_INVALID_ARGUMENT, "stage_state: buffer too small");
    if (c->s.cap > 0)
        GC_HIP(c, hipMemcpyAsync(dst, c->s.base, sizeof(double) * 10 * c->s.cap,
                                 hipMemcpyDeviceToHost, c->stream), "stage state");
    if (ev) GC_HIP(c, hipEventRecord(ev, c->stream), "stage state event");
    return GC_OK;
}

// The diagnostics kernels of the resident state, their 13 results copied
// (gc_diagnostics field order) into dst on the context's stream, then ev.
gc_status stage_diagnostics(gc_ctx* c, double* dst, hipEvent_t ev, size_t* n) {
    gc_status st = need_particles(c);
    if (st != GC_OK) return st;
    *n = c->s.n;
    if (c->s.n == 0) {
        std::memset(dst, 0, sizeof(gc_diagnostics));
    } else {
        launch_diagnostics(c->s, c->params.gravitational_constant, c->params.softening,
                           c->diag_scratch, c->diag_scratch_doubles, c->diag_out, c->stream);
        st = launch_status(c, "launch diagnostics kernels");
        if (st != GC_OK) return st;
        GC_HIP(c, hipMemcpyAsync(dst, c->diag_out, sizeof(gc_diagnostics), hipMemcpyDeviceToHost,
                                 c->stream), "stage diagnostics");
    }
    GC_HIP(c, hipEventRecord(ev, c->stream), "stage diagnostics event");
    return GC_OK;
}

}  // namespace gc

extern "C" {

void gc_default_params(gc_params* p) {
    if (!p) return;
    p->gravitational_constant = 1.0;
    p->softening = 0.02;
    p->solver = GC_SOLVER_DIRECT;
    p->precision = GC_PRECISION_FP64;
    p->theta = 0.6;
    p->leaf_capacity = 16;
    p->max_depth = 32;
    p->expansion_order = 4;
}

int gc_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int gc_solver_available(void) {
    const int n = gc_device_count();
    for (int d = 0; d < n; ++d)
        if (is_gfx950(d)) return 1;
    return 0;
}

const char* gc_status_string(gc_status s) {
    switch (s) {
    case GC_OK: return "ok";
    case GC_ERR_INVALID_ARGUMENT: return "invalid argument";
    case GC_ERR_NO_DEVICE: return "no gfx950 HIP device";
    case GC_ERR_HIP: return "HIP runtime error";
    case GC_ERR_OUT_OF_MEMORY: return "out of device memory";
    case GC_ERR_UNSUPPORTED: return "unsupported";
    case GC_ERR_STATE: return "invalid state";
    case GC_ERR_IO: return "output file error";
    }
    return "unknown status";
}

const char* gc_last_error(const gc_ctx* c) { return c ? c->err.c_str() : g_thread_err.c_str(); }

gc_status gc_create(gc_ctx** out, int device, const gc_params* p) {
    GC_ABI_BEGIN
    if (!out) return fail(nullptr, GC_ERR_INVALID_ARGUMENT, "gc_create: null output");
    *out = nullptr;
    gc_status st = check_params(nullptr, p);
    if (st != GC_OK) return st;
    const int count = gc_device_count();
    if (count <= 0) return fail(nullptr, GC_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= count)
        return fail(nullptr, GC_ERR_INVALID_ARGUMENT, "device ordinal out of range");
    if (!is_gfx950(device))
        return fail(nullptr, GC_ERR_NO_DEVICE, "device is not gfx950 (MI355X); kernels are built for gfx950 only");
    gc_ctx* c = new (std::nothrow) gc_ctx();
    if (!c) return fail(nullptr, GC_ERR_OUT_OF_MEMORY, "allocate context");
    c->device = device;
    c->params = *p;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->capture_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e == hipSuccess) e = hipMalloc(&c->diag_out, sizeof(double) * 16);
    if (e != hipSuccess) {
        std::string msg = std::string("create context: ") + hipGetErrorString(e);
        gc_destroy(c);
        return fail(nullptr, GC_ERR_HIP, msg);
    }
    c->stream = c->own_stream;
    *out = c;
    return GC_OK;
    GC_ABI_END("gc_create")
}

void gc_destroy(gc_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_step_graph(c);
    if (c->capture_stream) (void)hipStreamDestroy(c->capture_stream);
    delete c->fmm;
    free_state(c);
    (void)hipFree(c->vjp_lpack);
    (void)hipFree(c->vjp_part);
    (void)hipFree(c->vjp_sym_part);
    (void)hipFree(c->jp64);
    (void)hipFree(c->part64);
    (void)hipFree(c->diag_out);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
}

gc_status gc_set_params(gc_ctx* c, const gc_params* p) {
    GC_ABI_BEGIN
    if (!c) return fail(nullptr, GC_ERR_INVALID_ARGUMENT, "null context");
    gc_status st = check_params(c, p);
    if (st != GC_OK) return st;
    c->params = *p;
    return GC_OK;
    GC_ABI_END("gc_set_params")
}

gc_status gc_set_stream(gc_ctx* c, void* stream) {
    GC_ABI_BEGIN
    if (!c) return fail(nullptr, GC_ERR_INVALID_ARGUMENT, "null context");
    c->stream = static_cast<hipStream_t>(stream);
    return GC_OK;
    GC_ABI_END("gc_set_stream")
}

gc_status gc_use_own_stream(gc_ctx* c) {
    GC_ABI_BEGIN
    if (!c) return fail(nullptr, GC_ERR_INVALID_ARGUMENT, "null context");
    c->stream = c->own_stream;
    return GC_OK;
    GC_ABI_END("gc_use_own_stream")
}

void* gc_get_stream(gc_ctx* c) { return c ? static_cast<void*>(c->stream) : nullptr; }

gc_status gc_upload(gc_ctx* c, size_t n, const double* pos, const double* vel,
                    const double* mass) {
    GC_ABI_BEGIN
    gc_status st = need_particles(c);
    if (st != GC_OK) return st;
    if (n > 0 && (!pos || !mass)) return fail(c, GC_ERR_INVALID_ARGUMENT, "upload: null positions or masses");
    if (n > static_cast<size_t>(1) << 30) return fail(c, GC_ERR_INVALID_ARGUMENT, "upload: particle count exceeds 2^30");
    st = ensure_capacity(c, n);
    if (st != GC_OK) return st;
    if (n == 0) return GC_OK;
    const size_t b3 = sizeof(double) * 3 * n;
    GC_HIP(c, hipMemcpyAsync(c->stage, pos, b3, hipMemcpyHostToDevice, c->stream), "upload positions");
    launch_scatter_xyz(c->stage, c->s.x, c->s.y, c->s.z, n, c->stream);
    // the staging buffer is reused: order the copies on the stream
    GC_HIP(c, hipStreamSynchronize(c->stream), "upload positions");
    if (vel) {
        GC_HIP(c, hipMemcpyAsync(c->stage, vel, b3, hipMemcpyHostToDevice, c->stream), "upload velocities");
        launch_scatter_xyz(c->stage, c->s.vx, c->s.vy, c->s.vz, n, c->stream);
    } else {
        GC_HIP(c, hipMemsetAsync(c->s.vx, 0, sizeof(double) * 3 * c->s.cap, c->stream), "clear velocities");
    }
    GC_HIP(c, hipMemcpyAsync(c->s.m, mass, sizeof(double) * n, hipMemcpyHostToDevice, c->stream), "upload masses");
    GC_HIP(c, hipMemsetAsync(c->s.ax, 0, sizeof(double) * 3 * c->s.cap, c->stream), "clear accelerations");
    GC_HIP(c, hipStreamSynchronize(c->stream), "upload");
    return launch_status(c, "upload kernels");
    GC_ABI_END("gc_upload")
}

gc_status gc_set_accelerations(gc_ctx* c, const double* acc) {
    GC_ABI_BEGIN
    gc_status st = need_particles(c);
    if (st != GC_OK) return st;
    if (!acc) return fail(c, GC_ERR_INVALID_ARGUMENT, "set_accelerations: null array");
    const size_t n = c->s.n;
    if (n == 0) return GC_OK;
    GC_HIP(c, hipMemcpyAsync(c->stage, acc, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream),
           "upload accelerations");
    launch_scatter_xyz(c->stage, c->s.ax, c->s.ay, c->s.az, n, c->stream);
    GC_HIP(c, hipStreamSynchronize(c->stream), "upload accelerations");
    return launch_status(c, "upload accelerations");
    GC_ABI_END("gc_set_accelerations")
}

gc_status gc_download(gc_ctx* c, double* pos, double* vel, double* acc) {
    GC_ABI_BEGIN
    gc_status st = need_particles(c);
    if (st != GC_OK) return st;
    const size_t n = c->s.n;
    if (n == 0) return GC_OK;
    const size_t b3 = sizeof(double) * 3 * n;
    struct {
        double* host;
        double *x, *y, *z;
    } items[3] = {{pos, c->s.x, c->s.y, c->s.z}, {vel, c->s.vx, c->s.vy, c->s.vz},
                  {acc, c->s.ax, c->s.ay, c->s.az}};
    for (auto& it : items) {
        if (!it.host) continue;
        launch_gather_xyz(it.x, it.y, it.z, c->stage, n, c->stream);
        GC_HIP(c, hipMemcpyAsync(it.host, c->stage, b3, hipMemcpyDeviceToHost, c->stream), "download");
        GC_HIP(c, hipStreamSynchronize(c->stream), "download");
    }
    return launch_status(c, "download kernels");
    GC_ABI_END("gc_download")
}

size_t gc_particle_count(const gc_ctx* c) { return c ? c->s.n : 0; }

gc_status gc_device_state(gc_ctx* c, gc_device_view* v) {
    GC_ABI_BEGIN
    if (!c || !v) return fail(c, GC_ERR_INVALID_ARGUMENT, "device_state: null argument");
    v->n = c->s.n;
    v->x = c->s.x;
    v->y = c->s.y;
    v->z = c->s.z;
    v->vx = c->s.vx;
    v->vy = c->s.vy;
    v->vz = c->s.vz;
    v->ax = c->s.ax;
    v->ay = c->s.ay;
    v->az = c->s.az;
    v->mass = c->s.m;
    v->stream = c->stream;
    return GC_OK;
    GC_ABI_END("gc_device_state")
}

gc_status gc_compute_accelerations(gc_ctx* c) {
    GC_ABI_BEGIN
    gc_status st = need_particles(c);
    if (st != GC_OK) return st;
    if (c->s.n == 0) return GC_OK;
    return force(c, 0, c->s.n, false, 0.0);
    GC_ABI_END("gc_compute_accelerations")
}

gc_status gc_compute_accelerations_for_targets(gc_ctx* c, size_t begin, size_t end) {
    GC_ABI_BEGIN
    gc_status st = need_particles(c);
    if (st != GC_OK) return st;
    end = std::min(end, c->s.n);
    if (begin >= end) return GC_OK;
    return force(c, begin, end, false, 0.0);
    GC_ABI_END("gc_compute_accelerations_for_targets")
}

gc_status gc_step(gc_ctx* c, double dt, int nsteps) {
    GC_ABI_BEGIN
    gc_status st = need_particles(c);
    if (st != GC_OK) return st;
    if (!(dt > 0.0)) return fail(c, GC_ERR_INVALID_ARGUMENT, "dt must be positive");
    if (nsteps < 0) return fail(c, GC_ERR_INVALID_ARGUMENT, "steps must be non-negative");
    if (c->s.n == 0 || nsteps == 0) return GC_OK;
    const double half = 0.5 * dt;
    if (c->params.solver == GC_SOLVER_DIRECT && (st = prepare_direct(c)) != GC_OK) return st;
    const int chunk = c->graph_steps;
    const bool graph = use_step_graph(c);
    if (c->params.solver == GC_SOLVER_DIRECT && c->params.precision == GC_PRECISION_MIXED &&
        symmetric_pairs(c, 0, c->s.n) && nsteps > 1) {
        // each step's merge also opens the next step and packs its sources;
        // the middle steps (force + that merge) are what a chunk repeats
        launch_open_step(c->s, half, dt, c->stream);
        launch_pack_mixed(c->s, c->mixed, c->params.gravitational_constant, c->stream);
        launch_sym_weights(c->s, c->mixed, c->sym, c->params.softening, c->stream);
        int middle = nsteps - 1;
        if (graph && middle >= chunk &&
            replay_chunks(c, dt, middle / chunk, [&] {
                launch_sym_middle(c, dt, chunk);
                return GC_OK;
            }, &st)) {
            if (st != GC_OK) return st;
            middle %= chunk;
        }
        launch_sym_middle(c, dt, middle);
        launch_direct_sym(c->s, c->mixed, c->sym, c->params.softening, sym_ipt(c), c->stream, true);
        launch_merge_sym(c->s, c->sym, nullptr, true, half, c->stream);
        return launch_status(c, "launch step kernels");
    }
    int left = nsteps;
    if (graph && left >= chunk &&
        replay_chunks(c, dt, left / chunk, [&] { return launch_plain_steps(c, dt, chunk); }, &st)) {
        if (st != GC_OK) return st;
        left %= chunk;
    }
    if ((st = launch_plain_steps(c, dt, left)) != GC_OK) return st;
    return launch_status(c, "launch step kernels");
    GC_ABI_END("gc_step")
}

gc_status gc_set_graph_steps(gc_ctx* c, int steps) {
    GC_ABI_BEGIN
    if (!c) return fail(nullptr, GC_ERR_INVALID_ARGUMENT, "null context");
    if (steps < 0 || steps > 1000)
        return fail(c, GC_ERR_INVALID_ARGUMENT, "graph steps must be in [0, 1000]");
    GC_HIP(c, hipSetDevice(c->device), "select device");
    if (steps != c->graph_steps) drop_step_graph(c);
    c->graph_steps = steps;
    return GC_OK;
    GC_ABI_END("gc_set_graph_steps")
}

int gc_get_graph_steps(const gc_ctx* c) { return c ? c->graph_steps : -1; }

gc_status gc_generate_galaxies_device(gc_ctx* c, const gc_galaxy_config* galaxies, int count,
                                      uint64_t seed) {
    GC_ABI_BEGIN
    if (!c) return fail(nullptr, GC_ERR_INVALID_ARGUMENT, "null context");
    if (count < 0 || (count > 0 && !galaxies))
        return fail(c, GC_ERR_INVALID_ARGUMENT, "generate_galaxies_device: bad galaxy list");
    std::vector<IcSegment> segs;
    size_t n = 0;
    for (int k = 0; k < count; ++k) {
        if (galaxies[k].n_particles > kMaxIcParticles)  // bounds the sum below
            return fail(c, GC_ERR_INVALID_ARGUMENT, "generate_galaxies_device: particle count exceeds 2^30");
        if (ic::ref_disk_empty(galaxies[k])) continue;
        IcSegment sg;
        sg.model = kIcRefDisk;
        sg.stream = static_cast<unsigned>(k);
        sg.seed = seed;
        sg.begin = sg.dst = static_cast<long long>(n);
        sg.count = static_cast<long long>(galaxies[k].n_particles);
        ic::ref_disk_params(galaxies[k], c->params.gravitational_constant, c->params.softening, sg.p);
        segs.push_back(sg);
        n += galaxies[k].n_particles;
    }
    return generate_state(c, segs, n, "generate_galaxies_device");
    GC_ABI_END("gc_generate_galaxies_device")
}

gc_status gc_generate_disks_device(gc_ctx* c, const gc_disk_galaxy* galaxies, int count,
                                   uint64_t seed) {
    GC_ABI_BEGIN
    if (!c) return fail(nullptr, GC_ERR_INVALID_ARGUMENT, "null context");
    if (count < 0 || (count > 0 && !galaxies))
        return fail(c, GC_ERR_INVALID_ARGUMENT, "generate_disks_device: bad galaxy list");
    std::vector<IcSegment> segs;
    size_t n = 0;
    for (int k = 0; k < count; ++k) {
        if (galaxies[k].n_particles > kMaxIcParticles)  // bounds the sum below
            return fail(c, GC_ERR_INVALID_ARGUMENT, "generate_disks_device: particle count exceeds 2^30");
        if (galaxies[k].n_particles == 0) continue;
        IcSegment sg;
        sg.model = kIcTwoDisk;
        sg.stream = static_cast<unsigned>(k);
        sg.seed = seed;
        sg.begin = sg.dst = static_cast<long long>(n);
        sg.count = static_cast<long long>(galaxies[k].n_particles);
        ic::two_disk_params(galaxies[k], c->params.gravitational_constant, c->params.softening, sg.p);
        segs.push_back(sg);
        n += galaxies[k].n_particles;
    }
    return generate_state(c, segs, n, "generate_disks_device");
    GC_ABI_END("gc_generate_disks_device")
}

gc_status gc_generate_plummer_device(gc_ctx* c, size_t n, double total_mass, double scale,
                                     uint64_t seed) {
    GC_ABI_BEGIN
    if (!c) return fail(nullptr, GC_ERR_INVALID_ARGUMENT, "null context");
    if (!(scale > 0.0)) return fail(c, GC_ERR_INVALID_ARGUMENT, "generate_plummer_device: scale must be positive");
    std::vector<IcSegment> segs;
    if (n > 0) {
        IcSegment sg;
        sg.model = kIcPlummer;
        sg.seed = seed;
        sg.count = static_cast<long long>(n);
        sg.p[ic::kPlN] = static_cast<double>(n);
        sg.p[ic::kPlMass] = total_mass;
        sg.p[ic::kPlScale] = scale;
        sg.p[ic::kPlG] = c->params.gravitational_constant;
        segs.push_back(sg);
    }
    return generate_state(c, segs, n, "generate_plummer_device");
    GC_ABI_END("gc_generate_plummer_device")
}

gc_status gc_synchronize(gc_ctx* c) {
    GC_ABI_BEGIN
    if (!c) return fail(nullptr, GC_ERR_INVALID_ARGUMENT, "null context");
    GC_HIP(c, hipStreamSynchronize(c->stream), "synchronize");
    return GC_OK;
    GC_ABI_END("gc_synchronize")
}

gc_status gc_diagnostics_compute(gc_ctx* c, gc_diagnostics* out) {
    GC_ABI_BEGIN
    gc_status st = need_particles(c);
    if (st != GC_OK) return st;
    if (!out) return fail(c, GC_ERR_INVALID_ARGUMENT, "diagnostics: null output");
    std::memset(out, 0, sizeof(*out));
    if (c->s.n == 0) return GC_OK;
    launch_diagnostics(c->s, c->params.gravitational_constant, c->params.softening,
                       c->diag_scratch, c->diag_scratch_doubles, c->diag_out, c->stream);
    st = launch_status(c, "launch diagnostics kernels");
    if (st != GC_OK) return st;
    double host[13];
    GC_HIP(c, hipMemcpyAsync(host, c->diag_out, sizeof(host), hipMemcpyDeviceToHost, c->stream), "read diagnostics");
    GC_HIP(c, hipStreamSynchronize(c->stream), "diagnostics");
    out->kinetic_energy = host[0];
    out->potential_energy = host[1];
    out->total_energy = host[2];
    out->total_mass = host[3];
    for (int k = 0; k < 3; ++k) {
        out->momentum[k] = host[4 + k];
        out->center_of_mass[k] = host[7 + k];
        out->angular_momentum[k] = host[10 + k];
    }
    return GC_OK;
    GC_ABI_END("gc_diagnostics_compute")
}

gc_status gc_time_force_kernel(gc_ctx* c, int reps, double* ms, double* inter) {
    GC_ABI_BEGIN
    gc_status st = need_particles(c);
    if (st != GC_OK) return st;
    if (reps <= 0 || !ms) return fail(c, GC_ERR_INVALID_ARGUMENT, "time_force_kernel: bad arguments");
    const size_t n = c->s.n;
    if (n == 0) return fail(c, GC_ERR_STATE, "time_force_kernel: no particles");
    if (c->params.solver == GC_SOLVER_FMM || c->params.solver == GC_SOLVER_TREE) {
        if (!c->fmm) c->fmm = new (std::nothrow) gc::FmmEngine();
        if (!c->fmm) return fail(c, GC_ERR_OUT_OF_MEMORY, "allocate FMM engine");
        std::string msg;
        const int rc = c->fmm->time_kernels(c->s, c->params, reps, c->stream, ms, inter, &msg);
        if (rc != GC_OK) return fail(c, static_cast<gc_status>(rc), "hip-fmm timing: " + msg);
        return GC_OK;
    }
    int ipt = 1, js = 1, per = 8;
    const bool sym = c->params.precision == GC_PRECISION_MIXED && symmetric_pairs(c, 0, n);
    const bool fast64 = c->params.precision == GC_PRECISION_FP64_FAST;
    if (fast64) {
        fp64_fast_shape(n, n, c->shape.ipt, c->shape.jsplit, &ipt, &js, &per);
        st = ensure_fp64_fast(c, js);
        if (st != GC_OK) return st;
        launch_pack_fp64(c->s, c->jp64, c->params.gravitational_constant, c->stream);
    }
    if (c->params.precision == GC_PRECISION_MIXED) {
        if (sym) {
            st = ensure_sym(c);
        } else {
            mixed_shape(c, n, &ipt, &js);
            st = ensure_partials(c, js);
        }
        if (st != GC_OK) return st;
        launch_pack_mixed(c->s, c->mixed, c->params.gravitational_constant, c->stream);
        // the tile weights once, as the stepping merge keeps them: the
        // timed launches are the force kernel's alone
        if (sym) launch_sym_weights(c->s, c->mixed, c->sym, c->params.softening, c->stream);
    }
    GC_HIP(c, hipEventRecord(c->ev0, c->stream), "record start event");
    for (int r = 0; r < reps; ++r) {
        if (sym) {
            launch_direct_sym(c->s, c->mixed, c->sym, c->params.softening, sym_ipt(c),
                              c->stream, true);
        } else if (fast64) {
            launch_direct_fp64_fast(c->s, c->jp64, c->part64, c->params.softening, ipt, js, per, 0,
                                    n, c->stream);
        } else if (c->params.precision == GC_PRECISION_MIXED) {
            launch_direct_mixed(c->s, c->mixed, c->params.softening, ipt, js, scalar_sources(c), 0, n,
                                c->stream);
        } else {
            launch_direct_fp64(c->s, c->params.gravitational_constant, c->params.softening, 0, n,
                               false, 0.0, c->stream);
        }
    }
    GC_HIP(c, hipEventRecord(c->ev1, c->stream), "record stop event");
    GC_HIP(c, hipEventSynchronize(c->ev1), "wait stop event");
    st = launch_status(c, "timed force kernels");
    if (st != GC_OK) return st;
    float total = 0.f;
    GC_HIP(c, hipEventElapsedTime(&total, c->ev0, c->ev1), "elapsed time");
    *ms = static_cast<double>(total) / reps;
    if (inter) *inter = static_cast<double>(n) * static_cast<double>(n - 1);
    if (sym) {
        // restore accelerations consistent with the state
        launch_merge_sym(c->s, c->sym, nullptr, false, 0.0, c->stream);
    } else if (fast64) {
        launch_merge_fp64(c->s, c->part64, js, 0, n, false, 0.0, c->stream);
    } else if (c->params.precision == GC_PRECISION_MIXED) {
        launch_merge_mixed(c->s, c->mixed, js, 0, n, false, 0.0, c->stream);
    }
    return launch_status(c, "timed force kernels");
    GC_ABI_END("gc_time_force_kernel")
}

gc_status gc_set_state_device(gc_ctx* c, size_t n, const double* pos, const double* vel,
                              const double* mass, const double* acc) {
    GC_ABI_BEGIN
    gc_status st = need_particles(c);
    if (st != GC_OK) return st;
    if (n > 0 && (!pos || !mass)) return fail(c, GC_ERR_INVALID_ARGUMENT, "set_state_device: null positions or masses");
    if (n > static_cast<size_t>(1) << 30) return fail(c, GC_ERR_INVALID_ARGUMENT, "set_state_device: too many particles");
    st = ensure_capacity(c, n);
    if (st != GC_OK) return st;
    if (n == 0) return GC_OK;
    const size_t b = sizeof(double) * n;
    double* dst[3][3] = {{c->s.x, c->s.y, c->s.z}, {c->s.vx, c->s.vy, c->s.vz}, {c->s.ax, c->s.ay, c->s.az}};
    const double* src[3] = {pos, vel, acc};
    for (int k = 0; k < 3; ++k) {
        for (int d = 0; d < 3; ++d) {
            if (src[k]) {
                GC_HIP(c, hipMemcpyAsync(dst[k][d], src[k] + d * n, b, hipMemcpyDeviceToDevice, c->stream),
                       "copy device state in");
            } else {
                GC_HIP(c, hipMemsetAsync(dst[k][d], 0, b, c->stream), "clear device state");
            }
        }
    }
    GC_HIP(c, hipMemcpyAsync(c->s.m, mass, b, hipMemcpyDeviceToDevice, c->stream), "copy masses in");
    return GC_OK;
    GC_ABI_END("gc_set_state_device")
}

gc_status gc_get_state_device(gc_ctx* c, double* pos, double* vel, double* acc) {
    GC_ABI_BEGIN
    gc_status st = need_particles(c);
    if (st != GC_OK) return st;
    const size_t n = c->s.n;
    if (n == 0) return GC_OK;
    const size_t b = sizeof(double) * n;
    const double* src[3][3] = {{c->s.x, c->s.y, c->s.z}, {c->s.vx, c->s.vy, c->s.vz}, {c->s.ax, c->s.ay, c->s.az}};
    double* dst[3] = {pos, vel, acc};
    for (int k = 0; k < 3; ++k) {
        if (!dst[k]) continue;
        for (int d = 0; d < 3; ++d)
            GC_HIP(c, hipMemcpyAsync(dst[k] + d * n, src[k][d], b, hipMemcpyDeviceToDevice, c->stream),
                   "copy device state out");
    }
    return GC_OK;
    GC_ABI_END("gc_get_state_device")
}

gc_status gc_step_reverse(gc_ctx* c, double dt, int nsteps) {
    GC_ABI_BEGIN
    gc_status st = need_particles(c);
    if (st != GC_OK) return st;
    if (!(dt > 0.0)) return fail(c, GC_ERR_INVALID_ARGUMENT, "dt must be positive");
    if (nsteps < 0) return fail(c, GC_ERR_INVALID_ARGUMENT, "steps must be non-negative");
    if (c->s.n == 0) return GC_OK;
    const double half = -0.5 * dt;
    for (int k = 0; k < nsteps; ++k) {
        launch_open_step(c->s, half, -dt, c->stream);
        st = force(c, 0, c->s.n, true, half);
        if (st != GC_OK) return st;
    }
    return launch_status(c, "launch reverse step kernels");
    GC_ABI_END("gc_step_reverse")
}

gc_status gc_force_vjp(gc_ctx* c, const double* lambda, double* xbar, double* mbar,
                       int accumulate) {
    GC_ABI_BEGIN
    gc_status st = need_particles(c);
    if (st != GC_OK) return st;
    if (!lambda || !xbar || !mbar) return fail(c, GC_ERR_INVALID_ARGUMENT, "force_vjp: null array");
    if (c->params.solver != GC_SOLVER_DIRECT)
        return fail(c, GC_ERR_UNSUPPORTED, "force_vjp: only the direct solver has an adjoint");
    const size_t n = c->s.n;
    if (n == 0) return GC_OK;
    const int stride = static_cast<int>(n);
    if (c->params.precision != GC_PRECISION_MIXED) {
        launch_vjp_fp64(c->s, c->params.gravitational_constant, c->params.softening, lambda, stride,
                        xbar, mbar, stride, accumulate != 0, c->stream);
        return launch_status(c, "launch FP64 force VJP");
    }
    if (symmetric_pairs(c, 0, n)) {
        // each unordered pair once, on the force's work list
        st = ensure_sym(c);
        if (st != GC_OK) return st;
        const size_t ifl = static_cast<size_t>(c->sym.nitems) * 8 * kTile;
        const size_t jfl = c->sym.jpart_floats / 3 * 4;  // 4 sums per source, not 3
        if (c->vjp_lpack_n < c->s.cap || c->vjp_sym_part_n < ifl + jfl) {
            GC_HIP(c, hipStreamSynchronize(c->stream), "synchronize before VJP scratch");
            (void)hipFree(c->vjp_lpack);
            (void)hipFree(c->vjp_sym_part);
            c->vjp_lpack = nullptr;
            c->vjp_sym_part = nullptr;
            c->vjp_lpack_n = c->vjp_sym_part_n = 0;
            GC_HIP(c, hipMalloc(&c->vjp_lpack, sizeof(float4) * c->s.cap), "allocate VJP scratch");
            GC_HIP(c, hipMalloc(&c->vjp_sym_part, sizeof(float) * (ifl + jfl)),
                   "allocate VJP scratch");
            c->vjp_lpack_n = c->s.cap;
            c->vjp_sym_part_n = ifl + jfl;
        }
        launch_pack_mixed(c->s, c->mixed, c->params.gravitational_constant, c->stream);
        launch_pack_lambda(n, c->mixed.ntiles * kTile, lambda, stride, c->vjp_lpack, c->stream);
        const int ipt = (c->shape.ipt == 1 || c->shape.ipt == 2 || c->shape.ipt == 4) ? c->shape.ipt
                                                                                        : kVjpSymIpt;
        launch_vjp_sym(c->s, c->mixed, c->sym, c->vjp_lpack, c->vjp_sym_part,
                       c->vjp_sym_part + ifl, c->params.gravitational_constant, c->params.softening,
                       ipt, xbar, mbar, stride, accumulate != 0, c->stream);
        return launch_status(c, "launch pair-once force VJP");
    }
    const int js = vjp_mixed_splits(n);
    const size_t need_part = static_cast<size_t>(js) * 4 * c->s.cap;
    if (c->vjp_lpack_n < c->s.cap || c->vjp_part_n < need_part) {
        GC_HIP(c, hipStreamSynchronize(c->stream), "synchronize before VJP scratch");
        (void)hipFree(c->vjp_lpack);
        (void)hipFree(c->vjp_part);
        c->vjp_lpack = nullptr;
        c->vjp_part = nullptr;
        c->vjp_lpack_n = c->vjp_part_n = 0;
        GC_HIP(c, hipMalloc(&c->vjp_lpack, sizeof(float4) * c->s.cap), "allocate VJP scratch");
        GC_HIP(c, hipMalloc(&c->vjp_part, sizeof(double) * need_part), "allocate VJP scratch");
        c->vjp_lpack_n = c->s.cap;
        c->vjp_part_n = need_part;
    }
    launch_vjp_mixed(c->s, c->mixed, c->vjp_lpack, c->vjp_part, js, c->params.gravitational_constant,
                     c->params.softening, lambda, stride, xbar, mbar, stride, accumulate != 0,
                     c->stream);
    return launch_status(c, "launch mixed force VJP");
    GC_ABI_END("gc_force_vjp")
}

gc_status gc_fmm_last_stats(gc_ctx* c, gc_fmm_stats* out) {
    GC_ABI_BEGIN
    if (!c || !out) return fail(c, GC_ERR_INVALID_ARGUMENT, "fmm_last_stats: null argument");
    if (!c->fmm) return fail(c, GC_ERR_STATE, "fmm_last_stats: no hip-fmm evaluation yet");
    *out = c->fmm->stats();
    float b = 0.f, e = 0.f;
    if (c->fmm->phase_times(&b, &e) == GC_OK) {
        out->build_ms = b;
        out->eval_ms = e;
    }
    return GC_OK;
    GC_ABI_END("gc_fmm_last_stats")
}

gc_status gc_calibration_copy(int device, size_t bytes, int reps, double* ms) {
    GC_ABI_BEGIN
    if (!ms || bytes == 0 || bytes % 16 != 0 || reps <= 0)
        return fail(nullptr, GC_ERR_INVALID_ARGUMENT, "calibration_copy: bad arguments");
    if (device < 0 || device >= gc_device_count())
        return fail(nullptr, GC_ERR_NO_DEVICE, "calibration_copy: no such device");
    const hipError_t e = run_calibration_copy(device, bytes, reps, ms);
    if (e != hipSuccess)
        return fail(nullptr, e == hipErrorOutOfMemory ? GC_ERR_OUT_OF_MEMORY : GC_ERR_HIP,
                    std::string("calibration_copy: ") + hipGetErrorString(e));
    return GC_OK;
    GC_ABI_END("gc_calibration_copy")
}

gc_status gc_set_launch_shape(gc_ctx* c, int ipt, int jsplit) {
    GC_ABI_BEGIN
    if (!c) return fail(nullptr, GC_ERR_INVALID_ARGUMENT, "null context");
    if (ipt != 0 && ipt != 1 && ipt != 2 && ipt != 4)
        return fail(c, GC_ERR_INVALID_ARGUMENT, "targets per thread must be 0, 1, 2 or 4");
    if (jsplit < 0 || jsplit > 1024) return fail(c, GC_ERR_INVALID_ARGUMENT, "jsplit out of range");
    c->shape.ipt = ipt;
    c->shape.jsplit = jsplit;
    return GC_OK;
    GC_ABI_END("gc_set_launch_shape")
}

gc_status gc_get_launch_shape(gc_ctx* c, int* ipt, int* jsplit) {
    GC_ABI_BEGIN
    if (!c || !ipt || !jsplit) return fail(c, GC_ERR_INVALID_ARGUMENT, "get_launch_shape: null argument");
    if (c->params.precision == GC_PRECISION_FP64_FAST) {
        int per = 0;
        fp64_fast_shape(c->s.n, c->s.n, c->shape.ipt, c->shape.jsplit, ipt, jsplit, &per);
        return GC_OK;
    }
    if (c->params.precision == GC_PRECISION_FP64) {  // one target per lane, no splits
        *ipt = 1;
        *jsplit = 0;
        return GC_OK;
    }
    if (c->s.n > 0 && symmetric_pairs(c, 0, c->s.n)) {
        gc_status st = ensure_sym(c);
        if (st != GC_OK) return st;
        *ipt = sym_ipt(c);
        *jsplit = 0;
        return GC_OK;
    }
    mixed_shape(c, c->s.n, ipt, jsplit);
    return GC_OK;
    GC_ABI_END("gc_get_launch_shape")
}

gc_status gc_set_pair_items(gc_ctx* c, int tiles) {
    GC_ABI_BEGIN
    if (!c) return fail(nullptr, GC_ERR_INVALID_ARGUMENT, "null context");
    if (tiles < 0 || tiles > kSymMaxTiles)
        return fail(c, GC_ERR_INVALID_ARGUMENT, "source tiles per item out of range");
    c->shape.items = tiles;
    return GC_OK;
    GC_ABI_END("gc_set_pair_items")
}

gc_status gc_set_sources(gc_ctx* c, int mode) {
    GC_ABI_BEGIN
    if (!c) return fail(nullptr, GC_ERR_INVALID_ARGUMENT, "null context");
    if (mode != GC_SOURCES_AUTO && mode != GC_SOURCES_LDS && mode != GC_SOURCES_SCALAR)
        return fail(c, GC_ERR_INVALID_ARGUMENT, "source mode must be 0, 1 or 2");
    c->shape.sources = mode;
    return GC_OK;
    GC_ABI_END("gc_set_sources")
}

int gc_get_sources(const gc_ctx* c) {
    if (!c) return -1;
    return scalar_sources(c) ? GC_SOURCES_SCALAR : GC_SOURCES_LDS;
}

gc_status gc_set_pairs(gc_ctx* c, int mode) {
    GC_ABI_BEGIN
    if (!c) return fail(nullptr, GC_ERR_INVALID_ARGUMENT, "null context");
    if (mode != GC_PAIRS_AUTO && mode != GC_PAIRS_TARGETS && mode != GC_PAIRS_SYMMETRIC)
        return fail(c, GC_ERR_INVALID_ARGUMENT, "pair mode must be 0, 1 or 2");
    c->shape.pairs = mode;
    return GC_OK;
    GC_ABI_END("gc_set_pairs")
}

int gc_get_pairs(const gc_ctx* c) {
    if (!c) return -1;
    if (c->params.precision != GC_PRECISION_MIXED) return GC_PAIRS_TARGETS;
    return symmetric_pairs(c, 0, c->s.n) ? GC_PAIRS_SYMMETRIC : GC_PAIRS_TARGETS;
}

}  // extern "C"
