// batch.hip — many independent collisions of uneven N advanced together.
//
// Layout: runs are concatenated, each padded to a multiple of kTile, so no
// source tile spans two runs; padding slots have zero mass and are never
// targets.  Per step (all in one stream, capturable):
//   open   kick dt/2 + drift dt on every slot (padding stays at rest),
//   pack   per-tile FP64 centre + FP32 offsets (valid slots only),
//   force  ONE persistent launch: `blocks` workgroups walk a static list of
//          work items (target tile, source-tile chunk of `chunk` tiles),
//          longest first, item = blockIdx.x + k * gridDim.x; each item runs
//          the same tile sweep as the single-run kernel and writes its
//          (hi, lo) partials to its own slot,
//   merge  per particle: partials of its tile's chunks in chunk order (FP64),
//          closing kick dt/2.
// Because the sweep, the tile centres and the chunk order equal those of a
// single-run launch whose source split is `chunk` tiles, a batched run is
// bit-identical to that run stepped alone.  Multi-step chunks are captured
// once into a hipGraph and replayed.  The (chunk, blocks) schedule is chosen
// by timing candidates on the batch itself (gc_batch_autotune).
#include "gc_internal.hpp"
#include "abi_guard.hpp"
#include "ic_models.hpp"
#include "mixed_tiles.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <map>
#include <mutex>
#include <new>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

namespace gc {
namespace {

__global__ __launch_bounds__(256) void batch_open_kernel(
    int total, double* __restrict__ x, double* __restrict__ y, double* __restrict__ z,
    double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
    const double* __restrict__ ax, const double* __restrict__ ay, const double* __restrict__ az,
    double half_dt, double dt) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= total) return;
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

template <typename T>
__device__ __forceinline__ T wmin(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

template <typename T>
__device__ __forceinline__ T wmax(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Same arithmetic as pack_mixed_kernel, with the valid count per tile.
__global__ __launch_bounds__(kTile) void batch_pack_kernel(
    const int* __restrict__ tile_valid, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, const double* __restrict__ m,
    double G, int ntiles, float4* __restrict__ jp, double* __restrict__ centre) {
    __shared__ double red[6][kTile / 64];
    const int t = threadIdx.x;
    const int i = blockIdx.x * kTile + t;
    const bool valid = t < tile_valid[blockIdx.x];
    const double px = valid ? x[i] : 0.0, py = valid ? y[i] : 0.0, pz = valid ? z[i] : 0.0;
    double v[6] = {valid ? px : INFINITY, valid ? py : INFINITY, valid ? pz : INFINITY,
                   valid ? px : -INFINITY, valid ? py : -INFINITY, valid ? pz : -INFINITY};
    for (int d = 0; d < 3; ++d) {
        v[d] = wmin(v[d]);
        v[3 + d] = wmax(v[3 + d]);
    }
    const int w = t >> 6;
    if ((t & 63) == 0)
        for (int k = 0; k < 6; ++k) red[k][w] = v[k];
    __syncthreads();
    double c[3];
    for (int d = 0; d < 3; ++d) {
        double lo = red[d][0], hi = red[3 + d][0];
        for (int k = 1; k < kTile / 64; ++k) {
            lo = fmin(lo, red[d][k]);
            hi = fmax(hi, red[3 + d][k]);
        }
        c[d] = (lo + hi) * 0.5;
    }
    jp[i] = valid ? make_float4(static_cast<float>(px - c[0]), static_cast<float>(py - c[1]),
                                static_cast<float>(pz - c[2]), static_cast<float>(G * m[i]))
                  : make_float4(0.f, 0.f, 0.f, 0.f);
    if (t == 0) {
        centre[blockIdx.x] = c[0];
        centre[ntiles + blockIdx.x] = c[1];
        centre[2 * ntiles + blockIdx.x] = c[2];
    }
}

struct Item {
    int target_tile, j0, j1, slot;
};

template <bool SAFE>
__global__ __launch_bounds__(kTile) void batch_force_kernel(
    int nitems, const int4* __restrict__ items, const int* __restrict__ tile_valid,
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const float4* __restrict__ jp, const double* __restrict__ centre, int ntiles, float eps2,
    float* __restrict__ part) {
    __shared__ float4 tile[kTile];
    const int t = threadIdx.x;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const int4 w = items[it];  // target tile, j0, j1, slot
        const int i = w.x * kTile + t;
        const bool a = t < tile_valid[w.x];
        double xi[1] = {a ? x[i] : 0.0}, yi[1] = {a ? y[i] : 0.0}, zi[1] = {a ? z[i] : 0.0};
        float hx[1] = {0.f}, hy[1] = {0.f}, hz[1] = {0.f}, lx[1] = {0.f}, ly[1] = {0.f}, lz[1] = {0.f};
        tile_sweep<1, SAFE>(tile, w.y, w.z, jp, centre, ntiles, eps2, xi, yi, zi, hx, hy, hz, lx,
                            ly, lz);
        float* b = part + static_cast<size_t>(w.w) * 6 * kTile;
        b[t] = hx[0];
        b[kTile + t] = hy[0];
        b[2 * kTile + t] = hz[0];
        b[3 * kTile + t] = lx[0];
        b[4 * kTile + t] = ly[0];
        b[5 * kTile + t] = lz[0];
    }
}

__global__ __launch_bounds__(256) void batch_merge_kernel(
    int total, const int* __restrict__ tile_valid, const int* __restrict__ tile_slot0,
    const int* __restrict__ tile_nchunk, const float* __restrict__ part, double* __restrict__ ax,
    double* __restrict__ ay, double* __restrict__ az, double* __restrict__ vx,
    double* __restrict__ vy, double* __restrict__ vz, int close_kick, double half_dt) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= total) return;
    const int tl = i / kTile, l = i % kTile;
    if (l >= tile_valid[tl]) return;
    double a[3] = {0.0, 0.0, 0.0};
    const int s0 = tile_slot0[tl];
    for (int c = 0; c < tile_nchunk[tl]; ++c) {
        const float* b = part + static_cast<size_t>(s0 + c) * 6 * kTile;
        for (int d = 0; d < 3; ++d)
            a[d] += static_cast<double>(b[d * kTile + l]) + static_cast<double>(b[(d + 3) * kTile + l]);
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

// process-wide autotune cache: (runs, tiles, items at chunk 1, requested
// pair form) -> tuned schedule
struct Tuned {
    int form;                  // GC_PAIRS_TARGETS or GC_PAIRS_SYMMETRIC
    int chunk, blocks;         // per-target form
    int ipt, sym_blocks, sym_len;  // symmetric form
};
std::mutex g_tune_mutex;
std::map<std::tuple<int, int, long long, int>, Tuned> g_tune_cache;

// The autotune results also persist in a text file, one line per result:
//   gcsim-tune 1 <arch> <CUs> <runs> <tiles> <items at chunk 1> <requested
//   form> <form> <chunk> <blocks> <ipt> <symmetric blocks> <item length>
// so a new process on the same device model reuses them.  The file is
// $GCSIM_TUNE_CACHE, else $HOME/.cache/gcsim/batch_tune.txt; an empty
// GCSIM_TUNE_CACHE turns the file off.  The last line of a key wins; lines
// that do not parse, or hold values out of range, are ignored.
std::string tune_file_path() {
    if (const char* e = std::getenv("GCSIM_TUNE_CACHE")) return e;
    const char* h = std::getenv("HOME");
    if (!h || !*h) return std::string();
    return std::string(h) + "/.cache/gcsim/batch_tune.txt";
}

bool tune_file_lookup(const std::string& arch, int cus,
                      const std::tuple<int, int, long long, int>& key, Tuned* out) {
    const std::string path = tune_file_path();
    if (path.empty()) return false;
    std::ifstream f(path);
    if (!f) return false;
    std::string line;
    bool found = false;
    while (std::getline(f, line)) {
        char tag[32] = {0}, a[128] = {0};
        int ver = 0, c = 0, runs = 0, tiles = 0, req = 0;
        long long items = 0;
        Tuned t{};
        const int got = std::sscanf(line.c_str(), "%31s %d %127s %d %d %d %lld %d %d %d %d %d %d %d", tag,
                                    &ver, a, &c, &runs, &tiles, &items, &req, &t.form, &t.chunk,
                                    &t.blocks, &t.ipt, &t.sym_blocks, &t.sym_len);
        if (got != 14 || std::strcmp(tag, "gcsim-tune") != 0 || ver != 1) continue;
        if (arch != a || c != cus || std::make_tuple(runs, tiles, items, req) != key) continue;
        const bool ok = (t.form == GC_PAIRS_TARGETS || t.form == GC_PAIRS_SYMMETRIC) && t.chunk >= 1 &&
                        t.blocks >= 0 && (t.ipt == 0 || t.ipt == 1 || t.ipt == 2 || t.ipt == 4) &&
                        t.sym_blocks >= 0 && t.sym_len >= 0;
        if (!ok) continue;
        *out = t;
        found = true;
    }
    return found;
}

void tune_file_store(const std::string& arch, int cus,
                     const std::tuple<int, int, long long, int>& key, const Tuned& t) {
    const std::string path = tune_file_path();
    if (path.empty()) return;
    std::error_code ec;
    const std::filesystem::path dir = std::filesystem::path(path).parent_path();
    if (!dir.empty()) std::filesystem::create_directories(dir, ec);
    std::ofstream f(path, std::ios::app);
    if (!f) return;  // a cache that cannot be written is only a cache
    f << "gcsim-tune 1 " << arch << ' ' << cus << ' ' << std::get<0>(key) << ' ' << std::get<1>(key)
      << ' ' << std::get<2>(key) << ' ' << std::get<3>(key) << ' ' << t.form << ' ' << t.chunk << ' '
      << t.blocks << ' ' << t.ipt << ' ' << t.sym_blocks << ' ' << t.sym_len << '\n';
}

}  // namespace
}  // namespace gc

using namespace gc;

struct gc_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    gc_params params{};
    int runs = 0;
    std::vector<size_t> counts, offsets;  // per run, offsets in padded slots
    std::vector<int> run_tile0, run_ntiles;
    int ntiles = 0, total = 0;  // padded slots
    DeviceState s;              // n = cap = total
    float4* jp = nullptr;
    double* centre = nullptr;
    int* tile_valid = nullptr;
    int* tile_slot0 = nullptr;
    int* tile_nchunk = nullptr;
    int4* items = nullptr;
    int nitems = 0;
    float* part = nullptr;
    size_t part_floats = 0;
    int chunk = 4, blocks = 1024, graph_steps = 10;
    // each unordered pair once within a run (direct_sym.hip): pair form,
    // targets per lane (0 = automatic), persistent blocks (0 = one per item)
    int pairs = GC_PAIRS_AUTO, sym_ipt = 0, sym_blocks = 0;
    int sym_len = 0, sym_built_len = -1;  // J tiles per item (0: default), plan's
    int tuned_form = 0;  // AUTO after autotune: GC_PAIRS_TARGETS or _SYMMETRIC
    SymPlan sym;
    hipGraphExec_t graph = nullptr;
    double graph_dt = 0.0;
    double* stage = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
};

namespace {

thread_local std::string g_batch_err;

gc_status bfail(gc_batch* b, gc_status code, const std::string& msg) {
    if (b) b->err = msg;
    g_batch_err = msg;
    return code;
}

#define GB_HIP(b, call, what)                                                                    \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return bfail(b, e_ == hipErrorOutOfMemory ? GC_ERR_OUT_OF_MEMORY : GC_ERR_HIP,       \
                         std::string(what) + ": " + hipGetErrorString(e_));                     \
    } while (0)

void drop_graph(gc_batch* b) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
}

// The symmetric form unless the per-target form is asked for (or chosen by
// the autotuner) or a run has more than kSymMaxTiles tiles.
bool batch_symmetric(const gc_batch* b) {
    if (b->pairs == GC_PAIRS_TARGETS) return false;
    if (b->pairs == GC_PAIRS_AUTO && b->tuned_form == GC_PAIRS_TARGETS) return false;
    for (int nt : b->run_ntiles)
        if (nt > kSymMaxTiles) return false;
    return true;
}

// The work list for the current J tiles per item; a new list drops the
// captured graph.
gc_status ensure_batch_sym(gc_batch* b) {
    if (b->sym.items && b->sym_built_len == b->sym_len) return GC_OK;
    drop_graph(b);
    b->sym_built_len = -1;
    std::string msg;
    const int rc = sym_prepare_runs(b->sym, b->run_tile0.data(), b->run_ntiles.data(), b->runs,
                                    b->sym_len, b->stream, &msg);
    if (rc != 0) return bfail(b, rc == 2 ? GC_ERR_OUT_OF_MEMORY : GC_ERR_HIP, msg);
    b->sym_built_len = b->sym_len;
    return GC_OK;
}

// Targets per lane: as set, else the largest of 4, 2, 1 that still gives
// >= 8192 waves.
int batch_sym_ipt(const gc_batch* b) {
    if (b->sym_ipt == 1 || b->sym_ipt == 2 || b->sym_ipt == 4) return b->sym_ipt;
    const long waves4 = b->sym.nitems;
    if (waves4 >= 8192) return 4;
    if (2 * waves4 >= 8192) return 2;
    return 1;
}

// Builds the work list for `chunk` source tiles per item (slots in
// (run, target tile, chunk) order; processing order longest first).
gc_status build_schedule(gc_batch* b, int chunk) {
    std::vector<int> slot0(b->ntiles), nchunk(b->ntiles);
    std::vector<int4> items;
    int slot = 0;
    for (int r = 0; r < b->runs; ++r) {
        // the run's source split of a single-run launch with
        // jsplit = ceil(nt / chunk): equal ranges of tps tiles
        const int t0 = b->run_tile0[r], nt = b->run_ntiles[r];
        const int js = (nt + chunk - 1) / chunk;
        const int tps = (nt + js - 1) / js;
        const int nc = (nt + tps - 1) / tps;
        for (int ti = t0; ti < t0 + nt; ++ti) {
            slot0[ti] = slot;
            nchunk[ti] = nc;
            for (int c = 0; c < nc; ++c) {
                const int j0 = t0 + c * tps, j1 = std::min(j0 + tps, t0 + nt);
                items.push_back(make_int4(ti, j0, j1, slot++));
            }
        }
    }
    std::stable_sort(items.begin(), items.end(),
                     [](const int4& a, const int4& c) { return (a.z - a.y) > (c.z - c.y); });
    GB_HIP(b, hipStreamSynchronize(b->stream), "synchronize before schedule");
    // a failure below leaves no schedule (nitems 0): the next launch
    // rebuilds it (schedule_ready)
    drop_graph(b);
    (void)hipFree(b->items);
    b->items = nullptr;
    b->nitems = 0;
    GB_HIP(b, hipMalloc(&b->items, sizeof(int4) * std::max<size_t>(1, items.size())), "allocate work list");
    GB_HIP(b, hipMemcpy(b->items, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice), "upload work list");
    GB_HIP(b, hipMemcpy(b->tile_slot0, slot0.data(), sizeof(int) * b->ntiles, hipMemcpyHostToDevice), "upload slots");
    GB_HIP(b, hipMemcpy(b->tile_nchunk, nchunk.data(), sizeof(int) * b->ntiles, hipMemcpyHostToDevice), "upload slots");
    const size_t need = static_cast<size_t>(slot) * 6 * kTile;
    if (need > b->part_floats) {
        (void)hipFree(b->part);
        b->part = nullptr;
        b->part_floats = 0;
        GB_HIP(b, hipMalloc(&b->part, sizeof(float) * need), "allocate batch partials");
        b->part_floats = need;
    }
    b->nitems = static_cast<int>(items.size());
    b->chunk = chunk;
    return GC_OK;
}

// The per-target form needs its work list; one lost to a failed
// reallocation is rebuilt for the requested chunk before any launch.
gc_status schedule_ready(gc_batch* b) {
    if (batch_symmetric(b) || b->nitems > 0 || b->ntiles == 0) return GC_OK;
    return build_schedule(b, std::max(1, b->chunk));
}

void launch_force(gc_batch* b) {
    const float eps2 = static_cast<float>(b->params.softening * b->params.softening);
    const bool safe = !(eps2 >= kTinySeparation2);
    hipLaunchKernelGGL(batch_pack_kernel, dim3(b->ntiles), dim3(kTile), 0, b->stream, b->tile_valid,
                       b->s.x, b->s.y, b->s.z, b->s.m, b->params.gravitational_constant, b->ntiles,
                       b->jp, b->centre);
    if (batch_symmetric(b)) {
        launch_direct_sym_tiles(b->s, b->jp, b->centre, b->tile_valid, b->ntiles, b->sym,
                                b->params.softening, batch_sym_ipt(b), b->sym_blocks, b->stream);
        return;
    }
    const dim3 grid(static_cast<unsigned>(std::max(1, std::min(b->blocks, b->nitems))));
    if (safe) {
        hipLaunchKernelGGL(batch_force_kernel<true>, grid, dim3(kTile), 0, b->stream, b->nitems,
                           b->items, b->tile_valid, b->s.x, b->s.y, b->s.z, b->jp, b->centre,
                           b->ntiles, eps2, b->part);
    } else {
        hipLaunchKernelGGL(batch_force_kernel<false>, grid, dim3(kTile), 0, b->stream, b->nitems,
                           b->items, b->tile_valid, b->s.x, b->s.y, b->s.z, b->jp, b->centre,
                           b->ntiles, eps2, b->part);
    }
}

void launch_merge(gc_batch* b, bool kick, double half_dt) {
    if (batch_symmetric(b)) {
        launch_merge_sym(b->s, b->sym, b->tile_valid, kick, half_dt, b->stream);
        return;
    }
    hipLaunchKernelGGL(batch_merge_kernel, dim3((b->total + 255) / 256), dim3(256), 0, b->stream,
                       b->total, b->tile_valid, b->tile_slot0, b->tile_nchunk, b->part, b->s.ax,
                       b->s.ay, b->s.az, b->s.vx, b->s.vy, b->s.vz, kick ? 1 : 0, half_dt);
}

void launch_step(gc_batch* b, double dt) {
    const double half = 0.5 * dt;
    hipLaunchKernelGGL(batch_open_kernel, dim3((b->total + 255) / 256), dim3(256), 0, b->stream,
                       b->total, b->s.x, b->s.y, b->s.z, b->s.vx, b->s.vy, b->s.vz, b->s.ax, b->s.ay,
                       b->s.az, half, dt);
    launch_force(b);
    launch_merge(b, true, half);
}

// copies between host run-concatenated n x 3 arrays and padded SoA slots
gc_status transfer(gc_batch* b, const double* in, double* out, double* x, double* y, double* z) {
    size_t n_total = 0;
    for (size_t c : b->counts) n_total += c;
    std::vector<double> host(3 * static_cast<size_t>(b->total), 0.0);
    if (in) {
        size_t src = 0;
        for (int r = 0; r < b->runs; ++r)
            for (size_t k = 0; k < b->counts[r]; ++k, ++src)
                for (int d = 0; d < 3; ++d) host[d * b->total + b->offsets[r] + k] = in[3 * src + d];
        GB_HIP(b, hipMemcpy(b->stage, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice), "upload");
        GB_HIP(b, hipMemcpyAsync(x, b->stage, sizeof(double) * b->total, hipMemcpyDeviceToDevice, b->stream), "upload");
        GB_HIP(b, hipMemcpyAsync(y, b->stage + b->total, sizeof(double) * b->total, hipMemcpyDeviceToDevice, b->stream), "upload");
        GB_HIP(b, hipMemcpyAsync(z, b->stage + 2 * static_cast<size_t>(b->total), sizeof(double) * b->total, hipMemcpyDeviceToDevice, b->stream), "upload");
        GB_HIP(b, hipStreamSynchronize(b->stream), "upload");
    } else {
        GB_HIP(b, hipMemcpyAsync(b->stage, x, sizeof(double) * b->total, hipMemcpyDeviceToDevice, b->stream), "download");
        GB_HIP(b, hipMemcpyAsync(b->stage + b->total, y, sizeof(double) * b->total, hipMemcpyDeviceToDevice, b->stream), "download");
        GB_HIP(b, hipMemcpyAsync(b->stage + 2 * static_cast<size_t>(b->total), z, sizeof(double) * b->total, hipMemcpyDeviceToDevice, b->stream), "download");
        GB_HIP(b, hipStreamSynchronize(b->stream), "download");
        GB_HIP(b, hipMemcpy(host.data(), b->stage, sizeof(double) * host.size(), hipMemcpyDeviceToHost), "download");
        size_t dst = 0;
        for (int r = 0; r < b->runs; ++r)
            for (size_t k = 0; k < b->counts[r]; ++k, ++dst)
                for (int d = 0; d < 3; ++d) out[3 * dst + d] = host[d * b->total + b->offsets[r] + k];
    }
    (void)n_total;
    return GC_OK;
}

}  // namespace

extern "C" {

const char* gc_batch_last_error(const gc_batch* b) { return b ? b->err.c_str() : g_batch_err.c_str(); }

gc_status gc_batch_create(gc_batch** out, int device, const gc_params* p, int runs,
                          const size_t* counts) {
    GC_ABI_BEGIN
    if (!out || !p || runs <= 0 || !counts) return bfail(nullptr, GC_ERR_INVALID_ARGUMENT, "batch_create: bad arguments");
    *out = nullptr;
    if (p->solver != GC_SOLVER_DIRECT || p->precision != GC_PRECISION_MIXED)
        return bfail(nullptr, GC_ERR_UNSUPPORTED, "batched stepping runs the mixed-precision direct solver");
    if (!(p->softening >= 0.0)) return bfail(nullptr, GC_ERR_INVALID_ARGUMENT, "softening must be non-negative");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return bfail(nullptr, GC_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= ndev) return bfail(nullptr, GC_ERR_INVALID_ARGUMENT, "device ordinal out of range");
    gc_batch* b = new (std::nothrow) gc_batch();
    if (!b) return bfail(nullptr, GC_ERR_OUT_OF_MEMORY, "allocate batch");
    b->device = device;
    b->params = *p;
    b->runs = runs;
    size_t off = 0;
    for (int r = 0; r < runs; ++r) {
        if (counts[r] == 0 || counts[r] > (1u << 26)) {
            delete b;
            return bfail(nullptr, GC_ERR_INVALID_ARGUMENT, "batch_create: run sizes must be in [1, 2^26]");
        }
        b->counts.push_back(counts[r]);
        b->offsets.push_back(off);
        b->run_tile0.push_back(static_cast<int>(off / kTile));
        const size_t padded = round_up(counts[r], kTile);
        b->run_ntiles.push_back(static_cast<int>(padded / kTile));
        off += padded;
    }
    if (off > (static_cast<size_t>(1) << 30)) {
        delete b;
        return bfail(nullptr, GC_ERR_INVALID_ARGUMENT, "batch_create: too many particles");
    }
    b->total = static_cast<int>(off);
    b->ntiles = b->total / kTile;
    auto cleanup_fail = [&](const char* what, hipError_t e) {
        std::string msg = std::string(what) + ": " + hipGetErrorString(e);
        gc_batch_destroy(b);
        return bfail(nullptr, e == hipErrorOutOfMemory ? GC_ERR_OUT_OF_MEMORY : GC_ERR_HIP, msg);
    };
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    const size_t T = static_cast<size_t>(b->total);
    if (e == hipSuccess) e = hipMalloc(&b->s.base, sizeof(double) * 10 * T);
    if (e == hipSuccess) e = hipMalloc(&b->jp, sizeof(float4) * T);
    if (e == hipSuccess) e = hipMalloc(&b->centre, sizeof(double) * 3 * b->ntiles);
    if (e == hipSuccess) e = hipMalloc(&b->tile_valid, sizeof(int) * b->ntiles);
    if (e == hipSuccess) e = hipMalloc(&b->tile_slot0, sizeof(int) * b->ntiles);
    if (e == hipSuccess) e = hipMalloc(&b->tile_nchunk, sizeof(int) * b->ntiles);
    if (e == hipSuccess) e = hipMalloc(&b->stage, sizeof(double) * 3 * T);
    if (e != hipSuccess) return cleanup_fail("create batch", e);
    double* base = b->s.base;
    b->s.x = base;
    b->s.y = base + T;
    b->s.z = base + 2 * T;
    b->s.vx = base + 3 * T;
    b->s.vy = base + 4 * T;
    b->s.vz = base + 5 * T;
    b->s.ax = base + 6 * T;
    b->s.ay = base + 7 * T;
    b->s.az = base + 8 * T;
    b->s.m = base + 9 * T;
    b->s.n = b->s.cap = T;
    std::vector<int> valid(b->ntiles, 0);
    for (int r = 0; r < runs; ++r)
        for (int t = 0; t < b->run_ntiles[r]; ++t)
            valid[b->run_tile0[r] + t] =
                static_cast<int>(std::min<size_t>(kTile, counts[r] - static_cast<size_t>(t) * kTile));
    e = hipMemcpy(b->tile_valid, valid.data(), sizeof(int) * b->ntiles, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(b->s.base, 0, sizeof(double) * 10 * T);
    if (e != hipSuccess) return cleanup_fail("initialise batch", e);
    gc_status st = build_schedule(b, 4);
    if (st == GC_OK && batch_symmetric(b)) st = ensure_batch_sym(b);
    if (st != GC_OK) {
        std::string msg = b->err;
        gc_batch_destroy(b);
        return bfail(nullptr, st, msg);
    }
    *out = b;
    return GC_OK;
    GC_ABI_END("gc_batch_create")
}

void gc_batch_destroy(gc_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    sym_release(b->sym);
    void* ptrs[] = {b->s.base, b->jp, b->centre, b->tile_valid, b->tile_slot0, b->tile_nchunk,
                    b->items, b->part, b->stage};
    for (void* q : ptrs) (void)hipFree(q);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

size_t gc_batch_particle_count(const gc_batch* b) {
    if (!b) return 0;
    size_t n = 0;
    for (size_t c : b->counts) n += c;
    return n;
}

gc_status gc_batch_upload(gc_batch* b, const double* pos, const double* vel, const double* mass) {
    GC_ABI_BEGIN
    if (!b || !pos || !mass) return bfail(b, GC_ERR_INVALID_ARGUMENT, "batch_upload: null argument");
    GB_HIP(b, hipSetDevice(b->device), "select device");
    gc_status st = transfer(b, pos, nullptr, b->s.x, b->s.y, b->s.z);
    if (st != GC_OK) return st;
    if (vel) {
        st = transfer(b, vel, nullptr, b->s.vx, b->s.vy, b->s.vz);
        if (st != GC_OK) return st;
    } else {
        GB_HIP(b, hipMemset(b->s.vx, 0, sizeof(double) * 3 * b->total), "clear velocities");
    }
    std::vector<double> m(b->total, 0.0);
    size_t src = 0;
    for (int r = 0; r < b->runs; ++r)
        for (size_t k = 0; k < b->counts[r]; ++k) m[b->offsets[r] + k] = mass[src++];
    GB_HIP(b, hipMemcpy(b->s.m, m.data(), sizeof(double) * b->total, hipMemcpyHostToDevice), "upload masses");
    GB_HIP(b, hipMemset(b->s.ax, 0, sizeof(double) * 3 * b->total), "clear accelerations");
    return GC_OK;
    GC_ABI_END("gc_batch_upload")
}

gc_status gc_batch_generate_disks(gc_batch* b, const gc_disk_galaxy* galaxies,
                                  int galaxies_per_run, const uint64_t* seeds) {
    GC_ABI_BEGIN
    if (!b || !galaxies || !seeds || galaxies_per_run <= 0)
        return bfail(b, GC_ERR_INVALID_ARGUMENT, "batch_generate_disks: bad arguments");
    std::vector<IcSegment> segs;
    long long begin = 0;
    for (int r = 0; r < b->runs; ++r) {
        size_t in_run = 0;
        for (int k = 0; k < galaxies_per_run; ++k) {
            const gc_disk_galaxy& g = galaxies[static_cast<size_t>(r) * galaxies_per_run + k];
            if (g.n_particles == 0) continue;
            if (g.n_particles > b->counts[r] - in_run)  // also keeps in_run from wrapping
                return bfail(b, GC_ERR_INVALID_ARGUMENT,
                             "batch_generate_disks: run " + std::to_string(r) +
                                 "'s galaxies hold more than its " +
                                 std::to_string(b->counts[r]) + " particles");
            IcSegment sg;
            sg.model = kIcTwoDisk;
            sg.stream = static_cast<unsigned>(k);
            sg.seed = seeds[r];
            sg.begin = begin;
            sg.count = static_cast<long long>(g.n_particles);
            sg.dst = static_cast<long long>(b->offsets[r] + in_run);
            ic::two_disk_params(g, b->params.gravitational_constant, b->params.softening, sg.p);
            segs.push_back(sg);
            begin += sg.count;
            in_run += g.n_particles;
        }
        if (in_run != b->counts[r])
            return bfail(b, GC_ERR_INVALID_ARGUMENT,
                         "batch_generate_disks: run " + std::to_string(r) + " has " +
                             std::to_string(b->counts[r]) + " particles, its galaxies " +
                             std::to_string(in_run));
    }
    GB_HIP(b, hipSetDevice(b->device), "select device");
    // padding slots stay massless and at rest
    GB_HIP(b, hipMemsetAsync(b->s.base, 0, sizeof(double) * 10 * b->total, b->stream), "clear state");
    IcSegment* d = nullptr;
    GB_HIP(b, hipMalloc(&d, sizeof(IcSegment) * segs.size()), "batch_generate_disks");
    hipError_t e = hipMemcpyAsync(d, segs.data(), sizeof(IcSegment) * segs.size(),
                                  hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) {
        launch_ic(d, static_cast<int>(segs.size()), begin, b->s, b->stream);
        e = hipGetLastError();
    }
    const hipError_t e2 = hipStreamSynchronize(b->stream);
    (void)hipFree(d);
    if (e == hipSuccess) e = e2;
    GB_HIP(b, e, "batch_generate_disks");
    return GC_OK;
    GC_ABI_END("gc_batch_generate_disks")
}

gc_status gc_batch_download(gc_batch* b, double* pos, double* vel, double* acc) {
    GC_ABI_BEGIN
    if (!b) return bfail(nullptr, GC_ERR_INVALID_ARGUMENT, "batch_download: null batch");
    GB_HIP(b, hipSetDevice(b->device), "select device");
    gc_status st = GC_OK;
    if (pos && (st = transfer(b, nullptr, pos, b->s.x, b->s.y, b->s.z)) != GC_OK) return st;
    if (vel && (st = transfer(b, nullptr, vel, b->s.vx, b->s.vy, b->s.vz)) != GC_OK) return st;
    if (acc && (st = transfer(b, nullptr, acc, b->s.ax, b->s.ay, b->s.az)) != GC_OK) return st;
    return GC_OK;
    GC_ABI_END("gc_batch_download")
}

gc_status gc_batch_compute_accelerations(gc_batch* b) {
    GC_ABI_BEGIN
    if (!b) return bfail(nullptr, GC_ERR_INVALID_ARGUMENT, "null batch");
    GB_HIP(b, hipSetDevice(b->device), "select device");
    gc_status st = schedule_ready(b);
    if (st != GC_OK) return st;
    launch_force(b);
    launch_merge(b, false, 0.0);
    GB_HIP(b, hipGetLastError(), "launch batch force");
    return GC_OK;
    GC_ABI_END("gc_batch_compute_accelerations")
}

gc_status gc_batch_step(gc_batch* b, double dt, int nsteps) {
    GC_ABI_BEGIN
    if (!b) return bfail(nullptr, GC_ERR_INVALID_ARGUMENT, "null batch");
    if (!(dt > 0.0) || nsteps < 0) return bfail(b, GC_ERR_INVALID_ARGUMENT, "batch_step: dt must be positive, steps non-negative");
    GB_HIP(b, hipSetDevice(b->device), "select device");
    gc_status st = schedule_ready(b);
    if (st != GC_OK) return st;
    int left = nsteps;
    if (b->graph_steps > 1 && left >= b->graph_steps && (!b->graph || b->graph_dt != dt)) {
        // capture one chunk on the batch's own stream (nothing runs); a chunk
        // that cannot be captured or instantiated is launched eagerly below
        drop_graph(b);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            for (int k = 0; k < b->graph_steps; ++k) launch_step(b, dt);
            const hipError_t le = hipGetLastError();
            e = hipStreamEndCapture(b->stream, &g);
            if (e == hipSuccess) e = le;
        }
        if (e == hipSuccess && g) {
            e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
            if (e != hipSuccess) b->graph = nullptr;
        }
        if (g) (void)hipGraphDestroy(g);
        if (e == hipSuccess && b->graph) {
            b->graph_dt = dt;
        } else {
            (void)hipGetLastError();
            drop_graph(b);
        }
    }
    if (b->graph_steps > 1 && b->graph && b->graph_dt == dt) {
        while (left >= b->graph_steps) {
            GB_HIP(b, hipGraphLaunch(b->graph, b->stream), "launch graph");
            left -= b->graph_steps;
        }
    }
    for (; left > 0; --left) launch_step(b, dt);
    GB_HIP(b, hipGetLastError(), "launch batch step");
    return GC_OK;
    GC_ABI_END("gc_batch_step")
}

gc_status gc_batch_synchronize(gc_batch* b) {
    GC_ABI_BEGIN
    if (!b) return bfail(nullptr, GC_ERR_INVALID_ARGUMENT, "null batch");
    GB_HIP(b, hipStreamSynchronize(b->stream), "synchronize");
    return GC_OK;
    GC_ABI_END("gc_batch_synchronize")
}

gc_status gc_batch_set_schedule(gc_batch* b, int chunk_tiles, int blocks, int graph_steps) {
    GC_ABI_BEGIN
    if (!b) return bfail(nullptr, GC_ERR_INVALID_ARGUMENT, "null batch");
    const bool sym = batch_symmetric(b);
    if (chunk_tiles < (sym ? 0 : 1) || chunk_tiles > 4096 || blocks < (sym ? 0 : 1) ||
        blocks > 1 << 20 || graph_steps < 0 || graph_steps > 1000)
        return bfail(b, GC_ERR_INVALID_ARGUMENT, "batch_set_schedule: out of range");
    GB_HIP(b, hipSetDevice(b->device), "select device");
    drop_graph(b);
    if (sym) {
        b->sym_len = chunk_tiles;
        b->sym_blocks = blocks;
        gc_status st = ensure_batch_sym(b);
        if (st != GC_OK) return st;
    } else {
        if (chunk_tiles != b->chunk) {
            gc_status st = build_schedule(b, chunk_tiles);
            if (st != GC_OK) return st;
        }
        b->blocks = blocks;
    }
    b->graph_steps = graph_steps;
    return GC_OK;
    GC_ABI_END("gc_batch_set_schedule")
}

gc_status gc_batch_get_schedule(gc_batch* b, int* chunk_tiles, int* blocks, int* graph_steps) {
    GC_ABI_BEGIN
    if (!b || !chunk_tiles || !blocks || !graph_steps) return bfail(b, GC_ERR_INVALID_ARGUMENT, "null argument");
    const bool sym = batch_symmetric(b);
    *chunk_tiles = sym ? b->sym_len : b->chunk;
    *blocks = sym ? b->sym_blocks : b->blocks;
    *graph_steps = b->graph_steps;
    return GC_OK;
    GC_ABI_END("gc_batch_get_schedule")
}

gc_status gc_batch_time_force_kernel(gc_batch* b, int reps, double* ms, double* interactions) {
    GC_ABI_BEGIN
    if (!b || reps <= 0 || !ms) return bfail(b, GC_ERR_INVALID_ARGUMENT, "batch_time: bad arguments");
    GB_HIP(b, hipSetDevice(b->device), "select device");
    gc_status st0 = schedule_ready(b);
    if (st0 != GC_OK) return st0;
    const float eps2 = static_cast<float>(b->params.softening * b->params.softening);
    const bool safe = !(eps2 >= kTinySeparation2);
    hipLaunchKernelGGL(batch_pack_kernel, dim3(b->ntiles), dim3(kTile), 0, b->stream, b->tile_valid,
                       b->s.x, b->s.y, b->s.z, b->s.m, b->params.gravitational_constant, b->ntiles,
                       b->jp, b->centre);
    const dim3 grid(static_cast<unsigned>(std::max(1, std::min(b->blocks, b->nitems))));
    const bool sym = batch_symmetric(b);
    GB_HIP(b, hipEventRecord(b->ev0, b->stream), "record");
    for (int r = 0; r < reps; ++r) {
        if (sym) {
            launch_direct_sym_tiles(b->s, b->jp, b->centre, b->tile_valid, b->ntiles, b->sym,
                                    b->params.softening, batch_sym_ipt(b), b->sym_blocks,
                                    b->stream);
        } else if (safe) {
            hipLaunchKernelGGL(batch_force_kernel<true>, grid, dim3(kTile), 0, b->stream, b->nitems,
                               b->items, b->tile_valid, b->s.x, b->s.y, b->s.z, b->jp, b->centre,
                               b->ntiles, eps2, b->part);
        } else {
            hipLaunchKernelGGL(batch_force_kernel<false>, grid, dim3(kTile), 0, b->stream, b->nitems,
                               b->items, b->tile_valid, b->s.x, b->s.y, b->s.z, b->jp, b->centre,
                               b->ntiles, eps2, b->part);
        }
    }
    GB_HIP(b, hipEventRecord(b->ev1, b->stream), "record");
    GB_HIP(b, hipEventSynchronize(b->ev1), "wait");
    GB_HIP(b, hipGetLastError(), "timed batch kernels");
    float t = 0.f;
    GB_HIP(b, hipEventElapsedTime(&t, b->ev0, b->ev1), "elapsed");
    *ms = static_cast<double>(t) / reps;
    if (interactions) {
        double pairs = 0.0;
        for (size_t c : b->counts) pairs += static_cast<double>(c) * static_cast<double>(c - 1);
        *interactions = pairs;
    }
    launch_merge(b, false, 0.0);  // leave accelerations consistent with the state
    GB_HIP(b, hipGetLastError(), "merge after timing");
    return GC_OK;
    GC_ABI_END("gc_batch_time_force_kernel")
}

gc_status gc_batch_autotune(gc_batch* b, int* chunk_out, int* blocks_out) {
    GC_ABI_BEGIN
    if (!b) return bfail(nullptr, GC_ERR_INVALID_ARGUMENT, "null batch");
    GB_HIP(b, hipSetDevice(b->device), "select device");
    long long items1 = 0;
    for (int r = 0; r < b->runs; ++r) items1 += static_cast<long long>(b->run_ntiles[r]) * b->run_ntiles[r];
    const auto key = std::make_tuple(b->runs, b->ntiles, items1, b->pairs);
    Tuned best_t{batch_symmetric(b) ? GC_PAIRS_SYMMETRIC : GC_PAIRS_TARGETS, b->chunk, b->blocks,
                 b->sym_ipt, b->sym_blocks, b->sym_len};
    bool cached = false;
    {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            best_t = it->second;
            cached = true;
        }
    }
    const int requested = b->pairs;
    hipDeviceProp_t prop;
    GB_HIP(b, hipGetDeviceProperties(&prop, b->device), "device properties");
    const int cus = std::max(1, prop.multiProcessorCount);
    std::string arch(prop.gcnArchName);
    if (arch.empty() || arch.find(' ') != std::string::npos) arch = "unknown";
    if (!cached && tune_file_lookup(arch, cus, key, &best_t)) {
        cached = true;
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        g_tune_cache[key] = best_t;
    }
    if (!cached) {
        double best = 1e30;
        auto time_it = [&](double* ms) -> gc_status {
            gc_status st = gc_batch_time_force_kernel(b, 1, ms, nullptr);  // warm
            if (st != GC_OK) return st;
            return gc_batch_time_force_kernel(b, 3, ms, nullptr);
        };
        if (requested != GC_PAIRS_SYMMETRIC) {
            b->pairs = GC_PAIRS_TARGETS;
            const int chunks[] = {1, 2, 4, 8, 16, 32};
            const int per_cu[] = {2, 4, 8, 16};
            for (int c : chunks) {
                gc_status st = build_schedule(b, c);
                if (st != GC_OK) return b->pairs = requested, st;
                for (int pc : per_cu) {
                    b->blocks = cus * pc;
                    double ms = 0.0;
                    st = time_it(&ms);
                    if (st != GC_OK) return b->pairs = requested, st;
                    if (ms < best) {
                        best = ms;
                        best_t = Tuned{GC_PAIRS_TARGETS, c, b->blocks, 0, 0, 0};
                    }
                }
            }
        }
        b->pairs = requested;
        if (requested != GC_PAIRS_TARGETS) {
            b->pairs = GC_PAIRS_SYMMETRIC;
            if (batch_symmetric(b)) {
                gc_status st = GC_OK;
                const int ipts[] = {1, 2, 4};
                const int lens[] = {0, 1, 2, 4};
                const int per_cu[] = {0, 16};
                for (int len : lens) {
                    b->sym_len = len;
                    st = ensure_batch_sym(b);
                    if (st != GC_OK) return b->pairs = requested, st;
                    for (int ipt : ipts) {
                        for (int pc : per_cu) {
                            b->sym_ipt = ipt;
                            b->sym_blocks = cus * pc;
                            double ms = 0.0;
                            st = time_it(&ms);
                            if (st != GC_OK) return b->pairs = requested, st;
                            if (ms < best) {
                                best = ms;
                                best_t = Tuned{GC_PAIRS_SYMMETRIC, b->chunk, b->blocks, ipt,
                                               b->sym_blocks, len};
                            }
                        }
                    }
                }
            }
            b->pairs = requested;
        }
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        g_tune_cache[key] = best_t;
        tune_file_store(arch, cus, key, best_t);
    }
    b->tuned_form = best_t.form;
    gc_status st = build_schedule(b, best_t.chunk);
    if (st != GC_OK) return st;
    b->blocks = best_t.blocks;
    b->sym_ipt = best_t.ipt;
    b->sym_blocks = best_t.sym_blocks;
    b->sym_len = best_t.sym_len;
    if (batch_symmetric(b) && (st = ensure_batch_sym(b)) != GC_OK) return st;
    drop_graph(b);
    const bool sym = batch_symmetric(b);
    if (chunk_out) *chunk_out = sym ? best_t.sym_len : best_t.chunk;
    if (blocks_out) *blocks_out = sym ? best_t.sym_blocks : best_t.blocks;
    return GC_OK;
    GC_ABI_END("gc_batch_autotune")
}

gc_status gc_batch_set_pairs(gc_batch* b, int mode, int targets_per_lane) {
    GC_ABI_BEGIN
    if (!b) return bfail(nullptr, GC_ERR_INVALID_ARGUMENT, "null batch");
    if (mode != GC_PAIRS_AUTO && mode != GC_PAIRS_TARGETS && mode != GC_PAIRS_SYMMETRIC)
        return bfail(b, GC_ERR_INVALID_ARGUMENT, "pair mode must be 0, 1 or 2");
    if (targets_per_lane != 0 && targets_per_lane != 1 && targets_per_lane != 2 &&
        targets_per_lane != 4)
        return bfail(b, GC_ERR_INVALID_ARGUMENT, "targets per lane must be 0, 1, 2 or 4");
    GB_HIP(b, hipSetDevice(b->device), "select device");
    b->pairs = mode;
    b->tuned_form = 0;
    b->sym_ipt = targets_per_lane;
    if (batch_symmetric(b)) {
        gc_status st = ensure_batch_sym(b);
        if (st != GC_OK) return st;
    }
    drop_graph(b);
    return GC_OK;
    GC_ABI_END("gc_batch_set_pairs")
}

gc_status gc_batch_get_pairs(gc_batch* b, int* mode, int* targets_per_lane) {
    GC_ABI_BEGIN
    if (!b || !mode || !targets_per_lane) return bfail(b, GC_ERR_INVALID_ARGUMENT, "null argument");
    const bool sym = batch_symmetric(b);
    *mode = sym ? GC_PAIRS_SYMMETRIC : GC_PAIRS_TARGETS;
    *targets_per_lane = sym ? batch_sym_ipt(b) : 1;
    return GC_OK;
    GC_ABI_END("gc_batch_get_pairs")
}

}  // extern "C"
