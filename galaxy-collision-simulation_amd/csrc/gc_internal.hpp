// gc_internal.hpp — device-side layout and kernel entry points shared by the
// gcsim translation units.  Host code only; kernels live in *.hip files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>

#include "gcsim.h"
#include "sym_plan.hpp"

namespace gc {

// Source tile: the LDS j-tile of the all-pairs kernels, and the group of
// particles that shares one FP64 centre in the mixed-precision layout.
constexpr int kTile = 256;
// Sources the scalar-load sweep reads ahead; the source array carries this
// many entries of padding past its last tile.
constexpr int kStreamBatch = 8;

constexpr size_t round_up(size_t v, size_t m) { return (v + m - 1) / m * m; }

// FP64 source record of the FP64 kernels: position and G * mass.
struct alignas(32) Src64 {
    double x, y, z, gm;
};

// Resident SoA particle state in HBM (FP64), capacity padded to kTile.
struct DeviceState {
    size_t n = 0, cap = 0;
    double* base = nullptr;  // one allocation: 10 * cap doubles
    double *x = nullptr, *y = nullptr, *z = nullptr;
    double *vx = nullptr, *vy = nullptr, *vz = nullptr;
    double *ax = nullptr, *ay = nullptr, *az = nullptr;
    double* m = nullptr;
};

// Mixed-precision source layout: per particle (x - c, y - c, z - c, G*m) in
// FP32 relative to its tile's FP64 centre c; per tile the FP64 centre.
struct MixedLayout {
    size_t ntiles = 0;
    float4* jp = nullptr;      // ntiles * kTile entries (zero-mass padding) + kStreamBatch
    double* centre = nullptr;  // 3 * ntiles (SoA: cx | cy | cz)
    float* part = nullptr;     // jsplit * 6 * cap (hi x,y,z | lo x,y,z)
    size_t part_floats = 0;
};

// Symmetric (each unordered pair once) mixed form: the work list over the
// upper triangle of source-tile pairs and its partial sums.
constexpr int kSymMaxTiles = 1024;  // N <= 262,144: pair partials <= 1.6 GB
struct SymPlan {
    int ntiles = 0;             // plan built for this tile count
    int len = 0;                // ... and this item length (0: the default)
    int nitems = 0;
    SymItem* items = nullptr;   // longest first
    int* row_start = nullptr;   // ntiles + 1 slot offsets; a row's slots in J order
    int* tile_t0 = nullptr;     // first tile of each tile's run
    long long* tile_pbase = nullptr;  // pair-partial base of each tile's run
    float* ipart = nullptr;     // nitems * 3 * kTile row partials (x | y | z), item slots
    float* jpart = nullptr;     // per run nt (nt - 1) / 2 pairs * 3 * kTile
    float* wt = nullptr;        // ntiles: the G m that every particle of a full tile
                                // shares, or 0 (written before each evaluation)
    size_t ipart_floats = 0, jpart_floats = 0;
};

struct LaunchShape {
    int ipt = 0;      // targets per thread (1, 2, 4); 0 = automatic
    int jsplit = 0;   // source-range splits; 0 = automatic
    int sources = 0;  // GC_SOURCES_*: 0 automatic, 1 LDS tiles, 2 scalar loads
    int pairs = 0;    // GC_PAIRS_*: 0 automatic, 1 per target, 2 each pair once
    int items = 0;    // symmetric form: source tiles per work item; 0 = automatic
};

// One model over a run of particles for the device generator (ic.hip).
enum IcModel { kIcRefDisk = 0, kIcTwoDisk = 1, kIcPlummer = 2 };
constexpr int kIcParams = 20;
constexpr unsigned kIcPlummerTries = 1000;  // rejection pairs per particle, at most
struct IcSegment {
    int model = kIcRefDisk;
    unsigned stream = 0;             // Philox counter word 2
    unsigned long long seed = 0;     // Philox key
    long long begin = 0, count = 0;  // particles [begin, begin + count) of the launch
    long long dst = 0;               // state slot of the segment's first particle
    double p[kIcParams] = {};        // model parameters (ic_models.hpp layouts)
};

// ---- kernel launchers (defined in the .hip files) ------------------------
// direct_fp64.hip
void launch_direct_fp64(const DeviceState& s, double G, double eps, size_t begin, size_t end,
                        bool close_kick, double half_dt, hipStream_t st);
// direct_fp64.hip, throughput form: sources packed as Src64 (G m) with
// zero-mass padding (fp64_fast_sources(n) entries); launch shape for
// ntargets targets (ipt / jsplit 0: automatic): targets per lane, splits and
// sources per split; partials jsplit * 3 * s.cap doubles
size_t fp64_fast_sources(size_t n);
void launch_pack_fp64(const DeviceState& s, Src64* jp, double G, hipStream_t st);
void fp64_fast_shape(size_t n, size_t ntargets, int ipt, int jsplit, int* ipt_out,
                     int* jsplit_out, int* per_split_out);
void launch_direct_fp64_fast(const DeviceState& s, const Src64* jp, double* part, double eps,
                             int ipt, int jsplit, int per_split, size_t begin, size_t end,
                             hipStream_t st);
void launch_merge_fp64(const DeviceState& s, const double* part, int jsplit, size_t begin,
                       size_t end, bool close_kick, double half_dt, hipStream_t st);
// direct_mixed.hip
void launch_pack_mixed(const DeviceState& s, MixedLayout& L, double G, hipStream_t st);
// stream: sources by wave-uniform (scalar) loads instead of LDS tiles
void launch_direct_mixed(const DeviceState& s, MixedLayout& L, double eps, int ipt, int jsplit,
                         bool stream, size_t begin, size_t end, hipStream_t st);
void launch_merge_mixed(const DeviceState& s, const MixedLayout& L, int jsplit, size_t begin,
                        size_t end, bool close_kick, double half_dt, hipStream_t st);
// direct_sym.hip: 0 on success, 2 out of memory, 1/3 other errors (err set)
// len: source tiles per work item (0: the default of sym_plan.hpp)
int sym_prepare(SymPlan& plan, int ntiles, int len, hipStream_t st, std::string* err);
// independent runs of tiles [run_t0[r], run_t0[r] + run_nt[r]) (pairs only
// within a run)
// len: J tiles per item (0: the default of sym_plan.hpp)
int sym_prepare_runs(SymPlan& plan, const int* run_t0, const int* run_nt, int nruns, int len,
                     hipStream_t st, std::string* err);
void sym_release(SymPlan& plan);
// ipt: targets per lane (1, 2, 4)
// weights_ready: plan.wt already holds the tiles' shared G m for these
// sources (launch_sym_weights, or the merge that packed them); otherwise a
// weights kernel runs first
void launch_direct_sym(const DeviceState& s, const MixedLayout& L, const SymPlan& plan, double eps,
                       int ipt, hipStream_t st, bool weights_ready = false);
// tile_valid: valid targets per tile, or null for the first s.n slots;
// blocks > 0: persistent grid of that many blocks over the items
void launch_direct_sym_tiles(const DeviceState& s, const float4* jp, const double* centre,
                             const int* tile_valid, int ntiles, const SymPlan& plan, double eps,
                             int ipt, int blocks, hipStream_t st, bool weights_ready = false);
// plan.wt for the packed sources of L (first s.n slots)
void launch_sym_weights(const DeviceState& s, const MixedLayout& L, const SymPlan& plan, double eps,
                        hipStream_t st);
// tile_valid: valid targets per tile, or null for the first s.n slots
void launch_merge_sym(const DeviceState& s, const SymPlan& plan, const int* tile_valid,
                      bool close_kick, double half_dt, hipStream_t st);
// merge with the closing kick, then the next step's opening kick and drift
// and the source tiles of the next force evaluation (first s.n slots)
// (and plan.wt for those tiles)
void launch_merge_sym_open(const DeviceState& s, const SymPlan& plan, MixedLayout& L, double G,
                           double eps, double half_dt, double dt, hipStream_t st);
// integrate.hip
void launch_open_step(const DeviceState& s, double half_dt, double dt, hipStream_t st);
void launch_kick(const DeviceState& s, double dt, hipStream_t st);
void launch_scatter_xyz(const double* interleaved, double* x, double* y, double* z, size_t n,
                        hipStream_t st);
void launch_gather_xyz(const double* x, const double* y, const double* z, double* interleaved,
                       size_t n, hipStream_t st);
hipError_t run_calibration_copy(int device, size_t bytes, int reps, double* ms);
// ic.hip: segs in device memory, sorted by begin, covering [0, total)
void launch_ic(const IcSegment* segs, int nseg, long long total, const DeviceState& s,
               hipStream_t st);
// adjoint.hip: lam / xbar are SoA with strides lstride / ostride, mbar n.
void launch_vjp_fp64(const DeviceState& s, double G, double eps, const double* lam, int lstride,
                     double* xbar, double* mbar, int ostride, bool accumulate, hipStream_t st);
void launch_vjp_mixed(const DeviceState& s, MixedLayout& L, float4* lpack, double* part,
                      int jsplit, double G, double eps, const double* lam, int lstride,
                      double* xbar, double* mbar, int ostride, bool accumulate, hipStream_t st);
int vjp_mixed_splits(size_t n);
// FP32 cotangents (lam SoA, stride lstride) packed as float4 over `padded`
// slots, zero past n
void launch_pack_lambda(size_t n, size_t padded, const double* lam, int lstride, float4* out,
                        hipStream_t st);
// direct_sym.hip: the VJP with each unordered pair once over plan's items
// (L and lpack packed for the current state); vipart: nitems * 8 * kTile
// floats, vjpart: 4 * kTile floats per tile pair; ipt 1, 2 or 4
void launch_vjp_sym(const DeviceState& s, const MixedLayout& L, const SymPlan& plan,
                    const float4* lpack, float* vipart, float* vjpart, double G, double eps,
                    int ipt, double* xbar, double* mbar, int ostride, bool accumulate,
                    hipStream_t st);
// diagnostics.hip: out[13] in gc_diagnostics field order (KE, PE, E, M, P3, C3, L3)
void launch_diagnostics(const DeviceState& s, double G, double eps, double* scratch,
                        size_t scratch_doubles, double* out_dev, hipStream_t st);
size_t diagnostics_scratch_doubles(size_t n);

}  // namespace gc
