// gc_fmm.hpp — hip-fmm and hip-tree engine: octree, P2M/M2M, then either
// far/near lists with M2P + P2P (FMM) or a per-target Barnes-Hut walk
// (p.solver == GC_SOLVER_TREE), on the resident particle state.
#pragma once

#include "gc_internal.hpp"

#include <string>

namespace gc {

class FmmEngine {
public:
    FmmEngine();
    ~FmmEngine();
    FmmEngine(const FmmEngine&) = delete;
    FmmEngine& operator=(const FmmEngine&) = delete;

    // Accelerations of targets [begin, end) into s.ax/ay/az (and the closing
    // half kick when close_kick).  Returns a gc_status.
    int compute(const DeviceState& s, const gc_params& p, size_t begin, size_t end,
                bool close_kick, double half_dt, hipStream_t st, std::string* err);
    int time_kernels(const DeviceState& s, const gc_params& p, int reps, hipStream_t st,
                     double* ms, double* interactions, std::string* err);
    // Tree and list statistics of the last compute() (waits for a
    // Barnes-Hut evaluation's work counts).
    const gc_fmm_stats& stats();
    // Build (tree + lists) and evaluation times of the last compute() (ms).
    int phase_times(float* build_ms, float* eval_ms) const;

private:
    struct Impl;
    Impl* impl_;
};

}  // namespace gc
