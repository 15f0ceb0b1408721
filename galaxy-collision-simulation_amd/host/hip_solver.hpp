// hip_solver.hpp — C++ force-solver interface of the gfx950 engine, shaped
// like the reference's solver free functions so that `hip-direct`,
// `hip-fmm` and `hip-tree` slot into its `solver = ...` dispatch
// (cpp/core/main.cpp:83-139) next to `direct`, `fmm` and `tree`.
//
// Same conventions as the reference (cpp/direct/direct_solver.hpp:22-28,
// cpp/fmm/fmm_solver.hpp:71-83, cpp/cuda/cuda_solver.hpp:10-12):
//   * std::vector<Particle>& is mutated in place (acceleration for force
//     calls; position, velocity and acceleration for step calls);
//   * errors throw std::runtime_error("context: message");
//   * the caller owns the host vector.
// Differences: there is no silent CPU fallback (hip_solver_available()
// reports whether a gfx950 device is visible and every call without one
// throws), and HipForceSolver keeps the particles resident in HBM across
// steps instead of re-uploading per call.
//
// Built inside the reference, define GCSIM_REFERENCE_TYPES so these
// declarations use the reference's own Particle / PhysicsParams / FmmOptions;
// otherwise layout-identical local definitions are used.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "gcsim.h"

#if defined(GCSIM_REFERENCE_TYPES)
#include "core/diagnostics.hpp"
#include "core/particle.hpp"
#include "core/simulation_params.hpp"
#include "fmm/fmm_solver.hpp"
#else
namespace fmmgalaxy {

struct Vec3 {
    double x{0.0};
    double y{0.0};
    double z{0.0};
};
using Vec2 = Vec3;

struct Particle {
    Vec2 position{};
    Vec2 velocity{};
    Vec2 acceleration{};
    double mass{1.0};
    std::uint32_t group_id{0};
};

struct PhysicsParams {
    double gravitational_constant{1.0};
    double softening{0.02};
};

struct FmmOptions {
    double theta{0.6};
    std::size_t leaf_capacity{16};
    int max_depth{32};
    int expansion_order{4};
};

struct Diagnostics {
    double kinetic_energy{0.0};
    double potential_energy{0.0};
    double total_energy{0.0};
    double total_mass{0.0};
    Vec2 momentum{};
    Vec2 center_of_mass{};
    Vec2 angular_momentum{};
};

}  // namespace fmmgalaxy
#endif

namespace fmmgalaxy {

// fp64: bit-identical to the reference's FP64 targets form; mixed: FP32
// pairs, FP64 state; fp64_fast: the FP64 throughput form (rounding-level
// differences from fp64).
enum class HipPrecision { fp64, mixed, fp64_fast };

// Solver-name predicates in the style of cpp/core/main.cpp:62-73.
bool uses_hip_direct_solver(const std::string& solver);
bool uses_hip_fmm_solver(const std::string& solver);
bool uses_hip_tree_solver(const std::string& solver);
// "fp64" | "double" -> fp64, "mixed" | "fp32" -> mixed, "fp64-fast" ->
// fp64_fast; throws otherwise.
HipPrecision parse_hip_precision(const std::string& value);

bool hip_solver_available();

void compute_hip_direct_accelerations(std::vector<Particle>& particles, const PhysicsParams& params,
                                      HipPrecision precision = HipPrecision::fp64);
void compute_hip_direct_accelerations_for_targets(std::vector<Particle>& particles,
                                                  const PhysicsParams& params, std::size_t begin,
                                                  std::size_t end,
                                                  HipPrecision precision = HipPrecision::fp64);
void hip_direct_leapfrog_step(std::vector<Particle>& particles, double dt,
                              const PhysicsParams& params,
                              HipPrecision precision = HipPrecision::fp64);
void compute_hip_fmm_accelerations(std::vector<Particle>& particles, const PhysicsParams& params,
                                   FmmOptions options = {},
                                   HipPrecision precision = HipPrecision::mixed);
void compute_hip_fmm_accelerations_for_targets(std::vector<Particle>& particles,
                                               const PhysicsParams& params, std::size_t begin,
                                               std::size_t end, FmmOptions options = {},
                                               HipPrecision precision = HipPrecision::mixed);
// compute_tree_accelerations (cpp/fmm/quadtree.hpp:56-62): Barnes-Hut with
// the same defaults (theta 0.6, leaf 16, order 4; depth cap 32)
void compute_hip_tree_accelerations(std::vector<Particle>& particles, const PhysicsParams& params,
                                    double theta = 0.6, std::size_t leaf_capacity = 16,
                                    int expansion_order = 4,
                                    HipPrecision precision = HipPrecision::mixed);

// Resident-state solver: upload once, step many times, download on demand.
class HipForceSolver {
public:
    enum class Kind { direct, fmm, tree };
    HipForceSolver(const PhysicsParams& params, Kind kind, HipPrecision precision,
                   FmmOptions options = {}, int device = 0);
    ~HipForceSolver();
    HipForceSolver(const HipForceSolver&) = delete;
    HipForceSolver& operator=(const HipForceSolver&) = delete;

    // with_accelerations: also take each particle's acceleration as a(x)
    void upload(const std::vector<Particle>& particles, bool with_accelerations = false);
    void compute_accelerations();
    void compute_accelerations_for_targets(std::size_t begin, std::size_t end);
    void step(double dt, int steps = 1);
    void download(std::vector<Particle>& particles);
    Diagnostics diagnostics();
    std::size_t size() const;
    gc_ctx* handle() { return ctx_; }

private:
    gc_ctx* ctx_{nullptr};
};

// SnapshotWriter (cpp/io/snapshot_writer.hpp:13-27) over gc_writer: the
// same files, byte for byte, plus snapshots and diagnostics taken from a
// HipForceSolver's resident state (the snapshot is staged by one device
// copy and written by the writer's own thread).  Errors throw
// std::runtime_error; file errors of queued snapshots surface at flush()
// or destruction (the destructor swallows them: call flush() to see them).
class HipSnapshotWriter {
public:
    explicit HipSnapshotWriter(const std::string& directory);
    ~HipSnapshotWriter();
    HipSnapshotWriter(const HipSnapshotWriter&) = delete;
    HipSnapshotWriter& operator=(const HipSnapshotWriter&) = delete;

    void write_metadata(const gc_run_info& info);
    // SnapshotWriter::write_metadata(const SimulationConfig&, std::size_t):
    // any config type with the reference's SimulationConfig fields
    template <class Config>
    void write_metadata(const Config& config, std::size_t particle_count) {
        gc_run_info info{};
        info.name = config.name.c_str();
        info.solver = config.solver.c_str();
        info.particle_count = particle_count;
        info.steps = config.steps;
        info.dt = config.dt;
        info.snapshot_every = config.snapshot_every;
        info.dim = config.dim;
        info.gravitational_constant = config.physics.gravitational_constant;
        info.softening = config.physics.softening;
        info.tree_theta = config.tree_theta;
        info.tree_leaf_capacity = config.tree_leaf_capacity;
        info.fmm_expansion_order = config.fmm_expansion_order;
        write_metadata(info);
    }
    void write_snapshot(int step, double time, const std::vector<Particle>& particles);
    // group_ids: solver.size() entries, or null for all 0
    void write_snapshot(int step, double time, HipForceSolver& solver,
                        const std::vector<std::uint32_t>* group_ids = nullptr);
    void write_diagnostics(int step, double time, const Diagnostics& diagnostics,
                           std::size_t particle_count);
    void write_diagnostics(int step, double time, HipForceSolver& solver);
    void flush();

private:
    gc_writer* w_{nullptr};
};

}  // namespace fmmgalaxy
