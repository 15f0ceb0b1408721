// hip_solver.cpp — the C++ solver interface over the gcsim C ABI.
#include "hip_solver.hpp"

#include <algorithm>
#include <cctype>
#include <stdexcept>

namespace fmmgalaxy {

namespace {

void check(gc_status st, const gc_ctx* ctx, const char* what) {
    if (st != GC_OK) {
        throw std::runtime_error(std::string(what) + ": " + gc_last_error(ctx) + " (" +
                                 gc_status_string(st) + ")");
    }
}

gc_params make_params(const PhysicsParams& p, HipForceSolver::Kind kind, HipPrecision prec,
                      const FmmOptions& o) {
    gc_params g;
    gc_default_params(&g);
    g.gravitational_constant = p.gravitational_constant;
    g.softening = p.softening;
    g.solver = kind == HipForceSolver::Kind::fmm    ? GC_SOLVER_FMM
               : kind == HipForceSolver::Kind::tree ? GC_SOLVER_TREE
                                                    : GC_SOLVER_DIRECT;
    g.precision = prec == HipPrecision::mixed       ? GC_PRECISION_MIXED
                  : prec == HipPrecision::fp64_fast ? GC_PRECISION_FP64_FAST
                                                    : GC_PRECISION_FP64;
    g.theta = o.theta;
    g.leaf_capacity = static_cast<int>(std::min<std::size_t>(o.leaf_capacity, 1u << 30));
    g.max_depth = o.max_depth;
    g.expansion_order = o.expansion_order;
    return g;
}

void pack(const std::vector<Particle>& ps, std::vector<double>& pos, std::vector<double>& vel,
          std::vector<double>& mass) {
    const std::size_t n = ps.size();
    pos.resize(3 * n);
    vel.resize(3 * n);
    mass.resize(n);
    for (std::size_t i = 0; i < n; ++i) {
        pos[3 * i] = ps[i].position.x;
        pos[3 * i + 1] = ps[i].position.y;
        pos[3 * i + 2] = ps[i].position.z;
        vel[3 * i] = ps[i].velocity.x;
        vel[3 * i + 1] = ps[i].velocity.y;
        vel[3 * i + 2] = ps[i].velocity.z;
        mass[i] = ps[i].mass;
    }
}

void run_force(std::vector<Particle>& particles, const PhysicsParams& params,
               HipForceSolver::Kind kind, HipPrecision prec, const FmmOptions& options,
               std::size_t begin, std::size_t end) {
    end = std::min(end, particles.size());
    if (particles.empty() || begin >= end) return;
    HipForceSolver solver(params, kind, prec, options);
    solver.upload(particles);
    solver.compute_accelerations_for_targets(begin, end);
    std::vector<Particle> out(particles);
    solver.download(out);
    for (std::size_t i = begin; i < end; ++i) particles[i].acceleration = out[i].acceleration;
}

}  // namespace

bool uses_hip_direct_solver(const std::string& solver) {
    return solver == "hip-direct" || solver == "hip" || solver == "rocm-direct";
}

bool uses_hip_fmm_solver(const std::string& solver) {
    return solver == "hip-fmm" || solver == "rocm-fmm";
}

bool uses_hip_tree_solver(const std::string& solver) {
    return solver == "hip-tree" || solver == "rocm-tree";
}

HipPrecision parse_hip_precision(const std::string& value) {
    std::string v(value);
    std::transform(v.begin(), v.end(), v.begin(),
                   [](unsigned char ch) { return static_cast<char>(std::tolower(ch)); });
    if (v == "fp64" || v == "double") return HipPrecision::fp64;
    if (v == "mixed" || v == "fp32") return HipPrecision::mixed;
    if (v == "fp64-fast") return HipPrecision::fp64_fast;
    throw std::runtime_error("Unknown precision '" + value + "'. Use fp64, fp64-fast or mixed.");
}

bool hip_solver_available() { return gc_solver_available() != 0; }

void compute_hip_direct_accelerations(std::vector<Particle>& particles, const PhysicsParams& params,
                                      HipPrecision precision) {
    run_force(particles, params, HipForceSolver::Kind::direct, precision, FmmOptions{}, 0,
              particles.size());
}

void compute_hip_direct_accelerations_for_targets(std::vector<Particle>& particles,
                                                  const PhysicsParams& params, std::size_t begin,
                                                  std::size_t end, HipPrecision precision) {
    run_force(particles, params, HipForceSolver::Kind::direct, precision, FmmOptions{}, begin, end);
}

void hip_direct_leapfrog_step(std::vector<Particle>& particles, double dt,
                              const PhysicsParams& params, HipPrecision precision) {
    if (particles.empty()) return;
    // cpp/cuda/cuda_solver.cu:212-246: one KDK step on the device, starting
    // from the accelerations the particles carry
    HipForceSolver solver(params, HipForceSolver::Kind::direct, precision);
    solver.upload(particles, true);
    solver.step(dt, 1);
    solver.download(particles);
}

void compute_hip_fmm_accelerations(std::vector<Particle>& particles, const PhysicsParams& params,
                                   FmmOptions options, HipPrecision precision) {
    run_force(particles, params, HipForceSolver::Kind::fmm, precision, options, 0,
              particles.size());
}

void compute_hip_fmm_accelerations_for_targets(std::vector<Particle>& particles,
                                               const PhysicsParams& params, std::size_t begin,
                                               std::size_t end, FmmOptions options,
                                               HipPrecision precision) {
    run_force(particles, params, HipForceSolver::Kind::fmm, precision, options, begin, end);
}

void compute_hip_tree_accelerations(std::vector<Particle>& particles, const PhysicsParams& params,
                                    double theta, std::size_t leaf_capacity, int expansion_order,
                                    HipPrecision precision) {
    FmmOptions options;
    options.theta = theta;
    options.leaf_capacity = leaf_capacity;
    options.max_depth = 32;
    options.expansion_order = expansion_order;
    run_force(particles, params, HipForceSolver::Kind::tree, precision, options, 0,
              particles.size());
}

HipForceSolver::HipForceSolver(const PhysicsParams& params, Kind kind, HipPrecision precision,
                               FmmOptions options, int device) {
    const gc_params p = make_params(params, kind, precision, options);
    check(gc_create(&ctx_, device, &p), nullptr, "create HIP force solver");
}

HipForceSolver::~HipForceSolver() { gc_destroy(ctx_); }

void HipForceSolver::upload(const std::vector<Particle>& particles, bool with_accelerations) {
    std::vector<double> pos, vel, mass;
    pack(particles, pos, vel, mass);
    check(gc_upload(ctx_, particles.size(), pos.data(), vel.data(), mass.data()), ctx_,
          "upload particles to HIP device");
    if (with_accelerations) {
        for (std::size_t i = 0; i < particles.size(); ++i) {
            pos[3 * i] = particles[i].acceleration.x;
            pos[3 * i + 1] = particles[i].acceleration.y;
            pos[3 * i + 2] = particles[i].acceleration.z;
        }
        check(gc_set_accelerations(ctx_, pos.data()), ctx_, "upload accelerations to HIP device");
    }
}

void HipForceSolver::compute_accelerations() {
    check(gc_compute_accelerations(ctx_), ctx_, "compute HIP accelerations");
}

void HipForceSolver::compute_accelerations_for_targets(std::size_t begin, std::size_t end) {
    check(gc_compute_accelerations_for_targets(ctx_, begin, end), ctx_,
          "compute HIP accelerations for targets");
}

void HipForceSolver::step(double dt, int steps) {
    check(gc_step(ctx_, dt, steps), ctx_, "HIP leapfrog step");
}

void HipForceSolver::download(std::vector<Particle>& particles) {
    const std::size_t n = gc_particle_count(ctx_);
    std::vector<double> pos(3 * n), vel(3 * n), acc(3 * n);
    check(gc_download(ctx_, pos.data(), vel.data(), acc.data()), ctx_,
          "copy particles from HIP device");
    particles.resize(n);
    for (std::size_t i = 0; i < n; ++i) {
        particles[i].position = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
        particles[i].velocity = {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
        particles[i].acceleration = {acc[3 * i], acc[3 * i + 1], acc[3 * i + 2]};
    }
}

Diagnostics HipForceSolver::diagnostics() {
    gc_diagnostics d;
    check(gc_diagnostics_compute(ctx_, &d), ctx_, "HIP diagnostics");
    Diagnostics out;
    out.kinetic_energy = d.kinetic_energy;
    out.potential_energy = d.potential_energy;
    out.total_energy = d.total_energy;
    out.total_mass = d.total_mass;
    out.momentum = {d.momentum[0], d.momentum[1], d.momentum[2]};
    out.center_of_mass = {d.center_of_mass[0], d.center_of_mass[1], d.center_of_mass[2]};
    out.angular_momentum = {d.angular_momentum[0], d.angular_momentum[1], d.angular_momentum[2]};
    return out;
}

std::size_t HipForceSolver::size() const { return gc_particle_count(ctx_); }

HipSnapshotWriter::HipSnapshotWriter(const std::string& directory) {
    check(gc_writer_open(&w_, directory.c_str()), nullptr, "HIP snapshot writer");
}

HipSnapshotWriter::~HipSnapshotWriter() { (void)gc_writer_close(w_); }

void HipSnapshotWriter::write_metadata(const gc_run_info& info) {
    check(gc_writer_metadata(w_, &info), nullptr, "HIP snapshot writer");
}

void HipSnapshotWriter::write_snapshot(int step, double time,
                                       const std::vector<Particle>& particles) {
    std::vector<double> pos, vel, mass;
    pack(particles, pos, vel, mass);
    const std::size_t n = particles.size();
    std::vector<double> acc(3 * n);
    std::vector<std::uint32_t> groups(n);
    for (std::size_t i = 0; i < n; ++i) {
        acc[3 * i] = particles[i].acceleration.x;
        acc[3 * i + 1] = particles[i].acceleration.y;
        acc[3 * i + 2] = particles[i].acceleration.z;
        groups[i] = particles[i].group_id;
    }
    check(gc_writer_snapshot_host(w_, step, time, n, pos.data(), vel.data(), acc.data(),
                                  mass.data(), groups.data()),
          nullptr, "HIP snapshot writer");
}

void HipSnapshotWriter::write_snapshot(int step, double time, HipForceSolver& solver,
                                       const std::vector<std::uint32_t>* group_ids) {
    if (group_ids && group_ids->size() != solver.size())
        throw std::runtime_error("HIP snapshot writer: group_ids size differs from particle count");
    check(gc_writer_snapshot(w_, solver.handle(), step, time, group_ids ? group_ids->data() : nullptr),
          nullptr, "HIP snapshot writer");
}

void HipSnapshotWriter::write_diagnostics(int step, double time, const Diagnostics& d,
                                          std::size_t particle_count) {
    gc_diagnostics g;
    g.kinetic_energy = d.kinetic_energy;
    g.potential_energy = d.potential_energy;
    g.total_energy = d.total_energy;
    g.total_mass = d.total_mass;
    const Vec2* v[3] = {&d.momentum, &d.center_of_mass, &d.angular_momentum};
    double* o[3] = {g.momentum, g.center_of_mass, g.angular_momentum};
    for (int k = 0; k < 3; ++k) {
        o[k][0] = v[k]->x;
        o[k][1] = v[k]->y;
        o[k][2] = v[k]->z;
    }
    check(gc_writer_diagnostics_values(w_, step, time, particle_count, &g), nullptr,
          "HIP snapshot writer");
}

void HipSnapshotWriter::write_diagnostics(int step, double time, HipForceSolver& solver) {
    check(gc_writer_diagnostics(w_, solver.handle(), step, time), nullptr, "HIP snapshot writer");
}

void HipSnapshotWriter::flush() { check(gc_writer_flush(w_), nullptr, "HIP snapshot writer"); }

}  // namespace fmmgalaxy
