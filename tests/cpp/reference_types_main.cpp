// Built by tests/test_reference_types.py against the reference's own headers
// (-DGCSIM_REFERENCE_TYPES -I<reference>/cpp), as a maintainer would build
// host/hip_solver.cpp inside the reference: the solver functions take the
// reference's Particle / PhysicsParams / FmmOptions, and the writer's
// metadata template takes its SimulationConfig (cpp/core/config.hpp:19-34).
// Without a device every solver call must throw std::runtime_error.
#include "hip_solver.hpp"
#include "core/config.hpp"

#include <cstdio>
#include <stdexcept>
#include <vector>

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    std::vector<fmmgalaxy::Particle> ps(8);
    for (std::size_t i = 0; i < ps.size(); ++i) {
        ps[i].position = fmmgalaxy::Vec2{static_cast<double>(i), 0.5 * i, 0.0};
        ps[i].mass = 1.0;
    }
    fmmgalaxy::PhysicsParams physics;
    fmmgalaxy::FmmOptions options;
    const bool available = fmmgalaxy::hip_solver_available();
    int threw = 0;
    try {
        fmmgalaxy::compute_hip_direct_accelerations(ps, physics, fmmgalaxy::HipPrecision::mixed);
    } catch (const std::runtime_error&) {
        ++threw;
    }
    try {
        fmmgalaxy::compute_hip_fmm_accelerations(ps, physics, options);
    } catch (const std::runtime_error&) {
        ++threw;
    }
    try {
        fmmgalaxy::hip_direct_leapfrog_step(ps, 0.01, physics);
    } catch (const std::runtime_error&) {
        ++threw;
    }
    fmmgalaxy::SimulationConfig config;
    config.name = "reference_types";
    config.solver = "hip-direct";
    {
        fmmgalaxy::HipSnapshotWriter writer(argv[1]);
        writer.write_metadata(config, ps.size());
        writer.write_snapshot(0, 0.0, ps);
        writer.flush();
    }
    std::printf("available=%d threw=%d\n", available ? 1 : 0, threw);
    return 0;
}
