// ref_shim.cpp — extern "C" entry points over the reference's own C++ solvers.
//
// TEST INFRASTRUCTURE ONLY.  Compiled (by oracle/Makefile) together with the
// reference's unmodified sources under /root/reference/cpp into
// oracle/_ref/libgcref.so, which cross-checks the C restatement in
// gcoracle.c and can serve as the "reference" CPU baseline.  Nothing in the
// product links it.  Arrays are particle-major (n x 3) like gcoracle.h.
#include "core/diagnostics.hpp"
#include "core/initial_conditions.hpp"
#include "core/integrator.hpp"
#include "direct/direct_solver.hpp"
#include "fmm/fmm_solver.hpp"
#include "fmm/multipole.hpp"
#include "fmm/quadtree.hpp"
#include "io/snapshot_writer.hpp"

#include <cstddef>
#include <vector>

using namespace fmmgalaxy;

namespace {

std::vector<Particle> to_particles(std::size_t n, const double* pos, const double* vel,
                                   const double* acc, const double* mass) {
    std::vector<Particle> ps(n);
    for (std::size_t i = 0; i < n; ++i) {
        ps[i].position = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
        if (vel) ps[i].velocity = {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
        if (acc) ps[i].acceleration = {acc[3 * i], acc[3 * i + 1], acc[3 * i + 2]};
        ps[i].mass = mass[i];
    }
    return ps;
}

void from_particles(const std::vector<Particle>& ps, double* pos, double* vel, double* acc) {
    for (std::size_t i = 0; i < ps.size(); ++i) {
        if (pos) {
            pos[3 * i] = ps[i].position.x;
            pos[3 * i + 1] = ps[i].position.y;
            pos[3 * i + 2] = ps[i].position.z;
        }
        if (vel) {
            vel[3 * i] = ps[i].velocity.x;
            vel[3 * i + 1] = ps[i].velocity.y;
            vel[3 * i + 2] = ps[i].velocity.z;
        }
        if (acc) {
            acc[3 * i] = ps[i].acceleration.x;
            acc[3 * i + 1] = ps[i].acceleration.y;
            acc[3 * i + 2] = ps[i].acceleration.z;
        }
    }
}

PhysicsParams physics(double G, double eps) {
    PhysicsParams p;
    p.gravitational_constant = G;
    p.softening = eps;
    return p;
}

FmmOptions fmm_options(double theta, std::size_t leaf, int max_depth, int order) {
    FmmOptions o;
    o.theta = theta;
    o.leaf_capacity = leaf;
    o.max_depth = max_depth;
    o.expansion_order = order;
    return o;
}

}  // namespace

extern "C" {

void ref_direct_symmetric(std::size_t n, const double* pos, const double* mass, double G,
                          double eps, double* acc) {
    auto ps = to_particles(n, pos, nullptr, nullptr, mass);
    compute_direct_accelerations(ps, physics(G, eps));
    from_particles(ps, nullptr, nullptr, acc);
}

void ref_direct_targets(std::size_t n, const double* pos, const double* mass, double G,
                        double eps, std::size_t begin, std::size_t end, double* acc) {
    auto ps = to_particles(n, pos, nullptr, nullptr, mass);
    compute_direct_accelerations_for_targets(ps, physics(G, eps), begin, end);
    for (std::size_t i = begin; i < end && i < n; ++i) {
        acc[3 * i] = ps[i].acceleration.x;
        acc[3 * i + 1] = ps[i].acceleration.y;
        acc[3 * i + 2] = ps[i].acceleration.z;
    }
}

// solver: 0 direct (symmetric), 2 fmm, 3 tree
int ref_leapfrog(int solver, std::size_t n, double* pos, double* vel, double* acc,
                 const double* mass, double G, double eps, double theta, std::size_t leaf,
                 int order, double dt, int steps) {
    auto ps = to_particles(n, pos, vel, acc, mass);
    const PhysicsParams p = physics(G, eps);
    const FmmOptions o = fmm_options(theta, leaf, 32, order);
    auto compute = [&](std::vector<Particle>& s) {
        if (solver == 2) {
            compute_fmm_accelerations(s, p, o);
        } else if (solver == 3) {
            compute_tree_accelerations(s, p, theta, leaf, order);
        } else {
            compute_direct_accelerations(s, p);
        }
    };
    for (int k = 0; k < steps; ++k) leapfrog_step(ps, dt, compute);
    from_particles(ps, pos, vel, acc);
    return 0;
}

void ref_diagnostics(std::size_t n, const double* pos, const double* vel, const double* mass,
                     double G, double eps, double* out) {
    auto ps = to_particles(n, pos, vel, nullptr, mass);
    const Diagnostics d = compute_diagnostics(ps, physics(G, eps));
    out[0] = d.total_mass;
    out[1] = d.kinetic_energy;
    out[2] = d.potential_energy;
    out[3] = d.total_energy;
    out[4] = d.momentum.x;
    out[5] = d.momentum.y;
    out[6] = d.momentum.z;
    out[7] = d.center_of_mass.x;
    out[8] = d.center_of_mass.y;
    out[9] = d.center_of_mass.z;
    out[10] = d.angular_momentum.x;
    out[11] = d.angular_momentum.y;
    out[12] = d.angular_momentum.z;
    out[13] = out[14] = out[15] = 0.0;
}

int ref_fmm(std::size_t n, const double* pos, const double* mass, double G, double eps,
            double theta, std::size_t leaf, int max_depth, int order, std::size_t begin,
            std::size_t end, double* acc, std::size_t* stats4) {
    auto ps = to_particles(n, pos, nullptr, nullptr, mass);
    FastMultipoleSolver solver(physics(G, eps), fmm_options(theta, leaf, max_depth, order));
    solver.compute_targets(ps, begin, end);
    for (std::size_t i = begin; i < end && i < n; ++i) {
        acc[3 * i] = ps[i].acceleration.x;
        acc[3 * i + 1] = ps[i].acceleration.y;
        acc[3 * i + 2] = ps[i].acceleration.z;
    }
    if (stats4) {
        stats4[0] = solver.stats().node_count;
        stats4[1] = solver.stats().leaf_count;
        stats4[2] = solver.stats().far_interactions;
        stats4[3] = solver.stats().near_interactions;
    }
    return 0;
}

// compute_tree_accelerations (cpp/fmm/quadtree.cpp:245-254): every particle
void ref_tree(std::size_t n, const double* pos, const double* mass, double G, double eps,
              double theta, std::size_t leaf, int order, double* acc) {
    auto ps = to_particles(n, pos, nullptr, nullptr, mass);
    compute_tree_accelerations(ps, physics(G, eps), theta, leaf, order);
    from_particles(ps, nullptr, nullptr, acc);
}

void ref_p2m(double* moments, const double* offset, double mass) {
    CartesianMoments m;
    for (int i = 0; i < 35; ++i) m.values[i] = moments[i];
    add_multipole_point(m, {offset[0], offset[1], offset[2]}, mass);
    for (int i = 0; i < 35; ++i) moments[i] = m.values[i];
}

void ref_m2m(double* parent, const double* child, const double* offset, double child_mass) {
    CartesianMoments p, c;
    for (int i = 0; i < 35; ++i) {
        p.values[i] = parent[i];
        c.values[i] = child[i];
    }
    add_multipole_shifted_child(p, c, {offset[0], offset[1], offset[2]}, child_mass);
    for (int i = 0; i < 35; ++i) parent[i] = p.values[i];
}

void ref_m2p(const double* target, const double* com, double mass, const double* moments,
             double G, double eps, int order, double* out) {
    CartesianMoments m;
    for (int i = 0; i < 35; ++i) m.values[i] = moments[i];
    const Vec2 a = multipole_acceleration({target[0], target[1], target[2]},
                                          {com[0], com[1], com[2]}, mass, m, physics(G, eps),
                                          order);
    out[0] = a.x;
    out[1] = a.y;
    out[2] = a.z;
}

// Reference generator (cpp/core/initial_conditions.cpp) for two galaxies;
// gal is 2 x 13 doubles: n, mass, radius, pos[3], vel[3], orientation,
// group_id, thickness, inclination.  Returns the particle count written.
std::size_t ref_generate(const double* gal, int ngal, double G, double eps, unsigned long long seed,
                         double* pos, double* vel, double* mass, unsigned* group) {
    std::vector<GalaxyConfig> gs;
    for (int g = 0; g < ngal; ++g) {
        const double* c = gal + 13 * g;
        GalaxyConfig cfg;
        cfg.n_particles = static_cast<std::size_t>(c[0]);
        cfg.mass = c[1];
        cfg.radius = c[2];
        cfg.position = {c[3], c[4], c[5]};
        cfg.velocity = {c[6], c[7], c[8]};
        cfg.orientation = c[9];
        cfg.group_id = static_cast<std::uint32_t>(c[10]);
        cfg.thickness = c[11];
        cfg.inclination = c[12];
        gs.push_back(cfg);
    }
    auto ps = generate_galaxies(gs, physics(G, eps), seed);
    from_particles(ps, pos, vel, nullptr);
    for (std::size_t i = 0; i < ps.size(); ++i) {
        mass[i] = ps[i].mass;
        if (group) group[i] = ps[i].group_id;
    }
    return ps.size();
}

// SnapshotWriter (cpp/io/snapshot_writer.cpp): a writer on dir, then
// write_metadata, write_snapshot(step, time, particles) and one
// write_diagnostics row (diag: total_mass, kinetic, potential, total energy,
// momentum xyz, centre of mass xyz, angular momentum xyz).  snapshots > 1
// repeats the snapshot as steps step, step + 1, ... (timing).  Returns 0, or
// 1 when the writer throws.
int ref_write_run(const char* dir, const char* name, const char* solver, int steps, double dt,
                  int snapshot_every, double G, double eps, double theta, std::size_t leaf,
                  int order, int step, double time, std::size_t n, const double* pos,
                  const double* vel, const double* acc, const double* mass,
                  const unsigned* groups, const double* diag, int snapshots) {
    try {
        SimulationConfig cfg;
        cfg.name = name;
        cfg.solver = solver;
        cfg.steps = steps;
        cfg.dt = dt;
        cfg.snapshot_every = snapshot_every;
        cfg.physics = physics(G, eps);
        cfg.tree_theta = theta;
        cfg.tree_leaf_capacity = leaf;
        cfg.fmm_expansion_order = order;
        cfg.output.directory = dir;
        cfg.output.format = "csv";
        std::vector<Particle> ps = to_particles(n, pos, vel, acc, mass);
        for (std::size_t i = 0; i < n; ++i) ps[i].group_id = groups ? groups[i] : 0u;
        SnapshotWriter w(cfg);
        w.write_metadata(cfg, n);
        for (int k = 0; k < snapshots; ++k) w.write_snapshot(step + k, time, ps);
        Diagnostics d;
        d.total_mass = diag[0];
        d.kinetic_energy = diag[1];
        d.potential_energy = diag[2];
        d.total_energy = diag[3];
        d.momentum = {diag[4], diag[5], diag[6]};
        d.center_of_mass = {diag[7], diag[8], diag[9]};
        d.angular_momentum = {diag[10], diag[11], diag[12]};
        w.write_diagnostics(step, time, d, n);
    } catch (...) {
        return 1;
    }
    return 0;
}

}  // extern "C"
