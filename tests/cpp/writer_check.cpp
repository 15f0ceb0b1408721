// writer_check.cpp — the C++ SnapshotWriter mirror (HipSnapshotWriter) fed
// the way the reference's main loop feeds SnapshotWriter
// (cpp/core/main.cpp): a SimulationConfig-shaped config, a particle
// vector, Diagnostics values.  Reads a snapshot CSV (the format the writer
// produces), writes metadata.json, the snapshot again and one diagnostics
// row into <outdir>; the Python test compares the files with the
// reference's.  With a gfx950 device it also writes the same state from a
// HipForceSolver's resident copy as snapshot <step + 1>.
//
//   writer_check <snapshot.csv> <step> <outdir>
#include "hip_solver.hpp"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

using namespace fmmgalaxy;

namespace {

// the fields SnapshotWriter::write_metadata reads (cpp/core/config.hpp:19-34)
struct Config {
    std::string name{"benchmark_direct_250"};
    std::string solver{"direct"};
    int dim{3};
    int steps{20};
    double dt{0.01};
    int snapshot_every{20};
    double tree_theta{0.58};
    std::size_t tree_leaf_capacity{16};
    int fmm_expansion_order{4};
    PhysicsParams physics{1.0, 0.025};
};

std::vector<Particle> read_snapshot(const char* path, double* time) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error(std::string("cannot read ") + path);
    std::string line;
    std::getline(in, line);
    *time = std::strtod(line.c_str() + line.find('=') + 1, nullptr);
    std::getline(in, line);
    std::vector<Particle> ps;
    while (std::getline(in, line)) {
        std::stringstream ss(line);
        std::string f[12];
        for (auto& x : f) std::getline(ss, x, ',');
        Particle p;
        p.group_id = static_cast<std::uint32_t>(std::strtoul(f[1].c_str(), nullptr, 10));
        p.mass = std::strtod(f[2].c_str(), nullptr);
        p.position = {std::strtod(f[3].c_str(), nullptr), std::strtod(f[4].c_str(), nullptr),
                      std::strtod(f[5].c_str(), nullptr)};
        p.velocity = {std::strtod(f[6].c_str(), nullptr), std::strtod(f[7].c_str(), nullptr),
                      std::strtod(f[8].c_str(), nullptr)};
        p.acceleration = {std::strtod(f[9].c_str(), nullptr), std::strtod(f[10].c_str(), nullptr),
                          std::strtod(f[11].c_str(), nullptr)};
        ps.push_back(p);
    }
    return ps;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: writer_check <snapshot.csv> <step> <outdir>\n");
        return 2;
    }
    try {
        double time = 0.0;
        const std::vector<Particle> ps = read_snapshot(argv[1], &time);
        const int step = std::atoi(argv[2]);
        HipSnapshotWriter out(argv[3]);
        out.write_metadata(Config{}, ps.size());
        out.write_snapshot(step, time, ps);
        Diagnostics d;
        d.total_mass = 2.0;
        d.kinetic_energy = 0.25;
        d.potential_energy = -1.0 / 3.0;
        d.total_energy = 0.25 - 1.0 / 3.0;
        d.momentum = {1e-300, -0.0, 5e-324};
        d.center_of_mass = {0.1, 0.2, 0.3};
        d.angular_momentum = {-1.5, 2.5e10, 3.0};
        out.write_diagnostics(step, time, d, ps.size());
        if (hip_solver_available()) {
            HipForceSolver solver(PhysicsParams{1.0, 0.025}, HipForceSolver::Kind::direct,
                                  HipPrecision::fp64);
            solver.upload(ps, true);
            std::vector<std::uint32_t> groups;
            for (const auto& p : ps) groups.push_back(p.group_id);
            out.write_snapshot(step + 1, time, solver, &groups);
            out.write_diagnostics(step + 1, time, solver);
            std::printf("device snapshot written\n");
        }
        out.flush();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "writer_check: %s\n", e.what());
        return 1;
    }
    std::printf("writer_check done\n");
    return 0;
}
