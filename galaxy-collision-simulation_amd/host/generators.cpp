// generators.cpp — synthetic initial conditions for the benchmark workloads.
//
// All generators draw from std::mt19937_64 (as the reference's
// cpp/core/initial_conditions.cpp:94-114 does) with explicitly defined
// transforms, so a seed gives the same particles on every platform:
//   uniform01(u)  = (u >> 11) * 2^-53                      in [0, 1)
//   normal        = Box-Muller on two uniform01 draws (first draw mapped
//                   to (0, 1] to avoid log 0), cosine branch only.
//
// two_disk: a central point mass plus an exponential disk per galaxy.
//   R has pdf proportional to exp(-R/h) truncated to [r_min, r_max], sampled by
//   inverting its CDF; phi ~ U[0, 2 pi); z ~ N(0, z_sigma).
//   Circular speed uses the softened enclosed-mass form of the reference
//   (initial_conditions.cpp:67-72):  v_c = sqrt(G M(<R) R^2 / (R^2+eps^2)^1.5)
//   with M(<R) = M_central + M_disk * CDF(R); each velocity component then
//   gets N(0, v_sigma).  Galaxy B's disk is rotated by its inclination about
//   the x axis (positions and velocities) before its centre offset is added.
// plummer: Aarseth, Henon & Wielen (1974), G = M = a = 1 scaled by
//   (mass, scale): r = (U^(-2/3) - 1)^(-1/2); speed q * sqrt(2) (1+r^2)^(-1/4)
//   with q drawn by rejection from q^2 (1-q^2)^(7/2); isotropic directions.
//
// reference galaxies: generate_galaxies (cpp/core/initial_conditions.cpp
//   :37-114) restated: one std::mt19937_64(seed) shared by the galaxies in
//   order; per galaxy a fresh normal distribution.  The draws follow the
//   standard library the reference is built with (GNU libstdc++, the C++17
//   <random> of GCC 11):
//     uniform_real_distribution(0, 1) = generate_canonical<double, 53>:
//       one 64-bit draw u, u / 2^64 in double, 1.0 replaced by the largest
//       double below 1;
//     normal_distribution(0, 1) = Marsaglia's polar method: x, y = 2 c - 1
//       from two canonical draws until 0 < x^2 + y^2 <= 1, then
//       f = sqrt(-2 ln r2 / r2); returns y f and keeps x f for the next call.
//   With the same libm the particles equal the reference's bit for bit.
#include "gcsim.h"
#include "ic_models.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>

#pragma STDC FP_CONTRACT OFF

namespace {

using gc::ic::kPi;

struct Rng {
    std::mt19937_64 eng;
    explicit Rng(uint64_t seed) : eng(seed) {}
    double uniform01() { return static_cast<double>(eng() >> 11) * 0x1.0p-53; }
    double uniform(double a, double b) { return a + (b - a) * uniform01(); }
    double normal() {
        const double u1 = 1.0 - uniform01();  // (0, 1]
        const double u2 = uniform01();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    }
};

// One galaxy: central mass at index 0, then n_disk disk particles.
void one_disk(Rng& rng, const gc_disk_galaxy& g, double G, double eps, double* pos, double* vel,
              double* mass) {
    double p[gc::ic::kTdCount];
    gc::ic::two_disk_params(g, G, eps, p);
    for (size_t k = 0; k < g.n_particles; ++k) {
        double ur = 0, uphi = 0, nz = 0, nvx = 0, nvy = 0, nvz = 0;
        if (k > 0) {
            ur = rng.uniform01();
            uphi = rng.uniform01();
            nz = rng.normal();
            nvx = rng.normal();
            nvy = rng.normal();
            nvz = rng.normal();
        }
        gc::ic::two_disk_particle(p, k, ur, uphi, nz, nvx, nvy, nvz, pos + 3 * k, vel + 3 * k,
                                  mass + k);
    }
}

// generate_canonical<double, 53> over std::mt19937_64 (GNU libstdc++)
double canonical(std::mt19937_64& eng) {
    const double r = static_cast<double>(eng()) / 18446744073709551616.0;  // 2^64
    return r >= 1.0 ? std::nextafter(1.0, 0.0) : r;
}

// normal_distribution<double>(0, 1) (GNU libstdc++): Marsaglia's polar method
struct PolarNormal {
    bool saved_available = false;
    double saved = 0.0;
    double operator()(std::mt19937_64& eng) {
        if (saved_available) {
            saved_available = false;
            return saved;
        }
        double x, y, r2;
        do {
            x = 2.0 * canonical(eng) - 1.0;
            y = 2.0 * canonical(eng) - 1.0;
            r2 = x * x + y * y;
        } while (r2 > 1.0 || r2 == 0.0);
        const double f = std::sqrt(-2 * std::log(r2) / r2);
        saved = x * f;
        saved_available = true;
        return y * f;
    }
};

}  // namespace

extern "C" {

void gc_default_disk_galaxy(gc_disk_galaxy* g) {
    if (!g) return;
    g->n_particles = 500;
    g->central_mass = 1.0;
    g->disk_mass = 0.1;
    g->disk_particle_mass = 0.0;
    g->scale_length = 0.3;
    g->r_min = 0.05;
    g->r_max = 1.0;
    g->z_sigma = 0.02;
    g->v_sigma = 0.01;
    g->inclination = 0.0;
    for (int d = 0; d < 3; ++d) g->position[d] = g->velocity[d] = 0.0;
}

size_t gc_generate_disks(const gc_disk_galaxy* galaxies, int count, double G, double eps,
                         uint64_t seed, double* pos, double* vel, double* mass) {
    if (!galaxies || count <= 0 || !pos || !vel || !mass) return 0;
    Rng rng(seed);
    size_t off = 0;
    for (int g = 0; g < count; ++g) {
        one_disk(rng, galaxies[g], G, eps, pos + 3 * off, vel + 3 * off, mass + off);
        off += galaxies[g].n_particles;
    }
    return off;
}

size_t gc_generate_plummer(size_t n, double total_mass, double scale, double G, uint64_t seed,
                           double* pos, double* vel, double* mass) {
    if (n == 0 || !pos || !vel || !mass) return 0;
    Rng rng(seed);
    const double p[gc::ic::kPlCount] = {static_cast<double>(n), total_mass, scale, G};
    for (size_t i = 0; i < n; ++i) {
        double u = rng.uniform01();
        while (u <= 0.0) u = rng.uniform01();
        double ct = rng.uniform(-1.0, 1.0), ph = rng.uniform(0.0, 2.0 * kPi);
        double r;
        gc::ic::plummer_position(p, u, ct, ph, &r, pos + 3 * i);
        double q, g;
        do {
            q = rng.uniform01();
            g = 0.1 * rng.uniform01();
        } while (!gc::ic::plummer_accept(q, g));
        ct = rng.uniform(-1.0, 1.0);
        ph = rng.uniform(0.0, 2.0 * kPi);
        gc::ic::plummer_velocity(p, r, q, ct, ph, vel + 3 * i, mass + i);
    }
    return n;
}

void gc_default_galaxy_config(gc_galaxy_config* c) {
    if (!c) return;
    // GalaxyConfig defaults (cpp/core/initial_conditions.hpp:14-24)
    c->n_particles = 256;
    c->mass = 1.0;
    c->radius = 1.0;
    for (int d = 0; d < 3; ++d) c->position[d] = c->velocity[d] = 0.0;
    c->orientation = 0.0;
    c->group_id = 0;
    c->thickness = 0.0;
    c->inclination = 0.0;
}

size_t gc_generate_galaxies(const gc_galaxy_config* galaxies, int count, double G, double eps,
                            uint64_t seed, double* pos, double* vel, double* mass,
                            uint32_t* group_ids) {
    if (count < 0 || (count > 0 && !galaxies) || !pos || !vel || !mass) return 0;
    std::mt19937_64 eng(seed);
    size_t off = 0;
    for (int g = 0; g < count; ++g) {
        const gc_galaxy_config& c = galaxies[g];
        if (gc::ic::ref_disk_empty(c)) continue;  // no particles, no draws
        double p[gc::ic::kRdCount];
        gc::ic::ref_disk_params(c, G, eps, p);
        PolarNormal normal;
        for (size_t i = 0; i < c.n_particles; ++i, ++off) {
            const double ur = canonical(eng);
            const double ut = canonical(eng);
            const double nz = c.thickness > 0.0 ? normal(eng) : 0.0;
            const double nvx = normal(eng);
            const double nvy = normal(eng);
            const double nvz = normal(eng);
            gc::ic::ref_disk_particle(p, ur, ut, nz, nvx, nvy, nvz, pos + 3 * off, vel + 3 * off,
                                      mass + off);
            if (group_ids) group_ids[off] = c.group_id;
        }
    }
    return off;
}

int gc_sample_collisions(int runs, uint64_t seed, int n_min, int n_max, gc_collision* out) {
    if (runs <= 0 || !out || n_min < 2 || n_max < n_min) return -1;
    Rng rng(seed);
    const double lo = std::log(static_cast<double>(n_min)), hi = std::log(static_cast<double>(n_max));
    for (int r = 0; r < runs; ++r) {
        const double nn = std::exp(rng.uniform(lo, hi));
        long n = 2 * std::lround(nn / 2.0);
        if (n < n_min) n = n_min + (n_min & 1);
        if (n > n_max) n = n_max - (n_max & 1);
        out[r].n_particles = static_cast<size_t>(n);
        out[r].inclination = rng.uniform(0.0, 0.5 * kPi);
        out[r].impact_parameter = rng.uniform(0.5, 2.0);
        out[r].speed = rng.uniform(0.2, 0.5);
        out[r].seed = static_cast<uint64_t>(rng.eng());
    }
    return 0;
}

}  // extern "C"
