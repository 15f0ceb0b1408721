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
#include "gcsim.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>

namespace {

constexpr double kPi = 3.14159265358979323846;

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

void rotate_x(double angle, double* v) {
    const double c = std::cos(angle), s = std::sin(angle);
    const double y = c * v[1] - s * v[2];
    const double z = s * v[1] + c * v[2];
    v[1] = y;
    v[2] = z;
}

// One galaxy: central mass at index 0, then n_disk disk particles.
void one_disk(Rng& rng, const gc_disk_galaxy& g, double G, double eps, double* pos, double* vel,
              double* mass) {
    const double h = g.scale_length;
    const double ea = std::exp(-g.r_min / h), eb = std::exp(-g.r_max / h);
    const size_t nd = g.n_particles > 0 ? g.n_particles - 1 : 0;
    const double md = g.disk_particle_mass > 0.0 ? g.disk_particle_mass
                                                 : (nd > 0 ? g.disk_mass / static_cast<double>(nd) : 0.0);
    const double disk_total = md * static_cast<double>(nd);
    for (size_t k = 0; k < g.n_particles; ++k) {
        double p[3] = {0, 0, 0}, v[3] = {0, 0, 0}, m;
        if (k == 0) {
            m = g.central_mass;
        } else {
            m = md;
            const double u = rng.uniform01();
            const double R = -h * std::log(ea - u * (ea - eb));
            const double phi = 2.0 * kPi * rng.uniform01();
            const double zz = g.z_sigma * rng.normal();
            p[0] = R * std::cos(phi);
            p[1] = R * std::sin(phi);
            p[2] = zz;
            const double cdf = (ea - std::exp(-R / h)) / (ea - eb);
            const double menc = g.central_mass + disk_total * cdf;
            const double s2 = R * R + eps * eps;
            const double vc = std::sqrt(G * menc * R * R / (s2 * std::sqrt(s2)));
            v[0] = -std::sin(phi) * vc + g.v_sigma * rng.normal();
            v[1] = std::cos(phi) * vc + g.v_sigma * rng.normal();
            v[2] = g.v_sigma * rng.normal();
        }
        rotate_x(g.inclination, p);
        rotate_x(g.inclination, v);
        for (int d = 0; d < 3; ++d) {
            pos[3 * k + d] = g.position[d] + p[d];
            vel[3 * k + d] = g.velocity[d] + v[d];
        }
        mass[k] = m;
    }
}

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
    const double vscale = std::sqrt(G * total_mass / scale);
    for (size_t i = 0; i < n; ++i) {
        double u = rng.uniform01();
        while (u <= 0.0) u = rng.uniform01();
        const double r = 1.0 / std::sqrt(std::pow(u, -2.0 / 3.0) - 1.0);
        double ct = rng.uniform(-1.0, 1.0), ph = rng.uniform(0.0, 2.0 * kPi);
        double st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
        pos[3 * i] = scale * r * st * std::cos(ph);
        pos[3 * i + 1] = scale * r * st * std::sin(ph);
        pos[3 * i + 2] = scale * r * ct;
        double q, g;
        do {
            q = rng.uniform01();
            g = 0.1 * rng.uniform01();
        } while (g > q * q * std::pow(1.0 - q * q, 3.5));
        const double speed = q * std::sqrt(2.0) * std::pow(1.0 + r * r, -0.25);
        ct = rng.uniform(-1.0, 1.0);
        ph = rng.uniform(0.0, 2.0 * kPi);
        st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
        vel[3 * i] = vscale * speed * st * std::cos(ph);
        vel[3 * i + 1] = vscale * speed * st * std::sin(ph);
        vel[3 * i + 2] = vscale * speed * ct;
        mass[i] = total_mass / static_cast<double>(n);
    }
    return n;
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
