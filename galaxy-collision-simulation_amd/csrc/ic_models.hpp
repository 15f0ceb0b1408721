// ic_models.hpp — initial-condition models shared by the host generators
// and the device generator kernel (ic.hip): each model maps a particle's
// random draws to its position, velocity and mass, so host and device runs
// of a model differ only in where the draws come from (and in the last ulp
// of the math library).
//
// Models (parameter layouts in IcParam*):
//   reference disk  generate_disk_galaxy (cpp/core/initial_conditions.cpp
//                   :28-92): truncated exponential radius of scale R/3,
//                   uniform angle, Gaussian thickness, softened circular
//                   speed of the enclosed mass plus Gaussian dispersion,
//                   rotated by inclination (x) then orientation (z).  The
//                   arithmetic follows the reference step for step, so the
//                   host generator fed the reference's draws reproduces it
//                   bit for bit.
//   two-disk galaxy the benchmark collisions' galaxy: a central point mass
//                   plus an exponential disk truncated to [r_min, r_max]
//                   (host/generators.cpp documents it).
//   Plummer sphere  Aarseth, Henon & Wielen (1974).
//
// Counter-based draws (device generator): Philox4x32-10 (Salmon, Moraes,
// Dror & Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC11), key =
// the 64-bit seed, counter = (particle lo, particle hi, stream, block); a
// block gives two doubles of 53 bits, u = (a * 2^21 + (b >> 11)) * 2^-53.
#pragma once

#include <cmath>
#include <cstdint>

#include "gcsim.h"

#if defined(__HIP__)
#define GC_HD __host__ __device__
#else
#define GC_HD
#endif

namespace gc {
namespace ic {

constexpr double kPi = 3.141592653589793238462643383279502884;

// ---- Philox4x32-10 ---------------------------------------------------------
constexpr uint32_t kPhiloxM0 = 0xD2511F53u, kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u, kPhiloxW1 = 0xBB67AE85u;

GC_HD inline void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = static_cast<uint64_t>(kPhiloxM0) * c0;
        const uint64_t p1 = static_cast<uint64_t>(kPhiloxM1) * c2;
        const uint32_t n0 = static_cast<uint32_t>(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = static_cast<uint32_t>(p0 >> 32) ^ c3 ^ k1;
        c1 = static_cast<uint32_t>(p1);
        c3 = static_cast<uint32_t>(p0);
        c0 = n0;
        c2 = n2;
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// The two 53-bit uniforms in [0, 1) of block `block` of particle `i`.
GC_HD inline void philox_uniform2(uint64_t seed, uint64_t i, uint32_t stream, uint32_t block,
                                  double u[2]) {
    const uint32_t ctr[4] = {static_cast<uint32_t>(i), static_cast<uint32_t>(i >> 32), stream,
                             block};
    const uint32_t key[2] = {static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32)};
    uint32_t o[4];
    philox4x32_10(ctr, key, o);
    const double s = 1.0 / 9007199254740992.0;  // 2^-53
    u[0] = static_cast<double>((static_cast<uint64_t>(o[0]) << 21) ^ (o[1] >> 11)) * s;
    u[1] = static_cast<double>((static_cast<uint64_t>(o[2]) << 21) ^ (o[3] >> 11)) * s;
}

// Box-Muller pair from two uniforms in [0, 1) (the first mapped to (0, 1]).
GC_HD inline void box_muller(double ua, double ub, double* n0, double* n1) {
    const double r = std::sqrt(-2.0 * std::log(1.0 - ua));
    const double t = 2.0 * kPi * ub;
    *n0 = r * std::cos(t);
    *n1 = r * std::sin(t);
}

// ---- reference disk (generate_disk_galaxy) ---------------------------------
// Sine and cosine of one angle in one call.  The reference's build (GCC)
// merges each std::sin / std::cos pair of one argument into glibc's sincos,
// whose results differ from separate sin and cos in the last ulp for a few
// arguments; calling it here keeps the host generator bit-identical.
GC_HD inline void sin_cos(double a, double* s, double* c) { ::sincos(a, s, c); }

// p: n, mass, radius, pos[3], vel[3], orientation, thickness, inclination, G, eps
enum IcParamRefDisk {
    kRdN, kRdMass, kRdRadius, kRdPx, kRdPy, kRdPz, kRdVx, kRdVy, kRdVz, kRdOrient,
    kRdThick, kRdIncl, kRdG, kRdEps, kRdCount
};

// rotate_x then rotate_z (initial_conditions.cpp:12-26)
GC_HD inline void orient_disk(double orientation, double inclination, double v[3]) {
    double c, s;
    sin_cos(inclination, &s, &c);
    const double y1 = c * v[1] - s * v[2];
    const double z1 = s * v[1] + c * v[2];
    sin_cos(orientation, &s, &c);
    const double x2 = c * v[0] - s * y1;
    const double y2 = s * v[0] + c * y1;
    v[0] = x2;
    v[1] = y2;
    v[2] = z1;
}

// One particle from its draws: u_r, u_theta uniform in [0, 1); nz (used
// only when thickness > 0), nvx, nvy, nvz standard normal.
GC_HD inline void ref_disk_particle(const double* p, double u_r, double u_theta, double nz,
                                    double nvx, double nvy, double nvz, double x[3], double v[3],
                                    double* m) {
    const double R = p[kRdRadius], M = p[kRdMass], G = p[kRdG], eps = p[kRdEps];
    const double particle_mass = M / p[kRdN];
    const double dispersion = 0.025 * std::sqrt(G * M / R);
    // sample_truncated_exponential_radius (initial_conditions.cpp:28-33)
    const double scale = R / 3.0;
    const double max_cdf = 1.0 - std::exp(-R / scale);
    const double u = u_r < 1.0e-12 ? 1.0e-12 : (u_r > 1.0 - 1.0e-12 ? 1.0 - 1.0e-12 : u_r);
    const double r = -scale * std::log(1.0 - u * max_cdf);
    const double theta = 2.0 * kPi * u_theta;
    double ct, st;
    sin_cos(theta, &st, &ct);
    x[0] = r * ct;
    x[1] = r * st;
    x[2] = p[kRdThick] > 0.0 ? p[kRdThick] * nz : 0.0;
    const double frac_raw = (r * r) / (R * R);
    const double frac = frac_raw < 1.0 ? frac_raw : 1.0;
    const double menc_raw = M * frac;
    const double menc = menc_raw < particle_mass ? particle_mass : menc_raw;
    const double s2 = r * r + eps * eps;
    const double vc = std::sqrt(G * menc * r * r / std::pow(s2, 1.5));
    v[0] = -st * vc;
    v[1] = ct * vc;
    v[2] = 0.0 * vc;
    v[0] += dispersion * nvx;
    v[1] += dispersion * nvy;
    v[2] += 0.25 * dispersion * nvz;
    orient_disk(p[kRdOrient], p[kRdIncl], x);
    orient_disk(p[kRdOrient], p[kRdIncl], v);
    x[0] = p[kRdPx] + x[0];
    x[1] = p[kRdPy] + x[1];
    x[2] = p[kRdPz] + x[2];
    v[0] = p[kRdVx] + v[0];
    v[1] = p[kRdVy] + v[1];
    v[2] = p[kRdVz] + v[2];
    *m = particle_mass;
}

// ---- two-disk galaxy --------------------------------------------------------
// p: n (central mass + n - 1 disk), central_mass, disk particle mass, disk
// total mass, scale_length, r_min, r_max, z_sigma, v_sigma, inclination,
// pos[3], vel[3], G, eps
enum IcParamTwoDisk {
    kTdN, kTdCentral, kTdMd, kTdDiskTotal, kTdH, kTdRmin, kTdRmax, kTdZs, kTdVs, kTdIncl,
    kTdPx, kTdPy, kTdPz, kTdVx, kTdVy, kTdVz, kTdG, kTdEps, kTdCount
};

GC_HD inline void rotate_x(double angle, double v[3]) {
    const double c = std::cos(angle), s = std::sin(angle);
    const double y = c * v[1] - s * v[2];
    const double z = s * v[1] + c * v[2];
    v[1] = y;
    v[2] = z;
}

// Particle k of the galaxy (k = 0: the central mass, no draws used).
GC_HD inline void two_disk_particle(const double* p, uint64_t k, double u_r, double u_phi,
                                    double nz, double nvx, double nvy, double nvz, double x[3],
                                    double v[3], double* m) {
    x[0] = x[1] = x[2] = 0.0;
    v[0] = v[1] = v[2] = 0.0;
    if (k == 0) {
        *m = p[kTdCentral];
    } else {
        const double h = p[kTdH];
        const double ea = std::exp(-p[kTdRmin] / h), eb = std::exp(-p[kTdRmax] / h);
        *m = p[kTdMd];
        const double R = -h * std::log(ea - u_r * (ea - eb));
        const double phi = 2.0 * kPi * u_phi;
        x[0] = R * std::cos(phi);
        x[1] = R * std::sin(phi);
        x[2] = p[kTdZs] * nz;
        const double cdf = (ea - std::exp(-R / h)) / (ea - eb);
        const double menc = p[kTdCentral] + p[kTdDiskTotal] * cdf;
        const double s2 = R * R + p[kTdEps] * p[kTdEps];
        const double vc = std::sqrt(p[kTdG] * menc * R * R / (s2 * std::sqrt(s2)));
        v[0] = -std::sin(phi) * vc + p[kTdVs] * nvx;
        v[1] = std::cos(phi) * vc + p[kTdVs] * nvy;
        v[2] = p[kTdVs] * nvz;
    }
    rotate_x(p[kTdIncl], x);
    rotate_x(p[kTdIncl], v);
    x[0] = p[kTdPx] + x[0];
    x[1] = p[kTdPy] + x[1];
    x[2] = p[kTdPz] + x[2];
    v[0] = p[kTdVx] + v[0];
    v[1] = p[kTdVy] + v[1];
    v[2] = p[kTdVz] + v[2];
}

// ---- Plummer sphere ---------------------------------------------------------
// p: n, total_mass, scale, G
enum IcParamPlummer { kPlN, kPlMass, kPlScale, kPlG, kPlCount };

// Radius from u in [0, 1) (u = 0: the centre) and the isotropic direction (ct = cos theta in
// [-1, 1), ph in [0, 2 pi)).
GC_HD inline void plummer_position(const double* p, double u, double ct, double ph, double* r,
                                   double x[3]) {
    *r = 1.0 / std::sqrt(std::pow(u, -2.0 / 3.0) - 1.0);
    const double st0 = 1.0 - ct * ct;
    const double st = std::sqrt(st0 > 0.0 ? st0 : 0.0);
    const double a = p[kPlScale];
    x[0] = a * *r * st * std::cos(ph);
    x[1] = a * *r * st * std::sin(ph);
    x[2] = a * *r * ct;
}

// Rejection test for q (the speed in units of the escape speed): accept
// when g (uniform in [0, 0.1)) <= q^2 (1 - q^2)^(7/2).
GC_HD inline bool plummer_accept(double q, double g) {
    return !(g > q * q * std::pow(1.0 - q * q, 3.5));
}

GC_HD inline void plummer_velocity(const double* p, double r, double q, double ct, double ph,
                                   double v[3], double* m) {
    const double vscale = std::sqrt(p[kPlG] * p[kPlMass] / p[kPlScale]);
    const double speed = q * std::sqrt(2.0) * std::pow(1.0 + r * r, -0.25);
    const double st0 = 1.0 - ct * ct;
    const double st = std::sqrt(st0 > 0.0 ? st0 : 0.0);
    v[0] = vscale * speed * st * std::cos(ph);
    v[1] = vscale * speed * st * std::sin(ph);
    v[2] = vscale * speed * ct;
    *m = p[kPlMass] / p[kPlN];
}

// ---- parameter blocks of the models (host) ---------------------------------
inline void ref_disk_params(const gc_galaxy_config& c, double G, double eps, double* p) {
    p[kRdN] = static_cast<double>(c.n_particles);
    p[kRdMass] = c.mass;
    p[kRdRadius] = c.radius;
    for (int d = 0; d < 3; ++d) {
        p[kRdPx + d] = c.position[d];
        p[kRdVx + d] = c.velocity[d];
    }
    p[kRdOrient] = c.orientation;
    p[kRdThick] = c.thickness;
    p[kRdIncl] = c.inclination;
    p[kRdG] = G;
    p[kRdEps] = eps;
}

inline void two_disk_params(const gc_disk_galaxy& g, double G, double eps, double* p) {
    const size_t nd = g.n_particles > 0 ? g.n_particles - 1 : 0;
    const double md = g.disk_particle_mass > 0.0 ? g.disk_particle_mass
                                                 : (nd > 0 ? g.disk_mass / static_cast<double>(nd) : 0.0);
    p[kTdN] = static_cast<double>(g.n_particles);
    p[kTdCentral] = g.central_mass;
    p[kTdMd] = md;
    p[kTdDiskTotal] = md * static_cast<double>(nd);
    p[kTdH] = g.scale_length;
    p[kTdRmin] = g.r_min;
    p[kTdRmax] = g.r_max;
    p[kTdZs] = g.z_sigma;
    p[kTdVs] = g.v_sigma;
    p[kTdIncl] = g.inclination;
    for (int d = 0; d < 3; ++d) {
        p[kTdPx + d] = g.position[d];
        p[kTdVx + d] = g.velocity[d];
    }
    p[kTdG] = G;
    p[kTdEps] = eps;
}

// generate_disk_galaxy draws nothing for these (initial_conditions.cpp:45-47)
inline bool ref_disk_empty(const gc_galaxy_config& c) {
    return c.n_particles == 0 || c.mass <= 0.0 || c.radius <= 0.0;
}

}  // namespace ic
}  // namespace gc
