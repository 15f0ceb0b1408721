/*
 * gcoracle — CPU FP64 restatement of the galaxy-collision force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path (libgcsim) never links or calls it.
 *
 * Every function restates the reference's arithmetic in the same FP64
 * operation order (compiled with -ffp-contract=off), so that results are
 * bit-identical to the reference on the same inputs.  Arrays are
 * particle-major: pos/vel/acc are n x 3 (x,y,z interleaved), mass is n.
 *
 * Parity pins (all exercised by tests/test_oracle_*.py):
 *   direct (symmetric + targets)  — reference golden snapshots/diagnostics
 *                                   (direct_{250,500,1000}), closed-form
 *                                   two-body, momentum conservation,
 *                                   oracle/_ref on random inputs.
 *   kick/drift/leapfrog           — golden snapshot_000020 (20 KDK steps),
 *                                   time reversibility, circular orbit.
 *   diagnostics                   — golden diagnostics.csv rows.
 *   P2M/M2M                       — M2M(P2M(children)) == P2M(all),
 *                                   brute-force monomial sums.
 *   M2P                           — p=0 limit == point-mass force,
 *                                   error falls with p, _ref per call.
 *   FMM                           — golden readme snapshot_000000 FMM
 *                                   accelerations (bit-level), smoke-test
 *                                   80-particle threshold, _ref.
 */
#ifndef GCORACLE_H
#define GCORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- direct all-pairs (cpp/direct/direct_solver.cpp) ---------------- */
void orc_softened_acceleration(const double target[3], const double source[3],
                               double source_mass, double G, double eps,
                               double out[3]);
void orc_direct_symmetric(size_t n, const double* pos, const double* mass,
                          double G, double eps, double* acc);
void orc_direct_targets(size_t n, const double* pos, const double* mass,
                        double G, double eps, size_t begin, size_t end,
                        double* acc);
/* Same per-target arithmetic as orc_direct_targets, targets split over
 * OpenMP threads (bit-identical result).  Returns threads used. */
int orc_direct_targets_omp(size_t n, const double* pos, const double* mass,
                           double G, double eps, size_t begin, size_t end,
                           double* acc, int threads);

/* ---- integrator (cpp/core/integrator.cpp) --------------------------- */
void orc_kick(size_t n, double* vel, const double* acc, double dt);
void orc_drift(size_t n, double* pos, const double* vel, double dt);

/* solver: 0 = direct symmetric, 1 = direct targets, 2 = fmm, 3 = tree */
typedef struct {
    int solver;
    double G, eps;
    double theta;
    size_t leaf_capacity;
    int max_depth;
    int expansion_order;
} orc_solver;

int orc_compute_accelerations(const orc_solver* s, size_t n, const double* pos,
                              const double* mass, double* acc);
/* `steps` kick-drift-kick steps; acc must hold a(pos) on entry. */
int orc_leapfrog(const orc_solver* s, size_t n, double* pos, double* vel,
                 double* acc, const double* mass, double dt, int steps);

/* ---- diagnostics (cpp/core/diagnostics.cpp) ------------------------- */
/* out[16]: total_mass, KE, PE, E, P[3], COM[3], L[3], (pad) */
void orc_diagnostics(size_t n, const double* pos, const double* vel,
                     const double* mass, double G, double eps, double* out);

/* ---- Cartesian multipoles (cpp/fmm/multipole.cpp) ------------------- */
void orc_multipole_exponents(int* out /* 35*3 */);
void orc_p2m(double* moments /*35*/, const double offset[3], double mass);
void orc_m2m(double* parent, const double* child, const double offset[3],
             double child_mass);
void orc_m2p(const double target[3], const double com[3], double mass,
             const double* moments, double G, double eps, int order,
             double out[3]);

/* ---- FMM (cpp/fmm/fmm_solver.cpp) ----------------------------------- */
typedef struct {
    size_t node_count, leaf_count, far_interactions, near_interactions;
    int max_depth_reached;
} orc_fmm_stats;

int orc_fmm(size_t n, const double* pos, const double* mass, double G,
            double eps, double theta, size_t leaf_capacity, int max_depth,
            int expansion_order, size_t begin, size_t end, double* acc,
            orc_fmm_stats* stats);

/* ---- Barnes-Hut tree (cpp/fmm/quadtree.cpp) --------------------------- */
typedef struct {
    size_t node_count, leaf_count;  /* leaves holding particles */
    size_t m2p_evaluations, p2p_pairs;
    int max_depth_reached;
} orc_tree_stats;

int orc_tree(size_t n, const double* pos, const double* mass, double G, double eps,
             double theta, size_t leaf_capacity, int max_depth, int expansion_order,
             size_t begin, size_t end, double* acc, orc_tree_stats* stats);

/* ---- initial conditions -------------------------------------------- */
/* n-th output (1-based) of mt19937_64(seed). */
uint64_t orc_mt64_nth(uint64_t seed, size_t n);
/* generate_galaxies (cpp/core/initial_conditions.cpp:94-114) with GNU
 * libstdc++'s uniform and normal draws.  gal: count x 12 doubles (n, mass,
 * radius, pos[3], vel[3], orientation, thickness, inclination).  Returns
 * the particle count written. */
size_t orc_generate_galaxies(const double* gal, int count, double G, double eps, uint64_t seed,
                             double* pos, double* vel, double* mass);
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* The device generators' models on the Philox stream of (seed, stream):
 * model 0 reference disk (p: n, mass, radius, pos3, vel3, orientation,
 * thickness, inclination, G, eps), 1 two-disk galaxy (p: n, central mass,
 * disk particle mass, disk total, scale length, r_min, r_max, z_sigma,
 * v_sigma, inclination, pos3, vel3, G, eps), 2 Plummer (p: n, total mass,
 * scale, G).  Returns 0, or 1 for an unknown model. */
int orc_ic_generate(int model, const double* p, uint64_t seed, uint32_t stream, size_t n,
                    double* pos, double* vel, double* mass);

#ifdef __cplusplus
}
#endif
#endif
