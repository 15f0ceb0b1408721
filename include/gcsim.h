/*
 * gcsim.h — C ABI of the MI355X (gfx950) force engine for the galaxy
 * collision simulator's hot path: softened all-pairs gravity, the FMM
 * near/far field, and kick-drift-kick leapfrog, with particles resident in
 * HBM across steps.
 *
 * Plain C: opaque handles, int status codes, plain pointers and sizes, no
 * exceptions and no torch types.  Host arrays are particle-major n x 3
 * (x, y, z interleaved, the field order of the reference's Particle) and
 * are only read or written during the call.  Device state is SoA FP64 and
 * owned by the context.  One context per host thread; each context runs on
 * its own HIP stream (replaceable with gc_set_stream for interop).
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   cpp/cuda/cuda_solver.hpp:10   bool cuda_solver_available()
 *   cpp/cuda/cuda_solver.hpp:11   compute_cuda_direct_accelerations(particles, params)
 *   cpp/cuda/cuda_solver.hpp:12   cuda_direct_leapfrog_step(particles, dt, params)
 *   cpp/direct/direct_solver.hpp:22-28  compute_direct_accelerations[_for_targets]
 *   cpp/fmm/fmm_solver.hpp:71-83  compute_fmm_accelerations[_for_targets]
 *   cpp/core/integrator.hpp:13-21 kick / drift / leapfrog_step
 *   cpp/core/diagnostics.hpp:21   compute_diagnostics
 */
#ifndef GCSIM_H
#define GCSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GCSIM_ABI_VERSION 1

typedef enum {
    GC_OK = 0,
    GC_ERR_INVALID_ARGUMENT = 1,
    GC_ERR_NO_DEVICE = 2,
    GC_ERR_HIP = 3,
    GC_ERR_OUT_OF_MEMORY = 4,
    GC_ERR_UNSUPPORTED = 5,
    GC_ERR_STATE = 6,
    GC_ERR_IO = 7
} gc_status;

/* solver choice: the `solver = "hip-direct" | "hip-fmm" | "hip-tree"` config
 * values.  GC_SOLVER_TREE is the reference's Barnes-Hut solver
 * (compute_tree_accelerations, cpp/fmm/quadtree.cpp:245-254, selected by
 * solver = "tree"): the FMM's tree, walked per target particle with the
 * acceptance width / |com - x| < theta outside the cell, M2P of order
 * expansion_order, pair forces in opened leaves.  It uses theta,
 * leaf_capacity, max_depth and expansion_order below (the reference fixes
 * max_depth at 32). */
typedef enum { GC_SOLVER_DIRECT = 0, GC_SOLVER_FMM = 1, GC_SOLVER_TREE = 2 } gc_solver_kind;

/* GC_PRECISION_FP64: FP64 pairs, per target ascending-j sums, the same
 *   operation order as compute_direct_accelerations_for_targets
 *   (cpp/direct/direct_solver.cpp:59-83): bit-identical to the reference's
 *   targets form.
 * GC_PRECISION_MIXED: FP32 pair forces on FP32 offsets from FP64 tile
 *   centres, per-tile FP32 partials merged by two-sum, FP64 state.
 * GC_PRECISION_FP64_FAST: the direct solver's FP64 throughput form: FP64
 *   pairs with 1/sqrt from the hardware estimate plus one correction step,
 *   sources split into ranges whose FP64 partial sums are added in range
 *   order (deterministic); it differs from GC_PRECISION_FP64 by rounding
 *   only.  The FMM and tree solvers and the adjoint treat it as FP64. */
typedef enum {
    GC_PRECISION_FP64 = 0,
    GC_PRECISION_MIXED = 1,
    GC_PRECISION_FP64_FAST = 2
} gc_precision;

/* PhysicsParams (cpp/core/simulation_params.hpp:5-8) + FmmOptions
 * (cpp/fmm/fmm_solver.hpp:13-18) + solver/precision selection. */
typedef struct {
    double gravitational_constant; /* G, default 1 */
    double softening;              /* Plummer epsilon, default 0.02 */
    int solver;                    /* gc_solver_kind */
    int precision;                 /* gc_precision */
    double theta;                  /* FMM opening angle, default 0.6 */
    int leaf_capacity;             /* FMM leaf size, default 16 */
    int max_depth;                 /* FMM depth cap, default 32 */
    int expansion_order;           /* FMM order 0..4, default 4 */
} gc_params;

/* Diagnostics (cpp/core/diagnostics.hpp:11-19), FP64 */
typedef struct {
    double kinetic_energy;
    double potential_energy;
    double total_energy;
    double total_mass;
    double momentum[3];
    double center_of_mass[3];
    double angular_momentum[3];
} gc_diagnostics;

/* Device addresses of the resident SoA state (for torch interop). */
typedef struct {
    size_t n;
    double *x, *y, *z;
    double *vx, *vy, *vz;
    double *ax, *ay, *az;
    double* mass;
    void* stream;
} gc_device_view;

typedef struct gc_ctx gc_ctx;

/* Fills p with the reference defaults: G=1, softening=0.02, direct, fp64,
 * theta=0.6, leaf 16, depth 32, order 4. */
void gc_default_params(gc_params* p);

/* Number of visible HIP devices (0 without a GPU). */
int gc_device_count(void);

/* cuda_solver_available (cpp/cuda/cuda_solver.hpp:10): 1 iff a gfx950
 * device is visible.  There is no CPU fallback: callers must check. */
int gc_solver_available(void);

const char* gc_status_string(gc_status s);
/* Last error message of ctx (or of the calling thread when ctx is NULL).
 * No C++ exception leaves an entry point: a failed host allocation returns
 * GC_ERR_OUT_OF_MEMORY with its message in gc_last_error(NULL). */
const char* gc_last_error(const gc_ctx* ctx);

gc_status gc_create(gc_ctx** out, int device, const gc_params* p);
void gc_destroy(gc_ctx* ctx);
gc_status gc_set_params(gc_ctx* ctx, const gc_params* p);
/* Runs the context's work on hip_stream (NULL = the null stream, e.g.
 * torch's default stream); gc_use_own_stream returns to the context's own
 * non-blocking stream. */
gc_status gc_set_stream(gc_ctx* ctx, void* hip_stream);
gc_status gc_use_own_stream(gc_ctx* ctx);
void* gc_get_stream(gc_ctx* ctx);

/* Uploads n particles (pos, vel: n x 3; mass: n).  vel may be NULL (zero).
 * Accelerations are zeroed; call gc_compute_accelerations before stepping. */
gc_status gc_upload(gc_ctx* ctx, size_t n, const double* pos, const double* vel,
                    const double* mass);
/* Replaces the resident accelerations (n x 3), e.g. with a(x) the caller
 * already holds, as cuda_direct_leapfrog_step starts from the particles'
 * accelerations (cpp/cuda/cuda_solver.cu:212-246). */
gc_status gc_set_accelerations(gc_ctx* ctx, const double* acc);
/* Copies state back (any pointer may be NULL).  Synchronises the stream. */
gc_status gc_download(gc_ctx* ctx, double* pos, double* vel, double* acc);
size_t gc_particle_count(const gc_ctx* ctx);
gc_status gc_device_state(gc_ctx* ctx, gc_device_view* out);

/* a(x) on the resident state: compute_*_accelerations equivalent. */
gc_status gc_compute_accelerations(gc_ctx* ctx);
/* nsteps KDK steps (integrator.cpp:33-42): kick dt/2, drift dt, a(x),
 * kick dt/2.  Requires a(x) of the current positions (gc_compute_accelerations
 * or a previous step).  Asynchronous on the context's stream.  With the
 * direct solver, whole chunks of gc_get_graph_steps() steps are replayed on
 * the context's stream from a hipGraph (kernels only, captured once on a
 * private stream of the context; recaptured when dt, the parameters, the
 * launch shape or the particle buffers change); the remainder, a stream the
 * caller is capturing and a chunk that cannot be captured or instantiated
 * get eager launches of the same kernels, so the results are the same
 * either way. */
gc_status gc_step(gc_ctx* ctx, double dt, int nsteps);
/* Steps per captured gc_step chunk, 0..1000 (default 10; 0 or 1 launches
 * every kernel eagerly).  No reference counterpart: cuda_direct_leapfrog_step
 * (cpp/cuda/cuda_solver.cu:212-246) launches its kernels one by one.
 * gc_get_graph_steps returns -1 for a null context. */
gc_status gc_set_graph_steps(gc_ctx* ctx, int steps);
int gc_get_graph_steps(const gc_ctx* ctx);
gc_status gc_synchronize(gc_ctx* ctx);

/* FP64 diagnostics of the resident state on the device (wavefront
 * reductions; O(N^2/2) softened potential).  Synchronises. */
gc_status gc_diagnostics_compute(gc_ctx* ctx, gc_diagnostics* out);

/* Accelerations of an explicit target range [begin, end) only
 * (compute_*_accelerations_for_targets); other entries keep their values. */
gc_status gc_compute_accelerations_for_targets(gc_ctx* ctx, size_t begin, size_t end);

/* Measurement hooks: average duration (ms) of the dominant force kernel
 * over `reps` back-to-back launches on the context's stream (HIP events),
 * and the FP32-pair interactions one launch performs. */
gc_status gc_time_force_kernel(gc_ctx* ctx, int reps, double* ms_per_launch,
                               double* interactions_per_launch);

/* ---- no device allocation (no counterpart in the reference: for callers
 * that cannot allocate device memory, such as code under HIP graph
 * capture, tight serving loops and harnesses that forbid hipMalloc) ----
 * Mixed-precision direct accelerations of n particles held in caller-owned
 * device arrays: pos [3][n] (all x, then all y, then all z) and mass [n] in,
 * acc [3][n] out.  G and softening come from p, which must select
 * GC_SOLVER_DIRECT and GC_PRECISION_MIXED.  It runs the per-target kernels
 * of gc_compute_accelerations (GC_PAIRS_TARGETS, one target per lane,
 * scalar-load sources), asynchronously on `stream` (NULL: the null stream)
 * of the device that owns that stream.
 * The call performs no hipMalloc, hipFree or other allocation: its working
 * buffers (source tiles, tile centres, per-split partial sums, 16.8 MB in
 * all) are one static device array that each device holds until the process
 * ends.  Every call on a device shares that device's array, so calls that
 * overlap on one device race.  n above GC_NOALLOC_MAX_PARTICLES returns
 * GC_ERR_INVALID_ARGUMENT before any launch.  Source splits: the context
 * path's automatic count, min(64, ceil(16384 / T), T) for T = ceil(n / 256)
 * tiles, capped at GC_NOALLOC_PARTIAL_FLOATS / (6 * 256 * T); a context set
 * to the same split count (gc_set_launch_shape(ctx, 1, splits),
 * GC_PAIRS_TARGETS) gives bit-identical accelerations. */
#define GC_NOALLOC_MAX_PARTICLES 262144
#define GC_NOALLOC_PARTIAL_FLOATS 3145728
gc_status gc_direct_accelerations_noalloc(const gc_params* p, size_t n, const double* pos,
                                          const double* mass, double* acc, void* stream);

/* ---- differentiation (no reference counterpart: the reference has no
 * adjoint; these serve gradient-based callers such as torch.autograd) ----
 * Device arrays below are caller-owned, FP64, SoA "[3][n]" (all x, then all
 * y, then all z) and are read/written on the context's stream. */

/* Replaces the resident state from device memory; vel/acc may be NULL
 * (zeroed). */
gc_status gc_set_state_device(gc_ctx* ctx, size_t n, const double* pos, const double* vel,
                              const double* mass, const double* acc);
/* Copies the resident state to device memory (any pointer may be NULL). */
gc_status gc_get_state_device(gc_ctx* ctx, double* pos, double* vel, double* acc);
/* nsteps time-reversed KDK steps (the leapfrog with -dt), the recomputation
 * the adjoint uses to walk a trajectory backwards between checkpoints. */
gc_status gc_step_reverse(gc_ctx* ctx, double dt, int nsteps);
/* Vector-Jacobian product of the direct force at the resident positions and
 * masses: given lambda = dL/da ([3][n]), writes (accumulate == 0) or adds
 * xbar = (da/dx)^T lambda ([3][n]) and mbar = (da/dm)^T lambda ([n]).
 * FP64 or mixed per the context's precision; solver must be direct.  The
 * mixed form evaluates each unordered pair once for full-range states up to
 * 262,144 particles (the pair form gc_set_pairs selects; the per-target
 * form otherwise), with targets per lane from gc_set_launch_shape. */
gc_status gc_force_vjp(gc_ctx* ctx, const double* lambda, double* xbar, double* mbar,
                       int accumulate);

/* ---- batched runs (no reference counterpart: the reference runs one
 * simulation per process) ----
 * Many independent collisions of uneven N advanced together on one device
 * with the mixed-precision direct solver: one persistent force launch per
 * step over a static work list, multi-step chunks replayed from a captured
 * hipGraph.  Host arrays are the runs concatenated in run order (n_total x 3
 * and n_total). */
typedef struct gc_batch gc_batch;

gc_status gc_batch_create(gc_batch** out, int device, const gc_params* p, int runs,
                          const size_t* counts);
void gc_batch_destroy(gc_batch* b);
const char* gc_batch_last_error(const gc_batch* b);
size_t gc_batch_particle_count(const gc_batch* b);
gc_status gc_batch_upload(gc_batch* b, const double* pos, const double* vel, const double* mass);
gc_status gc_batch_download(gc_batch* b, double* pos, double* vel, double* acc);
gc_status gc_batch_compute_accelerations(gc_batch* b);
gc_status gc_batch_step(gc_batch* b, double dt, int nsteps);
gc_status gc_batch_synchronize(gc_batch* b);
/* Schedule of the pair form in use (gc_batch_set_pairs).  Per target:
 * chunk_tiles = source tiles (of 256 particles) per work item (>= 1),
 * blocks = persistent workgroups (>= 1).  Symmetric: chunk_tiles = source
 * tiles per work item of a row (0: about 16,384 items per run), blocks =
 * persistent workgroups (0: one per work item).  graph_steps: steps per
 * captured graph (0/1: none). */
gc_status gc_batch_set_schedule(gc_batch* b, int chunk_tiles, int blocks, int graph_steps);
gc_status gc_batch_get_schedule(gc_batch* b, int* chunk_tiles, int* blocks, int* graph_steps);
/* Times schedule candidates on this batch's force kernel and keeps the
 * fastest; results are cached by device model (architecture name and CU
 * count), batch shape and requested pair form (gc_batch_set_pairs), per
 * process and in a text file that later processes read:
 * $GCSIM_TUNE_CACHE, else $HOME/.cache/gcsim/batch_tune.txt (an empty
 * GCSIM_TUNE_CACHE turns the file off).  One line per result:
 *   gcsim-tune 1 <arch> <CUs> <runs> <tiles> <items at chunk 1>
 *   <requested form> <form> <chunk> <blocks> <ipt> <symmetric blocks>
 *   <item length>
 * the last line of a key wins. */
gc_status gc_batch_autotune(gc_batch* b, int* chunk_tiles, int* blocks);
gc_status gc_batch_time_force_kernel(gc_batch* b, int reps, double* ms_per_launch,
                                     double* interactions_per_launch);

/* FmmStats (cpp/fmm/fmm_solver.hpp:20-25) of the last hip-fmm force
 * evaluation, plus the particle-level work it implied.  After a hip-tree
 * evaluation: node_count, tree_nodes and max_depth as for hip-fmm;
 * leaf_count = leaves holding particles; far/near_interactions = 0 (no
 * lists); m2p_evaluations and p2p_pairs from the walks. */
typedef struct {
    size_t node_count;        /* the reference's count: 1 + 8 * internal nodes */
    size_t leaf_count;        /* leaves with mass > 0 (target leaves) */
    size_t far_interactions;  /* sum over target leaves of far-list length */
    size_t near_interactions; /* sum over target leaves of near-list length */
    size_t m2p_evaluations;   /* particle x far-cell evaluations */
    size_t p2p_pairs;         /* directed particle pairs in near leaves, self excluded */
    size_t tree_nodes;        /* non-empty nodes materialised on the device */
    int max_depth;            /* deepest level reached */
    float build_ms;           /* tree, moments and lists */
    float eval_ms;            /* M2P + P2P evaluation */
} gc_fmm_stats;

gc_status gc_fmm_last_stats(gc_ctx* ctx, gc_fmm_stats* out);

/* Calibration for HBM counters: copies `bytes` (multiple of 16) between two
 * fresh device buffers `reps` times with a plain float4 copy kernel
 * (calibration_copy_kernel) on `device`; returns ms per copy. */
gc_status gc_calibration_copy(int device, size_t bytes, int reps, double* ms_per_copy);

/* Tuning: mixed-kernel launch shape (targets per thread 1/2/4, j-splits),
 * also that of the FP64 throughput form (default 2 targets per lane).
 * 0 = automatic.  gc_get_launch_shape reports the shape a full-range force
 * of the resident particles uses: for the symmetric form its targets per
 * lane and jsplit 0 (it has no source splits). */
gc_status gc_set_launch_shape(gc_ctx* ctx, int targets_per_thread, int jsplit);
gc_status gc_get_launch_shape(gc_ctx* ctx, int* targets_per_thread, int* jsplit);

/* Tuning: how the mixed kernel's waves read source particles.  LDS: the
 * block stages each source tile in LDS (two barriers per tile) and lanes
 * read it by broadcast.  SCALAR: each wave reads the tile with wave-uniform
 * scalar loads into SGPRs, no LDS and no barrier.  Same arithmetic in the
 * same order, so both give identical sums.  AUTO (the default) is SCALAR,
 * the faster form on gfx950; gc_get_sources returns the form in effect
 * (1 or 2), -1 for a null context. */
enum { GC_SOURCES_AUTO = 0, GC_SOURCES_LDS = 1, GC_SOURCES_SCALAR = 2 };
gc_status gc_set_sources(gc_ctx* ctx, int mode);
int gc_get_sources(const gc_ctx* ctx);

/* Tuning: how the mixed path covers the pairs of a full-range force.
 * TARGETS: every target sums over all sources (the targets form,
 * cpp/direct/direct_solver.cpp:59-83), 2 rsq per unordered pair.
 * SYMMETRIC: each unordered pair once, its force added to one particle and
 * subtracted from the other (the reference's symmetric form,
 * cpp/direct/direct_solver.cpp:37-57), 1 rsq per unordered pair; used up to
 * 262,144 particles, target ranges always run TARGETS.  AUTO (the
 * default) is SYMMETRIC where it applies.  Both meet the same error bound;
 * sums differ by rounding.  gc_set_launch_shape's targets_per_thread also
 * sets the SYMMETRIC kernel's targets per lane (default 4, or 2 or 1 when
 * that leaves fewer than 8192 waves).  gc_get_pairs returns the form a
 * full-range force of the resident particles uses (1 or 2), -1 for a null
 * context. */
enum { GC_PAIRS_AUTO = 0, GC_PAIRS_TARGETS = 1, GC_PAIRS_SYMMETRIC = 2 };
gc_status gc_set_pairs(gc_ctx* ctx, int mode);
int gc_get_pairs(const gc_ctx* ctx);
/* Tuning: source tiles (of 256 particles) per work item of the SYMMETRIC
 * form, 0..1024 (0, the default: about 16,384 items, at least one tile
 * each).  The same quantity as a symmetric batch's chunk_tiles
 * (gc_batch_set_schedule), so a run of a batch and the same run on a
 * context give bit-identical sums when both use the same item length and
 * targets per lane.  Sums of different item lengths differ by rounding. */
gc_status gc_set_pair_items(gc_ctx* ctx, int tiles);

/* The same choice for a batch: SYMMETRIC covers each unordered pair of a
 * run once (runs of up to 262,144 particles; AUTO is SYMMETRIC when every
 * run fits), TARGETS is the scheduled per-target kernel above.
 * targets_per_lane: 1, 2, 4 or 0 (automatic, as for a context).
 * gc_batch_get_pairs reports the form in use and its targets per lane.
 * gc_batch_autotune times both forms unless one was set explicitly and
 * reports the winner's schedule as gc_batch_get_schedule does. */
gc_status gc_batch_set_pairs(gc_batch* b, int mode, int targets_per_lane);
gc_status gc_batch_get_pairs(gc_batch* b, int* mode, int* targets_per_lane);


/* ---- synthetic initial conditions (host side) -------------------------
 * No reference counterpart for these workloads: the reference's generator
 * (cpp/core/initial_conditions.cpp:37-114) has no central mass and no
 * Plummer sphere.  Deterministic per seed (std::mt19937_64 with explicit
 * uniform/normal transforms). */
typedef struct {
    size_t n_particles;        /* 1 central mass + (n_particles - 1) disk */
    double central_mass;       /* default 1.0 */
    double disk_mass;          /* total disk mass, used when disk_particle_mass == 0 */
    double disk_particle_mass; /* > 0 overrides disk_mass / (n_particles - 1) */
    double scale_length;       /* exponential scale, default 0.3 */
    double r_min, r_max;       /* radius truncation, default [0.05, 1.0] */
    double z_sigma;            /* vertical N(0, z_sigma), default 0.02 */
    double v_sigma;            /* per-component velocity dispersion, default 0.01 */
    double inclination;        /* rotation about x (radians) */
    double position[3], velocity[3];
} gc_disk_galaxy;

typedef struct {
    size_t n_particles;      /* even, split equally between the two galaxies */
    double inclination;      /* radians, of galaxy B about x */
    double impact_parameter; /* y offset between the centres */
    double speed;            /* relative approach speed along x */
    uint64_t seed;           /* per-run generator seed */
} gc_collision;

/* GalaxyConfig (cpp/core/initial_conditions.hpp:14-24) */
typedef struct {
    size_t n_particles;   /* default 256 */
    double mass;          /* total galaxy mass, default 1 */
    double radius;        /* disk radius (exponential scale radius / 3), default 1 */
    double position[3], velocity[3];
    double orientation;   /* rotation about z (radians), applied last */
    uint32_t group_id;
    double thickness;     /* Gaussian z scale; 0 = flat (no draw) */
    double inclination;   /* rotation about x (radians), applied first */
} gc_galaxy_config;

void gc_default_galaxy_config(gc_galaxy_config* c);
/* generate_galaxies (cpp/core/initial_conditions.cpp:94-114) on the host:
 * one mt19937_64(seed) over the galaxies in order, GNU libstdc++'s
 * uniform and normal draws; bit-identical to the reference built with the
 * same libm.  Writes sum(n_particles) particles (galaxies with n = 0,
 * mass <= 0 or radius <= 0 contribute none) and returns that count;
 * group_ids may be NULL.  G and eps are PhysicsParams. */
size_t gc_generate_galaxies(const gc_galaxy_config* galaxies, int count, double G, double eps,
                            uint64_t seed, double* pos, double* vel, double* mass,
                            uint32_t* group_ids);

void gc_default_disk_galaxy(gc_disk_galaxy* g);

/* ---- initial conditions in HBM (no reference counterpart: the reference
 * generates on the host with one sequential mt19937_64) ----
 * Each particle draws from the counter-based Philox4x32-10 stream keyed by
 * the seed, with counter (particle index within its galaxy, galaxy index,
 * draw block), so particles are independent and one launch writes the
 * resident state (accelerations zeroed) with nothing crossing PCIe.  The
 * models are the host generators' (positions and velocities from the same
 * formulas); the particles differ from the host generators' because the
 * draws do.  G and eps come from the context's (batch's) parameters.  A
 * context call takes at most 2^30 particles in all (GC_ERR_INVALID_ARGUMENT
 * beyond, before any launch). */
/* generate_disk_galaxy's model per gc_galaxy_config (empty galaxies skipped) */
gc_status gc_generate_galaxies_device(gc_ctx* ctx, const gc_galaxy_config* galaxies, int count,
                                      uint64_t seed);
/* the two-disk model of gc_generate_disks */
gc_status gc_generate_disks_device(gc_ctx* ctx, const gc_disk_galaxy* galaxies, int count,
                                   uint64_t seed);
/* the Plummer model of gc_generate_plummer */
gc_status gc_generate_plummer_device(gc_ctx* ctx, size_t n, double total_mass, double scale,
                                     uint64_t seed);
/* Run r's particles drawn on the device: galaxies[r * galaxies_per_run ..]
 * in order (the two-disk model of gc_generate_disks_device, stream = the
 * galaxy's index within its run) from seeds[r]; each run's galaxies must
 * hold exactly its particle count.  Replaces velocities and masses too and
 * zeroes accelerations. */
gc_status gc_batch_generate_disks(gc_batch* b, const gc_disk_galaxy* galaxies,
                                  int galaxies_per_run, const uint64_t* seeds);
/* Writes sum(n_particles) particles; returns that count (0 on bad input). */
size_t gc_generate_disks(const gc_disk_galaxy* galaxies, int count, double G, double eps,
                         uint64_t seed, double* pos, double* vel, double* mass);
size_t gc_generate_plummer(size_t n, double total_mass, double scale, double G, uint64_t seed,
                           double* pos, double* vel, double* mass);
/* Batch sampler: N log-uniform in [n_min, n_max] (even), inclination
 * U[0, pi/2], impact parameter U[0.5, 2], speed U[0.2, 0.5]. */
int gc_sample_collisions(int runs, uint64_t seed, int n_min, int n_max, gc_collision* out);

/* ---- run output (SnapshotWriter, cpp/io/snapshot_writer.hpp) ----
 * The reference's output directory, byte for byte: metadata.json,
 * diagnostics.csv (one row per call) and snapshot_NNNNNN.csv (id, group_id,
 * mass, x y z, vx vy vz, ax ay az at 17 significant digits), which its
 * Python tools (python/utils/snapshots.py) read.  gc_writer_snapshot
 * stages the resident state with one stream-ordered copy into a pinned
 * buffer and returns; the writer's own thread formats and writes the file
 * while the device steps on (two snapshots and 16 diagnostics rows in
 * flight at most).  A writer is
 * driven from one thread at a time, and its device snapshots come from
 * contexts on one device.  File errors of queued snapshots are reported
 * by the next gc_writer_flush or gc_writer_close (GC_ERR_IO). */
typedef struct gc_writer gc_writer;

/* SimulationConfig fields metadata.json records (cpp/core/config.hpp:19-34) */
typedef struct {
    const char* name;
    const char* solver;
    size_t particle_count;
    int steps;
    double dt;
    int snapshot_every;
    int dim;
    double gravitational_constant;
    double softening;
    double tree_theta;
    size_t tree_leaf_capacity;
    int fmm_expansion_order;
} gc_run_info;

/* SnapshotWriter::SnapshotWriter (snapshot_writer.cpp:31-48): creates the
 * directory and diagnostics.csv with its header line. */
gc_status gc_writer_open(gc_writer** out, const char* directory);
/* SnapshotWriter::write_metadata (snapshot_writer.cpp:50-70) */
gc_status gc_writer_metadata(gc_writer* w, const gc_run_info* info);
/* SnapshotWriter::write_snapshot (snapshot_writer.cpp:72-96) of the
 * context's resident state, asynchronously; group_ids (n entries, NULL =
 * all 0) are copied before the call returns. */
gc_status gc_writer_snapshot(gc_writer* w, gc_ctx* ctx, int step, double time,
                             const uint32_t* group_ids);
/* The same from host arrays (pos, vel, acc: n x 3; mass: n), copied before
 * the call returns. */
gc_status gc_writer_snapshot_host(gc_writer* w, int step, double time, size_t n,
                                  const double* pos, const double* vel, const double* acc,
                                  const double* mass, const uint32_t* group_ids);
/* SnapshotWriter::write_diagnostics (snapshot_writer.cpp:98-121) of the
 * context's resident state: the gc_diagnostics_compute kernels and a copy
 * of their results queued on the context's stream, the row written by the
 * writer's thread (rows keep call order). */
gc_status gc_writer_diagnostics(gc_writer* w, gc_ctx* ctx, int step, double time);
/* The same row from values the caller holds. */
gc_status gc_writer_diagnostics_values(gc_writer* w, int step, double time, size_t n,
                                       const gc_diagnostics* d);
/* Host threads that format one snapshot's rows (0 = min(8, hardware
 * threads), the default). */
gc_status gc_writer_set_threads(gc_writer* w, int threads);
/* Waits until every queued snapshot is on disk and flushes diagnostics.csv. */
gc_status gc_writer_flush(gc_writer* w);
/* Flushes, closes the files and frees the writer (NULL is a no-op). */
gc_status gc_writer_close(gc_writer* w);

#ifdef __cplusplus
}
#endif
#endif /* GCSIM_H */
