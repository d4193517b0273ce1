/*
 * mpm.h -- C ABI of the MI355X explicit MLS-MPM simulator (libmpm_hip.so).
 *
 * The reference simulator (JerryYan97/MPM_CUDA) exposes a C++ class, not a C
 * ABI.  Each entry point below names the reference member or free function it
 * replaces.  Conventions that differ from the reference on purpose:
 *   - every call returns an int status (MPM_OK = 0) and never exits the
 *     process; mpm_last_error() describes the last failure of this thread
 *     (the reference prints and calls exit(1), e.g. MPMCUDA.cu:800-803);
 *   - one HIP stream per simulator; mpm_step is asynchronous, every read
 *     synchronizes (the reference synchronizes after each kernel,
 *     MPMCUDA.cu:830, 855, 898);
 *   - plain pointers and sizes only; no C++ or torch types cross the ABI.
 * A simulator handle must not be used from two threads at once.
 */
#ifndef MPM_MPM_H
#define MPM_MPM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
enum {
    MPM_OK = 0,
    MPM_ERR_INVALID = 1,    /* bad argument (range, null pointer, unknown enum) */
    MPM_ERR_HIP = 2,        /* a HIP runtime call failed */
    MPM_ERR_IO = 3,         /* file could not be read or written */
    MPM_ERR_RANGE = 4,      /* object/particles outside the grid (MPMSimulator.cpp:210-222) */
    MPM_ERR_STATE = 5,      /* health check failed (see mpm_health) */
    MPM_ERR_NOMEM = 6
};

/* Material types, in the order of `enum MaterialType` (Elasiticity.cuh:10-12). */
enum { MPM_JELLO = 0, MPM_SNOW = 1, MPM_WATER = 2 };

/* Grid boundary conditions (reference: per-axis zeroing, MPMCUDA.cu:463-469). */
enum { MPM_BC_AXIS = 0, MPM_BC_STICKY = 1 };

/* Water affine matrix: APIC C (reference, MPMCUDA.cu:520-532) or C = grad v. */
enum { MPM_WATER_C_APIC = 0, MPM_WATER_C_GRADV = 1 };

/* Execution form of the step.
 * MPM_FORM_PLAIN: unsorted particles, P2G scatters with global atomics, the
 *   stress SVD runs in P2G and (snow) again in G2P -- the structure of
 *   MPMCUDA.cu:310-741 in FP32 SoA.
 * MPM_FORM_FAST: particles sorted by 4^3-node block (at most every
 *   `sort_every` steps, earlier when particles drift out of their block's
 *   arena), P2G accumulates per block in LDS without atomics and is staged
 *   for the grid update, G2P is fused with the stress evaluation so each
 *   step needs one SVD.
 * MPM_FORM_FUSED: as FAST, but the G2P of step n and the P2G of step n+1 run
 *   in one kernel per block (the grid velocities of the block's neighbourhood
 *   in LDS); the stress never goes to memory.  mpm_step_phase runs the FAST
 *   phases in this form. */
enum { MPM_FORM_PLAIN = 0, MPM_FORM_FAST = 1, MPM_FORM_FUSED = 2 };

/* Warm start for the polar decomposition in G2P. */
enum { MPM_WARM_NONE = 0, MPM_WARM_MLP = 1 };

/* Step-size control.  The reference computes a deformation factor
 * min(10, max_p max(J, 1/J)^3) after every step and never applies it
 * (MPMCUDA.cu:613-619, 700-708, 915-920): adp_dt stays max_dt.
 *   FIXED       every step takes the dt given to mpm_step (the reference);
 *   CFL         dt = min(max_dt, cfl h / max_p(|v_p| + c_p)), c_p the
 *               particle's elastic wave speed (jello sqrt((lam + 2 mu) / rho),
 *               snow with the hardening factor exp(xi (1 - J_p)), water
 *               sqrt(k gamma J^(1 - gamma) / rho));
 *   CFL_DEFORM  CFL, and at most max_dt divided by the reference's factor
 *               clamped to [1, 10] (jello det F, snow det F_p; water
 *               particles do not enter it, as in the reference).
 * The maxima are reduced on the device by each step's G2P and read by the
 * next step on the device (no host round trip); mpm_last_dt and
 * mpm_current_time then synchronize.  PLAIN and FAST forms only. */
enum { MPM_DT_FIXED = 0, MPM_DT_CFL = 1, MPM_DT_CFL_DEFORM = 2 };

/* Particle volume: reference quirk V_p = h^3 (MPMSimulator.cpp:263, occupancy
 * counts particles), or h^3 / ppc.  Dynamics do not depend on the choice. */
enum { MPM_VOLUME_REF = 0, MPM_VOLUME_PPC = 1 };

typedef struct mpm_config {
    double h;                 /* grid spacing, "gap" (MPMSimulator.cuh:92) */
    int32_t nx, ny, nz;       /* nodes per axis (MPMSimulator.cuh:94) */
    double origin[3];         /* grid origin; (0,0,0) in the reference (MPMSimulator.cpp:129) */
    double dt;                /* step used by mpm_step when dt <= 0 ("max_dt", :184) */
    double gravity[3];        /* (0,-9.8,0) (MPMSimulator.cpp:186) */
    int32_t bc_pad;           /* wall thickness in cells: 10 (MPMCUDA.cu:845-850) */
    int32_t bc_mode;          /* MPM_BC_* */
    int32_t ppc;              /* samples per cell for mpm_add_object (MPMSimulator.cuh:95) */
    int32_t volume_mode;      /* MPM_VOLUME_* */
    uint64_t seed;            /* jitter seed for mpm_add_object */
    int32_t form;             /* MPM_FORM_* */
    int32_t sort_every;       /* FAST/FUSED: most steps between re-sorts (24); a re-sort
                                 comes earlier once >0.01% of particles leave their arena */
    int32_t svd_sweeps;       /* FP32 Jacobi sweeps (plain solve) */
    int32_t warm_start;       /* MPM_WARM_* */
    int32_t warm_sweeps;      /* Jacobi sweeps after a warm start */
    int32_t tile;             /* FAST: P2G particles per pass over a tile: 256, 384 or 512; 0 = 384 */
    int32_t threads;          /* threads per workgroup for particle kernels, 0 = default */
    int32_t dt_mode;          /* MPM_DT_*; the dt given to mpm_step is max_dt */
    float cfl;                /* CFL number of the controller, in (0, 1] (0.5) */
    int32_t reserved[3];
} mpm_config;

typedef struct mpm_material {
    int32_t type;             /* MPM_JELLO / MPM_SNOW / MPM_WATER */
    int32_t water_c_mode;     /* MPM_WATER_C_* */
    double E, nu, rho;        /* Material(E, nu, rho, type) (Elasiticity.cuh:30) */
    double theta_c, theta_s;  /* snow clamp [1-theta_c, 1+theta_s]: 8e-2, 2e-2 (MPMCUDA.cu:658-660) */
    double xi;                /* snow hardening: 10 (MPMCUDA.cu:344) */
    double eos_k, eos_gamma;  /* water: 2.51, 7 (MPMCUDA.cu:353) */
} mpm_material;

/* Shapes for mpm_add_object. */
enum { MPM_SHAPE_MESH = 0, MPM_SHAPE_BOX = 1, MPM_SHAPE_SPHERE = 2 };

/* One object to sample, after `ObjInitInfo` (MPMSimulator.cuh:19-27). */
typedef struct mpm_object {
    int32_t shape;            /* MPM_SHAPE_* */
    int32_t reserved0;
    const char *obj_path;     /* MESH: Wavefront OBJ (v/f records) */
    double shape_param[6];    /* BOX: local min xyz, max xyz; SPHERE: local centre xyz, radius */
    double init_vel[3];       /* initVel */
    float translation[3];     /* initTranslation */
    float scale[3];           /* initScale */
    float rotation_axis[3];   /* initRotationAxis (normalised by the caller, main.cpp:146) */
    float rotation_deg;       /* initRotationDegree */
    mpm_material material;    /* mMaterial */
    int64_t max_particles;    /* >0: keep a seeded random subset of this size */
} mpm_object;

/* Particle fields for mpm_read_particles / mpm_write_particles.
 * Layout of the host buffer: n rows of the field's width, row-major 3x3. */
enum {
    MPM_P_POS = 0,      /* 3 floats  pPosVec */
    MPM_P_VEL = 1,      /* 3         pVelVec */
    MPM_P_C = 2,        /* 9         pAffineVel */
    MPM_P_F = 3,        /* 9         pElasticityDeformationGradient */
    MPM_P_JP = 4,       /* 1         snow: det Fp; water: J (pJVec) */
    MPM_P_TAU = 5       /* 6 (read-only) Kirchhoff stress xx,yy,zz,xy,xz,yz used by the next P2G */
};

/* Grid fields for mpm_read_grid, reference node order z*NX*NY + y*NX + x
 * (MPMCUDA.cu:375). */
enum {
    MPM_G_MASS = 0,     /* 1 float per node */
    MPM_G_MOMENTUM = 1, /* 3: after MPM_PHASE_P2G */
    MPM_G_VELOCITY = 2  /* 3: after MPM_PHASE_GRID */
};

enum { MPM_PHASE_P2G = 0, MPM_PHASE_GRID = 1, MPM_PHASE_G2P = 2 };

/* Health bits, accumulated on the device since the last mpm_health call
 * (the reference asserts these, MPMCUDA.cu:376-386, 431, 546, 614, 702, 722-739). */
enum {
    MPM_H_STENCIL_OOB = 1,  /* a stencil left the grid (clamped) */
    MPM_H_WEIGHT_SUM = 2,
    MPM_H_DET_NONPOS = 4,
    MPM_H_LEFT_DOMAIN = 8,
    MPM_H_NONFINITE = 16,
    MPM_H_ARENA_SPILL = 32  /* FAST: particles drifted beyond their tile arena (handled) */
};

typedef struct mpm_sim mpm_sim;

/* Fill `cfg` with the reference's defaults (MPMSimulator.cpp:184-186,
 * MPMCUDA.cu:845-850) and this library's tuned defaults. */
void mpm_config_default(mpm_config *cfg);

/* Material(E, nu, rho, type), Elasiticity.cuh:30-45: range checks of
 * :34-41 return MPM_ERR_INVALID instead of exiting. */
int mpm_material_init(mpm_material *m, double E, double nu, double rho, int type);

/* MPMSimulator::MPMSimulator grid part (MPMSimulator.cpp:178-197, initGrid :122-146). */
int mpm_create(const mpm_config *cfg, mpm_sim **out);

/* Per-object sampling of MPMSimulator.cpp:198-276 (mesh or procedural
 * shape, per-cell jittered samples, inside test, V_p and m_p). Returns the
 * new group's index through *group (may be NULL). */
int mpm_add_object(mpm_sim *sim, const mpm_object *obj, int32_t *group);

/* Most particles one simulator holds (all groups together): 2^30 - 1. */
#define MPM_MAX_PARTICLES ((1ll << 30) - 1)
/* Most particle groups one simulator holds. */
#define MPM_MAX_GROUPS 256

/* Add a group from explicit particles (positions xyz[3n]; vel[3n] or NULL
 * to use v0; particle_volume <= 0 means h^3).  initParticles,
 * MPMSimulator.cpp:10-120: F = I, Fp = I / J = 1, C = 0.  MPM_ERR_INVALID
 * when the simulator would hold more than MPM_MAX_PARTICLES particles or
 * MPM_MAX_GROUPS groups. */
int mpm_add_particles(mpm_sim *sim, const mpm_material *mat, int64_t n, const float *xyz,
                      const float *vel, const double v0[3], double particle_volume, int32_t *group);

/* MPMSimulator::step (MPMCUDA.cu:793-925); dt <= 0 uses cfg.dt. Asynchronous. */
int mpm_step(mpm_sim *sim, double dt);

/* nsteps calls of mpm_step, queued back to back on the simulator's stream (one
 * host call per step, no graph; in the sorted forms the host waits at most
 * until the device is two steps behind, to read the spill counts). */
int mpm_run(mpm_sim *sim, int32_t nsteps, double dt);

/* One phase of the step, for phase-by-phase parity checks. */
int mpm_step_phase(mpm_sim *sim, int32_t phase, double dt);

/* MPMSimulator::totalParticlesNum (MPMSimulator.cuh:99). */
int64_t mpm_total_particles(const mpm_sim *sim);
int32_t mpm_group_count(const mpm_sim *sim);
int64_t mpm_group_size(const mpm_sim *sim, int32_t group);

/* Particle state of one group, in the group's original particle order. */
int mpm_read_particles(mpm_sim *sim, int32_t group, int32_t field, float *out);
int mpm_write_particles(mpm_sim *sim, int32_t group, int32_t field, const float *in);

/* Positions of all groups concatenated: MPMSimulator::getGLParticlesPos
 * (MPMSimulator.cpp:362-371). out holds 3 * mpm_total_particles floats. */
int mpm_get_positions(mpm_sim *sim, float *out);

/* Dense grid field in reference node order; out holds width * nx*ny*nz floats. */
int mpm_read_grid(mpm_sim *sim, int32_t field, float *out);

/* Health bits since the last call (cleared on read), and counters. */
int mpm_health(mpm_sim *sim, uint32_t *bits);
/* stats[0] steps, [1] active grid blocks of the last step, [2] FP64 SVD
 * fallbacks, [3] warm-start rejections, [4] sorts, [5] time (s), [6] P2G
 * tiles of the last step, [7] particles of the last P2G that left their
 * tile's arena and took the atomic fallback. */
int mpm_stats(mpm_sim *sim, double stats[8]);

/* SVD work of the G2P since the simulator was created (32-bit device
 * counters, no reference counterpart): [0] particle updates that took the
 * Jacobi SVD (snow whose trial F left the clamp band, jello whose F has a
 * singular value outside (0.5, 2)); [1] of those, started from a network
 * quaternion (warm start); [2] warm results rejected for the cold SVD;
 * [3] FP64 fallbacks of every SVD (also those of the stress outside G2P). */
int mpm_svd_stats(mpm_sim *sim, double out[4]);

int mpm_synchronize(mpm_sim *sim);

/* Per-phase device timing with HIP events recorded on the simulator's
 * stream (no reference counterpart; the reference has only an FPS readout,
 * main.cpp:252-253).  Phases: [0] sort + tile build, [1] P2G, [2] grid
 * update (incl. active-block compaction), [3] G2P.  mpm_phase_times returns
 * milliseconds and launch counts accumulated since the previous call and
 * clears them.  stats_nodes (may be NULL) receives the number of grid nodes
 * with mass in the last grid update. */
int mpm_set_timing(mpm_sim *sim, int32_t enable);
int mpm_phase_times(mpm_sim *sim, double ms[4], int64_t counts[4], int64_t *stats_nodes);
/* The main kernel of each phase alone, timed like the phases (events around
 * the single launch): [0] P2G (k_p2g_block, or k_p2g_atomic in the PLAIN
 * form), [1] k_grid_update, [2] G2P (k_g2p, or k_g2p2g_block in the FUSED
 * form), [3] unused.  Accumulated since the previous call, then cleared. */
int mpm_kernel_times(mpm_sim *sim, double ms[4], int64_t counts[4]);

/* Simulation clock: current_time / current_frame (MPMSimulator.cuh:86-87).
 * With a dt controller the time accumulates on the device and reading it
 * synchronizes the simulator's stream. */
double mpm_current_time(const mpm_sim *sim);
int64_t mpm_current_frame(const mpm_sim *sim);
/* adp_dt (MPMSimulator.cuh:85): the dt of the last step (0 before the first);
 * synchronizes with a dt controller. */
double mpm_last_dt(const mpm_sim *sim);

/* OutputBego (FilesIO.h:10): write positions as a Houdini BGEO particle file
 * with one "position" attribute. */
int mpm_write_bgeo(const char *path, const float *xyz, int64_t n);

/* ---- learned warm start of the SVD (no reference counterpart; the
 * reference's SVD always starts from the identity, svd3_cuda.h) ----
 * A 9 -> 64 -> 64 -> 4 ReLU MLP maps F / |F|_F (row-major) to a quaternion
 * q; V0 = R(q / |q|) starts the Jacobi SVD of the G2P update (snow trial F,
 * jello F) with cfg.warm_sweeps sweeps.  A result whose columns of F V are
 * not orthogonal to 1e-6, or whose F is ill conditioned, is replaced by the
 * cold SVD (counted in mpm_stats [3]).  Products run in BF16 on the MFMA
 * units with FP32 accumulation.  Parameters: mpm_mlp_param_count() FP32
 * values, W1[9][64] b1[64] W2[64][64] b2[64] W3[64][4] b3[4], a layer being
 * y_j = sum_i W[i][j] x_i + b_j. */
int mpm_mlp_param_count(void);

/* Use `weights` for the warm start (cfg.warm_start = MPM_WARM_MLP). */
int mpm_set_warm_start_weights(mpm_sim *sim, const float *weights, int64_t count);

/* Network output q[n][4] for F[n][9] (row-major). */
int mpm_mlp_forward(const float *params, const float *F, int64_t n, float *q);

/* Mean training loss over F[n][9] and its gradient with respect to the
 * parameters.  loss = sum_{i!=j} (R^T A R)_ij^2 / |A|_F^2 with
 * A = F^T F / |F|_F^2 and R = R(q / |q|). */
int mpm_mlp_grad(const float *params, const float *F, int64_t n, float *grad, float *loss);

/* Adam training on the device: `steps` minibatches of `batch` samples drawn
 * (SplitMix64, `seed`) from F[n][9].  params: initial values in, trained
 * values out; loss_history[steps] (may be NULL): mean loss per step. */
int mpm_mlp_train(const float *F, int64_t n, int32_t steps, int32_t batch, float lr, uint64_t seed,
                  float *params, float *loss_history);

/* MPMSimulator::~MPMSimulator (MPMSimulator.cpp:288-352). */
void mpm_destroy(mpm_sim *sim);

/* ---- host-only sampling and device probes (no reference counterpart; they
 * expose single stages of the step for parity tests) ---- */

/* The sampling of mpm_add_object (MPMSimulator.cpp:198-265) on the host
 * only: with xyz == NULL stores the particle count in *n; otherwise writes
 * min(*n, count) positions and stores the count. */
int mpm_sample_object(const mpm_config *cfg, const mpm_object *obj, float *xyz, int64_t *n);

/* The sampler's point-in-mesh test (ray parity along +z, the rule of the
 * reference's point_inside_mesh, mesh_query0.1/mesh_query.cpp:287-308) for
 * n points in the OBJ file's own coordinates; inside[i] = 1 or 0. */
int mpm_probe_point_in_mesh(const char *obj_path, const double *pts, int64_t n, uint8_t *inside);

/* Achievable HBM bandwidth: a streaming device copy of `bytes` (multiple of
 * 16, 1 MiB to 8 GiB), read + write counted, median over `reps` timed
 * launches after one warm-up, in GB/s. */
int mpm_probe_copy_bandwidth(int64_t bytes, int32_t reps, double *gbs);

/* Bandwidth of a structure-of-arrays pass shaped like a particle update: n
 * elements (1 to 2^28), planes_in FP32 planes read and planes_out written
 * (1 to 64 each), one element per thread; median over `reps` timed launches
 * after one warm-up, bytes read + written per second in GB/s.  (A 1M-particle
 * FAST G2P moves 13 planes in and 31 out.) */
int mpm_probe_soa_bandwidth(int64_t n, int32_t planes_in, int32_t planes_out, int32_t reps, double *gbs);

/* LDS scatter rate on the device: 8,192 one-wave workgroups each add, for
 * reps x 64 particles, a 27-node x 4-float P2G stencil into an 8 KB LDS
 * arena; *ms is the launch's duration.  mode 0: float atomics (ds_add_f32),
 * every lane of an instruction on its own address; 1: float atomics, hashed
 * cells (collisions); 2: plain float4 read-modify-write, distinct cells; 3:
 * plain read-modify-write, hashed cells (a rate only: colliding lanes
 * overwrite each other). */
int mpm_probe_lds_scatter(int32_t mode, int32_t reps, double *ms);

/* Device FP32 base-node index max(0, floor((x - origin) * inv_h - 0.5)) of
 * n coordinates, as used by P2G / G2P / the sort keys. */
int mpm_probe_base_index(const float *x, int64_t n, float origin, float inv_h, int32_t *out);

/* Device SVD of n row-major 3x3 FP32 matrices: `sweeps` Jacobi sweeps plus
 * the FP64 fallback of the step kernels. U, V row-major, s[3] per matrix. */
int mpm_probe_svd(const float *F, int64_t n, int32_t sweeps, float *U, float *s, float *V);

/* Kirchhoff stress tau (xx, yy, zz, xy, xz, yz per matrix) of the fixed-
 * corotated model as the step kernels evaluate it for n row-major FP32
 * matrices F (MPMCUDA.cu:229-264): the polar rotation by Newton steps for F
 * within a factor 2 of a rotation, the Jacobi SVD (`sweeps`, FP64 fallback)
 * otherwise.  type MPM_JELLO or MPM_SNOW; snow scales mu and lam by
 * exp(xi (1 - Jp[i])) (MPMCUDA.cu:342-349), Jp may be NULL for jello. */
int mpm_probe_stress(const float *F, const float *Jp, int64_t n, int32_t type, double mu, double lam, double xi,
                     int32_t sweeps, float *tau);

/* Largest column orthogonality residual |b_p . b_q| / (|b_p| |b_q|) of
 * B = F V after `sweeps` Jacobi sweeps started from V0 = R(q / |q|) (q[n][4];
 * q == NULL: V0 = I), for n row-major FP32 matrices: the acceptance test of
 * the warm start. */
int mpm_probe_svd_residual(const float *F, const float *q, int64_t n, int32_t sweeps, float *off);

/* Device stable LSD radix sort of n keys on their low `bits` bits; writes
 * the sorted keys and the permutation (sorted[i] = keys[perm[i]]). */
int mpm_probe_sort(const uint32_t *keys, int64_t n, int32_t bits, uint32_t *sorted, uint32_t *perm);

/* The sort of mpm_probe_sort on device memory, allocating nothing: no
 * counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  keys, sorted and perm are device pointers to n words
 * (keys may equal sorted); the work is queued on `stream` (a hipStream_t,
 * NULL for the null stream) and is complete when the stream reaches it.
 * Its temporaries are one fixed device array per GPU, sized for
 * MPM_SORT_NOALLOC_MAX_KEYS keys: larger n is refused (MPM_ERR_RANGE), and
 * calls that overlap on one GPU (other streams, other threads) race on it. */
#define MPM_SORT_NOALLOC_MAX_KEYS (1 << 20)
int mpm_sort_keys_noalloc(const uint32_t *keys, int64_t n, int32_t bits, uint32_t *sorted, uint32_t *perm,
                          void *stream);

/* ---- Batches of independent scenes ------------------------------------
 * No counterpart in the reference, which steps one MPMSimulator per process
 * (main.cpp:202-221).  For callers with many small scenes: every scene of a
 * batch advances one step (each at its own cfg.dt) with one launch per stage
 * for all the scenes of a material -- P2G tiles, spill fallback, block
 * compaction, grid update, G2P; the materials side by side -- replayed from a
 * hipGraph; all scenes are re-sorted together, one launch per sort stage,
 * every sort_every steps.  A scene is an mpm_sim in the FAST form holding
 * exactly one particle group; it may still be read (mpm_read_particles,
 * mpm_stats, ...) and stepped on its own between batch runs, which are
 * ordered with its stream.  Destroy the batch before its scenes, and add no
 * particles to a scene of a live batch. */
typedef struct mpm_batch mpm_batch;

typedef struct mpm_batch_shape {
    int32_t p2g_workgroups;   /* width of the persistent P2G launch */
    int32_t grid_workgroups;  /* width of the persistent grid-update launch */
    int32_t g2p_threads;      /* G2P workgroup size: 64, 128, 192 or 256 */
    int32_t sort_every;       /* steps between re-sorts of every scene */
    int32_t use_graph;        /* 1: replay captured hipGraphs; 0: launch the stages directly */
    int32_t p2g_tile;         /* P2G particles per pass over a tile: 256, 384 or 512 (0 = 384) */
} mpm_batch_shape;

void mpm_batch_shape_default(mpm_batch_shape *shape);

/* Group n (1..256) scenes; shape NULL = defaults.  Sorts any scene that was
 * never sorted and computes stale stresses; does not advance time. */
int mpm_batch_create(mpm_sim *const *sims, int32_t n, const mpm_batch_shape *shape, mpm_batch **out);

/* Times 4 launch-width candidates of P2G, grid update and G2P on live steps
 * and keeps the fastest of each, then times the 3 P2G tile shapes (256, 384,
 * 512 particles per pass) with those widths and keeps the fastest; advances
 * every scene by 7 * reps steps.  ms_out (NULL or 15 floats): median ms per
 * stage (P2G, grid, G2P) and width candidate, then P2G ms per tile shape. */
int mpm_batch_autotune(mpm_batch *b, int32_t reps, float *ms_out);

int mpm_batch_get_shape(const mpm_batch *b, mpm_batch_shape *out);
int mpm_batch_set_shape(mpm_batch *b, const mpm_batch_shape *shape);

/* Advance every scene by nsteps steps (asynchronous unless step_ms is
 * given).  step_ms (NULL or nsteps floats): device time of each batch step,
 * a re-sort included in the step it precedes. */
int mpm_batch_run(mpm_batch *b, int32_t nsteps, float *step_ms);

int mpm_batch_synchronize(mpm_batch *b);
int64_t mpm_batch_particles(const mpm_batch *b);
int32_t mpm_batch_scenes(const mpm_batch *b);
void mpm_batch_destroy(mpm_batch *b);

/* ---- 2D snow ----------------------------------------------------------
 * The reference's 2D prototype MPMSnow.py (Taichi): substep() :77-150 --
 * P2G with the FixedCorotatedStress :28-32 and hardening exp(xi (1 - J_p))
 * :87, grid update with gravity and sticky walls :105-114, G2P with the
 * snow clamp [1 - theta_c, 1 + theta_s] :141-145 -- and reset() :154-162
 * (F_e = F_p = I, C = 0).  The grid is n_grid^2 nodes on the unit square,
 * dx = 1 / n_grid, particle volume (dx/2)^2 (:8-11).  FP32 on the device. */
typedef struct mpm2d_sim mpm2d_sim;

enum { MPM2D_POS = 0, MPM2D_VEL = 1, MPM2D_C = 2, MPM2D_F = 3, MPM2D_JP = 4 };  /* widths 2, 2, 4, 4, 1 */

typedef struct mpm2d_config {
    int32_t n_grid;           /* nodes per axis (MPMSnow.py:7): 128 */
    int32_t wall;             /* sticky walls: nodes i <= wall or i >= n - wall (:109): 5 */
    double dt;                /* step: 1e-4 */
    double E, nu, rho;        /* 1e3, 0.2 (:12), 1 (:10) */
    double theta_c, theta_s;  /* clamp band: 2.5e-2, 7.5e-3 (the script's :141-142 use 8e-2, 1.5e-2) */
    double xi;                /* hardening: 10 (:87) */
    double gravity[2];        /* acceleration (:108): (0, -50) */
    int32_t graph_steps;      /* steps per replayed hipGraph in mpm2d_run (0: plain launches): 50 */
    int32_t reserved[3];
} mpm2d_config;

void mpm2d_config_default(mpm2d_config *cfg);
int mpm2d_create(const mpm2d_config *cfg, mpm2d_sim **out);
/* reset(), :154-162: n particles at xy[n][2] (inside [dx, 1 - 2 dx]),
 * velocities vel[n][2] (NULL: 0); once per simulator. */
int mpm2d_add_particles(mpm2d_sim *s, int64_t n, const float *xy, const float *vel);
/* substep(), :77-150, nsteps times (asynchronous). */
int mpm2d_run(mpm2d_sim *s, int32_t nsteps);
/* Particle field (MPM2D_*) as n rows of its width, row-major 2x2. */
int mpm2d_read(mpm2d_sim *s, int32_t field, float *out);
/* Grid after the last update: mass[n_grid^2], vel[n_grid^2][2], node (i, j) at j * n_grid + i. */
int mpm2d_read_grid(mpm2d_sim *s, float *mass, float *vel);
int mpm2d_health(mpm2d_sim *s, uint32_t *bits);      /* MPM_H_* bits since creation */
int64_t mpm2d_particles(const mpm2d_sim *s);
int64_t mpm2d_steps(const mpm2d_sim *s);
int mpm2d_synchronize(mpm2d_sim *s);
void mpm2d_destroy(mpm2d_sim *s);

/* Message of the last failed call on this thread ("" if none). */
const char *mpm_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
