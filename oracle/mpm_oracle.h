/*
 * FP64 CPU oracle for the explicit MPM step (test infrastructure only).
 *
 * This is a plain-C restatement of the reference simulator's hot path,
 * JerryYan97/MPM_CUDA `MPMSimulator::step()` (src/simulator/MPMCUDA.cu:793-925):
 * grid clear -> P2G per particle group -> grid velocity update -> G2P + move.
 * Every routine cites the reference lines it restates.  All arithmetic is
 * FP64 (the reference mixes FP64 positions/weights with FP32 stress; the
 * oracle keeps FP64 everywhere so it is the accuracy yardstick).
 *
 * Only the tests, the smoke check and the CPU-baseline leg of the benchmark
 * use this library; the product path never links or calls it.
 */
#ifndef MPM_ORACLE_H
#define MPM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Material types, same order as `enum MaterialType` (Elasiticity.cuh:10-12). */
enum { ORC_JELLO = 0, ORC_SNOW = 1, ORC_WATER = 2 };

/* Grid boundary handling.
 * ORC_BC_AXIS:   reference behaviour (MPMCUDA.cu:463-469): zero velocity
 *                component a when the node lies within `pad` cells of a wall
 *                normal to axis a.
 * ORC_BC_STICKY: zero all components when the node lies within `pad` cells
 *                of any wall (the Taichi scripts, MPMSnow.py:109-111). */
enum { ORC_BC_AXIS = 0, ORC_BC_STICKY = 1 };

/* Water affine matrix: APIC C (reference, MPMCUDA.cu:520-532) or C = grad v. */
enum { ORC_WATER_C_APIC = 0, ORC_WATER_C_GRADV = 1 };

/* Health flags returned by the phase routines (the reference asserts these). */
enum {
    ORC_OK = 0,
    ORC_FLAG_STENCIL_OOB = 1,   /* a 3x3x3 stencil leaves the grid (MPMCUDA.cu:376-386) */
    ORC_FLAG_WEIGHT_SUM = 2,    /* |sum w - 1| too large (MPMCUDA.cu:431, 546) */
    ORC_FLAG_DET_NONPOS = 4,    /* det F <= 0 or det Fp <= 0 (MPMCUDA.cu:614, 702) */
    ORC_FLAG_LEFT_DOMAIN = 8,   /* particle left [o, o+(N-1)h] (MPMCUDA.cu:722-739) */
    ORC_FLAG_NONFINITE = 16     /* NaN / Inf in the particle state */
};

typedef struct orc_grid {
    int32_t nx, ny, nz, reserved;
    double h;
    double origin[3];
    double *m;   /* [nz][ny][nx]      node mass (MPMSimulator.cuh:60) */
    double *p;   /* [nz][ny][nx][3]   momentum after P2G, velocity after update */
} orc_grid;

typedef struct orc_material {
    int32_t type;
    int32_t water_c_mode;
    double mass, vol;          /* per-particle m_p and V_p (MPMSimulator.cpp:263-264) */
    double mu, lam;            /* Lame constants (Elasiticity.cuh:42-43) */
    double theta_c, theta_s;   /* snow clamp band [1-theta_c, 1+theta_s] (MPMCUDA.cu:658-660) */
    double xi;                 /* snow hardening (MPMCUDA.cu:344) */
    double eos_k, eos_gamma;   /* water p = k (J^-gamma - 1) (MPMCUDA.cu:353) */
} orc_material;

typedef struct orc_group {
    int64_t n;
    orc_material mat;
    /* AoS FP64, row-major 3x3 (F[3*r+c]), as ParticleGroup (MPMSimulator.cuh:30-52) */
    double *x, *v, *C, *Fe, *Fp, *J, *dgdiff;
} orc_group;

typedef struct orc_params {
    double gravity[3];         /* reference: (0, -9.8, 0) (MPMSimulator.cpp:186) */
    int32_t bc_pad;            /* reference: 10 cells (MPMCUDA.cu:845-850) */
    int32_t bc_mode;
} orc_params;

/* ---- scalar building blocks ---- */
double orc_N(double x);                         /* MPMCUDA.cu:279-290 */
double orc_dN(double x);                        /* MPMCUDA.cu:267-277 */
double orc_weight(const double xp[3], const double xi[3], double h);           /* :303-308 */
void   orc_weight_grad(const double xp[3], const double xi[3], double h, double g[3]); /* :292-301 */
int    orc_base_ref(double x, double o, double h);                             /* :356-358 */
int    orc_base_f32(float x, float o, float inv_h);
uint32_t orc_block_key(int bx, int by, int bz, int nbx, int nby);

/* ---- 3x3 linear algebra ---- */
double orc_det3(const double A[9]);                        /* MPMCUDA.cu:31-35 */
int    orc_inv3(const double A[9], double Ainv[9]);        /* MPMCUDA.cu:37-55 */
void   orc_svd3(const double F[9], double U[9], double s[3], double V[9]);
void   orc_fcr_tau(const double F[9], double mu, double lam, double tau[9]); /* :229-264 */
void   orc_snow_plastic(const double Ftr[9], const double Fp[9], double theta_c, double theta_s,
                        double Fe_new[9], double Fp_new[9]);                  /* :620-709 */

/* ---- step phases ---- */
void orc_clear_grid(orc_grid *g);                                           /* :798-808 */
int  orc_p2g(orc_grid *g, const orc_group *grp, double dt);                 /* :310-433 */
int  orc_grid_update(orc_grid *g, const orc_params *prm, double dt);        /* :435-472 */
int  orc_g2p(const orc_grid *g, orc_group *grp, double dt);                 /* :474-741 */
int  orc_step(orc_grid *g, orc_group *groups, int ngroups, const orc_params *prm, double dt); /* :793-925 */

/* Threads of the OpenMP loops (P2G with atomic node sums, grid update,
 * G2P); 1 by default.  Returns the count in effect. */
int orc_set_threads(int n);

/* ---- 2D snow (MPMSnow.py:77-150, mpm_oracle2d.c) ---- */
typedef struct orc2_params {
    int32_t n_grid;            /* nodes per axis (MPMSnow.py:7) */
    int32_t wall;              /* sticky band: i <= wall or i >= n - wall (:109) */
    double dx;                 /* 1 / n_grid (:8) */
    double p_vol, p_mass;      /* (dx/2)^2 and p_vol * rho (:10-11) */
    double mu0, lam0;          /* Lame constants (:13) */
    double xi;                 /* hardening h = exp(xi (1 - J_p)) (:87) */
    double theta_c, theta_s;   /* clamp band [1 - theta_c, 1 + theta_s] (:141-142) */
    double gravity[2];         /* acceleration added as dt * g (:108) */
} orc2_params;

typedef struct orc2_particles {
    int64_t n;
    double *x, *v;             /* [n][2] */
    double *C, *Fe, *Fp;       /* [n][4], row-major 2x2 */
} orc2_particles;

double orc2_N(double x);                                            /* :36-45 */
double orc2_dN(double x);                                           /* :55-65 */
void orc2_svd(const double F[4], double U[4], double s[2], double V[4]);
void orc2_fcr_tau(const double F[4], double mu, double lam, double tau[4]);   /* :28-32 */
int orc2_step(double *grid_m, double *grid_v, orc2_particles *p, const orc2_params *prm, double dt);

#ifdef __cplusplus
}
#endif
#endif
