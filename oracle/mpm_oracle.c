/*
 * FP64 CPU oracle for the explicit MPM step -- test infrastructure only.
 *
 * Restates JerryYan97/MPM_CUDA src/simulator/MPMCUDA.cu (P2G :310-433,
 * VelUpdate :435-472, InterpolateAndMove :474-741, step :793-925) in plain C
 * with FP64 arithmetic throughout.  Used by tests/, the smoke check and the
 * CPU-baseline leg of bench.py as the checker; never by the product path.
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off, so the FP32 base
 * index formula below rounds exactly like the GPU's).
 */
#include "mpm_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

#include <math.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Quadratic B-spline, branchy |x| form as in MPMCUDA.cu:279-290.            */
double orc_N(double x)
{
    double a = fabs(x);
    if (a < 0.5) return 0.75 - a * a;
    if (a < 1.5) return 0.5 * (1.5 - a) * (1.5 - a);
    return 0.0;
}

/* Derivative, MPMCUDA.cu:267-277 (note the half-open intervals there). */
double orc_dN(double x)
{
    if (x > -0.5 && x < 0.5) return -2.0 * x;
    if (x >= 0.5 && x < 1.5) return x - 1.5;
    if (x > -1.5 && x <= -0.5) return 1.5 + x;
    return 0.0;
}

/* w_ip = N((xp-xi)/h) per axis, MPMCUDA.cu:303-308. */
double orc_weight(const double xp[3], const double xi[3], double h)
{
    return orc_N((xp[0] - xi[0]) / h) * orc_N((xp[1] - xi[1]) / h) * orc_N((xp[2] - xi[2]) / h);
}

/* grad w_ip with respect to xp, MPMCUDA.cu:292-301. */
void orc_weight_grad(const double xp[3], const double xi[3], double h, double g[3])
{
    double hi = 1.0 / h;
    double d0 = hi * (xp[0] - xi[0]), d1 = hi * (xp[1] - xi[1]), d2 = hi * (xp[2] - xi[2]);
    double n0 = orc_N(d0), n1 = orc_N(d1), n2 = orc_N(d2);
    g[0] = hi * orc_dN(d0) * n1 * n2;
    g[1] = hi * n0 * orc_dN(d1) * n2;
    g[2] = hi * n0 * n1 * orc_dN(d2);
}

/* Base node index, reference FP64 form: max(0, int((x - o - h/2)/h)),
 * MPMCUDA.cu:356-358 (C truncation toward zero). */
int orc_base_ref(double x, double o, double h)
{
    int b = (int)((x - o - 0.5 * h) / h);
    return b < 0 ? 0 : b;
}

/* Base node index, FP32 form used by the HIP kernels on FP32 positions:
 * t = (x - o) * inv_h - 0.5f, each operation rounded to FP32 (no FMA
 * contraction), base = max(0, floor(t)).  Mathematically identical to the
 * reference form for x >= o; the two can differ only when x sits within an
 * FP32 rounding of a cell threshold, where the B-spline weights of the two
 * candidate stencils coincide. */
int orc_base_f32(float x, float o, float inv_h)
{
    volatile float d = x - o;        /* volatile: keep each FP32 rounding */
    volatile float s = d * inv_h;
    float t = s - 0.5f;
    int b = (int)floorf(t);
    return b < 0 ? 0 : b;
}

/* Cell-block sort key: linear index of the 4x4x4 block holding the base node. */
uint32_t orc_block_key(int bx, int by, int bz, int nbx, int nby)
{
    return ((uint32_t)bz * (uint32_t)nby + (uint32_t)by) * (uint32_t)nbx + (uint32_t)bx;
}

/* ------------------------------------------------------------------------ */
/* 3x3 helpers (row-major A[3*r+c]), MPMCUDA.cu:31-77.                       */
double orc_det3(const double A[9])
{
    return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
           A[2] * (A[3] * A[7] - A[4] * A[6]);
}

/* Inverse via adjugate.  The reference prints and leaves the output
 * untouched when det == 0 (MPMCUDA.cu:41-42); here that case returns -1. */
int orc_inv3(const double A[9], double B[9])
{
    double d = orc_det3(A);
    if (d == 0.0) return -1;
    double r = 1.0 / d;
    B[0] = r * (A[4] * A[8] - A[5] * A[7]);
    B[1] = r * (A[2] * A[7] - A[1] * A[8]);
    B[2] = r * (A[1] * A[5] - A[2] * A[4]);
    B[3] = r * (A[5] * A[6] - A[3] * A[8]);
    B[4] = r * (A[0] * A[8] - A[2] * A[6]);
    B[5] = r * (A[2] * A[3] - A[0] * A[5]);
    B[6] = r * (A[3] * A[7] - A[4] * A[6]);
    B[7] = r * (A[1] * A[6] - A[0] * A[7]);
    B[8] = r * (A[0] * A[4] - A[1] * A[3]);
    return 0;
}

static void mm3(const double A[9], const double B[9], double C[9])
{
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            C[3 * r + c] = A[3 * r] * B[c] + A[3 * r + 1] * B[3 + c] + A[3 * r + 2] * B[6 + c];
}

/* ------------------------------------------------------------------------ */
/* 3x3 SVD, F = U diag(s) V^T with U, V in SO(3), s0 >= s1 >= |s2|, and the
 * sign of det F carried by s2 -- the conventions of the reference's SVD
 * (McAdams et al. 2011, used through PolarSVD, MPMCUDA.cu:87-134).
 * Method: one-sided (Hestenes) Jacobi on the columns of F to convergence,
 * columns sorted by norm (each swap negates one column so det V stays +1),
 * then a Givens QR of B = F V gives U and the signed singular values.       */
void orc_svd3(const double F[9], double U[9], double s[3], double V[9])
{
    double B[9], W[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    memcpy(B, F, sizeof B);
    static const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
    for (int sweep = 0; sweep < 40; ++sweep) {
        int rotated = 0;
        for (int k = 0; k < 3; ++k) {
            int p = P[k], q = Q[k];
            double a = 0, b = 0, g = 0;
            for (int r = 0; r < 3; ++r) {
                a += B[3 * r + p] * B[3 * r + p];
                b += B[3 * r + q] * B[3 * r + q];
                g += B[3 * r + p] * B[3 * r + q];
            }
            if (g == 0.0 || fabs(g) <= 4e-16 * sqrt(a * b)) continue;
            rotated = 1;
            double zeta = (b - a) / (2.0 * g);
            double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
            double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
            for (int r = 0; r < 3; ++r) {
                double bp = B[3 * r + p], bq = B[3 * r + q];
                B[3 * r + p] = c * bp - sn * bq;
                B[3 * r + q] = sn * bp + c * bq;
                double vp = W[3 * r + p], vq = W[3 * r + q];
                W[3 * r + p] = c * vp - sn * vq;
                W[3 * r + q] = sn * vp + c * vq;
            }
        }
        if (!rotated) break;
    }
    /* sort columns by norm, descending */
    double nrm[3];
    for (int c = 0; c < 3; ++c)
        nrm[c] = B[c] * B[c] + B[3 + c] * B[3 + c] + B[6 + c] * B[6 + c];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2 - i; ++j)
            if (nrm[j] < nrm[j + 1]) {
                double tn = nrm[j]; nrm[j] = nrm[j + 1]; nrm[j + 1] = tn;
                for (int r = 0; r < 3; ++r) {
                    double tb = B[3 * r + j]; B[3 * r + j] = -B[3 * r + j + 1]; B[3 * r + j + 1] = tb;
                    double tw = W[3 * r + j]; W[3 * r + j] = -W[3 * r + j + 1]; W[3 * r + j + 1] = tw;
                }
            }
    /* Givens QR: R = G3 G2 G1 B, U = (G3 G2 G1)^T */
    double R[9], Qm[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    memcpy(R, B, sizeof R);
    static const int GR[3][3] = {{0, 1, 0}, {0, 2, 0}, {1, 2, 1}}; /* rows (i,j), zero column k */
    for (int gidx = 0; gidx < 3; ++gidx) {
        int i = GR[gidx][0], j = GR[gidx][1], k = GR[gidx][2];
        double a = R[3 * i + k], b = R[3 * j + k];
        double rr = sqrt(a * a + b * b);
        double c = 1.0, sn = 0.0;
        if (rr > 0.0) { c = a / rr; sn = b / rr; }
        for (int col = 0; col < 3; ++col) {
            double ri = R[3 * i + col], rj = R[3 * j + col];
            R[3 * i + col] = c * ri + sn * rj;
            R[3 * j + col] = -sn * ri + c * rj;
            /* Qm accumulates G^T on the right: Qm <- Qm G^T (columns i, j) */
            double qi = Qm[3 * col + i], qj = Qm[3 * col + j];
            Qm[3 * col + i] = c * qi + sn * qj;
            Qm[3 * col + j] = -sn * qi + c * qj;
        }
    }
    memcpy(U, Qm, sizeof Qm);
    memcpy(V, W, sizeof W);
    s[0] = R[0]; s[1] = R[4]; s[2] = R[8];
}

/* Fixed-corotated Kirchhoff stress tau = 2 mu (F - R) F^T + lam J (J-1) I,
 * R = U V^T, MPMCUDA.cu:229-264. */
void orc_fcr_tau(const double F[9], double mu, double lam, double tau[9])
{
    double U[9], s[3], V[9], R[9], Vt[9], D[9], Ft[9];
    orc_svd3(F, U, s, V);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) { Vt[3 * r + c] = V[3 * c + r]; Ft[3 * r + c] = F[3 * c + r]; }
    mm3(U, Vt, R);
    for (int i = 0; i < 9; ++i) D[i] = F[i] - R[i];
    mm3(D, Ft, tau);
    double J = orc_det3(F);
    for (int i = 0; i < 9; ++i) tau[i] *= 2.0 * mu;
    tau[0] += lam * J * (J - 1.0);
    tau[4] += lam * J * (J - 1.0);
    tau[8] += lam * J * (J - 1.0);
}

/* Snow return mapping, MPMCUDA.cu:631-678: SVD of the trial elastic
 * gradient, clamp singular values to [1-theta_c, 1+theta_s],
 * Fe = U S* V^T, Fp <- Fe^-1 Ftr Fp. */
void orc_snow_plastic(const double Ftr[9], const double Fp[9], double theta_c, double theta_s,
                      double Fe_new[9], double Fp_new[9])
{
    double U[9], s[3], V[9], US[9], Vt[9], Einv[9], T[9];
    orc_svd3(Ftr, U, s, V);
    for (int k = 0; k < 3; ++k) {
        double c = s[k];
        if (c < 1.0 - theta_c) c = 1.0 - theta_c;
        if (c > 1.0 + theta_s) c = 1.0 + theta_s;
        s[k] = c;
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) { US[3 * r + c] = U[3 * r + c] * s[c]; Vt[3 * r + c] = V[3 * c + r]; }
    mm3(US, Vt, Fe_new);
    if (orc_inv3(Fe_new, Einv) != 0) { memcpy(Fp_new, Fp, 9 * sizeof(double)); return; }
    mm3(Einv, Ftr, T);
    mm3(T, Fp, Fp_new);
}

/* ------------------------------------------------------------------------ */
static inline int64_t gidx(const orc_grid *g, int x, int y, int z)
{
    return ((int64_t)z * g->ny + y) * g->nx + x;    /* MPMCUDA.cu:375 */
}

void orc_clear_grid(orc_grid *g)                   /* MPMCUDA.cu:798-808 */
{
    int64_t G = (int64_t)g->nx * g->ny * g->nz;
    memset(g->m, 0, (size_t)G * sizeof(double));
    memset(g->p, 0, (size_t)G * 3 * sizeof(double));
}

/* P2G for one particle group, MPMCUDA.cu:310-433. */
int orc_p2g(orc_grid *g, const orc_group *grp, double dt)
{
    int flags = 0;
    const orc_material *M = &grp->mat;
    const double h = g->h, m = M->mass, vol = M->vol;
    /* with more than one thread (orc_set_threads) the node sums are atomic
     * and their order, hence the last bits, varies between runs */
#pragma omp parallel for schedule(static) reduction(| : flags)
    for (int64_t i = 0; i < grp->n; ++i) {
        const double *xp = grp->x + 3 * i, *vp = grp->v + 3 * i, *Cp = grp->C + 9 * i;
        double tau[9] = {0};
        double wscale = 0.0;  /* water: V_p * pressure * J (MPMCUDA.cu:413) */
        if (M->type == ORC_JELLO) {
            orc_fcr_tau(grp->Fe + 9 * i, M->mu, M->lam, tau);                    /* :329-333 */
        } else if (M->type == ORC_SNOW) {
            double Jp = orc_det3(grp->Fp + 9 * i);                              /* :342 */
            double hc = exp(M->xi * (1.0 - Jp));                                /* :345 */
            orc_fcr_tau(grp->Fe + 9 * i, M->mu * hc, M->lam * hc, tau);          /* :346-349 */
        } else {
            double J = grp->J[i];
            double pr = M->eos_k * (1.0 / pow(J, M->eos_gamma) - 1.0);          /* :352-353 */
            wscale = vol * pr * J;
        }
        int b[3];
        for (int a = 0; a < 3; ++a) b[a] = orc_base_ref(xp[a], g->origin[a], h);  /* :356-358 */
        double tw = 0.0;
        for (int ox = 0; ox < 3; ++ox)
            for (int oy = 0; oy < 3; ++oy)
                for (int oz = 0; oz < 3; ++oz) {
                    int ix = b[0] + ox, iy = b[1] + oy, iz = b[2] + oz;
                    if (ix >= g->nx || iy >= g->ny || iz >= g->nz) { flags |= ORC_FLAG_STENCIL_OOB; continue; }
                    double xi[3] = {g->origin[0] + ix * h, g->origin[1] + iy * h, g->origin[2] + iz * h};
                    double w = orc_weight(xp, xi, h);
                    int64_t n = gidx(g, ix, iy, iz);
#pragma omp atomic
                    g->m[n] += w * m;                                            /* :388 */
                    double d[3] = {xi[0] - xp[0], xi[1] - xp[1], xi[2] - xp[2]};
                    for (int c = 0; c < 3; ++c) {                                 /* :391-401 */
                        double cd = Cp[3 * c] * d[0] + Cp[3 * c + 1] * d[1] + Cp[3 * c + 2] * d[2];
#pragma omp atomic
                        g->p[3 * n + c] += w * m * (vp[c] + cd);
                    }
                    tw += w;
                    double gw[3];
                    orc_weight_grad(xp, xi, h, gw);                               /* :405-407 */
                    double f[3];
                    if (M->type == ORC_WATER) {                                   /* :412-414 */
                        for (int c = 0; c < 3; ++c) f[c] = wscale * gw[c];
                    } else {                                                      /* :409-411 */
                        for (int c = 0; c < 3; ++c)
                            f[c] = -vol * (tau[3 * c] * gw[0] + tau[3 * c + 1] * gw[1] + tau[3 * c + 2] * gw[2]);
                    }
                    for (int c = 0; c < 3; ++c) {                                 /* :419-421 */
#pragma omp atomic
                        g->p[3 * n + c] += dt * f[c];
                    }
                }
        if (fabs(tw - 1.0) >= 1e-3) flags |= ORC_FLAG_WEIGHT_SUM;                /* :431 */
    }
    return flags;
}

/* Grid velocity update, MPMCUDA.cu:435-472 (launch parameters :841-854).
 * The wall test uses node indices (i <= pad || i >= N - pad), which equals
 * the reference's FP64 coordinate comparison whenever (N-pad)*h is exact. */
int orc_grid_update(orc_grid *g, const orc_params *prm, double dt)
{
    const int pad = prm->bc_pad;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < g->nz; ++z)
        for (int y = 0; y < g->ny; ++y)
            for (int x = 0; x < g->nx; ++x) {
                int64_t n = gidx(g, x, y, z);
                double mass = g->m[n];
                double *v = g->p + 3 * n;
                if (fabs(mass) < 2.220446049250313e-16) {                        /* :444-449 */
                    g->m[n] = 0.0; v[0] = v[1] = v[2] = 0.0;
                    continue;
                }
                for (int c = 0; c < 3; ++c) v[c] = v[c] / mass + prm->gravity[c] * dt; /* :452-457 */
                int wall[3] = {x <= pad || x >= g->nx - pad, y <= pad || y >= g->ny - pad,
                               z <= pad || z >= g->nz - pad};
                if (prm->bc_mode == ORC_BC_STICKY) {
                    if (wall[0] || wall[1] || wall[2]) v[0] = v[1] = v[2] = 0.0;
                } else {
                    for (int c = 0; c < 3; ++c) if (wall[c]) v[c] = 0.0;           /* :467-469 */
                }
            }
    return ORC_OK;
}

/* G2P, deformation update and advection, MPMCUDA.cu:474-741. */
int orc_g2p(const orc_grid *g, orc_group *grp, double dt)
{
    int flags = 0;
    const orc_material *M = &grp->mat;
    const double h = g->h, h2i = 1.0 / (h * h);
#pragma omp parallel for schedule(static) reduction(| : flags)
    for (int64_t i = 0; i < grp->n; ++i) {
        double *xp = grp->x + 3 * i, *vp = grp->v + 3 * i, *Cp = grp->C + 9 * i;
        int b[3];
        for (int a = 0; a < 3; ++a) b[a] = orc_base_ref(xp[a], g->origin[a], h);  /* :486-488 */
        double vn[3] = {0, 0, 0}, Cn[9] = {0}, gv[9] = {0}, tw = 0.0, sdot = 0.0;
        for (int ox = 0; ox < 3; ++ox)
            for (int oy = 0; oy < 3; ++oy)
                for (int oz = 0; oz < 3; ++oz) {
                    int ix = b[0] + ox, iy = b[1] + oy, iz = b[2] + oz;
                    if (ix >= g->nx || iy >= g->ny || iz >= g->nz) { flags |= ORC_FLAG_STENCIL_OOB; continue; }
                    double xi[3] = {g->origin[0] + ix * h, g->origin[1] + iy * h, g->origin[2] + iz * h};
                    double w = orc_weight(xp, xi, h);
                    const double *vi = g->p + 3 * gidx(g, ix, iy, iz);
                    double d[3] = {xi[0] - xp[0], xi[1] - xp[1], xi[2] - xp[2]};
                    double gw[3];
                    orc_weight_grad(xp, xi, h, gw);
                    for (int r = 0; r < 3; ++r) {
                        vn[r] += w * vi[r];                                       /* :516-518 */
                        for (int c = 0; c < 3; ++c) {
                            Cn[3 * r + c] += 4.0 * w * h2i * vi[r] * d[c];        /* :520-532 */
                            gv[3 * r + c] += vi[r] * gw[c];                       /* :561-585 */
                        }
                    }
                    sdot += vi[0] * gw[0] + vi[1] * gw[1] + vi[2] * gw[2];        /* :535-537 */
                    tw += w;
                }
        if (fabs(tw - 1.0) >= 1e-4) flags |= ORC_FLAG_WEIGHT_SUM;                /* :546 */
        for (int r = 0; r < 3; ++r) vp[r] = vn[r];
        if (M->type == ORC_WATER && M->water_c_mode == ORC_WATER_C_GRADV)
            memcpy(Cp, gv, sizeof gv);
        else
            memcpy(Cp, Cn, sizeof Cn);                                            /* :548-556 */
        double L[9];                                                               /* I + dt grad v, :588-595 */
        for (int k = 0; k < 9; ++k) L[k] = dt * gv[k];
        L[0] += 1.0; L[4] += 1.0; L[8] += 1.0;
        if (M->type == ORC_JELLO) {                                                /* :596-619 */
            double Fn[9];
            mm3(L, grp->Fe + 9 * i, Fn);
            memcpy(grp->Fe + 9 * i, Fn, sizeof Fn);
            double d = orc_det3(Fn);
            if (!(d > 0)) flags |= ORC_FLAG_DET_NONPOS;
            if (grp->dgdiff) grp->dgdiff[i] = d > 1.0 ? d : 1.0 / d;
        } else if (M->type == ORC_SNOW) {                                          /* :620-709 */
            double Ftr[9], Fe[9], Fp[9];
            mm3(L, grp->Fe + 9 * i, Ftr);
            orc_snow_plastic(Ftr, grp->Fp + 9 * i, M->theta_c, M->theta_s, Fe, Fp);
            memcpy(grp->Fe + 9 * i, Fe, sizeof Fe);
            memcpy(grp->Fp + 9 * i, Fp, sizeof Fp);
            double d = orc_det3(Fp);
            if (!(d > 0)) flags |= ORC_FLAG_DET_NONPOS;
            if (grp->dgdiff) grp->dgdiff[i] = d > 1.0 ? d : 1.0 / d;
        } else {                                                                    /* :710-713 */
            grp->J[i] = (1.0 + dt * sdot) * grp->J[i];
        }
        for (int r = 0; r < 3; ++r) xp[r] += dt * vp[r];                           /* :717-719 */
        for (int a = 0; a < 3; ++a) {                                               /* :722-739 */
            int N = a == 0 ? g->nx : (a == 1 ? g->ny : g->nz);
            double lo = g->origin[a], hi = g->origin[a] + h * (N - 1);
            if (!(xp[a] >= lo && xp[a] <= hi)) flags |= ORC_FLAG_LEFT_DOMAIN;
        }
        for (int r = 0; r < 3; ++r)
            if (!isfinite(xp[r]) || !isfinite(vp[r])) flags |= ORC_FLAG_NONFINITE;
    }
    return flags;
}

/* One explicit step, MPMCUDA.cu:793-925 (host loop over particle groups). */
int orc_step(orc_grid *g, orc_group *groups, int ngroups, const orc_params *prm, double dt)
{
    int flags = 0;
    orc_clear_grid(g);
    for (int k = 0; k < ngroups; ++k) flags |= orc_p2g(g, groups + k, dt);
    flags |= orc_grid_update(g, prm, dt);
    for (int k = 0; k < ngroups; ++k) flags |= orc_g2p(g, groups + k, dt);
    return flags;
}

/* Threads of the OpenMP loops (1 = the sequential oracle the tests use). */
int orc_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
    return n > 0 ? n : 1;
#else
    (void)n;
    return 1;
#endif
}
