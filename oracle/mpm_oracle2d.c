/*
 * FP64 CPU oracle of the reference's 2D snow step (test infrastructure only).
 *
 * Restates `substep()` of JerryYan97/MPM_CUDA MPMSnow.py:77-150 (Taichi):
 * grid clear :80-82, P2G :84-103, grid velocity update :105-114, G2P with
 * snow plasticity and advection :116-150, with the stress of
 * FixedCorotatedStress :28-32 and the weights of BSpline1D :36-45,
 * BSplineDerivative :55-65, GradWip :69-74.  Every constant the script
 * hard-codes is a parameter (orc2_params).  FP64 throughout; the 2x2 SVD is
 * the polar decomposition followed by the symmetric eigen-decomposition.
 */
#include <math.h>
#include <string.h>

#include "mpm_oracle.h"

/* BSpline1D, MPMSnow.py:36-45 */
double orc2_N(double x)
{
    const double a = fabs(x);
    if (a < 0.5) return 0.75 - a * a;
    if (a < 1.5) return 0.5 * (1.5 - a) * (1.5 - a);
    return 0.0;
}

/* BSplineDerivative, MPMSnow.py:55-65 (interval ends as written there) */
double orc2_dN(double x)
{
    if (-0.5 < x && x < 0.5) return -2.0 * x;
    if (0.5 <= x && x < 1.5) return x - 1.5;
    if (-1.5 < x && x <= -0.5) return 1.5 + x;
    return 0.0;
}

/* ti.svd of a 2x2 (row-major F = U diag(s) V^T, U and V rotations): polar
 * rotation R of F, then the eigenvectors V of the symmetric S = R^T F,
 * U = R V; s[0] >= s[1] (s[1] < 0 only for det F < 0). */
void orc2_svd(const double F[4], double U[4], double s[2], double V[4])
{
    double c = F[0] + F[3], sn = F[2] - F[1];
    const double r = hypot(c, sn);
    if (r > 0.0) { c /= r; sn /= r; } else { c = 1.0; sn = 0.0; }
    const double a = c * F[0] + sn * F[2];                  /* S = R^T F */
    const double d = -sn * F[1] + c * F[3];
    const double b = 0.5 * ((c * F[1] + sn * F[3]) + (-sn * F[0] + c * F[2]));
    const double th = (b == 0.0 && a == d) ? 0.0 : 0.5 * atan2(2.0 * b, a - d);
    const double ct = cos(th), st = sin(th);
    V[0] = ct; V[1] = -st; V[2] = st; V[3] = ct;
    s[0] = a * ct * ct + 2.0 * b * ct * st + d * st * st;
    s[1] = a * st * st - 2.0 * b * ct * st + d * ct * ct;
    /* U = R V, R = [[c, -sn], [sn, c]] */
    U[0] = c * V[0] - sn * V[2]; U[1] = c * V[1] - sn * V[3];
    U[2] = sn * V[0] + c * V[2]; U[3] = sn * V[1] + c * V[3];
}

/* FixedCorotatedStress, MPMSnow.py:28-32: 2 mu (F - R) F^T + lam (J - 1) J I */
void orc2_fcr_tau(const double F[4], double mu, double lam, double tau[4])
{
    double U[4], s[2], V[4];
    orc2_svd(F, U, s, V);
    double R[4];
    R[0] = U[0] * V[0] + U[1] * V[1]; R[1] = U[0] * V[2] + U[1] * V[3];
    R[2] = U[2] * V[0] + U[3] * V[1]; R[3] = U[2] * V[2] + U[3] * V[3];
    const double J = F[0] * F[3] - F[1] * F[2];
    double D[4];
    for (int k = 0; k < 4; ++k) D[k] = F[k] - R[k];
    /* (F - R) F^T */
    tau[0] = D[0] * F[0] + D[1] * F[1]; tau[1] = D[0] * F[2] + D[1] * F[3];
    tau[2] = D[2] * F[0] + D[3] * F[1]; tau[3] = D[2] * F[2] + D[3] * F[3];
    for (int k = 0; k < 4; ++k) tau[k] *= 2.0 * mu;
    tau[0] += lam * (J - 1.0) * J;
    tau[3] += lam * (J - 1.0) * J;
}

static inline void mm2(const double A[4], const double B[4], double C[4])
{
    double T[4];
    T[0] = A[0] * B[0] + A[1] * B[2]; T[1] = A[0] * B[1] + A[1] * B[3];
    T[2] = A[2] * B[0] + A[3] * B[2]; T[3] = A[2] * B[1] + A[3] * B[3];
    memcpy(C, T, sizeof T);
}

static inline int base_of(double x, double inv_dx)
{
    return (int)(x * inv_dx - 0.5);                          /* .cast(int): truncation, :85 */
}

/* One substep (MPMSnow.py:77-150).  Grid arrays are [n][n] with index
 * j * n + i (i along x); grid_v holds momentum after P2G and velocity after
 * the update.  Returns health flags (stencil out of the grid, weight sum). */
int orc2_step(double *grid_m, double *grid_v, orc2_particles *pp, const orc2_params *prm, double dt)
{
    const int n = prm->n_grid;
    const double dx = prm->dx, inv_dx = 1.0 / prm->dx;
    int flags = 0;
    /* 1. clean grid, :80-82 */
    memset(grid_m, 0, sizeof(double) * (size_t)n * n);
    memset(grid_v, 0, sizeof(double) * 2 * (size_t)n * n);
    /* 2. P2G, :84-103 */
    for (int64_t p = 0; p < pp->n; ++p) {
        const double *x = pp->x + 2 * p, *v = pp->v + 2 * p, *Cp = pp->C + 4 * p;
        const double *Fe = pp->Fe + 4 * p, *Fp = pp->Fp + 4 * p;
        const int bx = base_of(x[0], inv_dx), by = base_of(x[1], inv_dx);
        if (bx < 0 || by < 0 || bx + 2 >= n || by + 2 >= n) { flags |= ORC_FLAG_STENCIL_OOB; continue; }
        const double Jp = Fp[0] * Fp[3] - Fp[1] * Fp[2];
        const double hh = exp(prm->xi * (1.0 - Jp));            /* :87 */
        double tau[4];
        orc2_fcr_tau(Fe, hh * prm->mu0, hh * prm->lam0, tau);
        double tw = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const double gx = (bx + i) * dx, gy = (by + j) * dx;
                const double ux = (x[0] - gx) / dx, uy = (x[1] - gy) / dx;
                const double w = orc2_N(ux) * orc2_N(uy);
                const double dpx = gx - x[0], dpy = gy - x[1];
                const double gwx = orc2_dN(ux) * orc2_N(uy) / dx, gwy = orc2_N(ux) * orc2_dN(uy) / dx;
                const int64_t k = (int64_t)(by + j) * n + (bx + i);
                grid_v[2 * k] += w * prm->p_mass * (v[0] + Cp[0] * dpx + Cp[1] * dpy);
                grid_v[2 * k + 1] += w * prm->p_mass * (v[1] + Cp[2] * dpx + Cp[3] * dpy);
                grid_m[k] += w * prm->p_mass;
                grid_v[2 * k] += -prm->p_vol * dt * (tau[0] * gwx + tau[1] * gwy);
                grid_v[2 * k + 1] += -prm->p_vol * dt * (tau[2] * gwx + tau[3] * gwy);
                tw += w;
            }
        if (fabs(tw - 1.0) > 1e-4) flags |= ORC_FLAG_WEIGHT_SUM;   /* :102-103 */
    }
    /* 3. velocity update, :105-114 */
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const int64_t k = (int64_t)j * n + i;
            if (grid_m[k] > 0.0) {
                grid_v[2 * k] = grid_v[2 * k] / grid_m[k] + dt * prm->gravity[0];
                grid_v[2 * k + 1] = grid_v[2 * k + 1] / grid_m[k] + dt * prm->gravity[1];
                if (i <= prm->wall || i >= n - prm->wall || j <= prm->wall || j >= n - prm->wall)
                    grid_v[2 * k] = grid_v[2 * k + 1] = 0.0;
            } else {
                grid_v[2 * k] = grid_v[2 * k + 1] = 0.0;
                grid_m[k] = 0.0;
            }
        }
    /* 4. G2P, :116-150 */
    for (int64_t p = 0; p < pp->n; ++p) {
        double *x = pp->x + 2 * p, *v = pp->v + 2 * p, *Cp = pp->C + 4 * p;
        double *Fe = pp->Fe + 4 * p, *Fp = pp->Fp + 4 * p;
        const int bx = base_of(x[0], inv_dx), by = base_of(x[1], inv_dx);
        if (bx < 0 || by < 0 || bx + 2 >= n || by + 2 >= n) { flags |= ORC_FLAG_STENCIL_OOB; continue; }
        double nv[2] = {0, 0}, nC[4] = {0, 0, 0, 0}, gv[4] = {0, 0, 0, 0}, tw = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const double gx = (bx + i) * dx, gy = (by + j) * dx;
                const double ux = (x[0] - gx) / dx, uy = (x[1] - gy) / dx;
                const double w = orc2_N(ux) * orc2_N(uy);
                const double dpx = gx - x[0], dpy = gy - x[1];
                const double gwx = orc2_dN(ux) * orc2_N(uy) / dx, gwy = orc2_N(ux) * orc2_dN(uy) / dx;
                const int64_t k = (int64_t)(by + j) * n + (bx + i);
                const double vx = grid_v[2 * k], vy = grid_v[2 * k + 1];
                nv[0] += w * vx; nv[1] += w * vy;
                const double f = 4.0 * w / (dx * dx);                 /* :130 */
                nC[0] += f * vx * dpx; nC[1] += f * vx * dpy;
                nC[2] += f * vy * dpx; nC[3] += f * vy * dpy;
                gv[0] += vx * gwx; gv[1] += vx * gwy;                 /* :131 */
                gv[2] += vy * gwx; gv[3] += vy * gwy;
                tw += w;
            }
        if (fabs(tw - 1.0) > 1e-4) flags |= ORC_FLAG_WEIGHT_SUM;
        /* Fe_trial = (I + dt grad_v) Fe, :137 */
        double A[4] = {1.0 + dt * gv[0], dt * gv[1], dt * gv[2], 1.0 + dt * gv[3]}, Ftr[4];
        mm2(A, Fe, Ftr);
        double U[4], s[2], V[4];
        orc2_svd(Ftr, U, s, V);
        double ss[2];
        for (int a = 0; a < 2; ++a) ss[a] = fmin(fmax(s[a], 1.0 - prm->theta_c), 1.0 + prm->theta_s);   /* :141-142 */
        /* Fe = U diag(ss) V^T, :144 */
        double Fn[4];
        Fn[0] = U[0] * ss[0] * V[0] + U[1] * ss[1] * V[1]; Fn[1] = U[0] * ss[0] * V[2] + U[1] * ss[1] * V[3];
        Fn[2] = U[2] * ss[0] * V[0] + U[3] * ss[1] * V[1]; Fn[3] = U[2] * ss[0] * V[2] + U[3] * ss[1] * V[3];
        /* Fp = Fe^-1 Fe_trial Fp, :145 */
        const double det = Fn[0] * Fn[3] - Fn[1] * Fn[2];
        const double Fi[4] = {Fn[3] / det, -Fn[1] / det, -Fn[2] / det, Fn[0] / det};
        double T[4];
        mm2(Fi, Ftr, T);
        mm2(T, Fp, Fp);
        memcpy(Fe, Fn, sizeof Fn);
        v[0] = nv[0]; v[1] = nv[1];                                   /* :146-148 */
        memcpy(Cp, nC, sizeof nC);
        x[0] += dt * v[0];
        x[1] += dt * v[1];
    }
    return flags;
}
