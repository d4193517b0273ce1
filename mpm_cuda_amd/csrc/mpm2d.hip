// mpm2d.hip -- the reference's 2D snow step on the device.
//
// Reference: JerryYan97/MPM_CUDA MPMSnow.py (Taichi): substep :77-150 --
// grid clear :80-82, P2G :84-103 (hardening :87, FixedCorotatedStress
// :28-32, force with the weight gradient GradWip :69-74), grid velocity
// update with gravity and sticky walls :105-114, G2P with the APIC C, the
// velocity gradient, the snow clamp and advection :116-150; particle
// initialisation reset :154-162 (F_e = F_p = I, C = 0).
//
// FP32 structure of arrays on the device.  The workload is small (thousands
// of particles on a 128^2 grid), so a step is three short kernels -- P2G
// with float atomics into a (m, p) grid, the grid update (which also clears
// the accumulation grid for the next step), G2P -- and the runs replay a
// hipGraph of `graph_steps` steps to take the launch cost off the host.
// J_p = det F_p is stored instead of F_p (the step reads only its
// determinant, :86): J_p <- J_p (s1 s2) / (s1* s2*), which equals
// det(F_e^-1 F_trial F_p) of :145.  F_e is written as F_trial + U (S* - S) V^T,
// equal to U S* V^T of :144 and exact when no singular value is clamped.
#include "mpm_launch.h"

#include <string.h>

#include <algorithm>
#include <vector>

using namespace mpm;

namespace {

enum Plane2 : int { Q_X = 0, Q_V = 2, Q_C = 4, Q_F = 8, Q_J = 12, Q_COUNT = 13 };

struct P2 {
    int n, wall;
    float dx, inv_dx, p_vol, p_mass, mu0, lam0, xi, theta_c, theta_s, gx, gy, dt;
};

__device__ __forceinline__ float ld2(const float *P, int64_t cap, int plane, int i) { return gp(P)[plane * cap + i]; }
__device__ __forceinline__ void st2(float *P, int64_t cap, int plane, int i, float v) { gp(P)[plane * cap + i] = v; }

// BSpline1D (:36-45) and BSplineDerivative (:55-65) at u = (x_p - x_i) / dx.
__device__ __forceinline__ float bs_N(float u)
{
    const float a = fabsf(u);
    if (a < 0.5f) return 0.75f - a * a;
    if (a < 1.5f) return 0.5f * (1.5f - a) * (1.5f - a);
    return 0.0f;
}
__device__ __forceinline__ float bs_dN(float u)
{
    if (-0.5f < u && u < 0.5f) return -2.0f * u;
    if (0.5f <= u && u < 1.5f) return u - 1.5f;
    if (-1.5f < u && u <= -0.5f) return 1.5f + u;
    return 0.0f;
}

// 2x2 SVD F = U diag(s) V^T (row-major): polar rotation of F, then the
// symmetric eigenproblem of S = R^T F; s[0] >= s[1].
__device__ __forceinline__ void svd2(const float F[4], float U[4], float s[2], float V[4])
{
    float c = F[0] + F[3], sn = F[2] - F[1];
    const float r = sqrtf(c * c + sn * sn);
    if (r > 0.0f) { c /= r; sn /= r; } else { c = 1.0f; sn = 0.0f; }
    const float a = c * F[0] + sn * F[2];
    const float d = -sn * F[1] + c * F[3];
    const float b = 0.5f * ((c * F[1] + sn * F[3]) + (-sn * F[0] + c * F[2]));
    const float th = (b == 0.0f && a == d) ? 0.0f : 0.5f * atan2f(2.0f * b, a - d);
    const float ct = cosf(th), st = sinf(th);
    V[0] = ct; V[1] = -st; V[2] = st; V[3] = ct;
    s[0] = a * ct * ct + 2.0f * b * ct * st + d * st * st;
    s[1] = a * st * st - 2.0f * b * ct * st + d * ct * ct;
    U[0] = c * V[0] - sn * V[2]; U[1] = c * V[1] - sn * V[3];
    U[2] = sn * V[0] + c * V[2]; U[3] = sn * V[1] + c * V[3];
}

__device__ __forceinline__ uint32_t base2(float x, float inv_dx, int n, int &b)
{
    b = (int)(x * inv_dx - 0.5f);                  // .cast(int), :85
    if (b >= 0 && b <= n - 3) return 0u;
    b = min(max(b, 0), n - 3);
    return MPM_H_STENCIL_OOB;
}

// P2G (:84-103): one thread per particle, 9 nodes x 3 float atomics.
__global__ void __launch_bounds__(256) k2_p2g(P2 p, const float *__restrict__ P, int64_t cap, int np, float4 *gmp,
                                              uint32_t *health)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float x[2] = {ld2(P, cap, Q_X, i), ld2(P, cap, Q_X + 1, i)};
    const float v[2] = {ld2(P, cap, Q_V, i), ld2(P, cap, Q_V + 1, i)};
    float C[4], F[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) { C[k] = ld2(P, cap, Q_C + k, i); F[k] = ld2(P, cap, Q_F + k, i); }
    const float Jp = ld2(P, cap, Q_J, i);
    int bx, by;
    uint32_t bits = base2(x[0], p.inv_dx, p.n, bx) | base2(x[1], p.inv_dx, p.n, by);
    // FixedCorotatedStress with h mu0, h lam0 (:87-89)
    const float hh = expf(p.xi * (1.0f - Jp));
    float U[4], s[2], V[4];
    svd2(F, U, s, V);
    const float R[4] = {U[0] * V[0] + U[1] * V[1], U[0] * V[2] + U[1] * V[3], U[2] * V[0] + U[3] * V[1],
                        U[2] * V[2] + U[3] * V[3]};
    const float J = F[0] * F[3] - F[1] * F[2];
    const float D[4] = {F[0] - R[0], F[1] - R[1], F[2] - R[2], F[3] - R[3]};
    const float mu2 = 2.0f * hh * p.mu0, lj = hh * p.lam0 * (J - 1.0f) * J;
    const float tau[4] = {mu2 * (D[0] * F[0] + D[1] * F[1]) + lj, mu2 * (D[0] * F[2] + D[1] * F[3]),
                          mu2 * (D[2] * F[0] + D[3] * F[1]), mu2 * (D[2] * F[2] + D[3] * F[3]) + lj};
    const float fs = -p.p_vol * p.dt;
    float tw = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float gxp = (float)(bx + a) * p.dx, gyp = (float)(by + b) * p.dx;
            const float ux = (x[0] - gxp) / p.dx, uy = (x[1] - gyp) / p.dx;
            const float nx = bs_N(ux), ny = bs_N(uy);
            const float w = nx * ny;
            const float dpx = gxp - x[0], dpy = gyp - x[1];
            const float gwx = bs_dN(ux) * ny / p.dx, gwy = nx * bs_dN(uy) / p.dx;
            const float mw = w * p.p_mass;
            const float px = mw * (v[0] + C[0] * dpx + C[1] * dpy) + fs * (tau[0] * gwx + tau[1] * gwy);
            const float py = mw * (v[1] + C[2] * dpx + C[3] * dpy) + fs * (tau[2] * gwx + tau[3] * gwy);
            float *dst = (float *)(gmp + (int64_t)(by + b) * p.n + (bx + a));
            gatomic_add(dst + 0, mw);
            gatomic_add(dst + 1, px);
            gatomic_add(dst + 2, py);
            tw += w;
        }
    if (fabsf(tw - 1.0f) > 1e-4f) bits |= MPM_H_WEIGHT_SUM;        // :102-103
    if (bits) gatomic_or(health, bits);
}

// Velocity update (:105-114), one thread per node; clears the accumulation
// grid for the next step.  gv = (vx, vy, m).
__global__ void __launch_bounds__(256) k2_grid(P2 p, float4 *gmp, float4 *gv)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= p.n * p.n) return;
    const float4 a = gload4(gmp, k);
    gstore4(gmp, k, make_float4(0.f, 0.f, 0.f, 0.f));
    float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
    if (a.x > 0.0f) {
        const int i = k % p.n, j = k / p.n;
        float vx = a.y / a.x + p.dt * p.gx, vy = a.z / a.x + p.dt * p.gy;
        if (i <= p.wall || i >= p.n - p.wall || j <= p.wall || j >= p.n - p.wall) vx = vy = 0.0f;
        out = make_float4(vx, vy, a.x, 0.f);
    }
    gstore4(gv, k, out);
}

// G2P, plasticity and advection (:116-150), one thread per particle.
__global__ void __launch_bounds__(256) k2_g2p(P2 p, float *P, int64_t cap, int np, const float4 *gv, uint32_t *health)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    float x[2] = {ld2(P, cap, Q_X, i), ld2(P, cap, Q_X + 1, i)};
    float F[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) F[k] = ld2(P, cap, Q_F + k, i);
    const float Jp = ld2(P, cap, Q_J, i);
    int bx, by;
    uint32_t bits = base2(x[0], p.inv_dx, p.n, bx) | base2(x[1], p.inv_dx, p.n, by);
    float nv[2] = {0.f, 0.f}, nC[4] = {0.f, 0.f, 0.f, 0.f}, gvel[4] = {0.f, 0.f, 0.f, 0.f}, tw = 0.f;
    const float c4 = 4.0f / (p.dx * p.dx);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float gxp = (float)(bx + a) * p.dx, gyp = (float)(by + b) * p.dx;
            const float ux = (x[0] - gxp) / p.dx, uy = (x[1] - gyp) / p.dx;
            const float nx = bs_N(ux), ny = bs_N(uy);
            const float w = nx * ny;
            const float dpx = gxp - x[0], dpy = gyp - x[1];
            const float gwx = bs_dN(ux) * ny / p.dx, gwy = nx * bs_dN(uy) / p.dx;
            const float4 g = gload4(gv, (int64_t)(by + b) * p.n + (bx + a));
            nv[0] += w * g.x; nv[1] += w * g.y;
            const float f = c4 * w;
            nC[0] += f * g.x * dpx; nC[1] += f * g.x * dpy;
            nC[2] += f * g.y * dpx; nC[3] += f * g.y * dpy;
            gvel[0] += g.x * gwx; gvel[1] += g.x * gwy;
            gvel[2] += g.y * gwx; gvel[3] += g.y * gwy;
            tw += w;
        }
    if (fabsf(tw - 1.0f) > 1e-4f) bits |= MPM_H_WEIGHT_SUM;
    // F_trial = (I + dt grad v) F_e (:137)
    const float A[4] = {1.0f + p.dt * gvel[0], p.dt * gvel[1], p.dt * gvel[2], 1.0f + p.dt * gvel[3]};
    const float Ft[4] = {A[0] * F[0] + A[1] * F[2], A[0] * F[1] + A[1] * F[3], A[2] * F[0] + A[3] * F[2],
                         A[2] * F[1] + A[3] * F[3]};
    float U[4], s[2], V[4];
    svd2(Ft, U, s, V);
    float ss[2], ds[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        ss[k] = fminf(fmaxf(s[k], 1.0f - p.theta_c), 1.0f + p.theta_s);   // :141-142
        ds[k] = ss[k] - s[k];
    }
    // F_e = F_trial + U (S* - S) V^T (:144); J_p update of :145
    const float Fn[4] = {Ft[0] + U[0] * ds[0] * V[0] + U[1] * ds[1] * V[1],
                         Ft[1] + U[0] * ds[0] * V[2] + U[1] * ds[1] * V[3],
                         Ft[2] + U[2] * ds[0] * V[0] + U[3] * ds[1] * V[1],
                         Ft[3] + U[2] * ds[0] * V[2] + U[3] * ds[1] * V[3]};
    const float Jn = Jp * ((s[0] * s[1]) / (ss[0] * ss[1]));
    x[0] += p.dt * nv[0];
    x[1] += p.dt * nv[1];
    if (!(isfinite(x[0]) && isfinite(x[1]) && isfinite(Jn))) bits |= MPM_H_NONFINITE;
    st2(P, cap, Q_X, i, x[0]); st2(P, cap, Q_X + 1, i, x[1]);
    st2(P, cap, Q_V, i, nv[0]); st2(P, cap, Q_V + 1, i, nv[1]);
#pragma unroll
    for (int k = 0; k < 4; ++k) { st2(P, cap, Q_C + k, i, nC[k]); st2(P, cap, Q_F + k, i, Fn[k]); }
    st2(P, cap, Q_J, i, Jn);
    if (bits) gatomic_or(health, bits);
}

__global__ void k2_init(float *P, int64_t cap, int np, const float *__restrict__ xy, const float *__restrict__ vel)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    st2(P, cap, Q_X, i, xy[2 * i]); st2(P, cap, Q_X + 1, i, xy[2 * i + 1]);
    st2(P, cap, Q_V, i, vel ? vel[2 * i] : 0.f); st2(P, cap, Q_V + 1, i, vel ? vel[2 * i + 1] : 0.f);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        st2(P, cap, Q_C + k, i, 0.f);
        st2(P, cap, Q_F + k, i, (k == 0 || k == 3) ? 1.f : 0.f);
    }
    st2(P, cap, Q_J, i, 1.f);
}

__global__ void k2_fill(float4 *p, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) gstore4(p, k, make_float4(0.f, 0.f, 0.f, 0.f));
}

inline int blocks(int64_t n, int t) { return (int)((n + t - 1) / t); }

}  // namespace

struct mpm2d_sim {
    mpm2d_config cfg;
    P2 p;
    hipStream_t stream = nullptr;
    float *P = nullptr;
    int64_t cap = 0;
    int np = 0;
    float4 *gmp = nullptr, *gv = nullptr;
    uint32_t *health = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_len = 0;
    int64_t steps = 0;
};

namespace {

int enqueue_steps(mpm2d_sim *s, int k)
{
    const int nn = s->p.n * s->p.n;
    for (int t = 0; t < k; ++t) {
        hipLaunchKernelGGL(k2_p2g, dim3(blocks(s->np, 256)), dim3(256), 0, s->stream, s->p, s->P, s->cap, s->np,
                           s->gmp, s->health);
        hipLaunchKernelGGL(k2_grid, dim3(blocks(nn, 256)), dim3(256), 0, s->stream, s->p, s->gmp, s->gv);
        hipLaunchKernelGGL(k2_g2p, dim3(blocks(s->np, 256)), dim3(256), 0, s->stream, s->p, s->P, s->cap, s->np,
                           s->gv, s->health);
    }
    MPM_HIP(hipGetLastError());
    return MPM_OK;
}

}  // namespace

extern "C" {

void mpm2d_config_default(mpm2d_config *c)
{
    if (!c) return;
    memset(c, 0, sizeof *c);
    c->n_grid = 128;          // MPMSnow.py:7
    c->wall = 5;              // :109
    c->dt = 1e-4;
    c->E = 1e3;
    c->nu = 0.2;              // :12
    c->rho = 1.0;             // :10
    c->theta_c = 2.5e-2;
    c->theta_s = 7.5e-3;
    c->xi = 10.0;             // :87
    c->gravity[1] = -50.0;
    c->graph_steps = 50;
}

int mpm2d_create(const mpm2d_config *cfg, mpm2d_sim **out)
{
    if (!cfg || !out) return set_error(MPM_ERR_INVALID, "null argument");
    *out = nullptr;
    if (cfg->n_grid < 8 || cfg->n_grid > 16384) return set_error(MPM_ERR_INVALID, "n_grid out of range [8, 16384]");
    if (cfg->wall < 0 || 2 * cfg->wall + 3 >= cfg->n_grid) return set_error(MPM_ERR_INVALID, "wall band too wide");
    if (!(cfg->dt > 0) || !(cfg->E > 0) || !(cfg->rho > 0) || !(cfg->nu > -1.0 && cfg->nu < 0.5))
        return set_error(MPM_ERR_INVALID, "dt, E, rho must be positive and -1 < nu < 0.5");
    if (cfg->graph_steps < 0 || cfg->graph_steps > 10000) return set_error(MPM_ERR_INVALID, "graph_steps out of range");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(MPM_ERR_HIP, "no HIP device");
    mpm2d_sim *s = new mpm2d_sim();
    s->cfg = *cfg;
    P2 &p = s->p;
    p.n = cfg->n_grid;
    p.wall = cfg->wall;
    const double dx = 1.0 / cfg->n_grid;                  // :8
    const double p_vol = 0.25 * dx * dx;                  // (dx / 2)^2, :10
    p.dx = (float)dx;
    p.inv_dx = (float)cfg->n_grid;
    p.p_vol = (float)p_vol;
    p.p_mass = (float)(p_vol * cfg->rho);                 // :11
    p.mu0 = (float)(cfg->E / (2.0 * (1.0 + cfg->nu)));    // :13
    p.lam0 = (float)(cfg->E * cfg->nu / ((1.0 + cfg->nu) * (1.0 - 2.0 * cfg->nu)));
    p.xi = (float)cfg->xi;
    p.theta_c = (float)cfg->theta_c;
    p.theta_s = (float)cfg->theta_s;
    p.gx = (float)cfg->gravity[0];
    p.gy = (float)cfg->gravity[1];
    p.dt = (float)cfg->dt;
    const size_t gbytes = (size_t)p.n * p.n * sizeof(float4);
    hipError_t e;
    auto fail = [&](hipError_t err, const char *w) { mpm2d_destroy(s); return hip_fail(err, w); };
    if ((e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking)) != hipSuccess) return fail(e, "stream");
    if ((e = hipMalloc(&s->gmp, gbytes)) != hipSuccess) return fail(e, "hipMalloc");
    if ((e = hipMalloc(&s->gv, gbytes)) != hipSuccess) return fail(e, "hipMalloc");
    if ((e = hipMalloc(&s->health, 4)) != hipSuccess) return fail(e, "hipMalloc");
    const int nn = p.n * p.n;
    hipLaunchKernelGGL(k2_fill, dim3(blocks(nn, 256)), dim3(256), 0, s->stream, s->gmp, nn);
    hipLaunchKernelGGL(k2_fill, dim3(blocks(nn, 256)), dim3(256), 0, s->stream, s->gv, nn);
    if ((e = hipMemsetAsync(s->health, 0, 4, s->stream)) != hipSuccess) return fail(e, "memset");
    if ((e = hipStreamSynchronize(s->stream)) != hipSuccess) return fail(e, "sync");
    *out = s;
    return MPM_OK;
}

int mpm2d_add_particles(mpm2d_sim *s, int64_t n, const float *xy, const float *vel)
{
    if (!s || !xy) return set_error(MPM_ERR_INVALID, "null argument");
    if (s->np) return set_error(MPM_ERR_INVALID, "particles were already added");
    if (n <= 0 || n > (1ll << 30)) return set_error(MPM_ERR_INVALID, "particle count out of range");
    const float lo = (float)s->p.dx, hi = 1.0f - 2.0f * (float)s->p.dx;
    for (int64_t i = 0; i < 2 * n; ++i)
        if (!(xy[i] >= lo && xy[i] <= hi)) return set_error(MPM_ERR_RANGE, "particle outside [dx, 1 - 2 dx]");
    s->cap = n;
    MPM_HIP(hipMalloc(&s->P, (size_t)Q_COUNT * n * sizeof(float)));
    float *dxy = nullptr, *dv = nullptr;
    MPM_HIP(hipMalloc(&dxy, (size_t)n * 2 * sizeof(float)));
    if (vel) MPM_HIP(hipMalloc(&dv, (size_t)n * 2 * sizeof(float)));
    MPM_HIP(hipMemcpyAsync(dxy, xy, (size_t)n * 2 * sizeof(float), hipMemcpyHostToDevice, s->stream));
    if (vel) MPM_HIP(hipMemcpyAsync(dv, vel, (size_t)n * 2 * sizeof(float), hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(k2_init, dim3(blocks(n, 256)), dim3(256), 0, s->stream, s->P, s->cap, (int)n, dxy, dv);
    MPM_HIP(hipGetLastError());
    MPM_HIP(hipStreamSynchronize(s->stream));
    (void)hipFree(dxy);
    if (dv) (void)hipFree(dv);
    s->np = (int)n;
    return MPM_OK;
}

int mpm2d_run(mpm2d_sim *s, int32_t nsteps)
{
    if (!s) return set_error(MPM_ERR_INVALID, "sim is null");
    if (nsteps < 0) return set_error(MPM_ERR_INVALID, "nsteps must be >= 0");
    if (!s->np) return set_error(MPM_ERR_INVALID, "no particles");
    const int G = s->cfg.graph_steps;
    int left = nsteps;
    if (G > 0 && left >= G) {
        if (!s->graph) {
            MPM_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeRelaxed));
            int rc = enqueue_steps(s, G);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(s->stream, &g);
            if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
            if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
            e = hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) { s->graph = nullptr; return hip_fail(e, "hipGraphInstantiate"); }
            s->graph_len = G;
        }
        for (; left >= s->graph_len; left -= s->graph_len) MPM_HIP(hipGraphLaunch(s->graph, s->stream));
    }
    int rc = enqueue_steps(s, left);
    if (rc) return rc;
    s->steps += nsteps;
    return MPM_OK;
}

int mpm2d_read(mpm2d_sim *s, int32_t field, float *out)
{
    if (!s || !out) return set_error(MPM_ERR_INVALID, "null argument");
    int p0, w;
    switch (field) {
    case MPM2D_POS: p0 = Q_X; w = 2; break;
    case MPM2D_VEL: p0 = Q_V; w = 2; break;
    case MPM2D_C: p0 = Q_C; w = 4; break;
    case MPM2D_F: p0 = Q_F; w = 4; break;
    case MPM2D_JP: p0 = Q_J; w = 1; break;
    default: return set_error(MPM_ERR_INVALID, "unknown field");
    }
    std::vector<float> tmp((size_t)w * s->np);
    for (int k = 0; k < w; ++k)
        MPM_HIP(hipMemcpyAsync(tmp.data() + (size_t)k * s->np, s->P + (int64_t)(p0 + k) * s->cap,
                               (size_t)s->np * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    MPM_HIP(hipStreamSynchronize(s->stream));
    for (int64_t i = 0; i < s->np; ++i)
        for (int k = 0; k < w; ++k) out[i * w + k] = tmp[(size_t)k * s->np + i];
    return MPM_OK;
}

int mpm2d_read_grid(mpm2d_sim *s, float *mass, float *vel)
{
    if (!s || !mass || !vel) return set_error(MPM_ERR_INVALID, "null argument");
    const size_t nn = (size_t)s->p.n * s->p.n;
    std::vector<float4> g(nn);
    MPM_HIP(hipMemcpyAsync(g.data(), s->gv, nn * sizeof(float4), hipMemcpyDeviceToHost, s->stream));
    MPM_HIP(hipStreamSynchronize(s->stream));
    for (size_t k = 0; k < nn; ++k) {
        mass[k] = g[k].z;
        vel[2 * k] = g[k].x;
        vel[2 * k + 1] = g[k].y;
    }
    return MPM_OK;
}

int mpm2d_health(mpm2d_sim *s, uint32_t *bits)
{
    if (!s || !bits) return set_error(MPM_ERR_INVALID, "null argument");
    MPM_HIP(hipMemcpyAsync(bits, s->health, 4, hipMemcpyDeviceToHost, s->stream));
    MPM_HIP(hipStreamSynchronize(s->stream));
    return MPM_OK;
}

int64_t mpm2d_particles(const mpm2d_sim *s) { return s ? s->np : 0; }
int64_t mpm2d_steps(const mpm2d_sim *s) { return s ? s->steps : 0; }

int mpm2d_synchronize(mpm2d_sim *s)
{
    if (!s) return set_error(MPM_ERR_INVALID, "sim is null");
    MPM_HIP(hipStreamSynchronize(s->stream));
    return MPM_OK;
}

void mpm2d_destroy(mpm2d_sim *s)
{
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->graph) (void)hipGraphExecDestroy(s->graph);
    void *ptrs[] = {s->P, s->gmp, s->gv, s->health};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

}  // extern "C"
