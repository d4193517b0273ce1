// mpm_kernels.h -- per-particle device routines shared by the P2G / G2P
// kernels of both execution forms.
#pragma once

#include "mpm_internal.h"

namespace mpm {

constexpr float kFallbackRatio = 1e-4f;   // sigma_min / sigma_max below this -> FP64 SVD

// Particle i of a group's plane: the plane's first element of the group is
// a uniform address and i < 2^30 (the particle limit), so the byte offset is
// a 32-bit lane value and a load takes the scalar-base form (no 64-bit
// address arithmetic per load).
__device__ __forceinline__ MPM_GLOBAL float *plane_at(const PartView &pv, int plane, uint32_t i)
{
    MPM_GLOBAL char *base = (MPM_GLOBAL char *)gp(pv.P + (int64_t)plane * pv.cap + pv.off);
    return (MPM_GLOBAL float *)(base + (i << 2));
}
__device__ __forceinline__ float ld(const PartView &pv, int plane, int i)
{
    return gp(pv.P)[(int64_t)plane * pv.cap + pv.off + i];
}
__device__ __forceinline__ void st(const PartView &pv, int plane, int i, float v)
{
    gp(pv.P)[(int64_t)plane * pv.cap + pv.off + i] = v;
}

// Stencil of one particle: base node (clamped so the 3x3x3 support stays in
// the grid) and per-axis B-spline weights / derivatives.
struct Stencil {
    int b[3];
    float fx[3];
    float w[3][3], dw[3][3];
};

// Flag the (up to 2x2x2) 4^3-node blocks the 3x3x3 stencil of base node b
// touches.
__device__ __forceinline__ void mark_blocks(const GridP &g, const int b[3], uint32_t *flags)
{
    int bx0 = b[0] >> 2, bx1 = (b[0] + 2) >> 2;
    int by0 = b[1] >> 2, by1 = (b[1] + 2) >> 2;
    int bz0 = b[2] >> 2, bz1 = (b[2] + 2) >> 2;
    for (int bz = bz0; bz <= bz1; ++bz)
        for (int by = by0; by <= by1; ++by)
            for (int bx = bx0; bx <= bx1; ++bx)
                gp(flags)[((int64_t)bz * g.nby + by) * g.nbx + bx] = 1u;
}

// The three axes' weights and derivatives: the closed form of
// bspline_axis_fast when every active lane's fx lies in [0.5, 1.5] (no base
// node clamped: the usual case), chosen once for the wave, so the common
// path carries none of the general form's per-lane selects; per lane
// otherwise.  Same values either way.
__device__ __forceinline__ void bspline3_wave(const float fx[3], float w[3][3], float dw[3][3])
{
    bool in = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) in = in && fx[a] >= 0.5f && fx[a] <= 1.5f;
    if (__all(in)) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float a0 = 1.5f - fx[a], d1 = fx[a] - 1.0f, a2 = fx[a] - 0.5f;
            w[a][0] = 0.5f * a0 * a0;
            w[a][1] = 0.75f - d1 * d1;
            w[a][2] = 0.5f * a2 * a2;
            dw[a][0] = -a0;
            dw[a][1] = -2.0f * d1;
            dw[a][2] = a2;
        }
    } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) bspline_axis_fast<float>(fx[a], w[a], dw[a]);
    }
}

// Base node (clamped so the 3x3x3 support stays in the grid) and the
// position relative to it in cells, fx = (x - o) * inv_h - b (one rounding).
__device__ __forceinline__ uint32_t base_fx(const GridP &g, const float x[3], int b[3], float fx[3])
{
    uint32_t bits = 0;
    const float o[3] = {g.ox, g.oy, g.oz};
    const int n[3] = {g.nx, g.ny, g.nz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int v = base_index_f32(x[a], o[a], g.inv_h);
        if (v > n[a] - 3) { v = n[a] - 3; bits |= MPM_H_STENCIL_OOB; }
        b[a] = v;
        fx[a] = fmaf(x[a] - o[a], g.inv_h, -(float)v);
    }
    return bits;
}

__device__ __forceinline__ uint32_t make_stencil(const GridP &g, const float x[3], Stencil &s)
{
    const uint32_t bits = base_fx(g, x, s.b, s.fx);
#pragma unroll
    for (int a = 0; a < 3; ++a) bspline_axis_fast<float>(s.fx[a], s.w[a], s.dw[a]);
    return bits;
}

// The base node alone (as make_stencil computes it, same clamp and flag).
__device__ __forceinline__ uint32_t base_only(const GridP &g, const float x[3], int b[3])
{
    uint32_t bits = 0;
    const float o[3] = {g.ox, g.oy, g.oz};
    const int n[3] = {g.nx, g.ny, g.nz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int v = base_index_f32(x[a], o[a], g.inv_h);
        if (v > n[a] - 3) { v = n[a] - 3; bits |= MPM_H_STENCIL_OOB; }
        b[a] = v;
    }
    return bits;
}

// FP32 SVD with the FP64 fallback for ill-conditioned / inverted F.
__device__ __forceinline__ void svd_checked(const float F[9], float U[9], float s[3], float V[9],
                                            int sweeps, uint32_t *counters)
{
    svd3<float>(F, U, s, V, sweeps);
    if (!(s[2] > kFallbackRatio * s[0])) {
        double Fd[9], Ud[9], sd[3], Vd[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) Fd[k] = F[k];
        svd3<double>(Fd, Ud, sd, Vd, 8);
#pragma unroll
        for (int k = 0; k < 9; ++k) { U[k] = (float)Ud[k]; V[k] = (float)Vd[k]; }
        s[0] = (float)sd[0]; s[1] = (float)sd[1]; s[2] = (float)sd[2];
        gatomic_add(counters + CT_FALLBACK, 1u);
    }
}

// SVD started from the warm-start quaternion stored for particle i: V0 =
// R(q / |q|), `warm_sweeps` Jacobi sweeps.  The result is accepted when the
// columns of F V are orthogonal to 1e-6 (largest |cos| between columns) and
// F is well conditioned; otherwise the cold SVD (with its FP64 fallback)
// replaces it and the rejection is counted.  A particle with no quaternion
// (all zero: not inferred since its planes were cleared) takes the cold SVD
// directly; the attempts with one are counted in CT_WARM.
constexpr float kWarmTol = 1e-6f;

__device__ __forceinline__ void svd_warm(const PartView &pv, int i, const float F[9], float U[9], float s[3],
                                         float V[9], int warm_sweeps, int sweeps, uint32_t *counters)
{
    float q[4] = {ld(pv, PL_Q, i), ld(pv, PL_Q + 1, i), ld(pv, PL_Q + 2, i), ld(pv, PL_Q + 3, i)};
    const float qn2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    const bool has_q = qn2 > 0.0f;
    wave_count(counters + CT_WARM, has_q);
    bool ok = false;
    if (has_q && qn2 > 1e-20f && qn2 < 1e20f) {
        const float inv = r_rsqrt(qn2);
#pragma unroll
        for (int k = 0; k < 4; ++k) q[k] *= inv;
        float V0[9];
        quat_to_mat<float>(q, V0);
        const float off = svd3<float>(F, U, s, V, warm_sweeps, V0);
        ok = off <= kWarmTol && s[2] > kFallbackRatio * s[0];
    }
    if (ok) return;
    wave_count(counters + CT_REJECT, has_q);
    svd_checked(F, U, s, V, sweeps, counters);
}

// The Jacobi SVD of a G2P update (particle i of pv, written to slot io):
// counted in CT_SVD, warm-started when the material's warm_sweeps > 0, and
// the particle listed (by its new slot) for the next step's inference when
// the material carries a list.
__device__ __forceinline__ void svd_g2p(const MatP &m, const PartView &pv, int i, int io, const float F[9],
                                        float U[9], float s[3], float V[9], int sweeps, uint32_t *counters)
{
    wave_count(counters + CT_SVD, true);
    if (m.warm_sweeps > 0) svd_warm(pv, i, F, U, s, V, m.warm_sweeps, sweeps, counters);
    else svd_checked(F, U, s, V, sweeps, counters);
    if (m.wlist) wave_append(m.wlist, m.wcount, (uint32_t)io);
}

// Kirchhoff stress of an elastic / snow particle from its deformation
// gradient (reference P2G, MPMCUDA.cu:329-349).
template <int MAT>
__device__ __forceinline__ void stress_from_F(const MatP &m, const float F[9], float Jp, int sweeps,
                                              float tau6[6], uint32_t *counters, const PartView *warm_pv = nullptr,
                                              int warm_i = 0, int warm_io = 0)
{
    float mu = m.mu, lam = m.lam;
    if (MAT == MPM_SNOW) {
        float hc = expf(m.xi * (1.0f - Jp));           // MPMCUDA.cu:345
        mu *= hc; lam *= hc;
    }
    // F within a factor 2 of a rotation (singular values in (0.5, 2)): the
    // stress needs only the polar rotation; five Newton steps reach FP32
    // there (a singular value of 2 or 1/2 goes 1.25, 1.025, 1 + 3e-4, 1 + 5e-8)
    if (sv_in_band<float>(F, 0.5f, 2.0f)) {
        float R[9];
        polar_newton<float>(F, R, 5);
        fcr_tau_polar<float>(F, R, mu, lam, tau6);
        return;
    }
    float U[9], s[3], V[9];
    if (warm_pv) svd_g2p(m, *warm_pv, warm_i, warm_io, F, U, s, V, sweeps, counters);
    else svd_checked(F, U, s, V, sweeps, counters);
    fcr_tau_from_svd<float>(U, s, mu, lam, tau6);
}

// The per-particle quantities P2G scatters: node momentum
//   p_i += w (a + B (x_i - x_p)) + G grad w_ip ,   node mass m_i += w m_p,
// with a = m_p v_p, B = m_p C_p (MPMCUDA.cu:391-401) and G the dt-scaled
// force matrix: -dt V_p tau (MPMCUDA.cu:409-411) or dt V_p p J I for water
// (MPMCUDA.cu:412-414).
struct Scatter {
    float m;
    float a[3];
    float B[9];
    float G[9];
};

template <int MAT>
__device__ __forceinline__ void load_scatter(const PartView &pv, int i, const MatP &mat, float dt,
                                             bool tau_stored, int sweeps, Scatter &sc,
                                             uint32_t *counters)
{
    sc.m = mat.mass;
#pragma unroll
    for (int r = 0; r < 3; ++r) sc.a[r] = mat.mass * ld(pv, PL_V + r, i);
#pragma unroll
    for (int k = 0; k < 9; ++k) sc.B[k] = mat.mass * ld(pv, PL_C + k, i);
    if (MAT == MPM_WATER) {
        float J = ld(pv, PL_S, i);
        float pr = mat.eos_k * (powf(J, -mat.eos_gamma) - 1.0f);   // MPMCUDA.cu:353
        float c = dt * mat.vol * pr * J;
#pragma unroll
        for (int k = 0; k < 9; ++k) sc.G[k] = (k % 4 == 0) ? c : 0.0f;
    } else {
        float t[6];
        if (tau_stored) {
#pragma unroll
            for (int k = 0; k < 6; ++k) t[k] = ld(pv, PL_TAU + k, i);
        } else {
            float F[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) F[k] = ld(pv, PL_F + k, i);
            float Jp = (MAT == MPM_SNOW) ? ld(pv, PL_S, i) : 1.0f;
            stress_from_F<MAT>(mat, F, Jp, sweeps, t, counters);
        }
        float c = -dt * mat.vol;
        sc.G[0] = c * t[0]; sc.G[4] = c * t[1]; sc.G[8] = c * t[2];
        sc.G[1] = sc.G[3] = c * t[3];
        sc.G[2] = sc.G[6] = c * t[4];
        sc.G[5] = sc.G[7] = c * t[5];
    }
}

// The FAST-form P2G inputs of one particle besides its position: v, C and
// the stored stress (water: J in s[0]).
struct RawP {
    float v[3], C[9], s[6];
};
// Plane p of particle i as (uniform plane base) + (32-bit lane offset): the
// loads then take the scalar-base form and share one offset register.
__device__ __forceinline__ float ld_plane(const PartView &pv, int plane, uint32_t i)
{
    return *plane_at(pv, plane, i);
}

// Contribution of one particle to stencil node (ox, oy, oz): out = (m, p).
__device__ __forceinline__ float4 node_contrib(const GridP &g, const Stencil &s, const Scatter &sc,
                                               int ox, int oy, int oz)
{
    float wx = s.w[0][ox], wy = s.w[1][oy], wz = s.w[2][oz];
    float w = wx * wy * wz;
    float gr[3] = {g.inv_h * s.dw[0][ox] * wy * wz, g.inv_h * wx * s.dw[1][oy] * wz,
                   g.inv_h * wx * wy * s.dw[2][oz]};
    float d[3] = {((float)ox - s.fx[0]) * g.h, ((float)oy - s.fx[1]) * g.h, ((float)oz - s.fx[2]) * g.h};
    float4 r;
    r.x = w * sc.m;
    float p[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float bd = sc.B[3 * c] * d[0] + sc.B[3 * c + 1] * d[1] + sc.B[3 * c + 2] * d[2];
        float gg = sc.G[3 * c] * gr[0] + sc.G[3 * c + 1] * gr[1] + sc.G[3 * c + 2] * gr[2];
        p[c] = w * (sc.a[c] + bd) + gg;
    }
    r.y = p[0]; r.z = p[1]; r.w = p[2];
    return r;
}

}  // namespace mpm
