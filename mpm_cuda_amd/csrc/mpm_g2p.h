// mpm_g2p.h -- G2P core shared by the stand-alone G2P kernel and the fused
// G2P2G kernel: interpolation of grid velocities, affine C, velocity
// gradient, deformation / plasticity / J update, advection (reference
// InterpolateAndMove, MPMCUDA.cu:474-741).  The grid-velocity source is a
// template parameter (global grid_v or a block's LDS copy).
#pragma once

#include "mpm_kernels.h"

namespace mpm {

// LDS slot of node (x, y, z) of an 8^3 block arena (float4 per node, dense
// 8 KB).  x is XOR-swizzled with 4 on odd y: the 8 lanes served per LDS cycle
// (cells x = 0..3, y = 0..1 of a block, shifted by the same stencil offset)
// then fall on 8 different 16-byte bank groups.
constexpr int A_SLOTS = ARENA_NODES;
__device__ __forceinline__ int a_slot(int x, int y, int z) { return (x ^ ((y & 1) << 2)) + ARENA * (y + ARENA * z); }

struct GlobalFetch {
    const float4 *gv;
    __device__ __forceinline__ float4 operator()(const GridP &g, int x, int y, int z) const
    {
        return gv[node_slot(g, x, y, z)];
    }
};

// Velocities of the block arena [a, a+8)^3 from LDS, other nodes from HBM.
struct ArenaFetch {
    const float4 *lds;
    const float4 *gv;
    int ax, ay, az;
    __device__ __forceinline__ float4 operator()(const GridP &g, int x, int y, int z) const
    {
        const unsigned lx = (unsigned)(x - ax), ly = (unsigned)(y - ay), lz = (unsigned)(z - az);
        if (lx < (unsigned)ARENA && ly < (unsigned)ARENA && lz < (unsigned)ARENA) return lds[a_slot(lx, ly, lz)];
        return gv[node_slot(g, x, y, z)];
    }
};

// Updated particle state kept in registers for a following scatter.
struct G2POut {
    float x[3], v[3], C[9], tau[6], S;
};

template <int MAT, class Fetch>
__device__ __forceinline__ uint32_t g2p_core(const GridP &g, const MatP &mat, const PartView &pv, int i, float dt,
                                             const Fetch &fetch, bool want_tau, int sweeps, uint32_t *counters,
                                             G2POut &out)
{
    float x[3] = {ld(pv, PL_X, i), ld(pv, PL_X + 1, i), ld(pv, PL_X + 2, i)};
    Stencil s;
    uint32_t bits = make_stencil(g, x, s);
    float v[3] = {0.f, 0.f, 0.f}, C[9], L[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) { C[k] = 0.f; L[k] = 0.f; }
    const float c4 = 4.0f * g.inv_h;       // 4 w / h^2 (x_i - x_p), with x_i - x_p = (o - fx) h
#pragma unroll
    for (int ox = 0; ox < 3; ++ox)
#pragma unroll
        for (int oy = 0; oy < 3; ++oy)
#pragma unroll
            for (int oz = 0; oz < 3; ++oz) {
                float4 nv = fetch(g, s.b[0] + ox, s.b[1] + oy, s.b[2] + oz);
                float wx = s.w[0][ox], wy = s.w[1][oy], wz = s.w[2][oz];
                float w = wx * wy * wz;
                float gr[3] = {g.inv_h * s.dw[0][ox] * wy * wz, g.inv_h * wx * s.dw[1][oy] * wz,
                               g.inv_h * wx * wy * s.dw[2][oz]};
                float dd[3] = {(float)ox - s.fx[0], (float)oy - s.fx[1], (float)oz - s.fx[2]};
                float vi[3] = {nv.x, nv.y, nv.z};
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    v[r] += w * vi[r];                                   // :516-518
                    float cw = c4 * w * vi[r];
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        C[3 * r + c] += cw * dd[c];                      // :520-532
                        L[3 * r + c] += vi[r] * gr[c];                   // :561-585
                    }
                }
            }
#pragma unroll
    for (int r = 0; r < 3; ++r) { st(pv, PL_V + r, i, v[r]); out.v[r] = v[r]; }
    const bool c_is_gradv = (MAT == MPM_WATER && mat.water_c_mode == MPM_WATER_C_GRADV);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        out.C[k] = c_is_gradv ? L[k] : C[k];
        st(pv, PL_C + k, i, out.C[k]);
    }
    float A[9];                                                          // I + dt grad v
#pragma unroll
    for (int k = 0; k < 9; ++k) A[k] = dt * L[k] + ((k % 4 == 0) ? 1.0f : 0.0f);
    if (MAT == MPM_JELLO) {                                              // :596-619
        float F[9], Fn[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) F[k] = ld(pv, PL_F + k, i);
        matmul3<float>(A, F, Fn);
#pragma unroll
        for (int k = 0; k < 9; ++k) st(pv, PL_F + k, i, Fn[k]);
        if (!(det3<float>(Fn) > 0.f)) bits |= MPM_H_DET_NONPOS;
        if (want_tau) stress_from_F<MPM_JELLO>(mat, Fn, 1.0f, sweeps, out.tau, counters);
        out.S = 1.0f;
    } else if (MAT == MPM_SNOW) {                                        // :620-709
        float F[9], Ft[9], U[9], sg[3], V[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) F[k] = ld(pv, PL_F + k, i);
        matmul3<float>(A, F, Ft);
        svd_checked(Ft, U, sg, V, sweeps, counters);
        float ratio = 1.0f;
        float sc[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            sc[k] = fminf(fmaxf(sg[k], 1.0f - mat.theta_c), 1.0f + mat.theta_s);   // :658-660
            ratio *= sg[k] / sc[k];
        }
        // Fe = U S* V^T, written as Ftr + U (S* - S) V^T: identical in exact
        // arithmetic, and exactly Ftr when no singular value is clamped, so
        // the FP32 SVD's rounding does not leak into F every step.
        float ds[3] = {sc[0] - sg[0], sc[1] - sg[1], sc[2] - sg[2]};
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                st(pv, PL_F + 3 * r + c, i,
                   Ft[3 * r + c] + (U[3 * r] * ds[0] * V[3 * c] + U[3 * r + 1] * ds[1] * V[3 * c + 1] +
                                    U[3 * r + 2] * ds[2] * V[3 * c + 2]));
        // det Fp_new = det(Fe_new^-1 Ftr Fp) = Jp * prod(s / s*)  (:671-678)
        float Jp = ld(pv, PL_S, i) * ratio;
        st(pv, PL_S, i, Jp);
        if (!(Jp > 0.f)) bits |= MPM_H_DET_NONPOS;
        if (want_tau) {
            float hc = expf(mat.xi * (1.0f - Jp));
            fcr_tau_from_svd<float>(U, sc, mat.mu * hc, mat.lam * hc, out.tau);
        }
        out.S = Jp;
    } else {                                                             // :710-713
        float J = (1.0f + dt * (L[0] + L[4] + L[8])) * ld(pv, PL_S, i);
        st(pv, PL_S, i, J);
        out.S = J;
    }
    const float lo[3] = {g.ox, g.oy, g.oz};
    const int nn[3] = {g.nx, g.ny, g.nz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float xn = x[a] + dt * v[a];                                     // :717-719
        st(pv, PL_X + a, i, xn);
        out.x[a] = xn;
        float hi = lo[a] + g.h * (float)(nn[a] - 1);
        if (!(xn >= lo[a] && xn <= hi)) bits |= MPM_H_LEFT_DOMAIN;       // :722-739
        if (!isfinite(xn) || !isfinite(v[a])) bits |= MPM_H_NONFINITE;
    }
    return bits;
}

// P2G quantities of an updated particle held in registers (see load_scatter).
template <int MAT>
__device__ __forceinline__ void scatter_from_out(const MatP &mat, const G2POut &o, float dt, Scatter &sc)
{
    sc.m = mat.mass;
#pragma unroll
    for (int r = 0; r < 3; ++r) sc.a[r] = mat.mass * o.v[r];
#pragma unroll
    for (int k = 0; k < 9; ++k) sc.B[k] = mat.mass * o.C[k];
    if (MAT == MPM_WATER) {
        float pr = mat.eos_k * (powf(o.S, -mat.eos_gamma) - 1.0f);      // MPMCUDA.cu:353
        float c = dt * mat.vol * pr * o.S;
#pragma unroll
        for (int k = 0; k < 9; ++k) sc.G[k] = (k % 4 == 0) ? c : 0.0f;
    } else {
        float c = -dt * mat.vol;
        sc.G[0] = c * o.tau[0]; sc.G[4] = c * o.tau[1]; sc.G[8] = c * o.tau[2];
        sc.G[1] = sc.G[3] = c * o.tau[3];
        sc.G[2] = sc.G[6] = c * o.tau[4];
        sc.G[5] = sc.G[7] = c * o.tau[5];
    }
}

}  // namespace mpm
