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
    static constexpr bool kArena = false, kGlobal = true;
    const float4 *gv;
    __device__ __forceinline__ float4 operator()(const GridP &g, int x, int y, int z) const
    {
        return gv[node_slot(g, x, y, z)];
    }
};

// Block-major slots are separable: slot(x, y, z) = X(x) + Y(y) + Z(z) with
// X(x) = 64 (x >> 2) + (x & 3), Y(y) = 64 nbx (y >> 2) + 4 (y & 3),
// Z(z) = 64 nbx nby (z >> 2) + 16 (z & 3) (node_slot expanded), so a
// stencil's 27 slots come from 9 per-axis terms and 27 adds (32-bit: a grid
// holds fewer than 2^32 node slots).
struct StencilSlots {
    uint32_t xy[9];     // X(b0 + ox) + Y(b1 + oy), index ox + 3 oy
    uint32_t z[3];      // Z(b2 + oz)
    __device__ __forceinline__ void init(const GridP &g, const int b[3])
    {
        uint32_t X[3], Y[3];
        const uint32_t ry = 64u * (uint32_t)g.nbx, rz = ry * (uint32_t)g.nby;
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            const uint32_t x = (uint32_t)(b[0] + o), y = (uint32_t)(b[1] + o), zz = (uint32_t)(b[2] + o);
            X[o] = 64u * (x >> 2) + (x & 3u);
            Y[o] = ry * (y >> 2) + 4u * (y & 3u);
            z[o] = rz * (zz >> 2) + 16u * (zz & 3u);
        }
#pragma unroll
        for (int oy = 0; oy < 3; ++oy)
#pragma unroll
            for (int ox = 0; ox < 3; ++ox) xy[ox + 3 * oy] = X[ox] + Y[oy];
    }
};

// Velocities of the block arena [a, a+8)^3 from LDS, other nodes from HBM.
struct ArenaFetch {
    static constexpr bool kArena = true, kGlobal = false;
    const float4 *lds;
    const float4 *gv;
    int ax, ay, az;
    __device__ __forceinline__ float4 operator()(const GridP &g, int x, int y, int z) const
    {
        const unsigned lx = (unsigned)(x - ax), ly = (unsigned)(y - ay), lz = (unsigned)(z - az);
        if (lx < (unsigned)ARENA && ly < (unsigned)ARENA && lz < (unsigned)ARENA) return lload4(lds, a_slot(lx, ly, lz));
        return gload4(gv, node_slot(g, x, y, z));
    }
    // all 27 stencil nodes of base b lie in the arena
    __device__ __forceinline__ bool covers(const int b[3]) const
    {
        return (unsigned)(b[0] - ax) <= (unsigned)(ARENA - 3) && (unsigned)(b[1] - ay) <= (unsigned)(ARENA - 3) &&
               (unsigned)(b[2] - az) <= (unsigned)(ARENA - 3);
    }
};

// The arena only, for stencils known to lie inside it: no per-node test,
// so the 27 LDS reads issue back to back.
struct ArenaOnlyFetch {
    static constexpr bool kArena = true, kGlobal = false;
    const float4 *lds;
    int ax, ay, az;
    __device__ __forceinline__ float4 operator()(const GridP &, int x, int y, int z) const
    {
        return lload4(lds, a_slot(x - ax, y - ay, z - az));
    }
};

// Interpolation (MPMCUDA.cu:516-532, 561-585), summed axis by axis:
//   v_r  = sum_o w_o v_or
//   M_rc = sum_o w_o o_c v_or          (C_rc = 4/h (M_rc - fx_c v_r))
//   L_rc = sum_o v_or dw_o/dx_c
// with w_o = Nx Ny Nz and D = dN/dx per axis; the inner sums over oz and oy
// are shared by all three outputs.
template <class Fetch>
__device__ __forceinline__ void gather_grid(const GridP &g, const Stencil &s, const Fetch &fetch, float v[3],
                                            float M[9], float L[9])
{
    float N[3][3], D[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int k = 0; k < 3; ++k) { N[a][k] = s.w[a][k]; D[a][k] = g.inv_h * s.dw[a][k]; }
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) v[r] = -0.0f;     // -0 + x == x: the first add folds away
#pragma unroll
    for (int k = 0; k < 9; ++k) { M[k] = -0.0f; L[k] = -0.0f; }
    StencilSlots sl;
    if constexpr (Fetch::kGlobal) sl.init(g, s.b);
#pragma unroll
    for (int ox = 0; ox < 3; ++ox) {
        float Y0[3] = {-0.f, -0.f, -0.f}, Y1[3] = {-0.f, -0.f, -0.f}, Y2[3] = {-0.f, -0.f, -0.f};
        float Y3[3] = {-0.f, -0.f, -0.f}, Y4[3] = {-0.f, -0.f, -0.f};
#pragma unroll
        for (int oy = 0; oy < 3; ++oy) {
            float S0[3] = {-0.f, -0.f, -0.f}, S1[3] = {-0.f, -0.f, -0.f}, S2[3] = {-0.f, -0.f, -0.f};
#pragma unroll
            for (int oz = 0; oz < 3; ++oz) {
                float4 nv;
                if constexpr (Fetch::kGlobal) nv = fetch.gv[sl.xy[ox + 3 * oy] + sl.z[oz]];
                else nv = fetch(g, s.b[0] + ox, s.b[1] + oy, s.b[2] + oz);
                const float vi[3] = {nv.x, nv.y, nv.z};
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    S0[r] += N[2][oz] * vi[r];
                    S1[r] += D[2][oz] * vi[r];
                    if (oz) S2[r] += ((float)oz * N[2][oz]) * vi[r];
                }
            }
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                Y0[r] += N[1][oy] * S0[r];
                Y1[r] += D[1][oy] * S0[r];
                if (oy) Y2[r] += ((float)oy * N[1][oy]) * S0[r];
                Y3[r] += N[1][oy] * S2[r];
                Y4[r] += N[1][oy] * S1[r];
            }
        }
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            v[r] += N[0][ox] * Y0[r];
            if (ox) M[3 * r] += ((float)ox * N[0][ox]) * Y0[r];
            M[3 * r + 1] += N[0][ox] * Y2[r];
            M[3 * r + 2] += N[0][ox] * Y3[r];
            L[3 * r] += D[0][ox] * Y0[r];
            L[3 * r + 1] += N[0][ox] * Y1[r];
            L[3 * r + 2] += N[0][ox] * Y4[r];
        }
    }
}

// Updated particle state kept in registers for a following scatter.
struct G2POut {
    float x[3], v[3], C[9], tau[6];
    float S;            // snow J_p, water J, jello det F
};

// Step-size controller inputs of one particle: |v| + c, with c the elastic
// wave speed at its deformation (snow: hardened by exp(xi (1 - J_p)); water:
// bulk modulus k gamma J^-gamma at density rho / J), and max(J, 1/J), the
// measure of the reference's unused factor (MPMCUDA.cu:613-619, 700-708;
// jello det F, snow det F_p; the reference never sets it for water, :710-713).
template <int MAT>
__device__ __forceinline__ void dt_measures(const MatP &m, const float v[3], float J, float &speed, float &deform)
{
    float c2 = m.c2;
    if (MAT == MPM_SNOW) c2 *= expf(m.xi * (1.0f - J));
    if (MAT == MPM_WATER) c2 *= powf(J, 1.0f - m.eos_gamma);
    speed = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) + sqrtf(fmaxf(c2, 0.0f));
    deform = MAT == MPM_WATER ? 0.0f : (J > 0.0f ? fmaxf(J, 1.0f / J) : 1e30f);
}

// Particle i of pv is updated into slot io of po: the same slot, or (the G2P
// that applies a re-sort) its slot in the new order in the other buffer,
// where every plane the state needs is written.
template <int MAT, class Fetch>
__device__ __forceinline__ uint32_t g2p_core(const GridP &g, const MatP &mat, const PartView &pv, int i,
                                             const PartView &po, int io, float dt, const Fetch &fetch, bool want_tau,
                                             int sweeps, uint32_t *counters, G2POut &out)
{
    float x[3] = {ld(pv, PL_X, i), ld(pv, PL_X + 1, i), ld(pv, PL_X + 2, i)};
    // every particle load is issued up front (one memory latency, not one
    // per load when the kernel runs near its register cap)
    float F[9], Sp = 1.0f;
    if (MAT != MPM_WATER) {
#pragma unroll
        for (int k = 0; k < 9; ++k) F[k] = ld(pv, PL_F + k, i);
    }
    if (MAT != MPM_JELLO) Sp = ld(pv, PL_S, i);
    __builtin_amdgcn_sched_group_barrier(0x0020, MAT == MPM_WATER ? 4 : (MAT == MPM_JELLO ? 12 : 13), 0);
    Stencil s;
    uint32_t bits = make_stencil(g, x, s);
    float v[3], M[9], L[9];
    if constexpr (Fetch::kArena) {
        // a particle scattered into this arena last step gathers from it
        // alone; one outside it falls back to per-node tests
        if (__all(fetch.covers(s.b))) gather_grid(g, s, ArenaOnlyFetch{fetch.lds, fetch.ax, fetch.ay, fetch.az}, v, M, L);
        else gather_grid(g, s, fetch, v, M, L);
    } else {
        gather_grid(g, s, fetch, v, M, L);
    }
    const float c4 = 4.0f * g.inv_h;
    float C[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) C[3 * r + c] = c4 * (M[3 * r + c] - s.fx[c] * v[r]);
#pragma unroll
    for (int r = 0; r < 3; ++r) { st(po, PL_V + r, io, v[r]); out.v[r] = v[r]; }
    const bool c_is_gradv = (MAT == MPM_WATER && mat.water_c_mode == MPM_WATER_C_GRADV);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        out.C[k] = c_is_gradv ? L[k] : C[k];
        st(po, PL_C + k, io, out.C[k]);
    }
    float A[9];                                                          // I + dt grad v
#pragma unroll
    for (int k = 0; k < 9; ++k) A[k] = dt * L[k] + ((k % 4 == 0) ? 1.0f : 0.0f);
    if (MAT == MPM_JELLO) {                                              // :596-619
        float Fn[9];
        matmul3<float>(A, F, Fn);
#pragma unroll
        for (int k = 0; k < 9; ++k) st(po, PL_F + k, io, Fn[k]);
        const float Jn = det3<float>(Fn);
        if (!(Jn > 0.f)) bits |= MPM_H_DET_NONPOS;
        if (want_tau) stress_from_F<MPM_JELLO>(mat, Fn, 1.0f, sweeps, out.tau, counters, &pv, i, io);
        out.S = Jn;
    } else if (MAT == MPM_SNOW) {                                        // :620-709
        float Ft[9];
        matmul3<float>(A, F, Ft);
        if (sv_in_band<float>(Ft, 1.0f - mat.theta_c, 1.0f + mat.theta_s)) {
            // no singular value reaches the clamp: F_e = F_tr, J_p unchanged
            // (its plane is left as it is); the stress needs only the polar
            // rotation, a few Newton steps instead of the SVD
#pragma unroll
            for (int k = 0; k < 9; ++k) st(po, PL_F + k, io, Ft[k]);
            if (po.P != pv.P) st(po, PL_S, io, Sp);        // moved: the plane goes along
            if (!(Sp > 0.f)) bits |= MPM_H_DET_NONPOS;
            if (want_tau) {
                float R[9];
                polar_newton<float>(Ft, R, mat.polar_steps);
                const float hc = expf(mat.xi * (1.0f - Sp));
                fcr_tau_polar<float>(Ft, R, mat.mu * hc, mat.lam * hc, out.tau);
            }
            out.S = Sp;
        } else {
            float U[9], sg[3], V[9];
            svd_g2p(mat, pv, i, io, Ft, U, sg, V, sweeps, counters);
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
                    st(po, PL_F + 3 * r + c, io,
                       Ft[3 * r + c] + (U[3 * r] * ds[0] * V[3 * c] + U[3 * r + 1] * ds[1] * V[3 * c + 1] +
                                        U[3 * r + 2] * ds[2] * V[3 * c + 2]));
            // det Fp_new = det(Fe_new^-1 Ftr Fp) = Jp * prod(s / s*)  (:671-678)
            float Jp = Sp * ratio;
            st(po, PL_S, io, Jp);
            if (!(Jp > 0.f)) bits |= MPM_H_DET_NONPOS;
            if (want_tau) {
                float hc = expf(mat.xi * (1.0f - Jp));
                fcr_tau_from_svd<float>(U, sc, mat.mu * hc, mat.lam * hc, out.tau);
            }
            out.S = Jp;
        }
    } else {                                                             // :710-713
        float J = (1.0f + dt * (L[0] + L[4] + L[8])) * Sp;
        st(po, PL_S, io, J);
        out.S = J;
    }
    const float lo[3] = {g.ox, g.oy, g.oz};
    const int nn[3] = {g.nx, g.ny, g.nz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float xn = x[a] + dt * v[a];                                     // :717-719
        st(po, PL_X + a, io, xn);
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
