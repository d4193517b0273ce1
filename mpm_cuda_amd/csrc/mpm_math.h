// mpm_math.h -- per-particle math of the MPM step, written once for FP32
// (the hot path) and FP64 (the SVD fallback).  Compiles as HIP device code
// and, with MPM_HD empty, as plain host C++ for CPU studies of the FP32
// algorithms.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MPM_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define MPM_HD inline
#endif

namespace mpm {

template <class T> MPM_HD T m_sqrt(T x) { return sqrt(x); }
template <> MPM_HD float m_sqrt<float>(float x) { return sqrtf(x); }
template <class T> MPM_HD T m_abs(T x) { return x < T(0) ? -x : x; }

// Reciprocal, reciprocal square root and square root for the rotations of
// the SVD, where a 1-ulp result is enough (the rotation stays orthogonal to
// working precision).  On the device the FP32 forms are the single hardware
// instructions (v_rcp_f32, v_rsq_f32, v_sqrt_f32) instead of the correctly
// rounded multi-instruction sequences; FP64 and host builds divide exactly.
template <class T> MPM_HD T r_rcp(T x) { return T(1) / x; }
template <class T> MPM_HD T r_rsqrt(T x) { return T(1) / m_sqrt(x); }
template <class T> MPM_HD T r_sqrt(T x) { return m_sqrt(x); }
#if defined(__HIP_DEVICE_COMPILE__)
template <> MPM_HD float r_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <> MPM_HD float r_rsqrt<float>(float x) { return __builtin_amdgcn_rsqf(x); }
template <> MPM_HD float r_sqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
#endif

// ---------------------------------------------------------------------------
// Quadratic B-spline N and N' (reference MPMCUDA.cu:267-290, branchy |x| form
// with the same half-open intervals), evaluated at the three stencil offsets
// of one axis.  fx is the particle coordinate relative to the base node in
// cell units, d_o = fx - o for o = 0, 1, 2.
template <class T>
MPM_HD void bspline_axis(T fx, T w[3], T dw[3])
{
#pragma unroll
    for (int o = 0; o < 3; ++o) {
        T d = fx - T(o);
        T a = m_abs(d);
        T inner = T(0.75) - a * a;
        T outer = T(0.5) * (T(1.5) - a) * (T(1.5) - a);
        w[o] = a < T(0.5) ? inner : (a < T(1.5) ? outer : T(0));
        T g = T(0);
        if (d > T(-0.5) && d < T(0.5)) g = T(-2) * d;
        else if (d >= T(0.5) && d < T(1.5)) g = d - T(1.5);
        else if (d > T(-1.5) && d <= T(-0.5)) g = T(1.5) + d;
        dw[o] = g;
    }
}

// bspline_axis for the usual case fx in [0.5, 1.5] (base node not clamped)
// in closed form without branches; any other fx takes the general form.
template <class T>
MPM_HD void bspline_axis_fast(T fx, T w[3], T dw[3])
{
    if (fx >= T(0.5) && fx <= T(1.5)) {
        const T a0 = T(1.5) - fx, d1 = fx - T(1), a2 = fx - T(0.5);
        w[0] = T(0.5) * a0 * a0;
        w[1] = T(0.75) - d1 * d1;
        w[2] = T(0.5) * a2 * a2;
        dw[0] = -a0;
        dw[1] = T(-2) * d1;
        dw[2] = a2;
    } else {
        bspline_axis<T>(fx, w, dw);
    }
}

// Base node index on FP32 positions: t = (x - o) * inv_h - 0.5f with every
// operation rounded to FP32 (no contraction), base = max(0, floor(t)).  The
// oracle's orc_base_f32 is the same formula; the reference form is
// max(0, int((x - o - h/2)/h)) in FP64 (MPMCUDA.cu:356-358).
MPM_HD int base_index_f32(float x, float o, float inv_h)
{
#pragma clang fp contract(off)
    float d = x - o;
    float s = d * inv_h;
    float t = s - 0.5f;
    int b = (int)floorf(t);
    return b < 0 ? 0 : b;
}

// ---------------------------------------------------------------------------
// 3x3 helpers, row-major A[3*r+c].
template <class T> MPM_HD T det3(const T A[9])
{
    return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
           A[2] * (A[3] * A[7] - A[4] * A[6]);
}

template <class T> MPM_HD void matmul3(const T A[9], const T B[9], T C[9])
{
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            C[3 * r + c] = A[3 * r] * B[c] + A[3 * r + 1] * B[3 + c] + A[3 * r + 2] * B[6 + c];
}

// Rotation matrix of a unit quaternion (w, x, y, z).
template <class T> MPM_HD void quat_to_mat(const T q[4], T R[9])
{
    T w = q[0], x = q[1], y = q[2], z = q[3];
    R[0] = T(1) - T(2) * (y * y + z * z); R[1] = T(2) * (x * y - w * z); R[2] = T(2) * (x * z + w * y);
    R[3] = T(2) * (x * y + w * z); R[4] = T(1) - T(2) * (x * x + z * z); R[5] = T(2) * (y * z - w * x);
    R[6] = T(2) * (x * z - w * y); R[7] = T(2) * (y * z + w * x); R[8] = T(1) - T(2) * (x * x + y * y);
}

// ---------------------------------------------------------------------------
// 3x3 SVD  F = U diag(s) V^T,  U, V in SO(3),  s0 >= s1 >= |s2|,  sign(s2) =
// sign(det F) -- the conventions of the reference's SVD (McAdams et al. 2011,
// used through PolarSVD, MPMCUDA.cu:87-134).  Method (own): one-sided
// Jacobi rotations on the columns of B = F V (starting from V = V0 when a
// warm start is given), up to `sweeps` cyclic sweeps (FP32 stops early once
// converged, kSvdExitCos below), columns sorted by
// norm with a sign flip per swap (keeps det V = +1), then a Givens QR of B
// gives U and the signed singular values.  Returns the largest remaining
// column non-orthogonality |b_p . b_q| / (|b_p| |b_q|) before sorting.
template <class T>
MPM_HD void jacobi_rotate(T B[9], T W[9], int p, int q)
{
    T a = B[p] * B[p] + B[3 + p] * B[3 + p] + B[6 + p] * B[6 + p];
    T b = B[q] * B[q] + B[3 + q] * B[3 + q] + B[6 + q] * B[6 + q];
    T g = B[p] * B[q] + B[3 + p] * B[3 + q] + B[6 + p] * B[6 + q];
    T zeta = (b - a) * (T(0.5) * r_rcp(g));
    T t = (zeta >= T(0) ? T(1) : T(-1)) * r_rcp(m_abs(zeta) + r_sqrt(T(1) + zeta * zeta));
    // g == 0 (already orthogonal, or 0/0) and overflowing zeta: no rotation
    if (!(g != T(0)) || !(t == t)) t = T(0);
    T c = r_rsqrt(T(1) + t * t), s = c * t;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        T bp = B[3 * r + p], bq = B[3 * r + q];
        B[3 * r + p] = c * bp - s * bq;
        B[3 * r + q] = s * bp + c * bq;
        T wp = W[3 * r + p], wq = W[3 * r + q];
        W[3 * r + p] = c * wp - s * wq;
        W[3 * r + q] = s * wp + c * wq;
    }
}

template <class T>
MPM_HD T col_offdiag(const T B[9])
{
    T n0 = B[0] * B[0] + B[3] * B[3] + B[6] * B[6];
    T n1 = B[1] * B[1] + B[4] * B[4] + B[7] * B[7];
    T n2 = B[2] * B[2] + B[5] * B[5] + B[8] * B[8];
    T g01 = B[0] * B[1] + B[3] * B[4] + B[6] * B[7];
    T g02 = B[0] * B[2] + B[3] * B[5] + B[6] * B[8];
    T g12 = B[1] * B[2] + B[4] * B[5] + B[7] * B[8];
    T r01 = g01 * g01 / (n0 * n1 + T(1e-30));
    T r02 = g02 * g02 / (n0 * n2 + T(1e-30));
    T r12 = g12 * g12 / (n1 * n2 + T(1e-30));
    T r = r01 > r02 ? r01 : r02;
    r = r > r12 ? r : r12;
    return m_sqrt(r);
}

template <class T>
MPM_HD void swap_cols_neg(T B[9], T W[9], int j)
{
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        T tb = B[3 * r + j]; B[3 * r + j] = -B[3 * r + j + 1]; B[3 * r + j + 1] = tb;
        T tw = W[3 * r + j]; W[3 * r + j] = -W[3 * r + j + 1]; W[3 * r + j + 1] = tw;
    }
}

template <class T>
MPM_HD void givens_rows(T R[9], T Q[9], int i, int j, int k)
{
    T a = R[3 * i + k], b = R[3 * j + k];
    T rr2 = a * a + b * b;
    T c = T(1), s = T(0);
    if (rr2 > T(0)) { T inv = r_rsqrt(rr2); c = a * inv; s = b * inv; }
#pragma unroll
    for (int col = 0; col < 3; ++col) {
        T ri = R[3 * i + col], rj = R[3 * j + col];
        R[3 * i + col] = c * ri + s * rj;
        R[3 * j + col] = -s * ri + c * rj;
        T qi = Q[3 * col + i], qj = Q[3 * col + j];
        Q[3 * col + i] = c * qi + s * qj;
        Q[3 * col + j] = -s * qi + c * qj;
    }
}

// Finish an SVD from B = F W with (nearly) orthogonal columns.
template <class T>
MPM_HD void svd3_finish(T B[9], T W[9], T U[9], T s[3], T V[9])
{
    T n0 = B[0] * B[0] + B[3] * B[3] + B[6] * B[6];
    T n1 = B[1] * B[1] + B[4] * B[4] + B[7] * B[7];
    T n2 = B[2] * B[2] + B[5] * B[5] + B[8] * B[8];
    if (n0 < n1) { swap_cols_neg(B, W, 0); T t = n0; n0 = n1; n1 = t; }
    if (n1 < n2) { swap_cols_neg(B, W, 1); T t = n1; n1 = n2; n2 = t; }
    if (n0 < n1) { swap_cols_neg(B, W, 0); }
    T Q[9] = {T(1), T(0), T(0), T(0), T(1), T(0), T(0), T(0), T(1)};
    givens_rows(B, Q, 0, 1, 0);
    givens_rows(B, Q, 0, 2, 0);
    givens_rows(B, Q, 1, 2, 1);
#pragma unroll
    for (int k = 0; k < 9; ++k) { U[k] = Q[k]; V[k] = W[k]; }
    s[0] = B[0]; s[1] = B[4]; s[2] = B[8];
}

// Column orthogonality (largest |cos|) at which the FP32 solve stops early.
// Two or three cyclic sweeps reach it for F near the snow band or within
// a factor 2 of a rotation (64-lane groups: 3 sweeps for >99% of snow-like
// and >93% of jello-like inputs); the residual 1e-6 is at the level of the
// FP32 rounding of the column dot products.
constexpr float kSvdExitCos = 1e-6f;

// True when c holds on any lane of the wave (device), or for this call (host).
MPM_HD bool mpm_any(bool c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __any(c);
#else
    return c;
#endif
}

template <class T>
MPM_HD T svd3(const T F[9], T U[9], T s[3], T V[9], int sweeps, const T *V0 = nullptr)
{
    T B[9], W[9];
    if (V0) {
#pragma unroll
        for (int k = 0; k < 9; ++k) W[k] = V0[k];
        matmul3(F, W, B);
    } else {
#pragma unroll
        for (int k = 0; k < 9; ++k) { B[k] = F[k]; W[k] = (k % 4 == 0) ? T(1) : T(0); }
    }
    // FP32: after three sweeps the rest run only while the columns of some
    // lane (on the device: of the wave, so the skip is uniform and saves the
    // issue slots) are not yet orthogonal to kSvdExitCos
    const int first = (sizeof(T) == 4 && sweeps > 3) ? 3 : sweeps;
    for (int it = 0; it < first; ++it) {
        jacobi_rotate(B, W, 0, 1);
        jacobi_rotate(B, W, 0, 2);
        jacobi_rotate(B, W, 1, 2);
    }
    if (first < sweeps && mpm_any(col_offdiag(B) > T(kSvdExitCos))) {
        for (int it = first; it < sweeps; ++it) {
            jacobi_rotate(B, W, 0, 1);
            jacobi_rotate(B, W, 0, 2);
            jacobi_rotate(B, W, 1, 2);
        }
    }
    T off = col_offdiag(B);
    svd3_finish(B, W, U, s, V);
    return off;
}

// ---------------------------------------------------------------------------
// Elastic snow update without the SVD.  A trial F whose singular values all
// lie strictly inside the plastic band [lo, hi] is not clamped
// (MPMCUDA.cu:658-678 leaves F_e = F_tr and F_p unchanged), and its stress
// needs only the polar rotation R (tau = 2 mu (F - R) F^T + lam J (J - 1) I,
// MPMCUDA.cu:229-264).

// All singular values of F strictly inside (lo, hi) and det F > 0: the
// eigenvalues of S = F^T F lie in (lo^2, hi^2) iff S - lo^2 I and
// hi^2 I - S are positive definite (leading principal minors > 0).
template <class T>
MPM_HD bool sv_in_band(const T F[9], T lo, T hi)
{
    T S[6];   // xx, yy, zz, xy, xz, yz of F^T F
    S[0] = F[0] * F[0] + F[3] * F[3] + F[6] * F[6];
    S[1] = F[1] * F[1] + F[4] * F[4] + F[7] * F[7];
    S[2] = F[2] * F[2] + F[5] * F[5] + F[8] * F[8];
    S[3] = F[0] * F[1] + F[3] * F[4] + F[6] * F[7];
    S[4] = F[0] * F[2] + F[3] * F[5] + F[6] * F[8];
    S[5] = F[1] * F[2] + F[4] * F[5] + F[7] * F[8];
    bool ok = det3(F) > T(0);
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        // side 0: S - lo^2 I; side 1: hi^2 I - S
        const T sg = side ? T(-1) : T(1), sh = side ? hi * hi : -(lo * lo);
        const T a = sg * S[0] + sh, b = sg * S[1] + sh, c = sg * S[2] + sh;
        const T d = sg * S[3], e = sg * S[4], f = sg * S[5];
        const T m2 = a * b - d * d;
        const T m3 = a * (b * c - f * f) - d * (d * c - f * e) + e * (d * f - b * e);
        ok = ok && a > T(0) && m2 > T(0) && m3 > T(0);
    }
    return ok;
}

// Polar rotation of F (det F > 0) by Newton's iteration R <- (R + R^-T) / 2
// from R = F (Higham 1986): quadratic convergence, the distance of the
// singular values from 1 squares (and halves) every step, so three steps
// reach FP32 precision for singular values within a few percent of 1.
template <class T>
MPM_HD void polar_newton(const T F[9], T R[9], int iters)
{
#pragma unroll
    for (int k = 0; k < 9; ++k) R[k] = F[k];
    for (int it = 0; it < iters; ++it) {
        // R^-T = cof(R) / det R
        T c[9];
        c[0] = R[4] * R[8] - R[5] * R[7]; c[1] = R[5] * R[6] - R[3] * R[8]; c[2] = R[3] * R[7] - R[4] * R[6];
        c[3] = R[2] * R[7] - R[1] * R[8]; c[4] = R[0] * R[8] - R[2] * R[6]; c[5] = R[1] * R[6] - R[0] * R[7];
        c[6] = R[1] * R[5] - R[2] * R[4]; c[7] = R[2] * R[3] - R[0] * R[5]; c[8] = R[0] * R[4] - R[1] * R[3];
        const T det = R[0] * c[0] + R[1] * c[1] + R[2] * c[2];
        const T hi = T(0.5) * r_rcp(det);
#pragma unroll
        for (int k = 0; k < 9; ++k) R[k] = T(0.5) * R[k] + hi * c[k];
    }
}

// tau = 2 mu (F - R) F^T + lam J (J - 1) I as (xx, yy, zz, xy, xz, yz).
template <class T>
MPM_HD void fcr_tau_polar(const T F[9], const T R[9], T mu, T lam, T tau6[6])
{
    const T J = det3(F);
    const T l = lam * J * (J - T(1));
    T D[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) D[k] = T(2) * mu * (F[k] - R[k]);
    // (D F^T)_rc = sum_k D_rk F_ck; the product is symmetric in exact arithmetic
    tau6[0] = D[0] * F[0] + D[1] * F[1] + D[2] * F[2] + l;
    tau6[1] = D[3] * F[3] + D[4] * F[4] + D[5] * F[5] + l;
    tau6[2] = D[6] * F[6] + D[7] * F[7] + D[8] * F[8] + l;
    tau6[3] = T(0.5) * (D[0] * F[3] + D[1] * F[4] + D[2] * F[5] + D[3] * F[0] + D[4] * F[1] + D[5] * F[2]);
    tau6[4] = T(0.5) * (D[0] * F[6] + D[1] * F[7] + D[2] * F[8] + D[6] * F[0] + D[7] * F[1] + D[8] * F[2]);
    tau6[5] = T(0.5) * (D[3] * F[6] + D[4] * F[7] + D[5] * F[8] + D[6] * F[3] + D[7] * F[4] + D[8] * F[5]);
}

// Kirchhoff stress of the fixed-corotated model from an SVD of F:
// tau = 2 mu (F - R) F^T + lam J (J - 1) I with R = U V^T (MPMCUDA.cu:229-264)
// equals U diag(2 mu (s_i - 1) s_i + lam J (J - 1)) U^T.  Output is the
// symmetric tensor as (xx, yy, zz, xy, xz, yz).
template <class T>
MPM_HD void fcr_tau_from_svd(const T U[9], const T s[3], T mu, T lam, T tau6[6])
{
    T J = s[0] * s[1] * s[2];
    T l = lam * J * (J - T(1));
    T d0 = T(2) * mu * (s[0] - T(1)) * s[0] + l;
    T d1 = T(2) * mu * (s[1] - T(1)) * s[1] + l;
    T d2 = T(2) * mu * (s[2] - T(1)) * s[2] + l;
    // tau_rc = sum_k U_rk d_k U_ck
    tau6[0] = U[0] * U[0] * d0 + U[1] * U[1] * d1 + U[2] * U[2] * d2;
    tau6[1] = U[3] * U[3] * d0 + U[4] * U[4] * d1 + U[5] * U[5] * d2;
    tau6[2] = U[6] * U[6] * d0 + U[7] * U[7] * d1 + U[8] * U[8] * d2;
    tau6[3] = U[0] * U[3] * d0 + U[1] * U[4] * d1 + U[2] * U[5] * d2;
    tau6[4] = U[0] * U[6] * d0 + U[1] * U[7] * d1 + U[2] * U[8] * d2;
    tau6[5] = U[3] * U[6] * d0 + U[4] * U[7] * d1 + U[5] * U[8] * d2;
}

}  // namespace mpm
