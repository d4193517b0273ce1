// mpm_mlp.hip -- the learned warm start of the polar decomposition.
//
// No counterpart in the reference (its SVD always starts cold, svd3_cuda.h).
// A 9 -> 64 -> 64 -> 4 MLP maps F / |F|_F (row-major) to a quaternion q;
// V0 = R(q / |q|) starts the one-sided Jacobi SVD of F (or of the trial F of
// the snow update), so that one sweep usually reaches the acceptance bound.
// Hidden layers use ReLU.  Matrix products run on the MFMA units in BF16
// with FP32 accumulation (v_mfma_f32_16x16x32_bf16); biases, activations,
// the loss and its gradient are FP32.
//
// Training is self-supervised: with A = F^T F / |F|_F^2 and R = R(q/|q|),
//   loss = sum_{i != j} (R^T A R)_ij^2 / |A|_F^2,
// which vanishes exactly when the columns of R are eigenvectors of A (right
// singular vectors of F), whatever their order or signs.  The forward,
// backward and Adam kernels below are the whole training path.
#include "mpm_kernels.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace mpm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 parameter vector (the layout of mpm_set_warm_start_weights):
//   W1[9][64] b1[64] W2[64][64] b2[64] W3[64][4] b3[4],  y_j = sum_i W[i][j] x_i + b_j
constexpr int MLP_IN = 9, MLP_H = 64, MLP_OUT = 4;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + MLP_IN * MLP_H, OFF_W2 = OFF_B1 + MLP_H;
constexpr int OFF_B2 = OFF_W2 + MLP_H * MLP_H, OFF_W3 = OFF_B2 + MLP_H, OFF_B3 = OFF_W3 + MLP_H * MLP_OUT;
constexpr int MLP_PARAMS = OFF_B3 + MLP_OUT;          // 5060

// LDS row strides (bf16 elements): K = 32 or 64 plus 8, so the 16 rows read
// by one MFMA operand load start on different banks.
constexpr int LD32 = 40, LD64 = 72;

// C(16x16 tile tr,tc) += A[rows][0..K) * Bt[cols][0..K)^T, both bf16 row-major in LDS.
__device__ __forceinline__ f32x4 mma_tile(const __bf16 *A, int lda, const __bf16 *Bt, int ldb, int K, int tr,
                                          int tc, f32x4 acc)
{
    const int l = threadIdx.x & 63;
    for (int k0 = 0; k0 < K; k0 += 32) {
        const bf16x8 a = *(const bf16x8 *)(A + (tr * 16 + (l & 15)) * lda + k0 + (l >> 4) * 8);
        const bf16x8 b = *(const bf16x8 *)(Bt + (tc * 16 + (l & 15)) * ldb + k0 + (l >> 4) * 8);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    return acc;
}

// Element (r, c) of the accumulator tile held by this lane: rows
// (l/16)*4 + j, column l%16.
__device__ __forceinline__ int tile_row(int tr, int j) { return tr * 16 + ((threadIdx.x & 63) >> 4) * 4 + j; }
__device__ __forceinline__ int tile_col(int tc) { return tc * 16 + (threadIdx.x & 15); }

struct MlpWeightsLds {
    alignas(16) __bf16 W1t[MLP_H][LD32];      // [hidden][input]   inputs 9..31 zero
    alignas(16) __bf16 W2t[MLP_H][LD64];      // [out][in]
    alignas(16) __bf16 W3t[16][LD64];         // [output][hidden]  outputs 4..15 zero
    float b1[MLP_H], b2[MLP_H], b3[16];
};

struct MlpTrainLds {
    alignas(16) __bf16 W2[MLP_H][LD64];       // [in][out]
    alignas(16) __bf16 W3[MLP_H][LD32];       // [hidden][output]  outputs 4..31 zero
};

struct MlpActLds {
    alignas(16) __bf16 X[64][LD32];           // [sample][input]
    alignas(16) __bf16 H1[64][LD64];          // [sample][hidden]
    alignas(16) __bf16 H2[64][LD64];
    float Q[64][4];
};

struct MlpBackLds {
    alignas(16) __bf16 Xt[32][LD64];          // [input][sample]   rows 9..31 zero
    alignas(16) __bf16 H1t[64][LD64];         // [hidden][sample]
    alignas(16) __bf16 H2t[64][LD64];
    alignas(16) __bf16 D3[64][LD32];          // [sample][output]  outputs 4..31 zero
    alignas(16) __bf16 D3t[16][LD64];         // [output][sample]  rows 4..15 zero
    alignas(16) __bf16 D2[64][LD64];          // [sample][hidden]
    alignas(16) __bf16 D2t[64][LD64];
    alignas(16) __bf16 D1t[64][LD64];
};

__device__ void load_weights(const float *__restrict__ p, MlpWeightsLds &W, MlpTrainLds *T)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int k = tid; k < MLP_H * LD32; k += nt) {
        const int j = k / LD32, i = k % LD32;
        W.W1t[j][i] = (__bf16)(i < MLP_IN ? p[OFF_W1 + i * MLP_H + j] : 0.0f);
    }
    for (int k = tid; k < MLP_H * LD64; k += nt) {
        const int a = k / LD64, b = k % LD64;
        W.W2t[a][b] = (__bf16)(b < MLP_H ? p[OFF_W2 + b * MLP_H + a] : 0.0f);      // W2t[out][in] = W2[in][out]
        if (T) T->W2[a][b] = (__bf16)(b < MLP_H ? p[OFF_W2 + a * MLP_H + b] : 0.0f);
    }
    for (int k = tid; k < 16 * LD64; k += nt) {
        const int o = k / LD64, j = k % LD64;
        W.W3t[o][j] = (__bf16)(o < MLP_OUT && j < MLP_H ? p[OFF_W3 + j * MLP_OUT + o] : 0.0f);
    }
    if (T)
        for (int k = tid; k < MLP_H * LD32; k += nt) {
            const int j = k / LD32, o = k % LD32;
            T->W3[j][o] = (__bf16)(o < MLP_OUT ? p[OFF_W3 + j * MLP_OUT + o] : 0.0f);
        }
    for (int k = tid; k < MLP_H; k += nt) { W.b1[k] = p[OFF_B1 + k]; W.b2[k] = p[OFF_B2 + k]; }
    for (int k = tid; k < 16; k += nt) W.b3[k] = k < MLP_OUT ? p[OFF_B3 + k] : 0.0f;
}

// Network input of one F: F / |F|_F, row-major (zero for F = 0).
__device__ __forceinline__ void mlp_input(const float F[9], float x[9])
{
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 9; ++k) s += F[k] * F[k];
    const float inv = s > 0.f ? 1.0f / sqrtf(s) : 0.0f;
#pragma unroll
    for (int k = 0; k < 9; ++k) x[k] = F[k] * inv;
}

// Forward pass of the 64 samples whose inputs are in A.X (one wave).
// Leaves H1, H2 (bf16, post-ReLU) and Q (fp32).  `B` (training) also gets
// the transposed activations.
__device__ void mlp_forward_wave(const MlpWeightsLds &W, MlpActLds &A, MlpBackLds *B)
{
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int tr = 0; tr < 4; ++tr)
#pragma unroll
        for (int tc = 0; tc < 4; ++tc) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mma_tile(&A.X[0][0], LD32, &W.W1t[0][0], LD32, 32, tr, tc, acc);
            const int c = tile_col(tc);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int r = tile_row(tr, j);
                const __bf16 h = (__bf16)fmaxf(acc[j] + W.b1[c], 0.0f);
                A.H1[r][c] = h;
                if (B) B->H1t[c][r] = h;
            }
        }
    __syncthreads();
#pragma unroll
    for (int tr = 0; tr < 4; ++tr)
#pragma unroll
        for (int tc = 0; tc < 4; ++tc) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mma_tile(&A.H1[0][0], LD64, &W.W2t[0][0], LD64, 64, tr, tc, acc);
            const int c = tile_col(tc);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int r = tile_row(tr, j);
                const __bf16 h = (__bf16)fmaxf(acc[j] + W.b2[c], 0.0f);
                A.H2[r][c] = h;
                if (B) B->H2t[c][r] = h;
            }
        }
    __syncthreads();
#pragma unroll
    for (int tr = 0; tr < 4; ++tr) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mma_tile(&A.H2[0][0], LD64, &W.W3t[0][0], LD64, 64, tr, 0, acc);
        const int c = tile_col(0);
        if (c < MLP_OUT) {
#pragma unroll
            for (int j = 0; j < 4; ++j) A.Q[tile_row(tr, j)][c] = acc[j] + W.b3[c];
        }
    }
    __syncthreads();
    (void)l;
}

// Loss of one sample and its gradient with respect to the raw output q.
__device__ __forceinline__ float warm_loss_grad(const float x[9], const float q[4], float dq[4])
{
    float A[9];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) A[3 * a + b] = x[a] * x[b] + x[3 + a] * x[3 + b] + x[6 + a] * x[6 + b];
    float nA = 0.f;
#pragma unroll
    for (int k = 0; k < 9; ++k) nA += A[k] * A[k];
    const float qn = fmaxf(sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]), 1e-12f);
    const float qh[4] = {q[0] / qn, q[1] / qn, q[2] / qn, q[3] / qn};
    float R[9];
    quat_to_mat<float>(qh, R);
    float AR[9], Bm[9];
    matmul3<float>(A, R, AR);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Bm[3 * i + j] = R[i] * AR[j] + R[3 + i] * AR[3 + j] + R[6 + i] * AR[6 + j];
    const float inv = nA > 0.f ? 1.0f / nA : 0.0f;
    float loss = 0.f, GB[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float v = i == j ? 0.f : Bm[3 * i + j];
            loss += v * v;
            GB[3 * i + j] = 2.0f * v * inv;
        }
    loss *= inv;
    // dL/dR = 2 A R G_B (A, G_B symmetric)
    float ARG[9], dR[9];
    matmul3<float>(AR, GB, ARG);
#pragma unroll
    for (int k = 0; k < 9; ++k) dR[k] = 2.0f * ARG[k];
    const float w = qh[0], X = qh[1], Y = qh[2], Z = qh[3];
    // dR/dq_k (each R entry is quadratic in q), contracted with dR
    const float gw = 2.0f * (dR[1] * -Z + dR[2] * Y + dR[3] * Z + dR[5] * -X + dR[6] * -Y + dR[7] * X);
    const float gx = 2.0f * (dR[1] * Y + dR[2] * Z + dR[3] * Y - 2.0f * X * dR[4] + dR[5] * -w + dR[6] * Z +
                             dR[7] * w - 2.0f * X * dR[8]);
    const float gy = 2.0f * (-2.0f * Y * dR[0] + dR[1] * X + dR[2] * w + dR[3] * X + dR[5] * Z + dR[6] * -w +
                             dR[7] * Z - 2.0f * Y * dR[8]);
    const float gz = 2.0f * (-2.0f * Z * dR[0] + dR[1] * -w + dR[2] * X + dR[3] * w - 2.0f * Z * dR[4] + dR[5] * Y +
                             dR[6] * X + dR[7] * Y);
    const float g[4] = {gw, gx, gy, gz};
    const float dot = g[0] * qh[0] + g[1] * qh[1] + g[2] * qh[2] + g[3] * qh[3];
#pragma unroll
    for (int k = 0; k < 4; ++k) dq[k] = (g[k] - qh[k] * dot) / qn;
    return loss;
}

// Inference activations of one wave: layer 2 overwrites layer 1 in place,
// row tile by row tile (row tile r of H2 depends on row tile r of H1 only).
struct MlpInferLds {
    alignas(16) __bf16 X[64][LD32];
    alignas(16) __bf16 H[64][LD64];
    float Q[64][4];
};

__device__ void mlp_infer_wave(const MlpWeightsLds &W, MlpInferLds &A)
{
#pragma unroll
    for (int tr = 0; tr < 4; ++tr)
#pragma unroll
        for (int tc = 0; tc < 4; ++tc) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mma_tile(&A.X[0][0], LD32, &W.W1t[0][0], LD32, 32, tr, tc, acc);
            const int c = tile_col(tc);
#pragma unroll
            for (int j = 0; j < 4; ++j) A.H[tile_row(tr, j)][c] = (__bf16)fmaxf(acc[j] + W.b1[c], 0.0f);
        }
    __syncthreads();
#pragma unroll
    for (int tr = 0; tr < 4; ++tr) {
        f32x4 acc[4];
#pragma unroll
        for (int tc = 0; tc < 4; ++tc) {
            acc[tc] = f32x4{0.f, 0.f, 0.f, 0.f};
            acc[tc] = mma_tile(&A.H[0][0], LD64, &W.W2t[0][0], LD64, 64, tr, tc, acc[tc]);
        }
        __asm__ volatile("" ::: "memory");     // every read of row tile tr precedes its overwrite
#pragma unroll
        for (int tc = 0; tc < 4; ++tc) {
            const int c = tile_col(tc);
#pragma unroll
            for (int j = 0; j < 4; ++j) A.H[tile_row(tr, j)][c] = (__bf16)fmaxf(acc[tc][j] + W.b2[c], 0.0f);
        }
    }
    __syncthreads();
#pragma unroll
    for (int tr = 0; tr < 4; ++tr) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mma_tile(&A.H[0][0], LD64, &W.W3t[0][0], LD64, 64, tr, 0, acc);
        const int c = tile_col(0);
        if (c < MLP_OUT) {
#pragma unroll
            for (int j = 0; j < 4; ++j) A.Q[tile_row(tr, j)][c] = acc[j] + W.b3[c];
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Inference: q for n deformation gradients.  F element k of sample i is
// F[i * fi + k * fk], q element k goes to q[i * qi + k * qk] (AoS or SoA).
// Four waves per workgroup share the weights; each wave takes 64 samples.
__global__ void __launch_bounds__(256) k_mlp_infer(const float *__restrict__ params, const float *__restrict__ F,
                                                   int64_t fi, int64_t fk, float *__restrict__ q, int64_t qi,
                                                   int64_t qk, int64_t n)
{
    __shared__ MlpWeightsLds W;
    __shared__ MlpInferLds act[4];
    load_weights(params, W, nullptr);
    __syncthreads();
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    MlpInferLds &A = act[w];
    // every wave runs the same number of chunks (barriers inside)
    const int64_t chunks = (n + 63) / 64, per = (int64_t)gridDim.x * 4;
    for (int64_t c0 = (int64_t)blockIdx.x * 4; c0 < chunks; c0 += per) {
        const int64_t c = c0 + w;
        const int64_t i = c * 64 + l;
        const bool ok = c < chunks && i < n;
        float Fv[9], x[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) Fv[k] = ok ? F[i * fi + k * fk] : 0.0f;
        mlp_input(Fv, x);
#pragma unroll
        for (int k = 0; k < LD32; ++k) A.X[l][k] = (__bf16)(k < MLP_IN ? x[k] : 0.0f);
        __syncthreads();
        mlp_infer_wave(W, A);
        if (ok) {
#pragma unroll
            for (int k = 0; k < 4; ++k) q[i * qi + k * qk] = A.Q[l][k];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Training: loss and gradient of a minibatch (sample indices idx[0..m)).
// One wave per workgroup; the weight gradients of all its chunks stay in
// MFMA accumulators and are added to grad (FP32, MLP_PARAMS) once.
__global__ void __launch_bounds__(64) k_mlp_grad(const float *__restrict__ params, const float *__restrict__ F,
                                                 const uint32_t *__restrict__ idx, int m, float scale,
                                                 float *__restrict__ grad, float *__restrict__ loss_out)
{
    __shared__ MlpWeightsLds W;
    __shared__ MlpTrainLds T;
    __shared__ MlpActLds A;
    __shared__ MlpBackLds B;
    load_weights(params, W, &T);
    const int l = threadIdx.x;
    for (int k = l; k < 32 * LD64; k += 64) (&B.Xt[0][0])[k] = (__bf16)0.0f;
    for (int k = l; k < 64 * LD32; k += 64) (&B.D3[0][0])[k] = (__bf16)0.0f;
    for (int k = l; k < 16 * LD64; k += 64) (&B.D3t[0][0])[k] = (__bf16)0.0f;
    __syncthreads();
    f32x4 gW1[4], gW2[4][4], gW3[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        gW1[a] = f32x4{0.f, 0.f, 0.f, 0.f};
        gW3[a] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int b = 0; b < 4; ++b) gW2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, lsum = 0.f;   // lane l: hidden unit l (b3: output l < 4)
    const int chunks = (m + 63) / 64;
    for (int c = blockIdx.x; c < chunks; c += gridDim.x) {
        const int s = c * 64 + l;
        const bool ok = s < m;
        float Fv[9], x[9];
        const int64_t i = ok ? (int64_t)idx[s] : 0;
#pragma unroll
        for (int k = 0; k < 9; ++k) Fv[k] = ok ? F[i * 9 + k] : 0.0f;
        mlp_input(Fv, x);
#pragma unroll
        for (int k = 0; k < LD32; ++k) A.X[l][k] = (__bf16)(k < MLP_IN ? x[k] : 0.0f);
#pragma unroll
        for (int k = 0; k < MLP_IN; ++k) B.Xt[k][l] = (__bf16)x[k];
        __syncthreads();
        mlp_forward_wave(W, A, &B);
        // output gradient, one lane per sample (bf16-rounded inputs, as the network saw them)
        float xb[9], qv[4], dq[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < 9; ++k) xb[k] = (float)A.X[l][k];
#pragma unroll
        for (int k = 0; k < 4; ++k) qv[k] = A.Q[l][k];
        if (ok) lsum += warm_loss_grad(xb, qv, dq);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const __bf16 d = (__bf16)(ok ? dq[k] * scale : 0.0f);
            B.D3[l][k] = d;
            B.D3t[k][l] = d;
        }
        __syncthreads();
        // layer 3: dW3 += H2^T D3, db3 += sum D3
#pragma unroll
        for (int tr = 0; tr < 4; ++tr) gW3[tr] = mma_tile(&B.H2t[0][0], LD64, &B.D3t[0][0], LD64, 64, tr, 0, gW3[tr]);
        if (l < MLP_OUT)
            for (int k = 0; k < 64; ++k) gb3 += (float)B.D3[k][l];
        // delta of hidden layer 2: (D3 W3^T) * relu'(H2)
#pragma unroll
        for (int tr = 0; tr < 4; ++tr)
#pragma unroll
            for (int tc = 0; tc < 4; ++tc) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                acc = mma_tile(&B.D3[0][0], LD32, &T.W3[0][0], LD32, 32, tr, tc, acc);
                const int col = tile_col(tc);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int r = tile_row(tr, j);
                    const __bf16 d = (__bf16)((float)A.H2[r][col] > 0.f ? acc[j] : 0.0f);
                    B.D2[r][col] = d;
                    B.D2t[col][r] = d;
                }
            }
        __syncthreads();
        for (int k = 0; k < 64; ++k) gb2 += (float)B.D2[k][l];
        // layer 2: dW2 += H1^T D2
#pragma unroll
        for (int tr = 0; tr < 4; ++tr)
#pragma unroll
            for (int tc = 0; tc < 4; ++tc)
                gW2[tr][tc] = mma_tile(&B.H1t[0][0], LD64, &B.D2t[0][0], LD64, 64, tr, tc, gW2[tr][tc]);
        // delta of hidden layer 1: (D2 W2^T) * relu'(H1)
#pragma unroll
        for (int tr = 0; tr < 4; ++tr)
#pragma unroll
            for (int tc = 0; tc < 4; ++tc) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                acc = mma_tile(&B.D2[0][0], LD64, &T.W2[0][0], LD64, 64, tr, tc, acc);
                const int col = tile_col(tc);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int r = tile_row(tr, j);
                    B.D1t[col][r] = (__bf16)((float)A.H1[r][col] > 0.f ? acc[j] : 0.0f);
                }
            }
        __syncthreads();
        for (int k = 0; k < 64; ++k) gb1 += (float)B.D1t[l][k];
        // layer 1: dW1 += X^T D1 (inputs 0..15 of the first row tile)
#pragma unroll
        for (int tc = 0; tc < 4; ++tc) gW1[tc] = mma_tile(&B.Xt[0][0], LD64, &B.D1t[0][0], LD64, 64, 0, tc, gW1[tc]);
        __syncthreads();
    }
    // flush: one atomic per parameter per workgroup
#pragma unroll
    for (int tc = 0; tc < 4; ++tc)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = tile_row(0, j), col = tile_col(tc);
            if (r < MLP_IN) atomicAdd(grad + OFF_W1 + r * MLP_H + col, gW1[tc][j]);
        }
#pragma unroll
    for (int tr = 0; tr < 4; ++tr)
#pragma unroll
        for (int tc = 0; tc < 4; ++tc)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                atomicAdd(grad + OFF_W2 + tile_row(tr, j) * MLP_H + tile_col(tc), gW2[tr][tc][j]);
#pragma unroll
    for (int tr = 0; tr < 4; ++tr)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = tile_col(0);
            if (col < MLP_OUT) atomicAdd(grad + OFF_W3 + tile_row(tr, j) * MLP_OUT + col, gW3[tr][j]);
        }
    atomicAdd(grad + OFF_B1 + l, gb1);
    atomicAdd(grad + OFF_B2 + l, gb2);
    if (l < MLP_OUT) atomicAdd(grad + OFF_B3 + l, gb3);
    // loss of the minibatch (sum over samples, unscaled)
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) lsum += __shfl_xor(lsum, d);
    if (l == 0) atomicAdd(loss_out, lsum);
}

// Adam step on the FP32 parameters (bias-corrected, step t >= 1).
__global__ void k_adam(float *p, const float *g, float *m1, float *m2, int n, float lr, float b1, float b2,
                       float eps, int t)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float a = b1 * m1[k] + (1.0f - b1) * gk, b = b2 * m2[k] + (1.0f - b2) * gk * gk;
    m1[k] = a;
    m2[k] = b;
    const float ah = a / (1.0f - powf(b1, (float)t)), bh = b / (1.0f - powf(b2, (float)t));
    p[k] -= lr * ah / (sqrtf(bh) + eps);
}

__global__ void k_fill_f32(float *p, int n, float v)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

// SplitMix64 sample indices for one minibatch.
__global__ void k_batch_indices(uint32_t *idx, int m, uint64_t seed, uint64_t step, uint32_t n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    uint64_t z = seed + 0x9e3779b97f4a7c15ull * (step * 1000003ull + (uint64_t)k + 1ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    idx[k] = (uint32_t)(z % n);
}

// Inference for the warm scenes of a batch that share one weight set
// (BatchSlice::warm_group == group): the particles each scene's last G2P
// listed (mat.wlist[0, warm_cnt[0])), in chunks of 64 numbered scene after
// scene; every wave maps its chunk to a scene with its own SliceScan.
// Writes q to the PL_Q planes.  Workgroups without a chunk leave before
// loading the weights.
__global__ void __launch_bounds__(256) k_mlp_infer_batch(const float *__restrict__ params,
                                                         const BatchSlice *__restrict__ tab, int nslices, int group)
{
    SliceScan sc;
    sc.build(nslices, [&](int s) {
        return tab[s].warm_group == group && tab[s].warm_cnt ? (gp(tab[s].warm_cnt)[0] + 63u) / 64u : 0u;
    });
    const uint32_t chunks = sc.total, per = gridDim.x * 4u;
    if (blockIdx.x * 4u >= chunks) return;       // whole workgroup
    __shared__ MlpWeightsLds W;
    __shared__ MlpInferLds act[4];
    load_weights(params, W, nullptr);
    __syncthreads();
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    MlpInferLds &A = act[w];
    // every wave runs the same number of chunks (barriers inside)
    for (uint32_t c0 = blockIdx.x * 4u; c0 < chunks; c0 += per) {
        const uint32_t c = c0 + (uint32_t)w;
        const uint32_t cc = c < chunks ? c : 0u;
        const uint32_t sid = (uint32_t)__builtin_amdgcn_readfirstlane((int)sc.find(cc));
        const BatchSlice &b = tab[sid];
        const uint32_t j = (cc - sc.start(sid)) * 64u + (uint32_t)l;
        const bool ok = c < chunks && j < min(gp(b.warm_cnt)[0], (uint32_t)b.pv.n);
        const int i = ok ? (int)gp(b.mat.wlist)[j] : 0;
        float Fv[9], x[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) Fv[k] = ok ? ld(b.pv, PL_F + k, i) : 0.0f;
        mlp_input(Fv, x);
#pragma unroll
        for (int k = 0; k < LD32; ++k) A.X[l][k] = (__bf16)(k < MLP_IN ? x[k] : 0.0f);
        __syncthreads();
        mlp_infer_wave(W, A);
        if (ok) {
#pragma unroll
            for (int k = 0; k < 4; ++k) st(b.pv, PL_Q + k, i, A.Q[l][k]);
        }
        __syncthreads();
    }
}

int launch_mlp_infer_batch(hipStream_t st, const float *params, const BatchSlice *tab, int nslices, int group,
                           int64_t chunks)
{
    if (chunks <= 0) return 0;
    const int wg = (int)std::min<int64_t>((chunks + 3) / 4, 256);
    hipLaunchKernelGGL(k_mlp_infer_batch, dim3(wg), dim3(256), 0, st, params, tab, nslices, group);
    return (int)hipGetLastError();
}

// Warm-start quaternions of the listed particles of a group: list[0, *count)
// holds group-local indices (the particles whose last G2P took the SVD).
// Block 0 first clears *next_count, the list the coming G2P appends to.
// Persistent: every workgroup loads the weights once and takes chunks of 64
// listed particles per wave; the count is read from device memory, so the
// grid is sized for the group and workgroups without a chunk leave before
// loading the weights.
__global__ void __launch_bounds__(256) k_mlp_infer_list(const float *__restrict__ params, PartView pv,
                                                        const uint32_t *__restrict__ list,
                                                        const uint32_t *__restrict__ count, uint32_t *next_count)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *gp(next_count) = 0u;
    const uint32_t n = min(*gp(count), (uint32_t)pv.n);
    const uint32_t chunks = (n + 63u) / 64u, per = gridDim.x * 4u;
    if (blockIdx.x * 4u >= chunks) return;       // whole workgroup
    __shared__ MlpWeightsLds W;
    __shared__ MlpInferLds act[4];
    load_weights(params, W, nullptr);
    __syncthreads();
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    MlpInferLds &A = act[w];
    // every wave runs the same number of chunks (barriers inside)
    for (uint32_t c0 = blockIdx.x * 4u; c0 < chunks; c0 += per) {
        const uint32_t j = (c0 + (uint32_t)w) * 64u + (uint32_t)l;
        const bool ok = j < n;
        const int i = ok ? (int)gp(list)[j] : 0;
        float Fv[9], x[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) Fv[k] = ok ? ld(pv, PL_F + k, i) : 0.0f;
        mlp_input(Fv, x);
#pragma unroll
        for (int k = 0; k < LD32; ++k) A.X[l][k] = (__bf16)(k < MLP_IN ? x[k] : 0.0f);
        __syncthreads();
        mlp_infer_wave(W, A);
        if (ok) {
#pragma unroll
            for (int k = 0; k < 4; ++k) st(pv, PL_Q + k, i, A.Q[l][k]);
        }
        __syncthreads();
    }
}

int launch_mlp_infer_list(hipStream_t st, const float *params, const PartView &pv, const uint32_t *list,
                          const uint32_t *count, uint32_t *next_count)
{
    if (pv.n <= 0) return 0;
    const int64_t chunks = (pv.n + 63) / 64;
    const int wg = (int)std::min<int64_t>((chunks + 3) / 4, 256);
    hipLaunchKernelGGL(k_mlp_infer_list, dim3(wg), dim3(256), 0, st, params, pv, list, count, next_count);
    return (int)hipGetLastError();
}

}  // namespace mpm

using namespace mpm;

// ---------------------------------------------------------------------------
// C ABI (standalone network entry points; own device buffers, synchronous).

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

extern "C" int mpm_mlp_param_count(void) { return MLP_PARAMS; }

extern "C" int mpm_mlp_forward(const float *params, const float *F, int64_t n, float *q)
{
    if (!params || (n > 0 && (!F || !q)) || n < 0) return set_error(MPM_ERR_INVALID, "bad argument");
    if (n == 0) return MPM_OK;
    DevBuf dp, dF, dq;
    MPM_HIP(hipMalloc(&dp.p, MLP_PARAMS * 4));
    MPM_HIP(hipMalloc(&dF.p, n * 9 * 4));
    MPM_HIP(hipMalloc(&dq.p, n * 4 * 4));
    MPM_HIP(hipMemcpy(dp.p, params, MLP_PARAMS * 4, hipMemcpyHostToDevice));
    MPM_HIP(hipMemcpy(dF.p, F, n * 9 * 4, hipMemcpyHostToDevice));
    const int64_t chunks = (n + 63) / 64;
    const int wg = (int)std::min<int64_t>((chunks + 3) / 4, 2048);
    hipLaunchKernelGGL(k_mlp_infer, dim3(wg), dim3(256), 0, 0, (const float *)dp.p, (const float *)dF.p, (int64_t)9,
                       (int64_t)1, (float *)dq.p, (int64_t)4, (int64_t)1, n);
    MPM_HIP(hipGetLastError());
    MPM_HIP(hipMemcpy(q, dq.p, n * 4 * 4, hipMemcpyDeviceToHost));
    return MPM_OK;
}

// Mean loss and its gradient over all n samples.
extern "C" int mpm_mlp_grad(const float *params, const float *F, int64_t n, float *grad, float *loss)
{
    if (!params || !F || !grad || !loss || n <= 0 || n > (1ll << 31)) return set_error(MPM_ERR_INVALID, "bad argument");
    DevBuf dp, dF, dg, dl, di;
    MPM_HIP(hipMalloc(&dp.p, MLP_PARAMS * 4));
    MPM_HIP(hipMalloc(&dF.p, n * 9 * 4));
    MPM_HIP(hipMalloc(&dg.p, MLP_PARAMS * 4));
    MPM_HIP(hipMalloc(&dl.p, 4));
    MPM_HIP(hipMalloc(&di.p, n * 4));
    MPM_HIP(hipMemcpy(dp.p, params, MLP_PARAMS * 4, hipMemcpyHostToDevice));
    MPM_HIP(hipMemcpy(dF.p, F, n * 9 * 4, hipMemcpyHostToDevice));
    std::vector<uint32_t> idx((size_t)n);
    for (int64_t k = 0; k < n; ++k) idx[(size_t)k] = (uint32_t)k;
    MPM_HIP(hipMemcpy(di.p, idx.data(), n * 4, hipMemcpyHostToDevice));
    MPM_HIP(hipMemset(dg.p, 0, MLP_PARAMS * 4));
    MPM_HIP(hipMemset(dl.p, 0, 4));
    const int chunks = (int)((n + 63) / 64);
    hipLaunchKernelGGL(k_mlp_grad, dim3(std::min(chunks, 1024)), dim3(64), 0, 0, (const float *)dp.p,
                       (const float *)dF.p, (const uint32_t *)di.p, (int)n, 1.0f / (float)n, (float *)dg.p,
                       (float *)dl.p);
    MPM_HIP(hipGetLastError());
    MPM_HIP(hipMemcpy(grad, dg.p, MLP_PARAMS * 4, hipMemcpyDeviceToHost));
    float l = 0.f;
    MPM_HIP(hipMemcpy(&l, dl.p, 4, hipMemcpyDeviceToHost));
    *loss = l / (float)n;
    return MPM_OK;
}

// Adam on minibatches drawn from the n samples; params in: initial values,
// out: trained.  loss_history (may be NULL) receives each step's mean loss.
extern "C" int mpm_mlp_train(const float *F, int64_t n, int32_t steps, int32_t batch, float lr, uint64_t seed,
                             float *params, float *loss_history)
{
    if (!F || !params || n <= 0 || n > (1ll << 31) || steps < 0 || batch <= 0 || !(lr > 0.f))
        return set_error(MPM_ERR_INVALID, "bad argument");
    DevBuf dp, dF, dg, dl, di, m1, m2;
    MPM_HIP(hipMalloc(&dp.p, MLP_PARAMS * 4));
    MPM_HIP(hipMalloc(&dF.p, n * 9 * 4));
    MPM_HIP(hipMalloc(&dg.p, MLP_PARAMS * 4));
    MPM_HIP(hipMalloc(&m1.p, MLP_PARAMS * 4));
    MPM_HIP(hipMalloc(&m2.p, MLP_PARAMS * 4));
    MPM_HIP(hipMalloc(&dl.p, (size_t)std::max(steps, 1) * 4));
    MPM_HIP(hipMalloc(&di.p, (size_t)batch * 4));
    MPM_HIP(hipMemcpy(dp.p, params, MLP_PARAMS * 4, hipMemcpyHostToDevice));
    MPM_HIP(hipMemcpy(dF.p, F, n * 9 * 4, hipMemcpyHostToDevice));
    const int nb = (MLP_PARAMS + 255) / 256;
    hipLaunchKernelGGL(k_fill_f32, dim3(nb), dim3(256), 0, 0, (float *)m1.p, MLP_PARAMS, 0.0f);
    hipLaunchKernelGGL(k_fill_f32, dim3(nb), dim3(256), 0, 0, (float *)m2.p, MLP_PARAMS, 0.0f);
    hipLaunchKernelGGL(k_fill_f32, dim3((std::max(steps, 1) + 255) / 256), dim3(256), 0, 0, (float *)dl.p,
                       std::max(steps, 1), 0.0f);
    const int chunks = (batch + 63) / 64;
    for (int t = 0; t < steps; ++t) {
        hipLaunchKernelGGL(k_fill_f32, dim3(nb), dim3(256), 0, 0, (float *)dg.p, MLP_PARAMS, 0.0f);
        hipLaunchKernelGGL(k_batch_indices, dim3((batch + 255) / 256), dim3(256), 0, 0, (uint32_t *)di.p, batch,
                           seed, (uint64_t)t, (uint32_t)n);
        hipLaunchKernelGGL(k_mlp_grad, dim3(std::min(chunks, 512)), dim3(64), 0, 0, (const float *)dp.p,
                           (const float *)dF.p, (const uint32_t *)di.p, batch, 1.0f / (float)batch, (float *)dg.p,
                           (float *)dl.p + t);
        hipLaunchKernelGGL(k_adam, dim3(nb), dim3(256), 0, 0, (float *)dp.p, (const float *)dg.p, (float *)m1.p,
                           (float *)m2.p, MLP_PARAMS, lr, 0.9f, 0.999f, 1e-8f, t + 1);
    }
    MPM_HIP(hipGetLastError());
    MPM_HIP(hipMemcpy(params, dp.p, MLP_PARAMS * 4, hipMemcpyDeviceToHost));
    if (loss_history && steps > 0) {
        std::vector<float> lh((size_t)steps);
        MPM_HIP(hipMemcpy(lh.data(), dl.p, (size_t)steps * 4, hipMemcpyDeviceToHost));
        for (int t = 0; t < steps; ++t) loss_history[t] = lh[(size_t)t] / (float)batch;
    }
    return MPM_OK;
}
