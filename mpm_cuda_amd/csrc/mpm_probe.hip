// mpm_probe.hip -- single-stage device probes and host-only sampling entry
// points of the C ABI (no reference counterpart; used by the parity tests).
#include <string.h>

#include <algorithm>
#include <vector>

#include "mpm_kernels.h"
#include "mpm_launch.h"

namespace mpm {

int probe_sort(const uint32_t *keys_h, int64_t n, int bits, uint32_t *sorted_h, uint32_t *perm_h);

__global__ void k_probe_base(const float *x, int n, float origin, float inv_h, int32_t *out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = base_index_f32(x[i], origin, inv_h);
}

__global__ void k_probe_svd(const float *F, int n, int sweeps, float *U, float *s, float *V, uint32_t *counters)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float f[9], u[9], sg[3], v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) f[k] = F[9 * i + k];
    svd_checked(f, u, sg, v, sweeps, counters);
#pragma unroll
    for (int k = 0; k < 9; ++k) { U[9 * i + k] = u[k]; V[9 * i + k] = v[k]; }
    s[3 * i] = sg[0]; s[3 * i + 1] = sg[1]; s[3 * i + 2] = sg[2];
}

// Kirchhoff stress of the step kernels (stress_from_F: polar rotation for F
// within a factor 2 of a rotation, the SVD otherwise) for n matrices.
template <int MAT>
__global__ void k_probe_stress(const float *F, const float *Jp, int n, MatP m, int sweeps, float *tau,
                               uint32_t *counters)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float f[9], t[6];
#pragma unroll
    for (int k = 0; k < 9; ++k) f[k] = F[9 * i + k];
    stress_from_F<MAT>(m, f, MAT == MPM_SNOW ? Jp[i] : 1.0f, sweeps, t, counters);
#pragma unroll
    for (int k = 0; k < 6; ++k) tau[6 * i + k] = t[k];
}

// Column orthogonality residual of F V after `sweeps` Jacobi sweeps from
// V0 = R(q / |q|) (q == nullptr: V0 = I).
__global__ void k_probe_svd_residual(const float *F, const float *q, int n, int sweeps, float *off)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float f[9], u[9], sg[3], v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) f[k] = F[9 * i + k];
    float r;
    if (q) {
        float qq[4] = {q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]};
        const float qn = sqrtf(qq[0] * qq[0] + qq[1] * qq[1] + qq[2] * qq[2] + qq[3] * qq[3]);
        if (qn > 0.f) {
#pragma unroll
            for (int k = 0; k < 4; ++k) qq[k] /= qn;
        } else {
            qq[0] = 1.f; qq[1] = qq[2] = qq[3] = 0.f;
        }
        float V0[9];
        quat_to_mat<float>(qq, V0);
        r = svd3<float>(f, u, sg, v, sweeps, V0);
    } else {
        r = svd3<float>(f, u, sg, v, sweeps);
    }
    off[i] = r;
}

// Plain streaming copy, one 16-byte element per thread (a grid of
// n4 / 256 workgroups: on a 1 GiB copy this form reached 6.2 TB/s where a
// grid-stride loop over 1,024-65,536 workgroups reached 4.3-5.6, measured).
__global__ void __launch_bounds__(256) k_probe_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, int64_t n4)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n4) gstore4(dst, i, gload4(src, i));
}

// Structure-of-arrays pass shaped like a particle update: element i reads
// plane p (< nin) at in[p * n + i] and writes plane q (< nout) at
// out[q * n + i], one element per thread.
__global__ void __launch_bounds__(256) k_probe_soa(const float *__restrict__ in, float *__restrict__ out, int64_t n,
                                                   int nin, int nout)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float acc = 0.f;
#pragma unroll 8
    for (int p = 0; p < nin; ++p) acc += gp(in)[p * n + i];
#pragma unroll 8
    for (int q = 0; q < nout; ++q) gp(out)[q * n + i] = acc + (float)q;
}

// LDS scatter probe: one-wave workgroups each add, for `reps` x 64
// particles, 27 node updates of 4 floats (a particle's P2G stencil in a 4^3
// block) into an 8 KB arena.  Lane cells: mode bit 0 clear -- lane l owns
// cell l (no two lanes on one address in an instruction); set -- a hashed
// cell (collisions).  Mode bit 1 clear: float atomics (ds_add_f32, SoA
// arena); set: plain float4 read-modify-write (lanes of one cell overwrite
// each other's adds: a rate measurement, the sums are not used).  Every
// arena index is below 4 * 512.
constexpr int LP_NODES = 512;
__device__ __forceinline__ uint32_t lp_hash(uint32_t x)
{
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}
__device__ __forceinline__ int lp_node(int cell, int o)
{
    const int cx = cell & 3, cy = (cell >> 2) & 3, cz = cell >> 4;
    return (cx + o % 3 + 1) + 8 * (cy + (o / 3) % 3 + 1) + 64 * (cz + o / 9 + 1);     // < 438
}
template <bool ATOMIC, bool HASHED>
__global__ void __launch_bounds__(64) k_probe_lds(float *out, int reps)
{
    __shared__ float A[4 * LP_NODES];
    const int lane = threadIdx.x;
    for (int k = lane; k < 4 * LP_NODES; k += 64) A[k] = 0.f;
    __syncthreads();
    float v = 1.0f + lane * 1e-3f;
    for (int r = 0; r < reps; ++r) {
        const int cell = HASHED ? (int)(lp_hash(lane + 64u * (r + reps * blockIdx.x)) & 63u) : lane;
#pragma unroll
        for (int o = 0; o < 27; ++o) {
            const int n = lp_node(cell, o);
            if (ATOMIC) {
                atomicAdd(&A[n], v);
                atomicAdd(&A[LP_NODES + n], 0.5f * v);
                atomicAdd(&A[2 * LP_NODES + n], 0.25f * v);
                atomicAdd(&A[3 * LP_NODES + n], 0.125f * v);
            } else {
                float4 *A4 = (float4 *)A;
                float4 a = A4[n];
                a.x += v; a.y += 0.5f * v; a.z += 0.25f * v; a.w += 0.125f * v;
                A4[n] = a;
                __asm__ volatile("" ::: "memory");
            }
        }
        v += 1e-6f;
    }
    __syncthreads();
    for (int k = lane; k < 4 * LP_NODES; k += 64) out[(int64_t)blockIdx.x * 4 * LP_NODES + k] = A[k];
}

// RAII set of device temporaries for the probes.
struct DevBufs {
    std::vector<void *> p;
    ~DevBufs() { for (void *q : p) if (q) (void)hipFree(q); }
    template <class T> hipError_t alloc(T **out, size_t bytes)
    {
        hipError_t e = hipMalloc((void **)out, bytes ? bytes : 4);
        if (e == hipSuccess) p.push_back(*out);
        return e;
    }
};

}  // namespace mpm

using namespace mpm;

extern "C" {

int mpm_sample_object(const mpm_config *cfg, const mpm_object *obj, float *xyz, int64_t *n)
{
    if (!cfg || !obj || !n) return set_error(MPM_ERR_INVALID, "null argument");
    std::vector<float> pts;
    int rc = sample_object(*cfg, *obj, pts);
    if (rc) return rc;
    int64_t cnt = (int64_t)pts.size() / 3;
    if (xyz) memcpy(xyz, pts.data(), (size_t)std::min(*n, cnt) * 3 * sizeof(float));
    *n = cnt;
    return MPM_OK;
}

int mpm_probe_copy_bandwidth(int64_t bytes, int32_t reps, double *gbs)
{
    if (!gbs) return set_error(MPM_ERR_INVALID, "null argument");
    if (bytes < (1 << 20) || bytes > (8ll << 30) || bytes % 16) return set_error(MPM_ERR_INVALID, "bytes must be a multiple of 16 in [1 MiB, 8 GiB]");
    if (reps < 1 || reps > 100) return set_error(MPM_ERR_INVALID, "reps out of range [1,100]");
    DevBufs b;
    float4 *src = nullptr, *dst = nullptr;
    MPM_HIP(b.alloc(&src, (size_t)bytes));
    MPM_HIP(b.alloc(&dst, (size_t)bytes));
    MPM_HIP(hipMemset(src, 0, (size_t)bytes));
    const int64_t n4 = bytes / 16;
    const unsigned wg = (unsigned)((n4 + 255) / 256);        // at most 2^21 for 8 GiB
    hipEvent_t e0 = nullptr, e1 = nullptr;
    MPM_HIP(hipEventCreate(&e0));
    hipError_t err = hipEventCreate(&e1);
    std::vector<float> ms;
    for (int r = 0; r <= reps && err == hipSuccess; ++r) {            // the first launch is a warm-up
        hipLaunchKernelGGL(k_probe_copy, dim3(wg), dim3(256), 0, 0, src, dst, n4);
        if ((err = hipGetLastError()) != hipSuccess) break;
        if ((err = hipEventRecord(e0, 0)) != hipSuccess) break;
        hipLaunchKernelGGL(k_probe_copy, dim3(wg), dim3(256), 0, 0, src, dst, n4);
        if ((err = hipEventRecord(e1, 0)) != hipSuccess) break;
        if ((err = hipEventSynchronize(e1)) != hipSuccess) break;
        float t = 0.f;
        if ((err = hipEventElapsedTime(&t, e0, e1)) != hipSuccess) break;
        if (r > 0) ms.push_back(t);
    }
    (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    MPM_HIP(err);
    std::sort(ms.begin(), ms.end());
    const double med = ms[ms.size() / 2];
    *gbs = 2.0 * (double)bytes / (med * 1e-3) / 1e9;
    return MPM_OK;
}

int mpm_probe_soa_bandwidth(int64_t n, int32_t planes_in, int32_t planes_out, int32_t reps, double *gbs)
{
    if (!gbs) return set_error(MPM_ERR_INVALID, "null argument");
    if (n < 1 || n > (1ll << 28)) return set_error(MPM_ERR_INVALID, "n out of range [1, 2^28]");
    if (planes_in < 1 || planes_in > 64 || planes_out < 1 || planes_out > 64)
        return set_error(MPM_ERR_INVALID, "planes out of range [1,64]");
    if (reps < 1 || reps > 100) return set_error(MPM_ERR_INVALID, "reps out of range [1,100]");
    DevBufs b;
    float *in = nullptr, *out = nullptr;
    MPM_HIP(b.alloc(&in, (size_t)n * planes_in * sizeof(float)));
    MPM_HIP(b.alloc(&out, (size_t)n * planes_out * sizeof(float)));
    MPM_HIP(hipMemset(in, 0, (size_t)n * planes_in * sizeof(float)));
    const unsigned wg = (unsigned)((n + 255) / 256);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    MPM_HIP(hipEventCreate(&e0));
    hipError_t err = hipEventCreate(&e1);
    std::vector<float> ms;
    for (int r = 0; r <= reps && err == hipSuccess; ++r) {            // the first launch is a warm-up
        if ((err = hipEventRecord(e0, 0)) != hipSuccess) break;
        hipLaunchKernelGGL(k_probe_soa, dim3(wg), dim3(256), 0, 0, in, out, n, planes_in, planes_out);
        if ((err = hipGetLastError()) != hipSuccess) break;
        if ((err = hipEventRecord(e1, 0)) != hipSuccess) break;
        if ((err = hipEventSynchronize(e1)) != hipSuccess) break;
        float t = 0.f;
        if ((err = hipEventElapsedTime(&t, e0, e1)) != hipSuccess) break;
        if (r > 0) ms.push_back(t);
    }
    (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    MPM_HIP(err);
    std::sort(ms.begin(), ms.end());
    const double med = ms[ms.size() / 2];
    *gbs = (double)n * (planes_in + planes_out) * sizeof(float) / (med * 1e-3) / 1e9;
    return MPM_OK;
}

int mpm_probe_lds_scatter(int32_t mode, int32_t reps, double *ms)
{
    if (!ms) return set_error(MPM_ERR_INVALID, "null argument");
    if (mode < 0 || mode > 3) return set_error(MPM_ERR_INVALID, "mode out of range [0,3]");
    if (reps < 1 || reps > 4096) return set_error(MPM_ERR_INVALID, "reps out of range [1,4096]");
    constexpr int wgs = 8192;
    DevBufs b;
    float *out = nullptr;
    MPM_HIP(b.alloc(&out, (size_t)wgs * 4 * LP_NODES * sizeof(float)));
    auto launch = [&](int rr) {
        switch (mode) {
        case 0: hipLaunchKernelGGL((k_probe_lds<true, false>), dim3(wgs), dim3(64), 0, 0, out, rr); break;
        case 1: hipLaunchKernelGGL((k_probe_lds<true, true>), dim3(wgs), dim3(64), 0, 0, out, rr); break;
        case 2: hipLaunchKernelGGL((k_probe_lds<false, false>), dim3(wgs), dim3(64), 0, 0, out, rr); break;
        default: hipLaunchKernelGGL((k_probe_lds<false, true>), dim3(wgs), dim3(64), 0, 0, out, rr); break;
        }
        return hipGetLastError();
    };
    MPM_HIP(launch(1));                                   // warm-up
    hipEvent_t e0 = nullptr, e1 = nullptr;
    MPM_HIP(hipEventCreate(&e0));
    hipError_t err = hipEventCreate(&e1);
    float t = 0.f;
    if (err == hipSuccess) err = hipEventRecord(e0, 0);
    if (err == hipSuccess) err = launch(reps);
    if (err == hipSuccess) err = hipEventRecord(e1, 0);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    if (err == hipSuccess) err = hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    MPM_HIP(err);
    *ms = t;
    return MPM_OK;
}

int mpm_probe_base_index(const float *x, int64_t n, float origin, float inv_h, int32_t *out)
{
    if ((!x || !out) && n > 0) return set_error(MPM_ERR_INVALID, "null argument");
    if (n <= 0) return MPM_OK;
    if (n > (1ll << 30)) return set_error(MPM_ERR_INVALID, "too many values");
    DevBufs b;
    float *dx = nullptr;
    int32_t *dout = nullptr;
    MPM_HIP(b.alloc(&dx, n * 4));
    MPM_HIP(b.alloc(&dout, n * 4));
    MPM_HIP(hipMemcpy(dx, x, n * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_probe_base, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, dx, (int)n, origin, inv_h,
                       dout);
    MPM_HIP(hipGetLastError());
    MPM_HIP(hipMemcpy(out, dout, n * 4, hipMemcpyDeviceToHost));
    return MPM_OK;
}

int mpm_probe_svd(const float *F, int64_t n, int32_t sweeps, float *U, float *s, float *V)
{
    if (n > 0 && (!F || !U || !s || !V)) return set_error(MPM_ERR_INVALID, "null argument");
    if (n <= 0) return MPM_OK;
    if (sweeps < 0 || sweeps > 16) return set_error(MPM_ERR_INVALID, "sweeps out of range [0,16]");
    if (n > (1ll << 26)) return set_error(MPM_ERR_INVALID, "too many matrices");
    DevBufs b;
    float *dF = nullptr, *dU = nullptr, *ds = nullptr, *dV = nullptr;
    uint32_t *cnt = nullptr;
    MPM_HIP(b.alloc(&dF, n * 36));
    MPM_HIP(b.alloc(&dU, n * 36));
    MPM_HIP(b.alloc(&dV, n * 36));
    MPM_HIP(b.alloc(&ds, n * 12));
    MPM_HIP(b.alloc(&cnt, CT_COUNT * 4));
    MPM_HIP(hipMemset(cnt, 0, CT_COUNT * 4));
    MPM_HIP(hipMemcpy(dF, F, n * 36, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_probe_svd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, dF, (int)n, sweeps, dU, ds,
                       dV, cnt);
    MPM_HIP(hipGetLastError());
    MPM_HIP(hipMemcpy(U, dU, n * 36, hipMemcpyDeviceToHost));
    MPM_HIP(hipMemcpy(V, dV, n * 36, hipMemcpyDeviceToHost));
    MPM_HIP(hipMemcpy(s, ds, n * 12, hipMemcpyDeviceToHost));
    return MPM_OK;
}

int mpm_probe_stress(const float *F, const float *Jp, int64_t n, int32_t type, double mu, double lam, double xi,
                     int32_t sweeps, float *tau)
{
    if (n > 0 && (!F || !tau || (type == MPM_SNOW && !Jp))) return set_error(MPM_ERR_INVALID, "null argument");
    if (type != MPM_JELLO && type != MPM_SNOW) return set_error(MPM_ERR_INVALID, "type must be JELLO or SNOW");
    if (n <= 0) return MPM_OK;
    if (sweeps < 1 || sweeps > 16) return set_error(MPM_ERR_INVALID, "sweeps out of range [1,16]");
    if (n > (1ll << 26)) return set_error(MPM_ERR_INVALID, "too many matrices");
    MatP m;
    memset(&m, 0, sizeof m);
    m.type = type;
    m.mu = (float)mu;
    m.lam = (float)lam;
    m.xi = (float)xi;
    DevBufs b;
    float *dF = nullptr, *dJ = nullptr, *dt = nullptr;
    uint32_t *cnt = nullptr;
    MPM_HIP(b.alloc(&dF, n * 36));
    MPM_HIP(b.alloc(&dJ, n * 4));
    MPM_HIP(b.alloc(&dt, n * 24));
    MPM_HIP(b.alloc(&cnt, CT_COUNT * 4));
    MPM_HIP(hipMemset(cnt, 0, CT_COUNT * 4));
    MPM_HIP(hipMemcpy(dF, F, n * 36, hipMemcpyHostToDevice));
    if (type == MPM_SNOW) MPM_HIP(hipMemcpy(dJ, Jp, n * 4, hipMemcpyHostToDevice));
    else MPM_HIP(hipMemset(dJ, 0, n * 4));
    const dim3 gr((unsigned)((n + 255) / 256)), bl(256);
    if (type == MPM_SNOW) hipLaunchKernelGGL(k_probe_stress<MPM_SNOW>, gr, bl, 0, 0, dF, dJ, (int)n, m, sweeps, dt, cnt);
    else hipLaunchKernelGGL(k_probe_stress<MPM_JELLO>, gr, bl, 0, 0, dF, dJ, (int)n, m, sweeps, dt, cnt);
    MPM_HIP(hipGetLastError());
    MPM_HIP(hipMemcpy(tau, dt, n * 24, hipMemcpyDeviceToHost));
    return MPM_OK;
}

int mpm_probe_svd_residual(const float *F, const float *q, int64_t n, int32_t sweeps, float *off)
{
    if (n > 0 && (!F || !off)) return set_error(MPM_ERR_INVALID, "null argument");
    if (n <= 0) return MPM_OK;
    if (sweeps < 0 || sweeps > 16) return set_error(MPM_ERR_INVALID, "sweeps out of range [0,16]");
    if (n > (1ll << 26)) return set_error(MPM_ERR_INVALID, "too many matrices");
    DevBufs b;
    float *dF = nullptr, *dq = nullptr, *doff = nullptr;
    MPM_HIP(b.alloc(&dF, n * 36));
    MPM_HIP(b.alloc(&doff, n * 4));
    MPM_HIP(hipMemcpy(dF, F, n * 36, hipMemcpyHostToDevice));
    if (q) {
        MPM_HIP(b.alloc(&dq, n * 16));
        MPM_HIP(hipMemcpy(dq, q, n * 16, hipMemcpyHostToDevice));
    }
    hipLaunchKernelGGL(k_probe_svd_residual, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, dF, dq, (int)n,
                       sweeps, doff);
    MPM_HIP(hipGetLastError());
    MPM_HIP(hipMemcpy(off, doff, n * 4, hipMemcpyDeviceToHost));
    return MPM_OK;
}

int mpm_probe_sort(const uint32_t *keys, int64_t n, int32_t bits, uint32_t *sorted, uint32_t *perm)
{
    if (n > 0 && (!keys || !sorted || !perm)) return set_error(MPM_ERR_INVALID, "null argument");
    return probe_sort(keys, n, bits, sorted, perm);
}

}  // extern "C"
