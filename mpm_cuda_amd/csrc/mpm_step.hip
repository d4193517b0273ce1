// mpm_step.hip -- HIP kernels of the explicit MPM step (gfx950).
//
// Reference: JerryYan97/MPM_CUDA src/simulator/MPMCUDA.cu -- P2G :310-433,
// VelUpdate :435-472, InterpolateAndMove :474-741, step :793-925.  Particles
// are FP32 structure-of-arrays; the grid is block-major (4^3 nodes per block)
// and only blocks touched by the current P2G are updated and read.
#include "mpm_g2p.h"
#include "mpm_launch.h"

namespace mpm {

// ---------------------------------------------------------------------------
// Small utilities.
__global__ void k_zero_u32(uint32_t *p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ---------------------------------------------------------------------------
// PLAIN form P2G: one thread per particle, 27 nodes x 4 FP32 global atomics,
// the stress SVD evaluated here (MPMCUDA.cu:310-433).
template <int MAT>
__global__ void __launch_bounds__(256) k_p2g_atomic(GridP g, MatP mat, PartView pv, float dt,
                                                    int sweeps, int tau_stored, float4 *grid,
                                                    uint32_t *flags, uint32_t *counters)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pv.n) return;
    dt = step_dt(g, dt);
    float x[3] = {ld(pv, PL_X, i), ld(pv, PL_X + 1, i), ld(pv, PL_X + 2, i)};
    Stencil s;
    uint32_t bits = make_stencil(g, x, s);
    Scatter sc;
    load_scatter<MAT>(pv, i, mat, dt, tau_stored != 0, sweeps, sc, counters);
#pragma unroll
    for (int ox = 0; ox < 3; ++ox)
#pragma unroll
        for (int oy = 0; oy < 3; ++oy)
#pragma unroll
            for (int oz = 0; oz < 3; ++oz) {
                float4 c = node_contrib(g, s, sc, ox, oy, oz);
                float *dst = (float *)(grid + node_slot(g, s.b[0] + ox, s.b[1] + oy, s.b[2] + oz));
                unsafeAtomicAdd(dst + 0, c.x);
                unsafeAtomicAdd(dst + 1, c.y);
                unsafeAtomicAdd(dst + 2, c.z);
                unsafeAtomicAdd(dst + 3, c.w);
            }
    mark_blocks(g, s.b, flags);
    if (bits) atomicOr(counters + CT_HEALTH, bits);
}

// ---------------------------------------------------------------------------
// Active-block list: compacts the block flags written by P2G (and clears them).
// One global atomic per workgroup (256 flags): wave counts are combined in LDS.
// Thread t of the workgroup handles flag b0 + t.  The workgroup of flag 0
// also moves the step's spill count (consumed by the spill kernel, which ran
// before) to CT_SPILLED and clears CT_SPILL for the next P2G, when
// clear_spill is set (else the grid update does it).
__device__ __forceinline__ void compact_flags(uint32_t *flags, int nblocks, uint32_t *list, uint32_t *counters,
                                              int b0, bool clear_spill)
{
    __shared__ uint32_t wave_cnt[4], wg_base;
    if (clear_spill && b0 == 0 && threadIdx.x == 0) {
        gp(counters)[CT_SPILLED] = gp(counters)[CT_SPILL];
        gp(counters)[CT_SPILL] = 0u;
    }
    int b = b0 + (int)threadIdx.x;
    bool f = b < nblocks && gp(flags)[b] != 0u;
    if (f) gp(flags)[b] = 0u;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t mask = __ballot(f);
    if (lane == 0) wave_cnt[w] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = wave_cnt[0] + wave_cnt[1] + wave_cnt[2] + wave_cnt[3];
        wg_base = tot ? gatomic_add(counters + CT_ACTIVE, tot) : 0u;
    }
    __syncthreads();
    uint32_t off = wg_base;
    for (int k = 0; k < w; ++k) off += wave_cnt[k];
    uint64_t below = lane ? (mask & ((~0ull) >> (64 - lane))) : 0ull;
    if (f) gp(list)[off + __popcll(below)] = (uint32_t)b;
}

__global__ void __launch_bounds__(256) k_compact_blocks(uint32_t *flags, int nblocks, uint32_t *list,
                                                        uint32_t *counters)
{
    compact_flags(flags, nblocks, list, counters, (int)(blockIdx.x * blockDim.x), true);
}

// ---------------------------------------------------------------------------
// Grid velocity update over active blocks (MPMCUDA.cu:435-472): v = p / m,
// gravity, walls.  One wave per active block, persistent over the active
// list.  Node (mass, momentum) is the sum of grid_mp (PLAIN-form scatter and
// FAST-form spills) and, in the FAST form, of the staged P2G arenas of the
// up to 2x2x2 particle blocks whose arenas cover the node.  grid_mp is
// cleared for the next step; the result goes to grid_v.
__device__ __forceinline__ void grid_block(const GridP &g, float dt, uint32_t b, float4 *gmp, float4 *gv,
                                           const float4 *__restrict__ staging, const int32_t *__restrict__ blk_tile,
                                           int ngroups, uint32_t staging_cap, int64_t nblocks)
{
    const int lane = threadIdx.x & 63;
    const int kx = lane & 3, ky = (lane >> 2) & 3, kz = lane >> 4;
    const int bx = (int)(b % (uint32_t)g.nbx);
    const int by = (int)((b / (uint32_t)g.nbx) % (uint32_t)g.nby);
    const int bz = (int)(b / ((uint32_t)g.nbx * (uint32_t)g.nby));
    const int64_t idx = (int64_t)b * BLK_NODES + lane;
    float4 a = gload4(gmp, idx);
    if (a.x != 0.f || a.y != 0.f || a.z != 0.f || a.w != 0.f) gstore4(gmp, idx, make_float4(0.f, 0.f, 0.f, 0.f));
    if (blk_tile) {
        for (int gi = 0; gi < ngroups; ++gi) {
            // lanes 0..26 look up the staging slots of the 27 neighbour
            // blocks at once; the gathers below are then independent
            const int32_t *bt = blk_tile + (int64_t)gi * nblocks;
            int32_t my_t = -1;
            if (lane < 27) {
                const int px = bx + lane % 3 - 1, py = by + (lane / 3) % 3 - 1, pz = bz + lane / 9 - 1;
                if (px >= 0 && py >= 0 && pz >= 0 && px < g.nbx && py < g.nby && pz < g.nbz) {
                    const int32_t t = gp(bt)[((int64_t)pz * g.nby + py) * g.nbx + px];
                    if (t >= 0 && ((uint32_t)t & TILE_SLOT_MASK) < staging_cap) my_t = t;
                }
            }
            // A node lies in exactly 8 of the 27 arenas: per axis its own
            // block's (d = 0) and the one below (d = -1, local k <= 2) or
            // above (d = +1, k = 3).  A neighbour's arena is read at index 7
            // (k = 2, below) or 0 (k = 3, above) only when its P2G wrote that
            // face (tile_entry); index 0 and 7 hold zero otherwise, and are
            // not written.  All loads are issued before any is used.
            const int sx = kx == 3 ? 1 : -1, sy = ky == 3 ? 1 : -1, sz = kz == 3 ? 1 : -1;
            float4 v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int dx = (e & 1) ? sx : 0, dy = (e & 2) ? sy : 0, dz = (e & 4) ? sz : 0;
                const int32_t t = __shfl(my_t, (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1));
                const int ax = kx + 1 - 4 * dx, ay = ky + 1 - 4 * dy, az = kz + 1 - 4 * dz;
                const uint32_t f = (uint32_t)t >> TILE_FACE_SHIFT;
                const bool face_ok = (ax != 0 || (f & 1u)) && (ax != 7 || (f & 2u)) && (ay != 0 || (f & 4u)) &&
                                     (ay != 7 || (f & 8u)) && (az != 0 || (f & 16u)) && (az != 7 || (f & 32u));
                const bool use = t >= 0 && face_ok;
                const int64_t ts = use ? (int64_t)((uint32_t)t & TILE_SLOT_MASK) : 0;
                v[e] = make_float4(0.f, 0.f, 0.f, 0.f);
                if (use) v[e] = gload4(staging, ts * ARENA_NODES + (az * ARENA + ay) * ARENA + ax);
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) { a.x += v[e].x; a.y += v[e].y; a.z += v[e].z; a.w += v[e].w; }
        }
    }
    float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!(fabsf(a.x) < 2.220446049250313e-16f)) {          // MPMCUDA.cu:444
        const int ix = bx * BLK + kx, iy = by * BLK + ky, iz = bz * BLK + kz;
        float vx = a.y / a.x + g.gx * dt;                  // MPMCUDA.cu:452-457
        float vy = a.z / a.x + g.gy * dt;
        float vz = a.w / a.x + g.gz * dt;
        bool wx = ix <= g.pad || ix >= g.nx - g.pad;       // MPMCUDA.cu:463-469
        bool wy = iy <= g.pad || iy >= g.ny - g.pad;
        bool wz = iz <= g.pad || iz >= g.nz - g.pad;
        if (g.bc_mode == MPM_BC_STICKY) {
            if (wx || wy || wz) vx = vy = vz = 0.f;
        } else {
            if (wx) vx = 0.f;
            if (wy) vy = 0.f;
            if (wz) vz = 0.f;
        }
        out = make_float4(vx, vy, vz, a.x);
    }
    gstore4(gv, idx, out);
}

__global__ void __launch_bounds__(64) k_grid_update(GridP g, float dt, const uint32_t *list, uint32_t *counters,
                                                    float4 *gmp, float4 *gv, const float4 *__restrict__ staging,
                                                    const int32_t *__restrict__ blk_tile, int ngroups,
                                                    uint32_t staging_cap, int64_t nblocks, int clear_spill)
{
    const uint32_t nact = counters[CT_ACTIVE];
    const uint32_t nact8 = round_xcd(nact);
    if (clear_spill && blockIdx.x == 0 && threadIdx.x == 0) {     // see compact_flags
        gp(counters)[CT_SPILLED] = gp(counters)[CT_SPILL];
        gp(counters)[CT_SPILL] = 0u;
    }
    dt = step_dt(g, dt);
    for (uint32_t k8 = blockIdx.x; k8 < nact8; k8 += gridDim.x) {
        const uint32_t k = xcd_spread(k8, nact8);
        if (k >= nact) continue;
        const uint32_t b = __builtin_amdgcn_readfirstlane(list[k]);
        grid_block(g, dt, b, gmp, gv, staging, blk_tile, ngroups, staging_cap, nblocks);
    }
}

// Nodes with mass after the grid update (statistics only; launched when
// timing is on).  One atomic per workgroup: same-address atomics from every
// wave of the grid update would serialize in L2.
__global__ void __launch_bounds__(256) k_count_nodes(const uint32_t *list, uint32_t *counters,
                                                     const float4 *__restrict__ gv)
{
    __shared__ uint32_t part[4];
    const uint32_t nact = counters[CT_ACTIVE];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t nodes = 0;
    for (uint32_t k = blockIdx.x * 4 + w; k < nact; k += gridDim.x * 4) {
        const float4 v = gv[(int64_t)list[k] * BLK_NODES + lane];
        nodes += (uint32_t)__popcll(__ballot(v.w != 0.f));
    }
    if (lane == 0) part[w] = nodes;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t t = part[0] + part[1] + part[2] + part[3];
        if (t) atomicAdd(counters + CT_NODES, t);
    }
}

int launch_count_nodes(hipStream_t st, const uint32_t *list, uint32_t *counters, const float4 *gv)
{
    hipLaunchKernelGGL(k_count_nodes, dim3(256), dim3(256), 0, st, list, counters, gv);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// G2P, deformation update and advection, one thread per particle
// (MPMCUDA.cu:474-741).  With write_tau the Kirchhoff stress of the updated
// state is stored for the next P2G, so a FAST-form step needs one SVD.
// With sd, the step-size controller's measures of the updated particle are
// returned in sd[0] (speed) and sd[1] (deformation).
// Particle i of pv goes to slot io of po (see g2p_core).
template <int MAT>
__device__ __forceinline__ void g2p_particle(const GridP &g, const MatP &mat, const PartView &pv, int i,
                                             const PartView &po, int io, float dt, const float4 *__restrict__ gv,
                                             bool write_tau, int sweeps, uint32_t *counters, float *sd = nullptr)
{
    G2POut o;
    GlobalFetch f{gv};
    uint32_t bits = g2p_core<MAT>(g, mat, pv, i, po, io, dt, f, write_tau && MAT != MPM_WATER, sweeps, counters, o);
    if (write_tau && MAT != MPM_WATER) {
#pragma unroll
        for (int k = 0; k < 6; ++k) st(po, PL_TAU + k, io, o.tau[k]);
    }
    if (bits) gatomic_or(counters + CT_HEALTH, bits);
    if (sd) dt_measures<MAT>(mat, o.v, o.S, sd[0], sd[1]);
}

// Jello and snow G2P fit 128 VGPRs without spills when asked for four waves
// per SIMD (the default allocation took 131-133, three waves; 1M snow G2P
// 45.1 vs 47.0 us, measured); water's fits by default, and asking spilled.
template <int MAT>
constexpr int kG2pWaves = MAT == MPM_WATER ? 1 : 4;

// With the step-size controller on (g.dtp), every wave also merges its
// particles' maxima into counters[CT_SPEED], counters[CT_DEFORM] for the
// next step's control.  With a permutation (the re-sort of this step, FAST
// form) thread j updates particle perm[j] of pv into slot j of po (the other
// particle buffer), with its id and, when copy_q, its warm-start quaternion:
// the G2P itself moves the particles into their new order.
template <int MAT, bool MOVE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kG2pWaves<MAT>))) k_g2p(GridP g, MatP mat, PartView pv, float dt,
                                             const float4 *__restrict__ gv, int write_tau,
                                             int sweeps, uint32_t *counters, const uint32_t *__restrict__ perm,
                                             PartView po, const uint32_t *__restrict__ pid, uint32_t *pid_out,
                                             int copy_q)
{
    const uint32_t nb = ((uint32_t)pv.n + blockDim.x - 1) / blockDim.x;
    const uint32_t blk = xcd_spread(blockIdx.x, round_xcd(nb));      // launched with round_xcd(nb) workgroups
    if (blk >= nb) return;                                             // whole workgroup
    const int j = (int)(blk * blockDim.x + threadIdx.x);
    int i = j;
    if (MOVE && j < pv.n) {
        const uint32_t src = gp(perm)[j];
        i = src < (uint32_t)pv.n ? (int)src : j;     // always a valid index for a permutation
        gp(pid_out)[pv.off + j] = gp(pid)[pv.off + i];
        if (copy_q) {
#pragma unroll
            for (int k = 0; k < 4; ++k) st(po, PL_Q + k, j, ld(pv, PL_Q + k, i));
        }
    }
    const PartView &out = MOVE ? po : pv;        // chosen at compile time (one instantiation per case)
    if (!g.dtp) {
        if (j < pv.n) g2p_particle<MAT>(g, mat, pv, i, out, j, dt, gv, write_tau != 0, sweeps, counters);
        return;
    }
    dt = step_dt(g, dt);
    float sd[2] = {0.0f, 0.0f};
    if (j < pv.n) g2p_particle<MAT>(g, mat, pv, i, out, j, dt, gv, write_tau != 0, sweeps, counters, sd);
    wave_max_to(counters + CT_SPEED, sd[0]);
    wave_max_to(counters + CT_DEFORM, sd[1]);
}

// Particles whose stencil left their block arena (listed by the block
// kernels): scatter with global atomics into grid_mp, which the grid update
// adds in.  Persistent over the list.
template <int MAT>
__device__ __forceinline__ void spill_particle(const GridP &g, const MatP &mat, const PartView &pv, int i, float dt,
                                               uint32_t *counters, float4 *grid, uint32_t *flags)
{
    float x[3] = {ld(pv, PL_X, i), ld(pv, PL_X + 1, i), ld(pv, PL_X + 2, i)};
    Stencil s;
    make_stencil(g, x, s);
    Scatter sc;
    load_scatter<MAT>(pv, i, mat, dt, true, 0, sc, counters);
#pragma unroll
    for (int ox = 0; ox < 3; ++ox)
#pragma unroll
        for (int oy = 0; oy < 3; ++oy)
#pragma unroll
            for (int oz = 0; oz < 3; ++oz) {
                float4 c = node_contrib(g, s, sc, ox, oy, oz);
                float *dst = (float *)(grid + node_slot(g, s.b[0] + ox, s.b[1] + oy, s.b[2] + oz));
                gatomic_add(dst + 0, c.x);
                gatomic_add(dst + 1, c.y);
                gatomic_add(dst + 2, c.z);
                gatomic_add(dst + 3, c.w);
            }
    if (flags) mark_blocks(g, s.b, flags);
}

// host_count (pinned host memory, or null) receives the spill count, then
// host_seq the tile launch number `seq` (a release store: a host that reads
// seq == its launch also reads that launch's count).  With clear_active the
// kernel also clears the active-block and node counters for the block
// compaction that follows (no kernel between P2G and it touches them), which
// saves the FAST form its step-start counter reset.
template <int MAT>
__global__ void __launch_bounds__(256) k_p2g_spill(GridP g, MatP mat, PartView pv, float dt,
                                                   const uint32_t *__restrict__ list, uint32_t *counters,
                                                   float4 *grid, uint32_t *flags, uint32_t list_cap,
                                                   uint32_t *host_count, uint32_t *host_seq, uint32_t seq,
                                                   int clear_active)
{
    const uint32_t listed = counters[CT_SPILL];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (host_count) {
            *host_count = listed;
            if (host_seq) __hip_atomic_store(host_seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        if (clear_active) {
            gp(counters)[CT_ACTIVE] = 0u;
            gp(counters)[CT_NODES] = 0u;
        }
    }
    const uint32_t n = min(listed, list_cap);
    dt = step_dt(g, dt);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int i = (int)list[k];
        if (i < 0 || i >= pv.n) continue;
        spill_particle<MAT>(g, mat, pv, i, dt, counters, grid, flags);
    }
}

// Spill scatter and block compaction in one launch (FAST form, one particle
// group): the tile P2G already flagged the blocks of the particles it listed,
// so the workgroups scatter their share of the list with atomics (into
// grid_mp, read only by the grid update after this kernel) and compact their
// 256 block flags without waiting on each other.  CT_SPILL is read by every
// workgroup here; the grid update clears it.
template <int MAT>
__global__ void __launch_bounds__(256) k_spill_compact(GridP g, MatP mat, PartView pv, float dt,
                                                       const uint32_t *__restrict__ spill, uint32_t spill_cap,
                                                       uint32_t *counters, float4 *grid, uint32_t *flags,
                                                       int nblocks, uint32_t *list, uint32_t *host_count,
                                                       uint32_t *host_seq, uint32_t seq)
{
    const uint32_t listed = counters[CT_SPILL];
    if (host_count && blockIdx.x == 0 && threadIdx.x == 0) {
        *host_count = listed;
        __hip_atomic_store(host_seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    const uint32_t n = min(listed, spill_cap);
    if (n) {
        dt = step_dt(g, dt);
        for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
            const int i = (int)spill[k];
            if (i < 0 || i >= pv.n) continue;
            spill_particle<MAT>(g, mat, pv, i, dt, counters, grid, nullptr);
        }
    }
    compact_flags(flags, nblocks, list, counters, (int)(blockIdx.x * blockDim.x), false);
}

// Kirchhoff stress of the current state, for the first P2G of the FAST form
// (and after the host rewrites F or J_p).
template <int MAT>
__global__ void k_compute_tau(MatP mat, PartView pv, int sweeps, uint32_t *counters)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pv.n) return;
    float F[9], t[6];
#pragma unroll
    for (int k = 0; k < 9; ++k) F[k] = ld(pv, PL_F + k, i);
    float Jp = (MAT == MPM_SNOW) ? ld(pv, PL_S, i) : 1.0f;
    stress_from_F<MAT>(mat, F, Jp, sweeps, t, counters);
#pragma unroll
    for (int k = 0; k < 6; ++k) st(pv, PL_TAU + k, i, t[k]);
}

// ---------------------------------------------------------------------------
// Batched kernels: the scenes of a batch (BatchSlice table in device memory,
// one particle group per scene) in one launch per stage.  A workgroup maps
// its work index to a scene through a SliceScan of the per-scene counts,
// reads that scene's parameters with uniform loads and runs the body of the
// single-scene kernel.  Materials are selected per scene at run time.

__global__ void k_batch_zero(const BatchSlice *__restrict__ tab, int nslices)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslices) return;
    uint32_t *c = tab[s].counters;
    gp(c)[CT_ACTIVE] = 0u;
    gp(c)[CT_NODES] = 0u;
    gp(c)[CT_SPILL] = 0u;
    // warm start: the particles the last G2P listed go to this step's
    // inference, and the list starts over for this step's G2P
    uint32_t *w = tab[s].warm_cnt;
    if (w) {
        gp(w)[0] = gp(w)[1];
        gp(w)[1] = 0u;
    }
}

template <int MAT>
__device__ __forceinline__ void spill_scene(const BatchSlice &b, uint32_t n, uint32_t tid, uint32_t stride)
{
    for (uint32_t k = tid; k < n; k += stride) {
        const int i = (int)gp(b.spill)[k];
        if (i < 0 || i >= b.pv.n) continue;
        spill_particle<MAT>(b.g, b.mat, b.pv, i, b.dt, b.counters, b.grid_mp, b.flags);
    }
}

// Spilled particles of every scene: the workgroups form groups of `per`,
// group q takes scenes q, q + groups, ... (the scenes' counter loads then
// overlap instead of following one another).
__global__ void __launch_bounds__(256) k_spill_batch(const BatchSlice *__restrict__ tab, int nslices)
{
    const uint32_t per = max(1u, gridDim.x / (uint32_t)nslices);
    const uint32_t groups = gridDim.x / per;
    if (blockIdx.x / per >= groups) return;
    const uint32_t tid = (blockIdx.x % per) * blockDim.x + threadIdx.x, stride = per * blockDim.x;
    for (uint32_t sid = blockIdx.x / per; sid < (uint32_t)nslices; sid += groups) {
        const BatchSlice &b = tab[sid];
        const uint32_t n = min(gp(b.counters)[CT_SPILL], (uint32_t)b.pv.n);
        if (n == 0u) continue;
        switch (b.mat.type) {
        case MPM_JELLO: spill_scene<MPM_JELLO>(b, n, tid, stride); break;
        case MPM_SNOW: spill_scene<MPM_SNOW>(b, n, tid, stride); break;
        default: spill_scene<MPM_WATER>(b, n, tid, stride); break;
        }
    }
}

// Active-block lists: workgroup w compacts 256 flags of one scene.
__global__ void __launch_bounds__(256) k_compact_batch(const BatchSlice *__restrict__ tab, int nslices)
{
    SliceScan sc;
    sc.build(nslices, [&](int s) { return (uint32_t)((tab[s].nblocks + 255) / 256); });
    if (blockIdx.x >= sc.total) return;
    const uint32_t sid = (uint32_t)__builtin_amdgcn_readfirstlane((int)sc.find(blockIdx.x));
    const BatchSlice &b = tab[sid];
    compact_flags(b.flags, (int)b.nblocks, b.list, b.counters, (int)((blockIdx.x - sc.start(sid)) * 256u), true);
}

// Grid update over the active blocks of every scene (one wave per block).
__global__ void __launch_bounds__(64) k_grid_batch(const BatchSlice *__restrict__ tab, int nslices)
{
    SliceScan sc;
    sc.build(nslices, [&](int s) { return gp(tab[s].counters)[CT_ACTIVE]; });
    const uint32_t n = sc.total, n8 = round_xcd(n);
    for (uint32_t k8 = blockIdx.x; k8 < n8; k8 += gridDim.x) {
        const uint32_t k = xcd_spread(k8, n8);
        if (k >= n) continue;
        const uint32_t sid = (uint32_t)__builtin_amdgcn_readfirstlane((int)sc.find(k));
        const BatchSlice &b = tab[sid];
        const uint32_t blk = (uint32_t)__builtin_amdgcn_readfirstlane((int)gp(b.list)[k - sc.start(sid)]);
        grid_block(b.g, b.dt, blk, b.grid_mp, b.grid_v, b.staging, b.blk_tile, 1, b.staging_cap, b.nblocks);
    }
}

// G2P of every particle of the scenes of material MAT (one launch per
// material, as the batched P2G); each scene's particles fill whole
// workgroups.  Launched with round_xcd(sum of ceil(n / blockDim)) workgroups
// over those scenes.
template <int MAT>
__global__ void __launch_bounds__(256) k_g2p_batch(const BatchSlice *__restrict__ tab, int nslices)
{
    SliceScan sc;
    sc.build(nslices, [&](int s) {
        return tab[s].mat.type == MAT ? ((uint32_t)tab[s].pv.n + blockDim.x - 1) / blockDim.x : 0u;
    });
    const uint32_t blk = xcd_spread(blockIdx.x, round_xcd(sc.total));
    if (blk >= sc.total) return;
    const uint32_t sid = (uint32_t)__builtin_amdgcn_readfirstlane((int)sc.find(blk));
    const BatchSlice &b = tab[sid];
    const int i = (int)((blk - sc.start(sid)) * blockDim.x + threadIdx.x);
    if (i >= b.pv.n) return;
    g2p_particle<MAT>(b.g, b.mat, b.pv, i, b.pv, i, b.dt, b.grid_v, true, b.sweeps, b.counters);
}

// A batch's re-sort move: every scene's particles from the buffers of table
// `src` into those of table `dst`, slot j taking particle perm[j] of the
// scene's last sort (all planes, the quaternions when copy_q, the ids); one
// thread per particle, each workgroup's 256 particles in one scene.
__global__ void __launch_bounds__(256) k_permute_batch(const BatchSlice *__restrict__ src,
                                                       const BatchSlice *__restrict__ dst, int nslices)
{
    SliceScan sc;
    sc.build(nslices, [&](int s) { return ((uint32_t)src[s].pv.n + 255u) / 256u; });
    const uint32_t blk = xcd_spread(blockIdx.x, round_xcd(sc.total));
    if (blk >= sc.total) return;
    const uint32_t sid = (uint32_t)__builtin_amdgcn_readfirstlane((int)sc.find(blk));
    const PartView a = src[sid].pv, b = dst[sid].pv;
    const int j = (int)((blk - sc.start(sid)) * 256u + threadIdx.x);
    if (j >= a.n) return;
    uint32_t from = gp(src[sid].perm)[j];
    if (from >= (uint32_t)a.n) from = (uint32_t)j;        // never taken for a valid permutation
    const bool q = src[sid].copy_q != 0;
    for (int p = 0; p < PL_COUNT; ++p)
        if (q || p < PL_Q || p >= PL_Q + 4) st(b, p, j, ld(a, p, (int)from));
    gp(b.pid)[b.off + j] = gp(a.pid)[a.off + (int64_t)from];
}

// ---------------------------------------------------------------------------
// Host-side launchers.
static inline int nblk(int64_t n, int t) { return (int)((n + t - 1) / t); }

int launch_zero_u32(hipStream_t st, uint32_t *p, int n)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_zero_u32, dim3(nblk(n, 256)), dim3(256), 0, st, p, n);
    return (int)hipGetLastError();
}

int launch_p2g_atomic(hipStream_t st, const GridP &g, const MatP &m, const PartView &pv, float dt,
                      int sweeps, int tau_stored, float4 *grid, uint32_t *flags, uint32_t *counters,
                      int threads)
{
    if (pv.n <= 0) return 0;
    dim3 gr(nblk(pv.n, threads)), bl(threads);
    switch (m.type) {
    case MPM_JELLO: hipLaunchKernelGGL(k_p2g_atomic<MPM_JELLO>, gr, bl, 0, st, g, m, pv, dt, sweeps, tau_stored, grid, flags, counters); break;
    case MPM_SNOW: hipLaunchKernelGGL(k_p2g_atomic<MPM_SNOW>, gr, bl, 0, st, g, m, pv, dt, sweeps, tau_stored, grid, flags, counters); break;
    default: hipLaunchKernelGGL(k_p2g_atomic<MPM_WATER>, gr, bl, 0, st, g, m, pv, dt, sweeps, tau_stored, grid, flags, counters); break;
    }
    return (int)hipGetLastError();
}

int launch_compact(hipStream_t st, uint32_t *flags, int nblocks, uint32_t *list, uint32_t *counters)
{
    hipLaunchKernelGGL(k_compact_blocks, dim3(nblk(nblocks, 256)), dim3(256), 0, st, flags, nblocks, list,
                       counters);
    return (int)hipGetLastError();
}

int launch_grid_update(hipStream_t st, const GridP &g, float dt, const uint32_t *list,
                       uint32_t *counters, float4 *gmp, float4 *gv, int nblocks, const float4 *staging,
                       const int32_t *blk_tile, int ngroups, uint32_t staging_cap, int clear_spill)
{
    int wg = nblocks < 8192 ? nblocks : 8192;
    if (wg < 1) wg = 1;
    hipLaunchKernelGGL(k_grid_update, dim3(wg), dim3(64), 0, st, g, dt, list, counters, gmp, gv, staging, blk_tile,
                       ngroups, staging_cap, (int64_t)nblocks, clear_spill);
    return (int)hipGetLastError();
}

int launch_g2p(hipStream_t st, const GridP &g, const MatP &m, const PartView &pv, float dt,
               const float4 *gv, int write_tau, int sweeps, uint32_t *counters, int threads, const G2PMove *mv)
{
    if (pv.n <= 0) return 0;
    dim3 gr(round_xcd((uint32_t)nblk(pv.n, threads))), bl(threads);
    const G2PMove none{nullptr, pv, nullptr, nullptr, 0};
    const G2PMove &v = mv ? *mv : none;
    // the plain update and the one that moves particles are separate
    // instantiations: the move's index and buffer selection left the plain
    // kernel's registers and load schedule worse (G2P 41.2 vs 40.2 us median)
#define MPM_G2P_LAUNCH(MAT_)                                                                                    \
    do {                                                                                                        \
        if (mv) hipLaunchKernelGGL((k_g2p<MAT_, true>), gr, bl, 0, st, g, m, pv, dt, gv, write_tau, sweeps,     \
                                   counters, v.perm, v.po, v.pid, v.pid_out, v.copy_q);                         \
        else hipLaunchKernelGGL((k_g2p<MAT_, false>), gr, bl, 0, st, g, m, pv, dt, gv, write_tau, sweeps,       \
                                counters, v.perm, v.po, v.pid, v.pid_out, v.copy_q);                            \
    } while (0)
    switch (m.type) {
    case MPM_JELLO: MPM_G2P_LAUNCH(MPM_JELLO); break;
    case MPM_SNOW: MPM_G2P_LAUNCH(MPM_SNOW); break;
    default: MPM_G2P_LAUNCH(MPM_WATER); break;
    }
#undef MPM_G2P_LAUNCH
    return (int)hipGetLastError();
}

int launch_p2g_spill(hipStream_t st, const GridP &g, const MatP &m, const PartView &pv, float dt,
                     const uint32_t *list, uint32_t list_cap, uint32_t *counters, float4 *grid, uint32_t *flags,
                     uint32_t *host_count, uint32_t *host_seq, uint32_t seq, int clear_active)
{
    if (pv.n <= 0) return 0;
    dim3 gr(64), bl(256);
    switch (m.type) {
    case MPM_JELLO: hipLaunchKernelGGL(k_p2g_spill<MPM_JELLO>, gr, bl, 0, st, g, m, pv, dt, list, counters, grid, flags, list_cap, host_count, host_seq, seq, clear_active); break;
    case MPM_SNOW: hipLaunchKernelGGL(k_p2g_spill<MPM_SNOW>, gr, bl, 0, st, g, m, pv, dt, list, counters, grid, flags, list_cap, host_count, host_seq, seq, clear_active); break;
    default: hipLaunchKernelGGL(k_p2g_spill<MPM_WATER>, gr, bl, 0, st, g, m, pv, dt, list, counters, grid, flags, list_cap, host_count, host_seq, seq, clear_active); break;
    }
    return (int)hipGetLastError();
}

int launch_compute_tau(hipStream_t st, const MatP &m, const PartView &pv, int sweeps, uint32_t *counters)
{
    if (pv.n <= 0 || m.type == MPM_WATER) return 0;
    dim3 gr(nblk(pv.n, 256)), bl(256);
    if (m.type == MPM_JELLO)
        hipLaunchKernelGGL(k_compute_tau<MPM_JELLO>, gr, bl, 0, st, m, pv, sweeps, counters);
    else
        hipLaunchKernelGGL(k_compute_tau<MPM_SNOW>, gr, bl, 0, st, m, pv, sweeps, counters);
    return (int)hipGetLastError();
}

int launch_batch_zero(hipStream_t st, const BatchSlice *tab, int nslices)
{
    hipLaunchKernelGGL(k_batch_zero, dim3(nblk(nslices, 64)), dim3(64), 0, st, tab, nslices);
    return (int)hipGetLastError();
}

int launch_permute_batch(hipStream_t st, const BatchSlice *src, const BatchSlice *dst, int nslices, int wgs)
{
    if (wgs <= 0) return 0;
    hipLaunchKernelGGL(k_permute_batch, dim3(wgs), dim3(256), 0, st, src, dst, nslices);
    return (int)hipGetLastError();
}

int launch_spill_batch(hipStream_t st, const BatchSlice *tab, int nslices, int wgs)
{
    hipLaunchKernelGGL(k_spill_batch, dim3(wgs), dim3(256), 0, st, tab, nslices);
    return (int)hipGetLastError();
}

int launch_compact_batch(hipStream_t st, const BatchSlice *tab, int nslices, int wgs)
{
    hipLaunchKernelGGL(k_compact_batch, dim3(wgs), dim3(256), 0, st, tab, nslices);
    return (int)hipGetLastError();
}

int launch_grid_batch(hipStream_t st, const BatchSlice *tab, int nslices, int wgs)
{
    hipLaunchKernelGGL(k_grid_batch, dim3(wgs), dim3(64), 0, st, tab, nslices);
    return (int)hipGetLastError();
}

int launch_g2p_batch(hipStream_t st, int mat, const BatchSlice *tab, int nslices, int wgs, int threads)
{
    switch (mat) {
    case MPM_JELLO: hipLaunchKernelGGL(k_g2p_batch<MPM_JELLO>, dim3(wgs), dim3(threads), 0, st, tab, nslices); break;
    case MPM_SNOW: hipLaunchKernelGGL(k_g2p_batch<MPM_SNOW>, dim3(wgs), dim3(threads), 0, st, tab, nslices); break;
    default: hipLaunchKernelGGL(k_g2p_batch<MPM_WATER>, dim3(wgs), dim3(threads), 0, st, tab, nslices); break;
    }
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// Step-size controller (no counterpart in the reference, which computes
// min(10, max_p max(J, 1/J)^3) after every step and never uses it,
// MPMCUDA.cu:915-920).  k_dt_scan reduces the controller's maxima over the
// current state (first step, or after the host rewrote particles); after
// that every G2P reduces them over the state it writes.  k_dt_control runs
// first in each step: it turns the maxima into the step's dt, advances the
// clock and clears the maxima for this step's G2P.
template <int MAT>
__global__ void __launch_bounds__(256) k_dt_scan(MatP mat, PartView pv, uint32_t *counters)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    float sd[2] = {0.0f, 0.0f};
    if (i < pv.n) {
        const float v[3] = {ld(pv, PL_V, i), ld(pv, PL_V + 1, i), ld(pv, PL_V + 2, i)};
        float J;
        if (MAT == MPM_JELLO) {
            float F[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) F[k] = ld(pv, PL_F + k, i);
            J = det3<float>(F);
        } else {
            J = ld(pv, PL_S, i);
        }
        dt_measures<MAT>(mat, v, J, sd[0], sd[1]);
    }
    wave_max_to(counters + CT_SPEED, sd[0]);
    wave_max_to(counters + CT_DEFORM, sd[1]);
}

__global__ void __launch_bounds__(64) k_dt_control(DtCtl *c, uint32_t *counters, float max_dt, float cfl_h, int mode)
{
    if (threadIdx.x != 0) return;
    const float speed = __uint_as_float(gp(counters)[CT_SPEED]);
    const float deform = __uint_as_float(gp(counters)[CT_DEFORM]);
    float dt = max_dt;
    if (speed > 0.0f) dt = fminf(dt, cfl_h / speed);
    if (mode == MPM_DT_CFL_DEFORM) {
        const float f = fminf(10.0f, fmaxf(1.0f, deform * deform * deform));
        dt = fminf(dt, max_dt / f);
    }
    gp(c)->dt = dt;
    gp(c)->last = (double)dt;
    gp(c)->time = gp(c)->time + (double)dt;
    gp(counters)[CT_SPEED] = 0u;
    gp(counters)[CT_DEFORM] = 0u;
}

int launch_spill_compact(hipStream_t st, const GridP &g, const MatP &m, const PartView &pv, float dt,
                         const uint32_t *spill, uint32_t spill_cap, uint32_t *counters, float4 *grid,
                         uint32_t *flags, int nblocks, uint32_t *list, uint32_t *host_count, uint32_t *host_seq,
                         uint32_t seq)
{
    if (!host_count || !host_seq) return (int)hipErrorInvalidValue;
    dim3 gr(nblk(nblocks, 256)), bl(256);
    switch (m.type) {
    case MPM_JELLO: hipLaunchKernelGGL(k_spill_compact<MPM_JELLO>, gr, bl, 0, st, g, m, pv, dt, spill, spill_cap, counters, grid, flags, nblocks, list, host_count, host_seq, seq); break;
    case MPM_SNOW: hipLaunchKernelGGL(k_spill_compact<MPM_SNOW>, gr, bl, 0, st, g, m, pv, dt, spill, spill_cap, counters, grid, flags, nblocks, list, host_count, host_seq, seq); break;
    default: hipLaunchKernelGGL(k_spill_compact<MPM_WATER>, gr, bl, 0, st, g, m, pv, dt, spill, spill_cap, counters, grid, flags, nblocks, list, host_count, host_seq, seq); break;
    }
    return (int)hipGetLastError();
}

int launch_dt_scan(hipStream_t st, const MatP &m, const PartView &pv, uint32_t *counters)
{
    if (pv.n <= 0) return 0;
    dim3 gr(nblk(pv.n, 256)), bl(256);
    switch (m.type) {
    case MPM_JELLO: hipLaunchKernelGGL(k_dt_scan<MPM_JELLO>, gr, bl, 0, st, m, pv, counters); break;
    case MPM_SNOW: hipLaunchKernelGGL(k_dt_scan<MPM_SNOW>, gr, bl, 0, st, m, pv, counters); break;
    default: hipLaunchKernelGGL(k_dt_scan<MPM_WATER>, gr, bl, 0, st, m, pv, counters); break;
    }
    return (int)hipGetLastError();
}

int launch_dt_control(hipStream_t st, DtCtl *c, uint32_t *counters, float max_dt, float cfl_h, int mode)
{
    hipLaunchKernelGGL(k_dt_control, dim3(1), dim3(64), 0, st, c, counters, max_dt, cfl_h, mode);
    return (int)hipGetLastError();
}

}  // namespace mpm
