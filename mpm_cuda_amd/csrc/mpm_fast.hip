// mpm_fast.hip -- FAST execution form: cell-block sort and LDS-blocked P2G.
//
// No counterpart in the reference (JerryYan97/MPM_CUDA scatters unsorted
// particles with global FP64 atomics, MPMCUDA.cu:388-421).  Pieces:
//   * exclusive scan (three-kernel, wave-shuffle based);
//   * stable LSD radix sort of 32-bit keys, 8-bit digits, ballot-matched
//     in-wave ranking;
//   * per-group sort of particles by (4x4x4 block, cell) of their base
//     node, payload permutation of every SoA plane, and one tile (block,
//     first particle, count, staging slot) per non-empty block;
//   * P2G over tiles without atomics: one wave per block scatters into an
//     8^3-node LDS arena by plain read-modify-write, lane l owning cell l,
//     and writes the arena to a staging slot that the grid update gathers.
// (On gfx950 an LDS float atomic runs ~20x slower than a plain LDS
// read-modify-write, and global float atomics at ~77 G/s would dominate a
// per-node flush.)
#include "mpm_g2p.h"
#include "mpm_launch.h"

namespace mpm {

PartView view_of(mpm_sim *s, int g);

// ---------------------------------------------------------------------------
// Scan helpers.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d);
        if (lane >= d) v += t;
    }
    return v;
}

// Exclusive scan of one value per thread across the workgroup (<= 1024 threads).
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t &total, uint32_t *lds)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) lds[wid] = inc;
    __syncthreads();
    if (threadIdx.x < 64) {
        uint32_t w = (int)threadIdx.x < nw ? lds[threadIdx.x] : 0u;
        uint32_t wi = wave_incl_scan(w);
        if ((int)threadIdx.x < nw) lds[16 + threadIdx.x] = wi - w;
        if ((int)threadIdx.x == nw - 1) lds[32] = wi;
    }
    __syncthreads();
    uint32_t res = lds[16 + wid] + inc - v;
    total = lds[32];
    __syncthreads();
    return res;
}

constexpr int SCAN_T = 256, SCAN_ITEMS = 8, SCAN_TILE = SCAN_T * SCAN_ITEMS;

__global__ void __launch_bounds__(SCAN_T) k_scan_reduce(const uint32_t *in, int n, uint32_t *partial)
{
    __shared__ uint32_t lds[40];
    int base = blockIdx.x * SCAN_TILE;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        int i = base + k * SCAN_T + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint32_t tot;
    block_excl_scan(s, tot, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(1024) k_scan_partials(uint32_t *partial, int np)
{
    __shared__ uint32_t lds[40];
    uint32_t carry = 0;
    for (int c = 0; c < np; c += 1024) {
        int i = c + threadIdx.x;
        uint32_t v = i < np ? partial[i] : 0u;
        uint32_t tot;
        uint32_t ex = block_excl_scan(v, tot, lds);
        if (i < np) partial[i] = carry + ex;
        carry += tot;
    }
}

__global__ void __launch_bounds__(SCAN_T) k_scan_apply(const uint32_t *in, uint32_t *out, int n,
                                                       const uint32_t *partial, uint32_t *total_out)
{
    __shared__ uint32_t lds[40];
    int i0 = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = (i0 + k < n) ? in[i0 + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = partial[blockIdx.x] + block_excl_scan(s, tot, lds);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        if (i0 + k < n) {
            out[i0 + k] = run;
            if (total_out && i0 + k == n - 1) *total_out = run + v[k];
        }
        run += v[k];
    }
}

// out = exclusive scan of in (in-place allowed); partial holds ceil(n / 2048) words.
static int scan_excl(hipStream_t st, const uint32_t *in, uint32_t *out, int n, uint32_t *partial,
                     uint32_t *total_out)
{
    int nt = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nt == 0) return 0;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nt), dim3(SCAN_T), 0, st, in, n, partial);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, st, partial, nt);
    hipLaunchKernelGGL(k_scan_apply, dim3(nt), dim3(SCAN_T), 0, st, in, out, n, partial, total_out);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort, 8-bit digits.  A tile of RS_TILE keys is handled by
// 4 waves, each owning a contiguous run of 1024 keys processed 64 at a time;
// equal digits inside a wave are matched with 8 ballots, so ranks follow
// input order and the sort is stable.
constexpr int RS_T = 256, RS_WAVE_KEYS = 1024, RS_ITERS = RS_WAVE_KEYS / 64, RS_TILE = 4 * RS_WAVE_KEYS;

__global__ void __launch_bounds__(RS_T) k_rs_hist(const uint32_t *keys, int n, int shift, uint32_t *hist,
                                                  int ntiles)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    int base = blockIdx.x * RS_TILE;
    for (int k = 0; k < RS_TILE / RS_T; ++k) {
        int i = base + k * RS_T + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(RS_T) k_rs_scatter(const uint32_t *kin, const uint32_t *vin, uint32_t *kout,
                                                     uint32_t *vout, int n, int shift, const uint32_t *hist,
                                                     int ntiles)
{
    __shared__ uint32_t cnt[4][256];
    for (int k = 0; k < 4; ++k) cnt[k][threadIdx.x] = 0;
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    uint32_t key[RS_ITERS], val[RS_ITERS], rank[RS_ITERS];
#pragma unroll
    for (int it = 0; it < RS_ITERS; ++it) {
        int i = blockIdx.x * RS_TILE + w * RS_WAVE_KEYS + it * 64 + lane;
        bool valid = i < n;
        uint32_t k = valid ? kin[i] : 0u;
        uint32_t d = (k >> shift) & 255u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            bool bit = (d >> b) & 1u;
            uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        uint32_t below = (uint32_t)__popcll(m & lt);
        uint32_t prev = 0;
        if (valid) prev = cnt[w][d];
        if (valid && below == 0) cnt[w][d] = prev + (uint32_t)__popcll(m);
        key[it] = k;
        val[it] = valid ? vin[i] : 0u;
        rank[it] = prev + below;
    }
    __syncthreads();
    {
        uint32_t run = 0;
        for (int k = 0; k < 4; ++k) {
            uint32_t t = cnt[k][threadIdx.x];
            cnt[k][threadIdx.x] = run;
            run += t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < RS_ITERS; ++it) {
        int i = blockIdx.x * RS_TILE + w * RS_WAVE_KEYS + it * 64 + lane;
        if (i < n) {
            uint32_t d = (key[it] >> shift) & 255u;
            uint32_t dst = hist[(int64_t)d * ntiles + blockIdx.x] + cnt[w][d] + rank[it];
            if (dst < (uint32_t)n) {      // always true for a consistent histogram
                kout[dst] = key[it];
                vout[dst] = val[it];
            }
        }
    }
}

// Sorts (keys, vals) of length n by the low `bits` key bits.  Returns through
// `in_alt` whether the result ended in the *_alt buffers.
static int radix_sort(hipStream_t st, uint32_t *keys, uint32_t *vals, uint32_t *keys_alt, uint32_t *vals_alt,
                      int n, int bits, uint32_t *hist, uint32_t *partial, bool &in_alt)
{
    in_alt = false;
    if (n <= 1) return 0;
    int ntiles = (n + RS_TILE - 1) / RS_TILE;
    uint32_t *ka = keys, *va = vals, *kb = keys_alt, *vb = vals_alt;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(RS_T), 0, st, ka, n, shift, hist, ntiles);
        int rc = scan_excl(st, hist, hist, 256 * ntiles, partial, nullptr);
        if (rc) return rc;
        hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(RS_T), 0, st, ka, va, kb, vb, n, shift, hist, ntiles);
        std::swap(ka, kb);
        std::swap(va, vb);
        in_alt = !in_alt;
    }
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// Keys, permutation and tiles.
//
// Sort key: (linear index of the 4^3-node block holding the particle's base
// node) << 6 | (cell of the base inside that block).  A stable sort leaves
// every block's particles grouped by cell, which P2G uses to give each lane
// of a wave its own cell.
constexpr int KEY_LOW = 6;

__device__ __forceinline__ void base_of(const GridP &g, const PartView &pv, int i, int b[3])
{
    const float o[3] = {g.ox, g.oy, g.oz};
    const int nn[3] = {g.nx, g.ny, g.nz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int v = base_index_f32(ld(pv, PL_X + a, i), o[a], g.inv_h);
        b[a] = v > nn[a] - 3 ? nn[a] - 3 : v;
    }
}

__global__ void k_block_keys(GridP g, PartView pv, uint32_t *keys, uint32_t *vals)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pv.n) return;
    int b[3];
    base_of(g, pv, i, b);
    uint32_t blk = ((uint32_t)(b[2] >> 2) * (uint32_t)g.nby + (uint32_t)(b[1] >> 2)) * (uint32_t)g.nbx +
                   (uint32_t)(b[0] >> 2);
    uint32_t cell = (uint32_t)(((b[2] & 3) * BLK + (b[1] & 3)) * BLK + (b[0] & 3));
    keys[i] = (blk << KEY_LOW) | cell;
    vals[i] = (uint32_t)i;
}

// Gathers every permuted SoA plane, the particle ids and the sort-time cell.
__global__ void k_permute(const float *__restrict__ P, float *__restrict__ Pa, int64_t cap, int64_t off, int n,
                          const uint32_t *__restrict__ perm, const uint32_t *__restrict__ skeys,
                          const uint32_t *__restrict__ pid, uint32_t *__restrict__ pida, uint8_t *__restrict__ cellkey,
                          uint32_t plane_mask)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t src = perm[j];
    if (src >= (uint32_t)n) src = (uint32_t)j;    // never taken for a valid permutation
    for (int p = 0; p < PL_COUNT; ++p)
        if ((plane_mask >> p) & 1u) Pa[(int64_t)p * cap + off + j] = P[(int64_t)p * cap + off + src];
    pida[off + j] = pid[off + src];
    cellkey[off + j] = (uint8_t)(skeys[j] & ((1u << KEY_LOW) - 1u));
}

// First and one-past-last sorted position of every block (untouched blocks
// keep the zeros written before).
__global__ void k_block_runs(const uint32_t *__restrict__ skeys, int n, uint32_t *blk_start, uint32_t *blk_end)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t b = skeys[i] >> KEY_LOW;
    if (i == 0 || (skeys[i - 1] >> KEY_LOW) != b) blk_start[b] = (uint32_t)i;
    if (i == n - 1 || (skeys[i + 1] >> KEY_LOW) != b) blk_end[b] = (uint32_t)(i + 1);
}

__global__ void k_tile_flags(const uint32_t *blk_start, const uint32_t *blk_end, uint32_t *has, int nblocks)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nblocks) has[b] = blk_end[b] > blk_start[b] ? 1u : 0u;
}

// One tile per non-empty block: (first particle, count, block, tile id).
// blk_tile maps the block to its global tile id (staging slot).
__global__ void k_emit_tiles(const uint32_t *blk_start, const uint32_t *blk_end, const uint32_t *tile_index,
                             int nblocks, uint4 *desc, uint32_t desc_cap, int32_t *blk_tile, uint32_t tile_base)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblocks) return;
    uint32_t c = blk_end[b] - blk_start[b];
    uint32_t t = tile_index[b];
    if (c > 0 && t < desc_cap) {
        desc[t] = make_uint4(blk_start[b], c, (uint32_t)b, tile_base + t);
        blk_tile[b] = (int32_t)(tile_base + t);
    } else {
        blk_tile[b] = -1;
    }
}

// ---------------------------------------------------------------------------
// P2G over blocks without atomics.
//
// One wave owns one block's particles.  Lane l takes cell l of the block;
// in round r it scatters the r-th particle of that cell, so the lanes of a
// round normally have 64 different base nodes and the 27 stencil updates of
// one instruction touch 64 different LDS slots: a plain read-modify-write is
// race-free.  Updates of different stencil offsets can alias across lanes,
// so they are issued in program order (compiler barrier between offsets; LDS
// executes a wave's operations in order).  Particles that drifted since the
// sort may share a base: lanes are matched on their base and same-base lanes
// take turns.  The arena (8^3 nodes: the block's 4 + 1 below + 3 above per
// axis) is written to the tile's staging slot with plain stores; the grid
// update gathers the staging slots overlapping each node block.  Particles
// whose stencil leaves the arena are listed for the atomic fallback kernel.
__device__ __forceinline__ uint64_t match_key9(uint32_t key)
{
    uint64_t m = ~0ull;
#pragma unroll
    for (int b = 0; b < 9; ++b) {
        bool bit = (key >> b) & 1u;
        uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        uint32_t o = __shfl_xor(v, d);
        v = v > o ? v : o;
    }
    return v;
}

// A tile (one block's particles) is processed by TILE_WAVES waves: wave w
// takes rounds w, w + TILE_WAVES, ... and scatters into its own LDS arena,
// so no two waves ever write the same slot; the arenas are summed when the
// tile is written back.  (Measured on 1M-particle snow: 1, 2 and 4 waves per
// tile run within 10% of each other; 2 is the fastest for both forms.)
constexpr int TILE_WAVES = 2;
constexpr int TILE_THREADS = 64 * TILE_WAVES;

// Per-tile set-up shared by the block kernels: clear the scatter arenas and
// find the [start, end) of every cell's particles (sorted by cell).
struct TileCtx {
    int bx, by, bz, ax, ay, az;
    uint32_t my0, myn, rounds;
};

__device__ __forceinline__ TileCtx tile_begin(const GridP &g, const PartView &pv, const uint8_t *__restrict__ cellkey,
                                              const uint4 &d, float4 (*arena)[A_SLOTS], uint32_t *cstart,
                                              uint32_t *cend)
{
    const int tid = threadIdx.x, lane = tid & 63;
    for (int k = tid; k < TILE_WAVES * A_SLOTS; k += TILE_THREADS)
        arena[k / A_SLOTS][k % A_SLOTS] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (tid < 64) { cstart[tid] = 0u; cend[tid] = 0u; }
    __syncthreads();
    for (uint32_t p = tid; p < d.y; p += TILE_THREADS) {
        uint32_t c = cellkey[pv.off + d.x + p];
        uint32_t cp = p > 0 ? cellkey[pv.off + d.x + p - 1] : 0xffu;
        uint32_t cn = p + 1 < d.y ? cellkey[pv.off + d.x + p + 1] : 0xffu;
        if (c != cp) cstart[c & 63] = p;
        if (c != cn) cend[c & 63] = p + 1;
    }
    __syncthreads();
    TileCtx t;
    t.my0 = min(cstart[lane], d.y);
    const uint32_t my1 = min(cend[lane], d.y);
    t.myn = my1 > t.my0 ? my1 - t.my0 : 0u;
    t.rounds = wave_max_u32(t.myn);
    t.bx = (int)(d.z % (uint32_t)g.nbx);
    t.by = (int)((d.z / (uint32_t)g.nbx) % (uint32_t)g.nby);
    t.bz = (int)(d.z / ((uint32_t)g.nbx * (uint32_t)g.nby));
    t.ax = t.bx * BLK - 1; t.ay = t.by * BLK - 1; t.az = t.bz * BLK - 1;
    return t;
}

// One round of the conflict-free scatter into this wave's LDS arena.  Lanes
// whose stencil leaves the arena (or whose tile has no staging slot) are
// appended to the spill list instead.
__device__ __forceinline__ void scatter_round(const GridP &g, float4 *arena, const TileCtx &t, bool active, int i,
                                              const Stencil &s, const Scatter &sc, bool staged, uint32_t *spill,
                                              uint32_t *counters)
{
    const int lane = threadIdx.x & 63;
    const int lx = s.b[0] - t.ax, ly = s.b[1] - t.ay, lz = s.b[2] - t.az;
    const bool inside = active && staged && lx >= 0 && lx <= ARENA - 3 && ly >= 0 && ly <= ARENA - 3 && lz >= 0 &&
                        lz <= ARENA - 3;
    const bool spill_me = active && !inside;
    const uint64_t sm = __ballot(spill_me);
    if (sm) {
        const int leader = __ffsll((unsigned long long)sm) - 1;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(counters + CT_SPILL, (uint32_t)__popcll(sm));
        base = __shfl(base, leader);
        const uint64_t below = lane ? (sm & ((~0ull) >> (64 - lane))) : 0ull;
        if (spill_me) spill[base + __popcll(below)] = (uint32_t)i;
    }
    const uint32_t key = inside ? (uint32_t)(lx + 6 * (ly + 6 * lz)) : 511u;
    uint64_t pending = __ballot(inside);
    const uint64_t same = match_key9(key);
    while (pending) {
        const uint64_t grp = same & pending;
        const bool winner = inside && ((pending >> lane) & 1ull) && (__ffsll((unsigned long long)grp) - 1 == lane);
        if (winner) {
            // 27 nodes in a fixed order; each contribution is computed right
            // before its read-modify-write.  The empty asm keeps the compiler
            // from reordering the LDS reads and writes across nodes.
            int ox = 0, oy = 0, oz = 0;
#pragma unroll
            for (int o = 0; o < 27; ++o) {
                const float wx = ox == 0 ? s.w[0][0] : (ox == 1 ? s.w[0][1] : s.w[0][2]);
                const float wy = oy == 0 ? s.w[1][0] : (oy == 1 ? s.w[1][1] : s.w[1][2]);
                const float wz = oz == 0 ? s.w[2][0] : (oz == 1 ? s.w[2][1] : s.w[2][2]);
                const float gx = ox == 0 ? s.dw[0][0] : (ox == 1 ? s.dw[0][1] : s.dw[0][2]);
                const float gy = oy == 0 ? s.dw[1][0] : (oy == 1 ? s.dw[1][1] : s.dw[1][2]);
                const float gz = oz == 0 ? s.dw[2][0] : (oz == 1 ? s.dw[2][1] : s.dw[2][2]);
                const float w = wx * wy * wz;
                const float gr[3] = {g.inv_h * gx * wy * wz, g.inv_h * wx * gy * wz, g.inv_h * wx * wy * gz};
                const float d[3] = {((float)ox - s.fx[0]) * g.h, ((float)oy - s.fx[1]) * g.h,
                                    ((float)oz - s.fx[2]) * g.h};
                float p[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float bd = sc.B[3 * c] * d[0] + sc.B[3 * c + 1] * d[1] + sc.B[3 * c + 2] * d[2];
                    const float gg = sc.G[3 * c] * gr[0] + sc.G[3 * c + 1] * gr[1] + sc.G[3 * c + 2] * gr[2];
                    p[c] = w * (sc.a[c] + bd) + gg;
                }
                const int a = a_slot(lx + ox, ly + oy, lz + oz);
                float4 v = arena[a];
                v.x += w * sc.m; v.y += p[0]; v.z += p[1]; v.w += p[2];
                arena[a] = v;
                __asm__ volatile("" ::: "memory");
                if (++ox == 3) { ox = 0; if (++oy == 3) { oy = 0; ++oz; } }
            }
        }
        pending &= ~__ballot(winner);
    }
}

// Sum the waves' arenas into staging slot `slot` (logical node order) and
// mark the node blocks that received mass.
__device__ __forceinline__ void tile_end(const GridP &g, const TileCtx &t, const float4 (*arena)[A_SLOTS],
                                         float4 *staging, uint32_t slot, bool staged, uint32_t *flags)
{
    __syncthreads();
    if (staged) {
        float4 *dst = staging + (int64_t)slot * ARENA_NODES;
        for (int k = threadIdx.x; k < ARENA_NODES; k += TILE_THREADS) {
            const int kx = k & 7, ky = (k >> 3) & 7, kz = k >> 6;
            const int a = a_slot(kx, ky, kz);
            float4 v = arena[0][a];
#pragma unroll
            for (int w = 1; w < TILE_WAVES; ++w) {
                const float4 u = arena[w][a];
                v.x += u.x; v.y += u.y; v.z += u.z; v.w += u.w;
            }
            dst[k] = v;
            const int nx_ = t.ax + kx, ny_ = t.ay + ky, nz_ = t.az + kz;
            if (v.x != 0.f && nx_ >= 0 && ny_ >= 0 && nz_ >= 0 && nx_ < g.nx && ny_ < g.ny && nz_ < g.nz)
                flags[((int64_t)(nz_ >> 2) * g.nby + (ny_ >> 2)) * g.nbx + (nx_ >> 2)] = 1u;
        }
    }
    __syncthreads();
}

template <int MAT>
__global__ void __launch_bounds__(TILE_THREADS) k_p2g_block(GridP g, MatP mat, PartView pv,
                                                            const uint8_t *__restrict__ cellkey,
                                                            const uint4 *__restrict__ desc,
                                                            const uint32_t *__restrict__ tile_count, uint32_t desc_cap,
                                                            uint32_t staging_cap, float dt,
                                                            float4 *__restrict__ staging, uint32_t *spill,
                                                            uint32_t *flags, uint32_t *counters)
{
    __shared__ float4 arena[TILE_WAVES][A_SLOTS];
    __shared__ uint32_t cstart[64], cend[64];
    const int wave = threadIdx.x >> 6;
    const uint32_t nt = min(*tile_count, desc_cap);
    uint32_t bits = 0;
    for (uint32_t tt = blockIdx.x; tt < nt; tt += gridDim.x) {
        const uint4 d = desc[tt];
        const TileCtx t = tile_begin(g, pv, cellkey, d, arena, cstart, cend);
        const bool staged = d.w < staging_cap;
        for (uint32_t r = wave; r < t.rounds; r += TILE_WAVES) {
            const bool active = r < t.myn;
            const int i = (int)(d.x + t.my0 + r);
            Stencil s;
            s.b[0] = s.b[1] = s.b[2] = 0;
            Scatter sc;
            if (active) {
                float x[3] = {ld(pv, PL_X, i), ld(pv, PL_X + 1, i), ld(pv, PL_X + 2, i)};
                bits |= make_stencil(g, x, s);
                load_scatter<MAT>(pv, i, mat, dt, true, 0, sc, counters);
            }
            scatter_round(g, arena[wave], t, active, i, s, sc, staged, spill, counters);
        }
        tile_end(g, t, arena, staging, d.w, staged, flags);
    }
    if (bits) atomicOr(counters + CT_HEALTH, bits);
    if (threadIdx.x == 0 && blockIdx.x == 0) counters[CT_TILES] = nt;
}

// ---------------------------------------------------------------------------
// FUSED form: G2P of step n and P2G of step n+1 in one pass over a block.
// The block's 8^3 grid-velocity neighbourhood is copied to LDS; each particle
// is gathered, updated (one SVD) and advected, its state written back, and
// scattered at its new position with the new v, C and stress into the
// scatter arenas, which are staged for the grid update of step n+1.
template <int MAT>
__global__ void __launch_bounds__(TILE_THREADS) k_g2p2g_block(GridP g, MatP mat, PartView pv,
                                                              const uint8_t *__restrict__ cellkey,
                                                              const uint4 *__restrict__ desc,
                                                              const uint32_t *__restrict__ tile_count,
                                                              uint32_t desc_cap, uint32_t staging_cap, float dt,
                                                              float dt_next, const float4 *__restrict__ gv,
                                                              float4 *__restrict__ staging, uint32_t *spill,
                                                              uint32_t *flags, uint32_t *counters, int sweeps)
{
    __shared__ float4 arena[TILE_WAVES][A_SLOTS];
    __shared__ float4 garena[A_SLOTS];
    __shared__ uint32_t cstart[64], cend[64];
    const int wave = threadIdx.x >> 6;
    const uint32_t nt = min(*tile_count, desc_cap);
    uint32_t bits = 0;
    for (uint32_t tt = blockIdx.x; tt < nt; tt += gridDim.x) {
        const uint4 d = desc[tt];
        const TileCtx t = tile_begin(g, pv, cellkey, d, arena, cstart, cend);
        // grid velocities of the arena nodes (zero outside the grid)
        for (int k = threadIdx.x; k < ARENA_NODES; k += TILE_THREADS) {
            const int kx = k & 7, ky = (k >> 3) & 7, kz = k >> 6;
            const int x = t.ax + kx, y = t.ay + ky, z = t.az + kz;
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (x >= 0 && y >= 0 && z >= 0 && x < g.nx && y < g.ny && z < g.nz) v = gv[node_slot(g, x, y, z)];
            garena[a_slot(kx, ky, kz)] = v;
        }
        __syncthreads();
        const bool staged = d.w < staging_cap;
        const ArenaFetch fetch{garena, gv, t.ax, t.ay, t.az};
        for (uint32_t r = wave; r < t.rounds; r += TILE_WAVES) {
            const bool active = r < t.myn;
            const int i = (int)(d.x + t.my0 + r);
            Stencil s;
            s.b[0] = s.b[1] = s.b[2] = 0;
            Scatter sc;
            if (active) {
                G2POut o;
                bits |= g2p_core<MAT>(g, mat, pv, i, dt, fetch, MAT != MPM_WATER, sweeps, counters, o);
                bits |= make_stencil(g, o.x, s);
                scatter_from_out<MAT>(mat, o, dt_next, sc);
                const int lx = s.b[0] - t.ax, ly = s.b[1] - t.ay, lz = s.b[2] - t.az;
                const bool inside = staged && lx >= 0 && lx <= ARENA - 3 && ly >= 0 && ly <= ARENA - 3 && lz >= 0 &&
                                    lz <= ARENA - 3;
                if (!inside && MAT != MPM_WATER) {
                    // the atomic fallback reads the stress from memory
#pragma unroll
                    for (int k = 0; k < 6; ++k) st(pv, PL_TAU + k, i, o.tau[k]);
                }
            }
            scatter_round(g, arena[wave], t, active, i, s, sc, staged, spill, counters);
        }
        tile_end(g, t, arena, staging, d.w, staged, flags);
    }
    if (bits) atomicOr(counters + CT_HEALTH, bits);
    if (threadIdx.x == 0 && blockIdx.x == 0) counters[CT_TILES] = nt;
}

// ---------------------------------------------------------------------------
// Host side.
static inline int nblk(int64_t n, int t) { return (int)((n + t - 1) / t); }

static int64_t tile_bound(const mpm_sim *s, int64_t n) { return std::min<int64_t>(s->nblocks, n); }

static int ensure_fast_buffers(mpm_sim *s)
{
    if (s->groups.size() > 256) return set_error(MPM_ERR_INVALID, "at most 256 particle groups");
    int64_t nmax = 0;
    for (auto &gi : s->groups) nmax = std::max(nmax, gi.n);
    if (!s->keys || s->sort_hist_len < nmax) {
        void *ptrs[] = {s->keys, s->keys_alt, s->vals, s->vals_alt, s->sort_hist, s->sort_scan};
        for (void *p : ptrs) if (p) (void)hipFree(p);
        int64_t ntiles = (nmax + RS_TILE - 1) / RS_TILE;
        int64_t hist_len = 256 * ntiles;
        int64_t scan_words = std::max<int64_t>((hist_len + SCAN_TILE - 1) / SCAN_TILE,
                                               (s->nblocks + SCAN_TILE - 1) / SCAN_TILE) + 16;
        MPM_HIP(hipMalloc(&s->keys, nmax * 4));
        MPM_HIP(hipMalloc(&s->keys_alt, nmax * 4));
        MPM_HIP(hipMalloc(&s->vals, nmax * 4));
        MPM_HIP(hipMalloc(&s->vals_alt, nmax * 4));
        MPM_HIP(hipMalloc(&s->sort_hist, hist_len * 4));
        MPM_HIP(hipMalloc(&s->sort_scan, scan_words * 4));
        s->sort_hist_len = nmax;
    }
    if (!s->blk_count) {
        MPM_HIP(hipMalloc(&s->blk_count, s->nblocks * 4));
        MPM_HIP(hipMalloc(&s->blk_start, s->nblocks * 4));
        MPM_HIP(hipMalloc(&s->scan_tmp, s->nblocks * 4));
    }
    if (!s->cellkey) MPM_HIP(hipMalloc(&s->cellkey, s->cap));
    if (!s->spill) MPM_HIP(hipMalloc(&s->spill, s->cap * sizeof(uint32_t)));
    int64_t need = 0;
    s->group_tile_off.assign(s->groups.size(), 0);
    for (size_t k = 0; k < s->groups.size(); ++k) {
        s->group_tile_off[k] = need;
        need += tile_bound(s, s->groups[k].n);
    }
    if (s->tile_cap < need) {
        if (s->tile_desc) (void)hipFree(s->tile_desc);
        if (s->staging) (void)hipFree(s->staging);
        if (s->blk_tile) (void)hipFree(s->blk_tile);
        s->tile_desc = nullptr; s->staging = nullptr; s->blk_tile = nullptr;
        MPM_HIP(hipMalloc(&s->tile_desc, need * sizeof(uint4)));
        // every non-empty block may need a staging slot; blocks beyond the cap scatter with atomics
        int64_t scap = std::min<int64_t>(need, (int64_t)1 << 19);
        MPM_HIP(hipMalloc(&s->staging, scap * ARENA_NODES * sizeof(float4)));
        MPM_HIP(hipMalloc(&s->blk_tile, (int64_t)s->groups.size() * s->nblocks * sizeof(int32_t)));
        s->staging_cap = scap;
        s->tile_cap = need;
    }
    if (!s->tile_counts) MPM_HIP(hipMalloc(&s->tile_counts, 256 * 4));
    return MPM_OK;
}

static int sort_all(mpm_sim *s)
{
    int rc = ensure_fast_buffers(s);
    if (rc) return rc;
    hipStream_t st = s->stream;
    int bits = KEY_LOW + 1;
    while ((1ll << (bits - KEY_LOW)) < s->nblocks) ++bits;
    uint32_t plane_mask = (1u << PL_COUNT) - 1u;
    if (!s->tau_valid) plane_mask &= ~(((1u << 6) - 1u) << PL_TAU);
    const int nb = (int)s->nblocks;
    for (size_t k = 0; k < s->groups.size(); ++k) {
        PartView pv = view_of(s, (int)k);
        int n = pv.n;
        uint32_t *blk_end = s->blk_count;
        if ((rc = launch_zero_u32(st, s->blk_start, nb)) || (rc = launch_zero_u32(st, blk_end, nb)))
            return hip_fail((hipError_t)rc, "zero");
        hipLaunchKernelGGL(k_block_keys, dim3(nblk(n, 256)), dim3(256), 0, st, s->g, pv, s->keys, s->vals);
        bool in_alt = false;
        if ((rc = radix_sort(st, s->keys, s->vals, s->keys_alt, s->vals_alt, n, bits, s->sort_hist, s->sort_scan,
                             in_alt)))
            return hip_fail((hipError_t)rc, "radix_sort");
        const uint32_t *perm = in_alt ? s->vals_alt : s->vals;
        const uint32_t *skeys = in_alt ? s->keys_alt : s->keys;
        hipLaunchKernelGGL(k_permute, dim3(nblk(n, 256)), dim3(256), 0, st, s->P, s->P_alt, s->cap, pv.off, n, perm,
                           skeys, s->pid, s->pid_alt, s->cellkey, plane_mask);
        // tiles: one per non-empty block, numbered in block order
        hipLaunchKernelGGL(k_block_runs, dim3(nblk(n, 256)), dim3(256), 0, st, skeys, n, s->blk_start, blk_end);
        hipLaunchKernelGGL(k_tile_flags, dim3(nblk(nb, 256)), dim3(256), 0, st, s->blk_start, blk_end, s->scan_tmp,
                           nb);
        if ((rc = scan_excl(st, s->scan_tmp, s->scan_tmp, nb, s->sort_scan, s->tile_counts + k)))
            return hip_fail((hipError_t)rc, "scan");
        hipLaunchKernelGGL(k_emit_tiles, dim3(nblk(nb, 256)), dim3(256), 0, st, s->blk_start, blk_end, s->scan_tmp,
                           nb, (uint4 *)s->tile_desc + s->group_tile_off[k], (uint32_t)tile_bound(s, n),
                           s->blk_tile + (int64_t)k * nb, (uint32_t)s->group_tile_off[k]);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "sort kernels");
    }
    std::swap(s->P, s->P_alt);
    std::swap(s->pid, s->pid_alt);
    s->sorted_valid = true;
    s->steps_since_sort = 0;
    s->sorts += 1;
    return MPM_OK;
}

int sort_if_due(mpm_sim *s)
{
    if (!s->sorted_valid || s->steps_since_sort >= s->cfg.sort_every) return sort_all(s);
    return MPM_OK;
}

int fast_prepare(mpm_sim *s)
{
    int rc = MPM_OK;
    if (!s->sorted_valid || s->steps_since_sort >= s->cfg.sort_every) {
        rc = sort_all(s);
        if (rc) return rc;
    }
    if (!s->tau_valid) {
        for (size_t k = 0; k < s->groups.size(); ++k) {
            rc = launch_compute_tau(s->stream, s->groups[k].mat, view_of(s, (int)k), s->cfg.svd_sweeps, s->counters);
            if (rc) return hip_fail((hipError_t)rc, "compute_tau");
        }
        s->tau_valid = true;
    }
    return MPM_OK;
}

int launch_p2g_tiles(mpm_sim *s, int group, float dt)
{
    PartView pv = view_of(s, group);
    if (pv.n <= 0) return MPM_OK;
    const MatP &m = s->groups[group].mat;
    int64_t bound = tile_bound(s, pv.n);
    int wg = (int)std::min<int64_t>(bound, 8192);
    const uint4 *desc = (const uint4 *)s->tile_desc + s->group_tile_off[group];
    const uint32_t *tc = s->tile_counts + group;
    const uint32_t cap = (uint32_t)bound;
    const uint32_t scap = (uint32_t)s->staging_cap;
    int rc = launch_zero_u32(s->stream, s->counters + CT_SPILL, 1);
    if (rc) return hip_fail((hipError_t)rc, "zero");
    dim3 gr(wg), bl(TILE_THREADS);
    switch (m.type) {
    case MPM_JELLO: hipLaunchKernelGGL(k_p2g_block<MPM_JELLO>, gr, bl, 0, s->stream, s->g, m, pv, s->cellkey, desc, tc, cap, scap, dt, s->staging, s->spill, s->flags, s->counters); break;
    case MPM_SNOW: hipLaunchKernelGGL(k_p2g_block<MPM_SNOW>, gr, bl, 0, s->stream, s->g, m, pv, s->cellkey, desc, tc, cap, scap, dt, s->staging, s->spill, s->flags, s->counters); break;
    default: hipLaunchKernelGGL(k_p2g_block<MPM_WATER>, gr, bl, 0, s->stream, s->g, m, pv, s->cellkey, desc, tc, cap, scap, dt, s->staging, s->spill, s->flags, s->counters); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "p2g_block");
    rc = launch_p2g_spill(s->stream, s->g, m, pv, dt, s->spill, (uint32_t)pv.n, s->counters, s->grid_mp, s->flags);
    return rc ? hip_fail((hipError_t)rc, "p2g_spill") : MPM_OK;
}

int launch_g2p2g(mpm_sim *s, int group, float dt, float dt_next)
{
    PartView pv = view_of(s, group);
    if (pv.n <= 0) return MPM_OK;
    const MatP &m = s->groups[group].mat;
    int64_t bound = tile_bound(s, pv.n);
    int wg = (int)std::min<int64_t>(bound, 8192);
    const uint4 *desc = (const uint4 *)s->tile_desc + s->group_tile_off[group];
    const uint32_t *tc = s->tile_counts + group;
    const uint32_t cap = (uint32_t)bound;
    const uint32_t scap = (uint32_t)s->staging_cap;
    const int sw = s->cfg.svd_sweeps;
    int rc = launch_zero_u32(s->stream, s->counters + CT_SPILL, 1);
    if (rc) return hip_fail((hipError_t)rc, "zero");
    dim3 gr(wg), bl(TILE_THREADS);
    switch (m.type) {
    case MPM_JELLO: hipLaunchKernelGGL(k_g2p2g_block<MPM_JELLO>, gr, bl, 0, s->stream, s->g, m, pv, s->cellkey, desc, tc, cap, scap, dt, dt_next, s->grid_v, s->staging, s->spill, s->flags, s->counters, sw); break;
    case MPM_SNOW: hipLaunchKernelGGL(k_g2p2g_block<MPM_SNOW>, gr, bl, 0, s->stream, s->g, m, pv, s->cellkey, desc, tc, cap, scap, dt, dt_next, s->grid_v, s->staging, s->spill, s->flags, s->counters, sw); break;
    default: hipLaunchKernelGGL(k_g2p2g_block<MPM_WATER>, gr, bl, 0, s->stream, s->g, m, pv, s->cellkey, desc, tc, cap, scap, dt, dt_next, s->grid_v, s->staging, s->spill, s->flags, s->counters, sw); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "g2p2g_block");
    rc = launch_p2g_spill(s->stream, s->g, m, pv, dt_next, s->spill, (uint32_t)pv.n, s->counters, s->grid_mp, s->flags);
    return rc ? hip_fail((hipError_t)rc, "p2g_spill") : MPM_OK;
}

// Stand-alone sort for the probe entry point (own temporaries).
int probe_sort(const uint32_t *keys_h, int64_t n, int bits, uint32_t *sorted_h, uint32_t *perm_h)
{
    if (n <= 0) return MPM_OK;
    if (n > (1ll << 30)) return set_error(MPM_ERR_INVALID, "too many keys");
    if (bits < 1 || bits > 32) return set_error(MPM_ERR_INVALID, "bits out of range [1,32]");
    int64_t ntiles = (n + RS_TILE - 1) / RS_TILE;
    int64_t hist_len = 256 * ntiles;
    int64_t scan_words = (hist_len + SCAN_TILE - 1) / SCAN_TILE + 16;
    uint32_t *k = nullptr, *ka = nullptr, *v = nullptr, *va = nullptr, *hist = nullptr, *scan = nullptr;
    hipStream_t st = nullptr;
    int rc = MPM_OK;
    auto chk = [&](hipError_t e, const char *w) { if (!rc && e != hipSuccess) rc = hip_fail(e, w); };
    chk(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream");
    chk(hipMalloc(&k, n * 4), "malloc"); chk(hipMalloc(&ka, n * 4), "malloc");
    chk(hipMalloc(&v, n * 4), "malloc"); chk(hipMalloc(&va, n * 4), "malloc");
    chk(hipMalloc(&hist, hist_len * 4), "malloc"); chk(hipMalloc(&scan, scan_words * 4), "malloc");
    std::vector<uint32_t> iota((size_t)n);
    for (int64_t i = 0; i < n; ++i) iota[i] = (uint32_t)i;
    chk(hipMemcpyAsync(k, keys_h, n * 4, hipMemcpyHostToDevice, st), "h2d");
    chk(hipMemcpyAsync(v, iota.data(), n * 4, hipMemcpyHostToDevice, st), "h2d");
    bool in_alt = false;
    if (!rc) {
        int e = radix_sort(st, k, v, ka, va, (int)n, bits, hist, scan, in_alt);
        if (e) rc = hip_fail((hipError_t)e, "radix_sort");
    }
    chk(hipMemcpyAsync(sorted_h, in_alt ? ka : k, n * 4, hipMemcpyDeviceToHost, st), "d2h");
    chk(hipMemcpyAsync(perm_h, in_alt ? va : v, n * 4, hipMemcpyDeviceToHost, st), "d2h");
    chk(hipStreamSynchronize(st), "sync");
    void *ptrs[] = {k, ka, v, va, hist, scan};
    for (void *p : ptrs) if (p) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}

int launch_g2p_fast(mpm_sim *s, int group, float dt)
{
    int rc = launch_g2p(s->stream, s->g, s->groups[group].mat, view_of(s, group), dt, s->grid_v, 1,
                        s->cfg.svd_sweeps, s->counters, s->cfg.threads);
    return rc ? hip_fail((hipError_t)rc, "g2p_fast") : MPM_OK;
}

}  // namespace mpm
