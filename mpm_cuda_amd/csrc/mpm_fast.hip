// mpm_fast.hip -- FAST execution form: block sort and LDS-blocked P2G.
//
// No counterpart in the reference (JerryYan97/MPM_CUDA scatters unsorted
// particles with global FP64 atomics, MPMCUDA.cu:388-421).  Pieces:
//   * exclusive scan (three-kernel, wave-shuffle based);
//   * stable LSD radix sort of 32-bit keys, 8- or 9-bit digits,
//     ballot-matched in-wave ranking;
//   * per-group sort of particles by the 4x4x4 node block of their base
//     node, payload permutation of every SoA plane, and one tile (block,
//     first particle, count, staging slot) per non-empty block;
//   * P2G over tiles without atomics: one wave per block bins its particles
//     by base node in LDS, sums each bin's particles in registers and adds
//     the sums to an 8^3-node LDS arena by plain read-modify-write, then
//     writes the arena's nodes with mass, and a mask of them, to a staging
//     slot that the grid update gathers.
// (On gfx950 an LDS float atomic add costs ~200 cycles per wave instruction
// even when the 64 lanes hit different addresses, ~15x a float4 plain LDS
// read-modify-write, and global float atomics at ~77 G/s would dominate a
// per-node flush.)
#include "mpm_g2p.h"
#include "mpm_launch.h"

#include <chrono>
#include <cstring>
#include <mutex>
#include <thread>

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

// Workgroup `wg` of a two-launch exclusive scan (SCAN_TILE values each):
// its sum into partial[wg]; then its values scanned from the sum of the
// partials before it (one wave adds them up).  lds: 40 words.
__device__ __forceinline__ void scan_reduce_wg(const uint32_t *in, int n, uint32_t *partial, int wg, uint32_t *lds)
{
    const int base = wg * SCAN_TILE;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const int i = base + k * SCAN_T + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint32_t tot;
    block_excl_scan(s, tot, lds);
    if (threadIdx.x == 0) partial[wg] = tot;
}

__device__ __forceinline__ void scan_apply_wg(const uint32_t *in, uint32_t *out, int n, const uint32_t *partial,
                                              uint32_t *total_out, int wg, uint32_t *lds, uint32_t *wg_base)
{
    if (threadIdx.x < 64) {
        uint32_t b = 0;
        for (int j = (int)threadIdx.x; j < wg; j += 64) b += partial[j];
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) b += __shfl_xor(b, d);
        if (threadIdx.x == 0) *wg_base = b;
    }
    __syncthreads();
    const int i0 = wg * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = (i0 + k < n) ? in[i0 + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = *wg_base + block_excl_scan(s, tot, lds);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        if (i0 + k < n) {
            out[i0 + k] = run;
            if (total_out && i0 + k == n - 1) *total_out = run + v[k];
        }
        run += v[k];
    }
}

__global__ void __launch_bounds__(SCAN_T) k_scan_reduce(const uint32_t *in, int n, uint32_t *partial)
{
    __shared__ uint32_t lds[40];
    scan_reduce_wg(in, n, partial, (int)blockIdx.x, lds);
}

__global__ void __launch_bounds__(SCAN_T) k_scan_apply(const uint32_t *in, uint32_t *out, int n,
                                                       const uint32_t *partial, uint32_t *total_out)
{
    __shared__ uint32_t lds[40];
    __shared__ uint32_t wg_base;
    scan_apply_wg(in, out, n, partial, total_out, (int)blockIdx.x, lds, &wg_base);
}

// out = exclusive scan of in (in-place allowed); partial holds ceil(n / 2048) words.
static int scan_excl(hipStream_t st, const uint32_t *in, uint32_t *out, int n, uint32_t *partial,
                     uint32_t *total_out)
{
    int nt = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nt == 0) return 0;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nt), dim3(SCAN_T), 0, st, in, n, partial);
    hipLaunchKernelGGL(k_scan_apply, dim3(nt), dim3(SCAN_T), 0, st, in, out, n, partial, total_out);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort with DB-bit digits (8, or 9 so that 17-18-bit block
// keys take two passes instead of three).  A tile of RS_TILE keys is handled
// by one workgroup of RS_WAVES waves, each owning a contiguous run of
// RS_WAVE_KEYS keys processed 64 at a time; equal digits inside a wave are
// matched with DB ballots, so ranks follow input order and the sort is
// stable.  A pass is three launches: the digit histogram of every tile
// (hist[d * ntiles + t]), the scan of each digit's column over the tiles
// with the digit totals (k_rs_spine), and the scatter, whose workgroups
// scan the totals into digit bases themselves.  Sixteen waves per tile keep
// each wave's rank chain short, and the ~250 tiles of a 1M-key sort put
// about four waves on every SIMD.
constexpr int RS_WAVES = 16, RS_T = 64 * RS_WAVES, RS_WAVE_KEYS = 256, RS_ITERS = RS_WAVE_KEYS / 64,
              RS_TILE = RS_WAVES * RS_WAVE_KEYS;
constexpr int RS_MAX_DIGITS = 512;      // digits of up to 9 bits: the size of the totals buffer

static inline int radix_digit_bits(int bits) { return (bits > 16 && bits <= 18) ? 9 : 8; }

// Adds every active lane's digit d to the LDS histogram h: the lanes that
// share the first active lane's digit (most of a wave, for nearly sorted
// keys) as one add, the others one each.
__device__ __forceinline__ void hist_add(uint32_t *h, bool active, uint32_t d)
{
    const uint64_t act = __ballot(active);
    if (!act) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)act) - 1;
    const uint32_t dl = (uint32_t)__shfl((int)d, leader);
    const uint64_t same = __ballot(active && d == dl);
    if (lane == leader) atomicAdd(&h[dl], (uint32_t)__popcll(same));
    else if (active && d != dl) atomicAdd(&h[d], 1u);
}

// Per-tile bodies of the radix kernels (`tile` is the workgroup's tile of
// RS_TILE keys; LDS comes from the kernel), shared by the single sort and
// the batched one.
template <int DB>
__device__ __forceinline__ void rs_hist_tile(const uint32_t *keys, int n, int shift, uint32_t dmask, uint32_t *hist,
                                             int ntiles, int tile, uint32_t *h)
{
    constexpr int NB = 1 << DB;
    const uint32_t MASK = dmask & (NB - 1);       // the last digit may be narrower
    for (int d = threadIdx.x; d < NB; d += RS_T) h[d] = 0;
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t dg[RS_ITERS];
#pragma unroll
    for (int it = 0; it < RS_ITERS; ++it) {
        const int i = tile * RS_TILE + w * RS_WAVE_KEYS + it * 64 + lane;
        dg[it] = i < n ? (keys[i] >> shift) & MASK : 0u;
    }
#pragma unroll
    for (int it = 0; it < RS_ITERS; ++it) hist_add(h, tile * RS_TILE + w * RS_WAVE_KEYS + it * 64 + lane < n, dg[it]);
    __syncthreads();
    for (int d = threadIdx.x; d < NB; d += RS_T) hist[(int64_t)d * ntiles + tile] = h[d];
}

template <int DB>
__global__ void __launch_bounds__(RS_T) k_rs_hist(const uint32_t *keys, int n, int shift, uint32_t dmask,
                                                  uint32_t *hist, int ntiles)
{
    __shared__ uint32_t h[1 << DB];
    rs_hist_tile<DB>(keys, n, shift, dmask, hist, ntiles, (int)blockIdx.x, h);
}

// One wave replaces digit d's column hist[d * ntiles + t] by its exclusive
// prefix over the tiles t and writes the digit's total.
__device__ __forceinline__ void rs_spine_column(uint32_t *hist, int ntiles, uint32_t *totals, int d)
{
    const int lane = threadIdx.x & 63;
    uint32_t *col = hist + (int64_t)d * ntiles;
    uint32_t carry = 0;
    for (int t0 = 0; t0 < ntiles; t0 += 64) {
        const int t = t0 + lane;
        const uint32_t v = t < ntiles ? col[t] : 0u;
        const uint32_t inc = wave_incl_scan(v);
        if (t < ntiles) col[t] = carry + inc - v;
        carry += __shfl(inc, 63);
    }
    if (lane == 0) totals[d] = carry;
}

// Workgroup d (one wave): digit d's column.
__global__ void __launch_bounds__(64) k_rs_spine(uint32_t *hist, int ntiles, uint32_t *totals)
{
    rs_spine_column(hist, ntiles, totals, (int)blockIdx.x);
}

// Scatter of one tile; vin null means the values are the input positions.
// cnt: RS_WAVES x NB words of LDS, lds: 40.
template <int DB>
__device__ __forceinline__ void rs_scatter_tile(const uint32_t *kin, const uint32_t *vin, uint32_t *kout, uint32_t *vout,
                                                int n, int shift, uint32_t dmask, const uint32_t *hist,
                                                const uint32_t *totals, int ntiles, int tile, uint32_t (*cnt)[1 << DB],
                                                uint32_t *lds)
{
    constexpr int NB = 1 << DB;
    static_assert(NB <= RS_T, "one digit per thread");
    const uint32_t MASK = dmask & (NB - 1);
    const int t = threadIdx.x;
    for (int j = t; j < RS_WAVES * NB; j += RS_T) (&cnt[0][0])[j] = 0;
    // this tile's first output position of digit t: the digit's base (the
    // totals of the smaller digits) plus its count in the tiles before
    const uint32_t tot_t = t < NB ? totals[t] : 0u;
    uint32_t all;
    uint32_t first = block_excl_scan(tot_t, all, lds);
    if (t < NB) first += hist[(int64_t)t * ntiles + tile];
    __syncthreads();
    const int w = t >> 6, lane = t & 63;
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    uint32_t key[RS_ITERS], val[RS_ITERS], rank[RS_ITERS];
#pragma unroll
    for (int it = 0; it < RS_ITERS; ++it) {
        const int i = tile * RS_TILE + w * RS_WAVE_KEYS + it * 64 + lane;
        key[it] = i < n ? kin[i] : 0u;
        val[it] = i < n ? (vin ? vin[i] : (uint32_t)i) : 0u;
    }
#pragma unroll
    for (int it = 0; it < RS_ITERS; ++it) {
        const int i = tile * RS_TILE + w * RS_WAVE_KEYS + it * 64 + lane;
        const bool valid = i < n;
        const uint32_t d = (key[it] >> shift) & MASK;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < DB; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        const uint32_t below = (uint32_t)__popcll(m & lt);
        uint32_t prev = 0;
        if (valid) prev = cnt[w][d];
        if (valid && below == 0) cnt[w][d] = prev + (uint32_t)__popcll(m);
        rank[it] = prev + below;
        __asm__ volatile("" ::: "memory");
    }
    __syncthreads();
    if (t < NB) {
        uint32_t run = first;
        for (int k = 0; k < RS_WAVES; ++k) {
            const uint32_t c = cnt[k][t];
            cnt[k][t] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < RS_ITERS; ++it) {
        const int i = tile * RS_TILE + w * RS_WAVE_KEYS + it * 64 + lane;
        if (i < n) {
            const uint32_t dst = cnt[w][(key[it] >> shift) & MASK] + rank[it];
            if (dst < (uint32_t)n) {      // always true for a consistent histogram
                kout[dst] = key[it];
                vout[dst] = val[it];
            }
        }
    }
}

template <int DB>
__global__ void __launch_bounds__(RS_T) k_rs_scatter(const uint32_t *kin, const uint32_t *vin, uint32_t *kout,
                                                     uint32_t *vout, int n, int shift, uint32_t dmask,
                                                     const uint32_t *hist, const uint32_t *totals, int ntiles)
{
    __shared__ uint32_t cnt[RS_WAVES][1 << DB];
    __shared__ uint32_t lds[40];
    rs_scatter_tile<DB>(kin, vin, kout, vout, n, shift, dmask, hist, totals, ntiles, (int)blockIdx.x, cnt, lds);
}

struct RadixResult {
    uint32_t *keys, *vals;
};

// Sorts (keys, vals) of length n by the low `bits` key bits, ping-ponging
// with the *_alt buffers.  `hist` holds 512 x ceil(n / RS_TILE) entries,
// `totals` RS_MAX_DIGITS.  With hist0_ready the first pass's digit
// histogram is already in `hist` (radix_digit_bits(bits)-bit digits, written
// by the caller); with iota_vals the values are the input positions (vals is
// not read); vals_out, when given (distinct from the other buffers), takes
// the last pass's values.  res names the sorted keys and values.
static int radix_sort_ex(hipStream_t st, uint32_t *keys, uint32_t *vals, uint32_t *keys_alt, uint32_t *vals_alt,
                         int n, int bits, uint32_t *hist, uint32_t *totals, bool hist0_ready, bool iota_vals,
                         uint32_t *vals_out, RadixResult &res)
{
    res = RadixResult{keys, vals};
    if (n <= 0) return 0;
    const int ntiles = (n + RS_TILE - 1) / RS_TILE;
    uint32_t *ka = keys, *va = vals, *kb = keys_alt, *vb = vals_alt;
    const int db = radix_digit_bits(bits);
    for (int shift = 0; shift < bits; shift += db) {
        const bool first = shift == 0, last = shift + db >= bits;
        const int w = std::min(db, bits - shift);
        const uint32_t dmask = (w >= 32) ? ~0u : ((1u << w) - 1u);
        uint32_t *vdst = (last && vals_out) ? vals_out : vb;
        const uint32_t *vsrc = (first && iota_vals) ? nullptr : va;
        if (db == 9) {
            if (!(first && hist0_ready))
                hipLaunchKernelGGL(k_rs_hist<9>, dim3(ntiles), dim3(RS_T), 0, st, ka, n, shift, dmask, hist, ntiles);
            hipLaunchKernelGGL(k_rs_spine, dim3(1 << 9), dim3(64), 0, st, hist, ntiles, totals);
            hipLaunchKernelGGL(k_rs_scatter<9>, dim3(ntiles), dim3(RS_T), 0, st, ka, vsrc, kb, vdst, n, shift, dmask,
                               hist, totals, ntiles);
        } else {
            if (!(first && hist0_ready))
                hipLaunchKernelGGL(k_rs_hist<8>, dim3(ntiles), dim3(RS_T), 0, st, ka, n, shift, dmask, hist, ntiles);
            hipLaunchKernelGGL(k_rs_spine, dim3(1 << 8), dim3(64), 0, st, hist, ntiles, totals);
            hipLaunchKernelGGL(k_rs_scatter<8>, dim3(ntiles), dim3(RS_T), 0, st, ka, vsrc, kb, vdst, n, shift, dmask,
                               hist, totals, ntiles);
        }
        res = RadixResult{kb, vdst};
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    return (int)hipGetLastError();
}

// The plain form: in_alt tells whether the result ended in the *_alt buffers.
static int radix_sort(hipStream_t st, uint32_t *keys, uint32_t *vals, uint32_t *keys_alt, uint32_t *vals_alt,
                      int n, int bits, uint32_t *hist, uint32_t *totals, bool &in_alt)
{
    RadixResult r{};
    const int rc = radix_sort_ex(st, keys, vals, keys_alt, vals_alt, n, bits, hist, totals, false, false, nullptr, r);
    in_alt = r.keys != keys;
    return rc;
}

// ---------------------------------------------------------------------------
// Keys, permutation and tiles.
//
// Sort key: linear index of the 4^3-node block holding the particle's base
// node, above KEY_LOW bits of the base node's position inside the block, so
// that a block's particles are stored cell by cell: the P2G's counting sort
// of a chunk by base node then finds them nearly in order, and the item
// gathers of its phase C read neighbouring particles.  P2G still re-bins
// every chunk by its particles' current base nodes in LDS.
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

// Block keys of the particles (linear index of the 4^3-node block holding
// the base node) of tile `tile` (RS_TILE particles) and the radix sort's
// first digit histogram (DB-bit digits, dmask the first digit's bits) of
// the tile, laid out as k_rs_hist writes it; h: 2^DB words of LDS.  Before
// that, the workgroup clears its share (from `first`, every `stride`) of the
// nb-entry block arrays z0, z1, z2 that the run and tile kernels fill.
template <int DB>
__device__ __forceinline__ void keys_hist_tile(const GridP &g, const PartView &pv, uint32_t *keys, uint32_t dmask,
                                               uint32_t *hist, int ntiles, int tile, uint32_t *z0, uint32_t *z1,
                                               uint32_t *z2, int nb, int first, int stride, uint32_t *h)
{
    constexpr int NB = 1 << DB;
    const uint32_t MASK = dmask & (NB - 1);
    for (int d = threadIdx.x; d < NB; d += RS_T) h[d] = 0;
    for (int j = first + (int)threadIdx.x; j < nb; j += stride) {
        z0[j] = 0u;
        z1[j] = 0u;
        z2[j] = 0u;
    }
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int it = 0; it < RS_ITERS; ++it) {
        const int i = tile * RS_TILE + w * RS_WAVE_KEYS + it * 64 + lane;
        const bool valid = i < pv.n;
        uint32_t key = 0;
        if (valid) {
            int b[3];
            base_of(g, pv, i, b);
            key = ((((uint32_t)(b[2] >> 2) * (uint32_t)g.nby + (uint32_t)(b[1] >> 2)) * (uint32_t)g.nbx +
                    (uint32_t)(b[0] >> 2)) << KEY_LOW) |
                  (uint32_t)((b[0] & 3) | ((b[1] & 3) << 2) | ((b[2] & 3) << 4));
            keys[i] = key;
        }
        hist_add(h, valid, key & MASK);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < NB; d += RS_T) hist[(int64_t)d * ntiles + tile] = h[d];
}

template <int DB>
__global__ void __launch_bounds__(RS_T) k_block_keys_hist(GridP g, PartView pv, uint32_t *keys, uint32_t dmask,
                                                          uint32_t *hist, int ntiles, uint32_t *z0, uint32_t *z1,
                                                          uint32_t *z2, int nb)
{
    __shared__ uint32_t h[1 << DB];
    keys_hist_tile<DB>(g, pv, keys, dmask, hist, ntiles, (int)blockIdx.x, z0, z1, z2, nb, (int)blockIdx.x * RS_T,
                       (int)gridDim.x * RS_T, h);
}

// Gathers every permuted SoA plane and the particle ids.
__global__ void k_permute(const float *__restrict__ P, float *__restrict__ Pa, int64_t cap, int64_t off, int n,
                          const uint32_t *__restrict__ perm, const uint32_t *__restrict__ pid,
                          uint32_t *__restrict__ pida, uint64_t plane_mask)
{
    // XCD-aware: a re-sort's permutation is close to the identity, so an
    // XCD that writes a contiguous eighth of the positions reads its sources
    // from about the same eighth, and each source line is fetched into one L2
    const uint32_t nb8 = gridDim.x;
    const int j = (int)(xcd_spread(blockIdx.x, nb8) * blockDim.x + threadIdx.x);
    if (j >= n) return;
    uint32_t src = perm[j];
    if (src >= (uint32_t)n) src = (uint32_t)j;    // never taken for a valid permutation
    for (int p = 0; p < PL_COUNT; ++p)
        if ((plane_mask >> p) & 1ull) Pa[(int64_t)p * cap + off + j] = P[(int64_t)p * cap + off + src];
    pida[off + j] = pid[off + src];
}

// Sorted position i: first and one-past-last sorted position of its block,
// and has[b] = 1 for every non-empty block (all three cleared before,
// untouched blocks keep the zeros).
__device__ __forceinline__ void block_run_at(const uint32_t *__restrict__ skeys, int n, int i, uint32_t *blk_start,
                                             uint32_t *blk_end, uint32_t *has)
{
    if (i >= n) return;
    const uint32_t b = skeys[i] >> KEY_LOW;
    if (i == 0 || (skeys[i - 1] >> KEY_LOW) != b) {
        blk_start[b] = (uint32_t)i;
        has[b] = 1u;
    }
    if (i == n - 1 || (skeys[i + 1] >> KEY_LOW) != b) blk_end[b] = (uint32_t)(i + 1);
}

__global__ void k_block_runs(const uint32_t *__restrict__ skeys, int n, uint32_t *blk_start, uint32_t *blk_end,
                             uint32_t *has)
{
    block_run_at(skeys, n, (int)(blockIdx.x * blockDim.x + threadIdx.x), blk_start, blk_end, has);
}

// Block b's tile: (first particle, count, block, tile id); blk_tile maps the
// block to its global tile id (staging slot), every arena face marked
// written until the block's first P2G, or -1 for an empty block.
__device__ __forceinline__ void emit_tile_at(const uint32_t *blk_start, const uint32_t *blk_end,
                                             const uint32_t *tile_index, int nblocks, int b, uint4 *desc,
                                             uint32_t desc_cap, int32_t *blk_tile, uint32_t tile_base)
{
    if (b >= nblocks) return;
    const uint32_t c = blk_end[b] - blk_start[b];
    const uint32_t t = tile_index[b];
    if (c > 0 && t < desc_cap) {
        desc[t] = make_uint4(blk_start[b], c, (uint32_t)b, tile_base + t);
        blk_tile[b] = tile_entry(tile_base + t, TILE_ALL_FACES);
    } else {
        blk_tile[b] = -1;
    }
}

__global__ void k_emit_tiles(const uint32_t *blk_start, const uint32_t *blk_end, const uint32_t *tile_index,
                             int nblocks, uint4 *desc, uint32_t desc_cap, int32_t *blk_tile, uint32_t tile_base)
{
    emit_tile_at(blk_start, blk_end, tile_index, nblocks, (int)(blockIdx.x * blockDim.x + threadIdx.x), desc,
                 desc_cap, blk_tile, tile_base);
}

// ---------------------------------------------------------------------------
// Batched re-sort: every scene of a SortSlice table through each stage of the
// sort above in one launch per stage (8-bit digits, the same number of
// passes for every scene).  A workgroup maps its index to a scene and to the
// scene's own tile, digit columns or block range with a SliceScan of the
// per-scene counts (the map is wave-uniform, so the bodies run as in the
// single sort).
constexpr int BS_DB = 8;

template <class Count>
__device__ __forceinline__ bool bsort_locate(const SortSlice *__restrict__ tab, int ns, Count count, uint32_t &sid,
                                             int &local)
{
    SliceScan sc;
    sc.build(ns, count);
    if (blockIdx.x >= sc.total) return false;
    sid = (uint32_t)__builtin_amdgcn_readfirstlane((int)sc.find(blockIdx.x));
    local = (int)(blockIdx.x - sc.start(sid));
    return true;
}

__global__ void __launch_bounds__(RS_T) k_bsort_keys(const SortSlice *__restrict__ tab, int ns)
{
    __shared__ uint32_t h[1 << BS_DB];
    uint32_t sid;
    int tile;
    if (!bsort_locate(tab, ns, [&](int s) { return (uint32_t)tab[s].ntiles; }, sid, tile)) return;
    const SortSlice &t = tab[sid];
    keys_hist_tile<BS_DB>(t.g, t.pv, t.keys, 0xFFu, t.hist, t.ntiles, tile, t.blk_start, t.blk_end, t.has, t.nblocks,
                          tile * RS_T, t.ntiles * RS_T, h);
}

__global__ void __launch_bounds__(RS_T) k_bsort_hist(const SortSlice *__restrict__ tab, int ns, int pass)
{
    __shared__ uint32_t h[1 << BS_DB];
    uint32_t sid;
    int tile;
    if (!bsort_locate(tab, ns, [&](int s) { return (uint32_t)tab[s].ntiles; }, sid, tile)) return;
    const SortSlice &t = tab[sid];
    rs_hist_tile<BS_DB>((pass & 1) ? t.keys_alt : t.keys, t.n, BS_DB * pass, 0xFFu, t.hist, t.ntiles, tile, h);
}

// Four digit columns per workgroup, one per wave.
__global__ void __launch_bounds__(256) k_bsort_spine(const SortSlice *__restrict__ tab, int ns)
{
    uint32_t sid;
    int q;
    if (!bsort_locate(tab, ns, [&](int) { return (uint32_t)((1 << BS_DB) / 4); }, sid, q)) return;
    const SortSlice &t = tab[sid];
    rs_spine_column(t.hist, t.ntiles, t.scan, q * 4 + (int)(threadIdx.x >> 6));
}

// Pass `pass` reads the keys (values: the positions in pass 0) from the
// buffers of its parity and writes the other ones; the last pass writes the
// values into the scene's permutation.
__global__ void __launch_bounds__(RS_T) k_bsort_scatter(const SortSlice *__restrict__ tab, int ns, int pass, int last)
{
    __shared__ uint32_t cnt[RS_WAVES][1 << BS_DB];
    __shared__ uint32_t lds[40];
    uint32_t sid;
    int tile;
    if (!bsort_locate(tab, ns, [&](int s) { return (uint32_t)tab[s].ntiles; }, sid, tile)) return;
    const SortSlice &t = tab[sid];
    const bool odd = (pass & 1) != 0;
    const uint32_t *vin = pass == 0 ? nullptr : (odd ? t.vals_alt : t.vals);
    uint32_t *vout = last ? t.perm : (odd ? t.vals : t.vals_alt);
    rs_scatter_tile<BS_DB>(odd ? t.keys_alt : t.keys, vin, odd ? t.keys : t.keys_alt, vout, t.n, BS_DB * pass, 0xFFu,
                           t.hist, t.scan, t.ntiles, tile, cnt, lds);
}

// keys_alt: the sorted keys ended in the *_alt buffers (odd pass count).
__global__ void __launch_bounds__(256) k_bsort_runs(const SortSlice *__restrict__ tab, int ns, int keys_alt)
{
    uint32_t sid;
    int q;
    if (!bsort_locate(tab, ns, [&](int s) { return (uint32_t)((tab[s].n + 255) / 256); }, sid, q)) return;
    const SortSlice &t = tab[sid];
    block_run_at(keys_alt ? t.keys_alt : t.keys, t.n, q * 256 + (int)threadIdx.x, t.blk_start, t.blk_end, t.has);
}

__global__ void __launch_bounds__(SCAN_T) k_bsort_scan_reduce(const SortSlice *__restrict__ tab, int ns)
{
    __shared__ uint32_t lds[40];
    uint32_t sid;
    int wg;
    if (!bsort_locate(tab, ns, [&](int s) { return (uint32_t)((tab[s].nblocks + SCAN_TILE - 1) / SCAN_TILE); }, sid,
                      wg))
        return;
    const SortSlice &t = tab[sid];
    scan_reduce_wg(t.has, t.nblocks, t.scan, wg, lds);
}

__global__ void __launch_bounds__(SCAN_T) k_bsort_scan_apply(const SortSlice *__restrict__ tab, int ns)
{
    __shared__ uint32_t lds[40];
    __shared__ uint32_t wg_base;
    uint32_t sid;
    int wg;
    if (!bsort_locate(tab, ns, [&](int s) { return (uint32_t)((tab[s].nblocks + SCAN_TILE - 1) / SCAN_TILE); }, sid,
                      wg))
        return;
    const SortSlice &t = tab[sid];
    scan_apply_wg(t.has, t.has, t.nblocks, t.scan, t.tile_count, wg, lds, &wg_base);
}

// Tiles, and the warm-start list counts cleared (their lists name old slots).
__global__ void __launch_bounds__(256) k_bsort_emit(const SortSlice *__restrict__ tab, int ns)
{
    uint32_t sid;
    int q;
    if (!bsort_locate(tab, ns, [&](int s) { return (uint32_t)((tab[s].nblocks + 255) / 256); }, sid, q)) return;
    const SortSlice &t = tab[sid];
    if (q == 0 && t.warm_cnt)
        for (int j = (int)threadIdx.x; j < 2 * MAX_GROUPS; j += 256) t.warm_cnt[j] = 0u;
    emit_tile_at(t.blk_start, t.blk_end, t.has, t.nblocks, q * 256 + (int)threadIdx.x, t.desc, t.desc_cap, t.blk_tile,
                 0u);
}

// ---------------------------------------------------------------------------
// P2G over blocks without atomics.
//
// A tile is one node block's particles (grouped by the last sort).  It is
// processed in chunks of at most CHUNK particles:
//   A. one lane per particle (dense and coalesced): FAST loads the position,
//      FUSED runs the particle's G2P update and writes its record (below)
//      to LDS; both find the base node.  The base, relative to the block's
//      8^3-node arena (the block's 4 nodes + 1 below + 3 above per axis),
//      is one of 6^3 bins.
//   B. the tile's waves (FAST: its one wave; FUSED: both) sort the chunk by
//      bin (stable counting sort in LDS) and cut every bin into items of at
//      most ITEM_CAP particles.
//   C. one lane per item: the item's particles share their base node, so the
//      lane sums their 27 node contributions in registers (FAST reloads each
//      particle and builds its record in registers; FUSED reads the LDS
//      record) and adds them to its wave's LDS arena once, in a fixed node
//      order.  In each of the 27 read-modify-writes the lanes of different
//      bins touch different slots; lanes of one bin take turns.
// The waves' arenas are summed into the tile's staging slot, which the grid
// update gathers.  Particles whose stencil leaves the arena (or whose tile
// has no staging slot) are listed for the atomic fallback kernel.
// (On gfx950 an LDS float atomic runs ~20x slower than a plain LDS
// read-modify-write; summing in registers first also divides the LDS traffic
// by the particles per item.)
constexpr int TILE_WAVES = 2;                    // FUSED form: waves per tile
constexpr int TILE_THREADS = 64 * TILE_WAVES;
constexpr int P2G_THREADS = 64;                  // FAST form: one wave per tile
static_assert(ARENA_NODES % TILE_THREADS == 0, "arena nodes per thread must be whole");
constexpr int CHUNK = 384;          // particles per pass over a tile (a full 4^3-cell block holds ~320 at 5 per cell)
static_assert(CHUNK % TILE_THREADS == 0, "whole phase-A passes");
// FAST form: particles per pass over a tile, chosen per simulator
// (mpm_config.tile) or per batch (its autotuner) among these
constexpr int TILE_SHAPES[3] = {256, 384, 512};
constexpr int TILE_DEFAULT = 384;
constexpr int ITEM_CAP = 4;         // particles per lane item
constexpr int NBIN = 216;           // 6^3 base positions inside an arena
constexpr uint32_t NO_BIN = 255u;
// Record of one particle for the register scatter (FP32):
//   [0..2]  fx: position relative to the base node, in cells
//   [3..5]  q0 = m v - h m C fx: the APIC momentum at the base node
//   [6..14] hB = h m C (row-major)
//   [15..20] G = -dt V tau (xx, yy, zz, xy, xz, yz); water: [15] = dt V p J
// so node o (offsets in {0,1,2}^3) receives
//   m w_o  and  w_o (q0 + hB o) + G grad w_o     (MPMCUDA.cu:391-414).
constexpr int REC = 21;

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        uint32_t o = __shfl_xor(v, d);
        v = v > o ? v : o;
    }
    return v;
}

// Mask of the lanes whose 8-bit key equals this lane's.
__device__ __forceinline__ uint64_t match_key8(uint32_t key)
{
    uint64_t m = ~0ull;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        bool bit = (key >> b) & 1u;
        uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

struct TileLds {
    float4 arena[TILE_WAVES][A_SLOTS];
    float rec[REC][CHUNK];
    uint32_t items[CHUNK];
    uint16_t perm[CHUNK];
    uint16_t rnk[CHUNK];
    uint16_t cnt[NBIN];         // phase B: wave 0's counts per bin
    uint16_t cnt1[NBIN];        // phase B: wave 1's counts, then each bin's first sorted position
    uint8_t key[CHUNK];
    uint32_t nitems;
};

// LDS of a FAST-form tile (one wave) of CH particles per pass: the
// particles' fx, bin keys, the sort and the arena (14.5 / 17.1 / 19.8 KB for
// CH = 256 / 384 / 512; the rest of the particle records stays in HBM /
// cache).
template <int CH>
struct TileKeys {
    static_assert(CH % 64 == 0 && CH <= 4095, "whole 64-particle passes; 12-bit item positions");
    float4 arena[A_SLOTS];
    float fx[3][CH];            // phase A: position relative to the base node, in cells
    uint32_t items[CH];
    uint16_t perm[CH];
    uint16_t rnk[CH];
    uint16_t cnt[NBIN];
    uint16_t start[NBIN];       // each bin's first sorted position
    uint8_t key[CH];
    uint32_t nitems;
};

struct TileGeo {
    int ax, ay, az;     // arena origin (node index of arena slot 0)
};

__device__ __forceinline__ TileGeo tile_geo(const GridP &g, uint32_t blk)
{
    const int bx = (int)(blk % (uint32_t)g.nbx);
    const int by = (int)((blk / (uint32_t)g.nbx) % (uint32_t)g.nby);
    const int bz = (int)(blk / ((uint32_t)g.nbx * (uint32_t)g.nby));
    return TileGeo{bx * BLK - 1, by * BLK - 1, bz * BLK - 1};
}

// Phase A tail: the particle's bin key, or NO_BIN and an entry in the list
// of the atomic fallback (whose node blocks are flagged here when `flags` is
// given).  Returns true when the particle was listed.
template <class TL>
__device__ __forceinline__ bool bin_key(const GridP &g, TL &L, const TileGeo &t, bool active, int slot, int i,
                                        const int b[3], bool staged, uint32_t *spill, uint32_t *counters,
                                        uint32_t *flags)
{
    const int lane = threadIdx.x & 63;
    const int lx = b[0] - t.ax, ly = b[1] - t.ay, lz = b[2] - t.az;
    const bool inside = active && staged && lx >= 0 && lx <= ARENA - 3 && ly >= 0 && ly <= ARENA - 3 && lz >= 0 &&
                        lz <= ARENA - 3;
    const bool spill_me = active && !inside;
    const uint64_t sm = __ballot(spill_me);
    if (sm) {
        const int leader = __ffsll((unsigned long long)sm) - 1;
        uint32_t base = 0;
        if (lane == leader) base = gatomic_add(counters + CT_SPILL, (uint32_t)__popcll(sm));
        base = __shfl(base, leader);
        const uint64_t below = lane ? (sm & ((~0ull) >> (64 - lane))) : 0ull;
        if (spill_me) {
            gp(spill)[base + __popcll(below)] = (uint32_t)i;
            if (flags) mark_blocks(g, b, flags);
        }
    }
    if (active) L.key[slot] = inside ? (uint8_t)(lx + 6 * (ly + 6 * lz)) : (uint8_t)NO_BIN;
    return spill_me;
}

// The particle's P2G record from its stencil and scatter quantities:
// fx, q0 = m v - h m C fx, hB = h m C, G (xx, yy, zz, xy, xz, yz).
template <int MAT>
__device__ __forceinline__ void make_record(const GridP &g, const Stencil &s, const Scatter &sc, float q0[3],
                                            float hB[9], float Gs[6])
{
    const float h = 1.0f / g.inv_h;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float bf = sc.B[3 * c] * s.fx[0] + sc.B[3 * c + 1] * s.fx[1] + sc.B[3 * c + 2] * s.fx[2];
        q0[c] = sc.a[c] - h * bf;
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) hB[k] = h * sc.B[k];
    Gs[0] = sc.G[0]; Gs[1] = sc.G[4]; Gs[2] = sc.G[8]; Gs[3] = sc.G[1]; Gs[4] = sc.G[2]; Gs[5] = sc.G[5];
}

// FUSED form phase A tail: bin the particle and write its record to LDS, or
// list it for the atomic fallback.  Returns true when it was listed.
template <int MAT>
__device__ __forceinline__ bool bin_particle(const GridP &g, TileLds &L, const TileGeo &t, bool active, int slot,
                                             int i, const Stencil &s, const Scatter &sc, bool staged,
                                             uint32_t *spill, uint32_t *counters)
{
    const bool spill_me = bin_key(g, L, t, active, slot, i, s.b, staged, spill, counters, nullptr);
    if (active && !spill_me) {
        float q0[3], hB[9], Gs[6];
        make_record<MAT>(g, s, sc, q0, hB, Gs);
#pragma unroll
        for (int a = 0; a < 3; ++a) { L.rec[a][slot] = s.fx[a]; L.rec[3 + a][slot] = q0[a]; }
#pragma unroll
        for (int k = 0; k < 9; ++k) L.rec[6 + k][slot] = hB[k];
#pragma unroll
        for (int k = 0; k < (MAT == MPM_WATER ? 1 : 6); ++k) L.rec[15 + k][slot] = Gs[k];
    }
    return spill_me;
}

// Phase B items of a chunk from its per-bin counts (one wave; lane l owns
// bins 4l .. 4l+3, counts c[4]).  A bin of c particles gives c / ITEM_CAP
// full items and one item of the remainder.  Full items come first, then
// remainders by decreasing size, so a batch of 64 items mostly holds items
// of equal length (fewer idle lanes per round) and the items of one bin
// mostly fall into different batches (fewer turns in the flush).  Each
// bin's first sorted position goes to start[].  Items pack
// bin | first sorted position << 8 | count << 20.
__device__ __forceinline__ void build_items(const uint32_t c[4], uint32_t *items, uint16_t *start, uint32_t *nitems)
{
    const int lane = threadIdx.x & 63;
    uint32_t nfull = 0, csum = 0, nrem[ITEM_CAP];
#pragma unroll
    for (int r = 0; r < ITEM_CAP; ++r) nrem[r] = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        csum += c[j];
        nfull += c[j] / ITEM_CAP;
#pragma unroll
        for (int r = 1; r < ITEM_CAP; ++r) nrem[r] += (c[j] % ITEM_CAP) == (uint32_t)r ? 1u : 0u;
    }
    uint32_t crun = wave_incl_scan(csum) - csum;
    const uint32_t fincl = wave_incl_scan(nfull);
    uint32_t frun = fincl - nfull;
    uint32_t total = __shfl(fincl, 63);
    uint32_t rrun[ITEM_CAP];
#pragma unroll
    for (int r = ITEM_CAP - 1; r >= 1; --r) {
        const uint32_t inc = wave_incl_scan(nrem[r]);
        rrun[r] = total + inc - nrem[r];
        total += __shfl(inc, 63);
    }
    if (lane == 0) *nitems = total;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int b = 4 * lane + j;
        if (b < NBIN) start[b] = (uint16_t)crun;
        const uint32_t nf = c[j] / ITEM_CAP, rm = c[j] % ITEM_CAP;
        for (uint32_t q = 0; q < nf; ++q)
            items[frun + q] = (uint32_t)b | ((crun + q * ITEM_CAP) << 8) | ((uint32_t)ITEM_CAP << 20);
        frun += nf;
#pragma unroll
        for (int r = 1; r < ITEM_CAP; ++r)
            if (rm == (uint32_t)r) items[rrun[r]++] = (uint32_t)b | ((crun + nf * ITEM_CAP) << 8) | (rm << 20);
        crun += c[j];
    }
}

// Phase B, FUSED form (both waves): counting sort of the chunk's nc
// particles by bin and the item list.  Wave w counts the 64-particle
// batches w, w + 2, ... into its own counts; wave 0 then builds the items
// from the summed counts; both waves place their particles.  Within a bin,
// wave 0's particles precede wave 1's, each in slot order, so the order
// (and every sum) is the same in every run.
__device__ __forceinline__ void sort_chunk(TileLds &L, int nc)
{
    static_assert(TILE_WAVES == 2, "two counting waves");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lt = lane ? ((~0ull) >> (64 - lane)) : 0ull;
    uint16_t *cw = wave ? L.cnt1 : L.cnt;
    for (int b = lane; b < NBIN; b += 64) cw[b] = 0;
    __asm__ volatile("" ::: "memory");
    for (int s0 = wave * 64; s0 < nc; s0 += TILE_THREADS) {
        const int sl = s0 + lane;
        const uint32_t key = sl < nc ? (uint32_t)L.key[sl] : NO_BIN;
        const bool v = key != NO_BIN;
        const uint64_t m = match_key8(key) & __ballot(v);
        const uint32_t below = (uint32_t)__popcll(m & lt);
        uint32_t prev = 0;
        if (v) prev = cw[key];
        if (v && below == 0) cw[key] = (uint16_t)(prev + (uint32_t)__popcll(m));
        if (v) L.rnk[sl] = (uint16_t)(prev + below);
        __asm__ volatile("" ::: "memory");
    }
    __syncthreads();
    if (wave == 0) {
        uint32_t c[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int b = 4 * lane + j;
            c[j] = b < NBIN ? (uint32_t)L.cnt[b] + (uint32_t)L.cnt1[b] : 0u;
        }
        __asm__ volatile("" ::: "memory");
        build_items(c, L.items, L.cnt1, &L.nitems);      // cnt1 becomes each bin's first position
    }
    __syncthreads();
    for (int s0 = wave * 64; s0 < nc; s0 += TILE_THREADS) {
        const int sl = s0 + lane;
        const uint32_t key = sl < nc ? (uint32_t)L.key[sl] : NO_BIN;
        if (key != NO_BIN) {
            const uint32_t off = wave ? (uint32_t)L.cnt[key] : 0u;     // wave 1 after wave 0's particles
            L.perm[(uint32_t)L.cnt1[key] + off + L.rnk[sl]] = (uint16_t)sl;
        }
    }
}

// Phase B, FAST form (the tile's one wave): the same sort, in slot order
// within each bin.  One wave's LDS operations execute in order, so the
// steps need no barrier; the compiler barriers keep their order.
template <class TK>
__device__ __forceinline__ void sort_chunk_wave(TK &L, int nc)
{
    const int lane = threadIdx.x & 63;
    const uint64_t lt = lane ? ((~0ull) >> (64 - lane)) : 0ull;
    for (int b = lane; b < NBIN; b += 64) L.cnt[b] = 0;
    __asm__ volatile("" ::: "memory");
    for (int s0 = 0; s0 < nc; s0 += 64) {
        const int sl = s0 + lane;
        const uint32_t key = sl < nc ? (uint32_t)L.key[sl] : NO_BIN;
        const bool v = key != NO_BIN;
        const uint64_t m = match_key8(key) & __ballot(v);
        const uint32_t below = (uint32_t)__popcll(m & lt);
        uint32_t prev = 0;
        if (v) prev = L.cnt[key];
        if (v && below == 0) L.cnt[key] = (uint16_t)(prev + (uint32_t)__popcll(m));
        if (v) L.rnk[sl] = (uint16_t)(prev + below);
        __asm__ volatile("" ::: "memory");
    }
    uint32_t c[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int b = 4 * lane + j;
        c[j] = b < NBIN ? (uint32_t)L.cnt[b] : 0u;
    }
    __asm__ volatile("" ::: "memory");
    build_items(c, L.items, L.start, &L.nitems);
    __asm__ volatile("" ::: "memory");
    for (int s0 = 0; s0 < nc; s0 += 64) {
        const int sl = s0 + lane;
        const uint32_t key = sl < nc ? (uint32_t)L.key[sl] : NO_BIN;
        if (key != NO_BIN) L.perm[(uint32_t)L.start[key] + L.rnk[sl]] = (uint16_t)sl;
    }
    __asm__ volatile("" ::: "memory");
}

// Phase C, one particle given its record in registers (fx, q0, hB and G as
// xx, yy, zz, xy, xz, yz; water: the diagonal value in Gs[0]): add its 27
// node contributions to acc (node ox + 3 oy + 9 oz).  With N the B-spline
// weights, D = dN/dx and w = Nx Ny Nz, node o receives m w and
//   p_c(o) = w (q0_c + hB_c . o) + G_c . grad w
//          = Nz alpha_c(ox, oy) + (Nx Ny) zeta_c(oz)
//   alpha_c = Ny (Nx (q0_c + ox hB_c0) + G_c0 Dx) + Nx (oy Ny hB_c1 + G_c1 Dy)
//   zeta_c  = oz Nz hB_c2 + G_c2 Dz
// so each node costs one FMA for the mass and two per momentum component.
// Offsets enter through kadd (no multiply by a constant 0, which the
// compiler must keep); water's G is diagonal.
__device__ __forceinline__ float kadd(int k, float a, float b) { return k == 0 ? a : (k == 1 ? a + b : a + (b + b)); }

template <int MAT>
__device__ __forceinline__ void accum_regs(const float fx[3], const float q0[3], const float hB[9], const float Gs[6],
                                           float inv_h, float m, float4 acc[27])
{
    constexpr bool W = MAT == MPM_WATER;
    float G[9];
    if (W) {
        G[0] = G[4] = G[8] = Gs[0];
        G[1] = G[2] = G[3] = G[5] = G[6] = G[7] = 0.0f;     // never read
    } else {
        G[0] = Gs[0]; G[4] = Gs[1]; G[8] = Gs[2];
        G[1] = G[3] = Gs[3];
        G[2] = G[6] = Gs[4];
        G[5] = G[7] = Gs[5];
    }
    // G_ck is structurally zero for water off the diagonal
#define GZ(c, k) (W && (c) != (k))
    float N[3][3], D[3][3];
    bspline3_wave(fx, N, D);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k) D[a][k] *= inv_h;
    // ey[oy][c] = oy Ny hB_c1 + G_c1 Dy and ez[oz][c] = oz Nz hB_c2 + G_c2 Dz
    // (zeta); eyz marks a structural zero
    float ey[3][3], ez[3][3];
#pragma unroll
    for (int o = 0; o < 3; ++o)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float gy = GZ(c, 1) ? 0.0f : G[3 * c + 1] * D[1][o];
            const float ny = o == 0 ? 0.0f : N[1][o] * (o == 1 ? hB[3 * c + 1] : hB[3 * c + 1] + hB[3 * c + 1]);
            ey[o][c] = o == 0 ? gy : (GZ(c, 1) ? ny : ny + gy);
            const float gz = GZ(c, 2) ? 0.0f : G[3 * c + 2] * D[2][o];
            const float nz = o == 0 ? 0.0f : N[2][o] * (o == 1 ? hB[3 * c + 2] : hB[3 * c + 2] + hB[3 * c + 2]);
            ez[o][c] = o == 0 ? gz : (GZ(c, 2) ? nz : nz + gz);
        }
    float mz[3];
#pragma unroll
    for (int oz = 0; oz < 3; ++oz) mz[oz] = m * N[2][oz];
#pragma unroll
    for (int ox = 0; ox < 3; ++ox) {
        float sx[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float t = N[0][ox] * kadd(ox, q0[c], hB[3 * c]);
            sx[c] = GZ(c, 0) ? t : fmaf(G[3 * c], D[0][ox], t);
        }
#pragma unroll
        for (int oy = 0; oy < 3; ++oy) {
            const float nxy = N[0][ox] * N[1][oy];
            float al[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const bool eyz = oy == 0 && GZ(c, 1);
                al[c] = eyz ? N[1][oy] * sx[c] : fmaf(N[1][oy], sx[c], N[0][ox] * ey[oy][c]);
            }
#pragma unroll
            for (int oz = 0; oz < 3; ++oz) {
                float4 &A = acc[ox + 3 * oy + 9 * oz];
                const float nz = N[2][oz];
                A.x = fmaf(nxy, mz[oz], A.x);
                float *Ac[3] = {&A.y, &A.z, &A.w};
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    float v = fmaf(nz, al[c], *Ac[c]);
                    if (!(oz == 0 && GZ(c, 2))) v = fmaf(nxy, ez[oz][c], v);
                    *Ac[c] = v;
                }
            }
        }
    }
#undef GZ
}

// The same from the particle's LDS record (FUSED form).
template <int MAT>
__device__ __forceinline__ void accum_particle(const TileLds &L, int slot, float inv_h, float m, float4 acc[27])
{
    float fx[3], q0[3], hB[9], Gs[6];
#pragma unroll
    for (int a = 0; a < 3; ++a) fx[a] = L.rec[a][slot];
#pragma unroll
    for (int c = 0; c < 3; ++c) q0[c] = L.rec[3 + c][slot];
#pragma unroll
    for (int k = 0; k < 9; ++k) hB[k] = L.rec[6 + k][slot];
#pragma unroll
    for (int k = 0; k < 6; ++k) Gs[k] = (MAT == MPM_WATER && k > 0) ? 0.0f : L.rec[15 + k][slot];
    accum_regs<MAT>(fx, q0, hB, Gs, inv_h, m, acc);
}

// Phase C flush: the lane's 27 accumulated nodes are added to its wave's
// arena in a fixed order (compiler barrier; LDS executes one wave's
// operations in order).  Lanes of different bins touch different slots in
// each instruction; lanes of the same bin take turns.
__device__ __forceinline__ void flush_item(float4 *arena, bool has, uint32_t bin, const float4 acc[27])
{
    const int lane = threadIdx.x & 63;
    const int bx = (int)(bin % 6u), by = (int)((bin / 6u) % 6u), bz = (int)(bin / 36u);
    uint64_t pending = __ballot(has);
    const uint64_t same = match_key8(bin);
    while (pending) {
        const uint64_t grp = same & pending;
        const bool winner = has && ((pending >> lane) & 1ull) && (__ffsll((unsigned long long)grp) - 1 == lane);
        if (winner) {
#pragma unroll
            for (int oz = 0; oz < 3; ++oz)
#pragma unroll
                for (int oy = 0; oy < 3; ++oy)
#pragma unroll
                    for (int ox = 0; ox < 3; ++ox) {
                        const int a = a_slot(bx + ox, by + oy, bz + oz);
                        const float4 &c = acc[ox + 3 * oy + 9 * oz];
                        float4 v = arena[a];
                        v.x += c.x; v.y += c.y; v.z += c.z; v.w += c.w;
                        arena[a] = v;
                        __asm__ volatile("" ::: "memory");
                    }
        }
        pending &= ~__ballot(winner);
    }
}

// Phase C for one wave, FUSED form: items [batch * 64, ...) of the chunk,
// records from LDS.
template <int MAT>
__device__ __forceinline__ void scatter_items(TileLds &L, float4 *arena, int batch, float inv_h, float m)
{
    const int lane = threadIdx.x & 63;
    const uint32_t it = (uint32_t)(batch * 64 + lane);
    const bool has = it < L.nitems;
    const uint32_t item = has ? L.items[it] : 0u;
    const uint32_t bin = has ? (item & 255u) : NO_BIN;
    const uint32_t first = (item >> 8) & 4095u, cnt = has ? (item >> 20) : 0u;
    float4 acc[27];
#pragma unroll
    for (int o = 0; o < 27; ++o) acc[o] = make_float4(0.f, 0.f, 0.f, 0.f);
    // item counts never increase along the item list (build_items), so the
    // batch's first lane holds its largest
    const uint32_t rounds = (uint32_t)__builtin_amdgcn_readfirstlane((int)cnt);
    for (uint32_t r = 0; r < rounds; ++r)
        if (r < cnt) accum_particle<MAT>(L, L.perm[first + r], inv_h, m, acc);
    flush_item(arena, has, bin, acc);
}

// The P2G record straight from a loaded particle (FAST form): fx, hB = hm C
// with hm = h m, q0 = m v - hB fx, and G (xx, yy, zz, xy, xz, yz) = -dt V
// tau, water: G[0] = dt V p J (MPMCUDA.cu:345-353, 391-414).
template <int MAT>
__device__ __forceinline__ void record_from_raw(const MatP &mat, float dt, float hm, const RawP &r,
                                                const float fx[3], float q0[3], float hB[9], float Gs[6])
{
#pragma unroll
    for (int k = 0; k < 9; ++k) hB[k] = hm * r.C[k];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float q = mat.mass * r.v[c];
        q = fmaf(-hB[3 * c], fx[0], q);
        q = fmaf(-hB[3 * c + 1], fx[1], q);
        q0[c] = fmaf(-hB[3 * c + 2], fx[2], q);
    }
    if (MAT == MPM_WATER) {
        const float J = r.s[0];
        const float pr = mat.eos_k * (powf(J, -mat.eos_gamma) - 1.0f);   // MPMCUDA.cu:353
        Gs[0] = dt * mat.vol * pr * J;
#pragma unroll
        for (int k = 1; k < 6; ++k) Gs[k] = 0.0f;
    } else {
        const float c = -dt * mat.vol;
#pragma unroll
        for (int k = 0; k < 6; ++k) Gs[k] = c * r.s[k];
    }
}

// v, C and the stored stress (water: J) of particle i: the FAST P2G's phase
// C inputs besides fx, which phase A left in LDS.
template <int MAT>
__device__ __forceinline__ void load_vcs(const PartView &pv, int i, RawP &r)
{
    const uint32_t u = (uint32_t)i;
#pragma unroll
    for (int a = 0; a < 3; ++a) r.v[a] = ld_plane(pv, PL_V + a, u);
#pragma unroll
    for (int k = 0; k < 9; ++k) r.C[k] = ld_plane(pv, PL_C + k, u);
    if (MAT == MPM_WATER) {
        r.s[0] = ld_plane(pv, PL_S, u);
    } else {
#pragma unroll
        for (int k = 0; k < 6; ++k) r.s[k] = ld_plane(pv, PL_TAU + k, u);
    }
}

// Phase C for one wave, FAST form: the item's particles are loaded again
// from the particle planes (the tile's lines are in cache from phase A or
// arrive now; no LDS records) and their records built in registers with the
// fx phase A computed.
template <int MAT, class TK>
__device__ __forceinline__ void scatter_items_reload(TK &L, float4 *arena, int batch, const GridP &g,
                                                     const MatP &mat, const PartView &pv, int i0, float dt, float hm)
{
    const int lane = threadIdx.x & 63;
    const uint32_t it = (uint32_t)(batch * 64 + lane);
    const bool has = it < L.nitems;
    const uint32_t item = has ? L.items[it] : 0u;
    const uint32_t bin = has ? (item & 255u) : NO_BIN;
    const uint32_t first = (item >> 8) & 4095u, cnt = has ? (item >> 20) : 0u;
    float4 acc[27];
#pragma unroll
    for (int o = 0; o < 27; ++o) acc[o] = make_float4(0.f, 0.f, 0.f, 0.f);
    // item counts never increase along the item list (build_items), so the
    // batch's first lane holds its largest
    const uint32_t rounds = (uint32_t)__builtin_amdgcn_readfirstlane((int)cnt);
    // the next round's particle is loaded while this one is accumulated;
    // lanes without a next particle reload a valid one (one basic block)
    auto slot_of = [&](uint32_t r) -> int { return cnt ? (int)L.perm[first + min(r, cnt - 1u)] : 0; };
    RawP nxt;
    float nfx[3];
    int sl = slot_of(0);
    load_vcs<MAT>(pv, i0 + sl, nxt);
#pragma unroll
    for (int a = 0; a < 3; ++a) nfx[a] = L.fx[a][sl];
    for (uint32_t r = 0; r < rounds; ++r) {
        const RawP raw = nxt;
        const float fx[3] = {nfx[0], nfx[1], nfx[2]};
        if (r + 1 < rounds) {
            sl = slot_of(r + 1);
            load_vcs<MAT>(pv, i0 + sl, nxt);
#pragma unroll
            for (int a = 0; a < 3; ++a) nfx[a] = L.fx[a][sl];
        }
        if (r < cnt) {
            float q0[3], hB[9], Gs[6];
            record_from_raw<MAT>(mat, dt, hm, raw, fx, q0, hB, Gs);
            accum_regs<MAT>(fx, q0, hB, Gs, g.inv_h, mat.mass, acc);
        }
    }
    flush_item(arena, has, bin, acc);
}

// Phases B and C of one chunk (all threads; ends synchronized).
template <int MAT>
__device__ __forceinline__ void scatter_chunk(TileLds &L, int nc, float inv_h, float m)
{
    __syncthreads();
    sort_chunk(L, nc);
    __syncthreads();
    const int wave = threadIdx.x >> 6;
    const int nb = (int)((L.nitems + 63u) >> 6);
    for (int b = wave; b < nb; b += TILE_WAVES) scatter_items<MAT>(L, L.arena[wave], b, inv_h, m);
    __syncthreads();
}

__device__ __forceinline__ void tile_clear(TileLds &L)
{
    for (int k = threadIdx.x; k < TILE_WAVES * A_SLOTS; k += TILE_THREADS)
        L.arena[k / A_SLOTS][k % A_SLOTS] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Sum the waves' arenas into staging slot `slot` (logical node order) and
// mark the node blocks that received mass.
__device__ __forceinline__ void tile_end(const GridP &g, const TileGeo &t, const TileLds &L, float4 *staging,
                                         uint32_t slot, bool staged, uint32_t *flags, int32_t *blk_entry)
{
    __syncthreads();
    if (staged) {
        // the whole arena is written: every face
        if (threadIdx.x == 0) *gp(blk_entry) = tile_entry(slot, TILE_ALL_FACES);
        const int64_t dst = (int64_t)slot * ARENA_NODES;
        for (int k = threadIdx.x; k < ARENA_NODES; k += TILE_THREADS) {
            const int kx = k & 7, ky = (k >> 3) & 7, kz = k >> 6;
            const int a = a_slot(kx, ky, kz);
            float4 v = L.arena[0][a];
#pragma unroll
            for (int w = 1; w < TILE_WAVES; ++w) {
                const float4 u = L.arena[w][a];
                v.x += u.x; v.y += u.y; v.z += u.z; v.w += u.w;
            }
            gstore4(staging, dst + k, v);
            const int nx_ = t.ax + kx, ny_ = t.ay + ky, nz_ = t.az + kz;
            if (v.x != 0.f && nx_ >= 0 && ny_ >= 0 && nz_ >= 0 && nx_ < g.nx && ny_ < g.ny && nz_ < g.nz)
                gp(flags)[((int64_t)(nz_ >> 2) * g.nby + (ny_ >> 2)) * g.nbx + (nx_ >> 2)] = 1u;
        }
    }
    __syncthreads();
}

// One tile (descriptor d = first particle, count, block, staging slot) of
// the FAST-form P2G, by one wave; returns health bits.  Phase A loads only
// positions (bin keys); phase C reloads x, v, C and the stored stress of
// each item's particles, so the tile needs no LDS records (12.5 KB instead
// of 53 KB) and no barrier, and the register file alone bounds the tiles
// in flight per CU (8 one-wave tiles instead of 3 two-wave ones with LDS
// records: 77 vs 106 us per 1M-particle snow step, measured).
template <int MAT, int CH>
__device__ __forceinline__ uint32_t p2g_tile(const GridP &g, const MatP &mat, const PartView &pv, const uint4 d,
                                             uint32_t staging_cap, float dt, float4 *__restrict__ staging,
                                             uint32_t *spill, uint32_t *flags, uint32_t *counters, int32_t *blk_tile,
                                             TileKeys<CH> &L)
{
    uint32_t bits = 0;
    const TileGeo t = tile_geo(g, d.z);
    const bool staged = d.w < staging_cap;
    const float hm = (1.0f / g.inv_h) * mat.mass;
    const int lane = threadIdx.x & 63;
    for (int k = lane; k < A_SLOTS; k += 64) L.arena[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    constexpr int PASSES = CH / 64;
    uint32_t lane_faces = 0;    // arena faces this lane's particles' stencils reach (see tile_entry)
    for (uint32_t c0 = 0; c0 < d.y; c0 += CH) {
        const int nc = (int)min((uint32_t)CH, d.y - c0);
        // A: all position loads of the chunk first (lanes past its end reload
        // its last particle; passes past its end load nothing), then bin keys
        float xs[PASSES][3];
#pragma unroll
        for (int it = 0; it < PASSES; ++it) {
            if (it * 64 >= nc) break;
            const uint32_t slot = (uint32_t)min(it * 64 + lane, nc - 1);
#pragma unroll
            for (int a = 0; a < 3; ++a) xs[it][a] = ld_plane(pv, PL_X + a, d.x + c0 + slot);
        }
#pragma unroll
        for (int it = 0; it < PASSES; ++it) {
            if (it * 64 >= nc) break;
            const int slot = it * 64 + lane;
            const bool active = slot < nc;
            int b[3] = {0, 0, 0};
            float f[3];
            if (active) {
                bits |= base_fx(g, xs[it], b, f);
#pragma unroll
                for (int a = 0; a < 3; ++a) L.fx[a][slot] = f[a];
            }
            const bool listed = bin_key(g, L, t, active, slot, (int)(d.x + c0) + slot, b, staged, spill, counters,
                                        flags);
            // a base on the arena's first (last) bin position reaches its
            // face at index 0 (7)
            if (active && !listed) {
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const int rel = b[a] - (a == 0 ? t.ax : (a == 1 ? t.ay : t.az));
                    lane_faces |= (rel == 0 ? 1u : 0u) << (2 * a);
                    lane_faces |= (rel == ARENA - 3 ? 2u : 0u) << (2 * a);
                }
            }
        }
        __asm__ volatile("" ::: "memory");
        // B and C
        sort_chunk_wave(L, nc);
        const int nb = (int)((L.nitems + 63u) >> 6);
        for (int b = 0; b < nb; ++b) scatter_items_reload<MAT>(L, L.arena, b, g, mat, pv, (int)(d.x + c0), dt, hm);
        __asm__ volatile("" ::: "memory");
    }
    // the arena into the tile's staging slot (logical node order): indices
    // 1..6 of each axis and the faces some stencil reached (every other node
    // is zero); mark the node blocks that received mass
    if (staged) {
        uint32_t faces = 0;
#pragma unroll
        for (int f = 0; f < 6; ++f)
            if (__ballot((lane_faces >> f) & 1u)) faces |= 1u << f;
        const int64_t dst = (int64_t)d.w * ARENA_NODES;
        const int lo[3] = {(faces & 1u) ? 0 : 1, (faces & 4u) ? 0 : 1, (faces & 16u) ? 0 : 1};
        const int hi[3] = {(faces & 2u) ? 7 : 6, (faces & 8u) ? 7 : 6, (faces & 32u) ? 7 : 6};
        const int kx = lane & 7, ky = lane >> 3;
        const bool row = kx >= lo[0] && kx <= hi[0] && ky >= lo[1] && ky <= hi[1];
        // all eight z planes are read from LDS back to back (the accumulators
        // are dead here), then the planes in range are stored
        float4 v[ARENA];
#pragma unroll
        for (int kz = 0; kz < ARENA; ++kz) v[kz] = L.arena[a_slot(kx, ky, kz)];
#pragma unroll
        for (int kz = 0; kz < ARENA; ++kz) {
            if (!row || kz < lo[2] || kz > hi[2]) continue;
            gstore4(staging, dst + kz * 64 + lane, v[kz]);
            const int nx_ = t.ax + kx, ny_ = t.ay + ky, nz_ = t.az + kz;
            if (v[kz].x != 0.f && nx_ >= 0 && ny_ >= 0 && nz_ >= 0 && nx_ < g.nx && ny_ < g.ny && nz_ < g.nz)
                gp(flags)[((int64_t)(nz_ >> 2) * g.nby + (ny_ >> 2)) * g.nbx + (nx_ >> 2)] = 1u;
        }
        if (lane == 0) gp(blk_tile)[d.z] = tile_entry(d.w, faces);
    }
    __asm__ volatile("" ::: "memory");
    return bits;
}

template <int MAT, int CH>
__global__ void __launch_bounds__(P2G_THREADS) __attribute__((amdgpu_waves_per_eu(2))) k_p2g_block(GridP g, MatP mat, PartView pv,
                                                            const uint4 *__restrict__ desc,
                                                            const uint32_t *__restrict__ tile_count, uint32_t desc_cap,
                                                            uint32_t staging_cap, float dt,
                                                            float4 *__restrict__ staging, uint32_t *spill,
                                                            uint32_t *flags, uint32_t *counters, int32_t *blk_tile,
                                                            int clear_active)
{
    __shared__ TileKeys<CH> L;
    const uint32_t nt = min(*tile_count, desc_cap);
    // the active-block and node counters are not used during P2G: clear
    // them here for the block compaction that follows
    if (clear_active && blockIdx.x == 0 && threadIdx.x == 0) {
        gp(counters)[CT_ACTIVE] = 0u;
        gp(counters)[CT_NODES] = 0u;
    }
    uint32_t bits = 0;
    const uint32_t nt8 = round_xcd(nt);
    dt = step_dt(g, dt);
    for (uint32_t k8 = blockIdx.x; k8 < nt8; k8 += gridDim.x) {
        const uint32_t tt = xcd_spread(k8, nt8);
        if (tt >= nt) continue;
        bits |= p2g_tile<MAT, CH>(g, mat, pv, desc[tt], staging_cap, dt, staging, spill, flags, counters, blk_tile,
                                  L);
    }
    if (bits) atomicOr(counters + CT_HEALTH, bits);
    if (threadIdx.x == 0 && blockIdx.x == 0) counters[CT_TILES] = nt;
}

// Batched P2G: the tiles of the scenes of material MAT among up to
// P2G_CHUNK scenes of a batch in one persistent launch (one launch per
// material -- a run-time material switch over the three tile bodies doubles
// the tile time at this register cap -- and per chunk of P2G_CHUNK scenes).
// Work index = the tiles of the first such scene, then of the next, ...;
// workgroups walk it with the XCD-aware stride of the single-scene kernel.
// The scene map (end of each scene's tile range) sits in LDS beside the
// tile's (256 B; LDS is allocated in granules, so any addition is checked
// against the workgroups per CU it leaves).
constexpr int P2G_CHUNK = 64;
template <int MAT, int CH>
__global__ void __launch_bounds__(P2G_THREADS) __attribute__((amdgpu_waves_per_eu(2))) k_p2g_batch(const BatchSlice *__restrict__ tab, int nslices)
{
    __shared__ TileKeys<CH> L;
    __shared__ uint32_t s_end[P2G_CHUNK];
    __shared__ uint32_t s_total;
    if (threadIdx.x < 64) {
        SliceScan sc;
        sc.build(nslices, [&](int s) {
            return tab[s].mat.type == MAT ? min(*gp(tab[s].tile_count), tab[s].desc_cap) : 0u;
        });
        s_end[threadIdx.x] = sc.end[0];
        if (threadIdx.x == 0) s_total = sc.total;
    }
    __syncthreads();
    const uint32_t nt = s_total, nt8 = round_xcd(nt);
    const int lane = threadIdx.x & 63;
    for (uint32_t k8 = blockIdx.x; k8 < nt8; k8 += gridDim.x) {
        const uint32_t tt = xcd_spread(k8, nt8);
        if (tt >= nt) continue;
        const uint32_t sid = (uint32_t)__builtin_amdgcn_readfirstlane((int)__popcll(__ballot(s_end[lane] <= tt)));
        const BatchSlice &b = tab[sid];
        const uint4 d = gload_u4(b.desc, tt - (sid ? s_end[sid - 1] : 0u));
        const uint32_t bits =
            p2g_tile<MAT, CH>(b.g, b.mat, b.pv, d, b.staging_cap, b.dt, b.staging, b.spill, b.flags, b.counters,
                              b.blk_tile, L);
        if (bits) gatomic_or(b.counters + CT_HEALTH, bits);
    }
    if (blockIdx.x == 0)
        for (int s = threadIdx.x; s < nslices; s += blockDim.x)
            if (tab[s].mat.type == MAT) gp(tab[s].counters)[CT_TILES] = min(*gp(tab[s].tile_count), tab[s].desc_cap);
}

// ---------------------------------------------------------------------------
// FUSED form: G2P of step n and P2G of step n+1 in one pass over a block.
// The block's 8^3 grid-velocity neighbourhood is copied to LDS; in phase A
// each particle is gathered, updated (one SVD) and advected, its state is
// written back, and its record for step n+1 is built from the new state;
// phases B and C scatter the chunk into the arenas, which are staged for the
// grid update of step n+1.
template <int MAT>
__global__ void __launch_bounds__(TILE_THREADS) __attribute__((amdgpu_waves_per_eu(2))) k_g2p2g_block(GridP g, MatP mat, PartView pv,
                                                              const uint4 *__restrict__ desc,
                                                              const uint32_t *__restrict__ tile_count,
                                                              uint32_t desc_cap, uint32_t staging_cap, float dt,
                                                              float dt_next, const float4 *__restrict__ gv,
                                                              float4 *__restrict__ staging, uint32_t *spill,
                                                              uint32_t *flags, uint32_t *counters, int sweeps,
                                                              int32_t *blk_tile)
{
    __shared__ TileLds L;
    __shared__ float4 garena[A_SLOTS];
    const uint32_t nt = min(*tile_count, desc_cap);
    uint32_t bits = 0;
    const uint32_t nt8 = round_xcd(nt);
    for (uint32_t k8 = blockIdx.x; k8 < nt8; k8 += gridDim.x) {
        const uint32_t tt = xcd_spread(k8, nt8);
        if (tt >= nt) continue;
        const uint4 d = desc[tt];
        const TileGeo t = tile_geo(g, d.z);
        const bool staged = d.w < staging_cap;
        tile_clear(L);
        // grid velocities of the arena nodes (zero outside the grid); all
        // loads are issued before the first LDS write
        {
            constexpr int R = ARENA_NODES / TILE_THREADS;
            float4 v[R];
            bool in[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int k = (int)threadIdx.x + r * TILE_THREADS;
                const int x = t.ax + (k & 7), y = t.ay + ((k >> 3) & 7), z = t.az + (k >> 6);
                in[r] = x >= 0 && y >= 0 && z >= 0 && x < g.nx && y < g.ny && z < g.nz;
                v[r] = gv[node_slot(g, min(max(x, 0), g.nx - 1), min(max(y, 0), g.ny - 1), min(max(z, 0), g.nz - 1))];
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int k = (int)threadIdx.x + r * TILE_THREADS;
                garena[a_slot(k & 7, (k >> 3) & 7, k >> 6)] = in[r] ? v[r] : make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }
        __syncthreads();
        const ArenaFetch fetch{garena, gv, t.ax, t.ay, t.az};
        for (uint32_t c0 = 0; c0 < d.y; c0 += CHUNK) {
            const int nc = (int)min((uint32_t)CHUNK, d.y - c0);
            for (int s0 = (int)(threadIdx.x & ~63u); s0 < nc; s0 += TILE_THREADS) {
                const int slot = s0 + (int)(threadIdx.x & 63u);
                const bool active = slot < nc;
                const int i = (int)(d.x + c0) + slot;
                Stencil s;
                s.b[0] = s.b[1] = s.b[2] = 0;
                Scatter sc;
                G2POut o;
                if (active) {
                    bits |= g2p_core<MAT>(g, mat, pv, i, pv, i, dt, fetch, MAT != MPM_WATER, sweeps, counters, o);
                    bits |= make_stencil(g, o.x, s);
                    scatter_from_out<MAT>(mat, o, dt_next, sc);
                }
                const bool listed = bin_particle<MAT>(g, L, t, active, slot, i, s, sc, staged, spill, counters);
                if (listed && MAT != MPM_WATER) {
                    // the atomic fallback reads the stress from memory
#pragma unroll
                    for (int k = 0; k < 6; ++k) st(pv, PL_TAU + k, i, o.tau[k]);
                }
            }
            scatter_chunk<MAT>(L, nc, g.inv_h, mat.mass);
        }
        tile_end(g, t, L, staging, d.w, staged, flags, blk_tile + d.z);
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
        int64_t hist_len = 512 * ntiles;     // digits of up to 9 bits
        // scan partials of the tile numbering, or the radix digit totals
        int64_t scan_words = std::max<int64_t>(RS_MAX_DIGITS, (s->nblocks + SCAN_TILE - 1) / SCAN_TILE + 16);
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

// Sort every group's particles by node block and build the tiles.  With
// `defer` the particles are not moved: the order is left in g2p_perm for the
// G2P that follows (FAST form), which writes its updated particles in that
// order -- the re-sort then costs no pass of its own over the 31 planes.
int sort_all(mpm_sim *s, bool defer)
{
    int rc = ensure_fast_buffers(s);
    if (rc) return rc;
    hipStream_t st = s->stream;
    int bits = KEY_LOW + 1;
    while ((1ll << (bits - KEY_LOW)) < s->nblocks) ++bits;
    uint64_t plane_mask = (1ull << PL_COUNT) - 1ull;
    if (!s->tau_valid) plane_mask &= ~(((1ull << 6) - 1ull) << PL_TAU);
    if (defer && s->g2p_perm_cap < s->cap) {
        if (s->g2p_perm) (void)hipFree(s->g2p_perm);
        s->g2p_perm = nullptr;
        s->g2p_perm_cap = 0;
        MPM_HIP(hipMalloc(&s->g2p_perm, (size_t)s->cap * sizeof(uint32_t)));
        s->g2p_perm_cap = s->cap;
    }
    // warm-start quaternions travel with their particles while the warm start
    // lists particles (the lists' indices are dropped: they name slots of the
    // old order); a deferred order keeps the lists until the G2P moves the
    // particles (its inference reads the old slots, its list the new ones)
    if (!s->warm_cnt) plane_mask &= ~(((1ull << 4) - 1ull) << PL_Q);
    else if (!defer && (rc = launch_zero_u32(st, s->warm_cnt, 2 * MAX_GROUPS))) return hip_fail((hipError_t)rc, "zero");
    const int nb = (int)s->nblocks;
    const int db = radix_digit_bits(bits);
    const uint32_t dmask0 = (1u << std::min(db, bits)) - 1u;
    for (size_t k = 0; k < s->groups.size(); ++k) {
        PartView pv = view_of(s, (int)k);
        int n = pv.n;
        uint32_t *blk_end = s->blk_count;
        // keys, first digit histogram, cleared block arrays; the sort's
        // values are the particles' current slots
        const int ntiles = nblk(n, RS_TILE);
        if (db == 9)
            hipLaunchKernelGGL(k_block_keys_hist<9>, dim3(ntiles), dim3(RS_T), 0, st, s->g, pv, s->keys, dmask0,
                               s->sort_hist, ntiles, s->blk_start, blk_end, s->scan_tmp, nb);
        else
            hipLaunchKernelGGL(k_block_keys_hist<8>, dim3(ntiles), dim3(RS_T), 0, st, s->g, pv, s->keys, dmask0,
                               s->sort_hist, ntiles, s->blk_start, blk_end, s->scan_tmp, nb);
        // a deferred order goes straight to the G2P's permutation
        RadixResult r{};
        if ((rc = radix_sort_ex(st, s->keys, s->vals, s->keys_alt, s->vals_alt, n, bits, s->sort_hist, s->sort_scan,
                                true, true, defer ? s->g2p_perm + pv.off : nullptr, r)))
            return hip_fail((hipError_t)rc, "radix_sort");
        const uint32_t *perm = r.vals;
        const uint32_t *skeys = r.keys;
        if (!defer)
            hipLaunchKernelGGL(k_permute, dim3(round_xcd((uint32_t)nblk(n, 256))), dim3(256), 0, st, s->P, s->P_alt,
                               s->cap, pv.off, n, perm, s->pid, s->pid_alt, plane_mask);
        // tiles: one per non-empty block, numbered in block order
        hipLaunchKernelGGL(k_block_runs, dim3(nblk(n, 256)), dim3(256), 0, st, skeys, n, s->blk_start, blk_end,
                           s->scan_tmp);
        if ((rc = scan_excl(st, s->scan_tmp, s->scan_tmp, nb, s->sort_scan, s->tile_counts + k)))
            return hip_fail((hipError_t)rc, "scan");
        hipLaunchKernelGGL(k_emit_tiles, dim3(nblk(nb, 256)), dim3(256), 0, st, s->blk_start, blk_end, s->scan_tmp,
                           nb, (uint4 *)s->tile_desc + s->group_tile_off[k], (uint32_t)tile_bound(s, n),
                           s->blk_tile + (int64_t)k * nb, (uint32_t)s->group_tile_off[k]);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "sort kernels");
    }
    if (defer) {
        s->g2p_move_pending = true;
    } else {
        std::swap(s->P, s->P_alt);
        std::swap(s->pid, s->pid_alt);
    }
    s->sorted_valid = true;
    s->steps_since_sort = 0;
    s->sorted_at_launch = s->tile_launches;
    s->sorts += 1;
    return MPM_OK;
}

// Re-sort when the interval is up, or earlier once a tile P2G listed more
// than 0.01% of the particles (at least 64) for the atomic fallback:
// particles drift out of their block arenas at a rate set by the flow, and
// every spilled particle costs 108 global atomics.  The count read is that of
// the tile launch two back; waiting for its copy never idles the device
// (a later step is still queued) and bounds how far the host runs ahead.
// The interval adapts: a re-sort forced by spills after k steps sets the
// interval, and the longest one known to stay clear of spills, to k - 3 (the
// count arrives late, and in a steady flow whole layers of particles leave
// their arenas together, so the burst would come back at the same age); the
// interval then stays at that bound, which every fourth interval-called
// re-sort raises by a step (up to sort_every), so a flow that slows down is
// followed and a burst is met about once per dozen re-sorts instead of once
// per three or four.
// `ahead`: steps that will have completed when the new order is first used
// (1 for a re-sort the coming G2P applies).
static bool sort_due(mpm_sim *s, int ahead = 0)
{
    if (!s->sorted_valid) return true;
    const int64_t cap = s->cfg.sort_every;
    if (s->sort_safe <= 0 || s->sort_safe > cap) s->sort_safe = cap;
    if (s->sort_interval <= 0 || s->sort_interval > s->sort_safe) s->sort_interval = s->sort_safe;
    if (s->steps_since_sort + ahead >= s->sort_interval) {
        if (++s->sort_clean % 4 == 0) s->sort_safe = std::min<int64_t>(cap, s->sort_safe + 1);
        s->sort_interval = std::min<int64_t>(s->sort_safe, s->sort_interval + 1);
        return true;
    }
    const int64_t want = s->tile_launches - mpm_sim::SPILL_SLOTS;
    if (want < s->sorted_at_launch || !s->spill_host) return false;    // counts of the current order only
    const int slot = (int)(want % mpm_sim::SPILL_SLOTS);
    if (s->spill_launch[slot] != want) return false;
    // wait until every group's spill kernel of that launch has published its
    // count (launch number written after the count with release order)
    const uint32_t *seq = s->spill_host + mpm_sim::SPILL_SLOTS * MAX_GROUPS + slot * MAX_GROUPS;
    const auto t0 = std::chrono::steady_clock::now();
    for (size_t k = 0; k < s->groups.size(); ++k) {
        while (__atomic_load_n(&seq[k], __ATOMIC_ACQUIRE) != (uint32_t)want) {
            if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(2)) return false;
            std::this_thread::yield();
        }
    }
    s->spill_launch[slot] = -1;
    uint64_t tot = 0;
    for (size_t k = 0; k < s->groups.size(); ++k) tot += s->spill_host[slot * MAX_GROUPS + k];
    if (tot <= std::max<uint64_t>(64, (uint64_t)s->total / 10000)) return false;
    s->sort_interval = s->sort_safe = std::max<int64_t>(1, s->steps_since_sort - 3);
    s->sort_clean = 0;
    return true;
}

// Pinned host word that receives the spill count of group `group`'s tile
// launch (written by the spill kernel itself: no copy in the stream).
static int spill_host_word(mpm_sim *s, int group, uint32_t **word)
{
    if (!s->spill_host) {
        constexpr size_t words = 2 * mpm_sim::SPILL_SLOTS * MAX_GROUPS;
        MPM_HIP(hipHostMalloc((void **)&s->spill_host, words * sizeof(uint32_t), hipHostMallocCoherent));
        std::memset(s->spill_host, 0, words / 2 * sizeof(uint32_t));
        std::memset(s->spill_host + words / 2, 0xff, words / 2 * sizeof(uint32_t));   // no launch yet
    }
    const int slot = (int)(s->tile_launches % mpm_sim::SPILL_SLOTS);
    *word = s->spill_host + slot * MAX_GROUPS + group;
    return MPM_OK;
}

// After group `group`'s tile launch: the launch counter advances after the
// last group.
static int note_spills(mpm_sim *s, int group)
{
    const int slot = (int)(s->tile_launches % mpm_sim::SPILL_SLOTS);
    if (group + 1 == (int)s->groups.size()) {
        s->spill_launch[slot] = s->tile_launches;
        s->tile_launches += 1;
    }
    return MPM_OK;
}

int sort_if_due(mpm_sim *s)
{
    if (sort_due(s)) return sort_all(s, false);
    return MPM_OK;
}

int fast_prepare(mpm_sim *s)
{
    // a re-sort whose G2P did not complete: apply its order before the tiles are used
    int rc = finish_g2p_move(s, false);
    if (rc) return rc;
    if (sort_due(s)) {
        rc = sort_all(s, false);
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

// Sorted at least once and the stored stress current; no re-sort schedule.
int fast_ready(mpm_sim *s)
{
    int rc = MPM_OK;
    if (!s->sorted_valid && (rc = sort_all(s, false))) return rc;
    if (!s->tau_valid) {
        for (size_t k = 0; k < s->groups.size(); ++k) {
            rc = launch_compute_tau(s->stream, s->groups[k].mat, view_of(s, (int)k), s->cfg.svd_sweeps, s->counters);
            if (rc) return hip_fail((hipError_t)rc, "compute_tau");
        }
        s->tau_valid = true;
    }
    return MPM_OK;
}

// Launchers of the FAST P2G kernels by material and tile shape.
using BlockLaunch = void (*)(dim3, dim3, hipStream_t, const GridP &, const MatP &, const PartView &, const uint4 *,
                             const uint32_t *, uint32_t, uint32_t, float, float4 *, uint32_t *, uint32_t *, uint32_t *,
                             int32_t *);
using BatchLaunch = void (*)(dim3, dim3, hipStream_t, const BatchSlice *, int);
template <int MAT, int CH>
static void launch_block(dim3 gr, dim3 bl, hipStream_t st, const GridP &g, const MatP &m, const PartView &pv,
                         const uint4 *desc, const uint32_t *tc, uint32_t cap, uint32_t scap, float dt,
                         float4 *staging, uint32_t *spill, uint32_t *flags, uint32_t *counters, int32_t *blk_tile)
{
    hipLaunchKernelGGL((k_p2g_block<MAT, CH>), gr, bl, 0, st, g, m, pv, desc, tc, cap, scap, dt, staging, spill,
                       flags, counters, blk_tile, 1);
}
template <int MAT, int CH>
static void launch_batch(dim3 gr, dim3 bl, hipStream_t st, const BatchSlice *tab, int n)
{
    hipLaunchKernelGGL((k_p2g_batch<MAT, CH>), gr, bl, 0, st, tab, n);
}
static int tile_index(int tile) { return tile == 256 ? 0 : (tile == 512 ? 2 : 1); }
static int mat_index(int mat) { return mat == MPM_JELLO ? 0 : (mat == MPM_SNOW ? 1 : 2); }
static BlockLaunch block_launcher(int mat, int tile)
{
    static const BlockLaunch t[3][3] = {
        {launch_block<MPM_JELLO, 256>, launch_block<MPM_JELLO, 384>, launch_block<MPM_JELLO, 512>},
        {launch_block<MPM_SNOW, 256>, launch_block<MPM_SNOW, 384>, launch_block<MPM_SNOW, 512>},
        {launch_block<MPM_WATER, 256>, launch_block<MPM_WATER, 384>, launch_block<MPM_WATER, 512>}};
    return t[mat_index(mat)][tile_index(tile)];
}
static BatchLaunch batch_launcher(int mat, int tile)
{
    static const BatchLaunch t[3][3] = {
        {launch_batch<MPM_JELLO, 256>, launch_batch<MPM_JELLO, 384>, launch_batch<MPM_JELLO, 512>},
        {launch_batch<MPM_SNOW, 256>, launch_batch<MPM_SNOW, 384>, launch_batch<MPM_SNOW, 512>},
        {launch_batch<MPM_WATER, 256>, launch_batch<MPM_WATER, 384>, launch_batch<MPM_WATER, 512>}};
    return t[mat_index(mat)][tile_index(tile)];
}

int tile_shape(int tile) { return tile == 0 ? TILE_DEFAULT : tile; }

bool tile_shape_valid(int tile)
{
    if (tile == 0) return true;
    for (int t : TILE_SHAPES)
        if (t == tile) return true;
    return false;
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
    int rc = MPM_OK;
    if (group > 0) {        // the step's counter reset cleared group 0's count
        rc = launch_zero_u32(s->stream, s->counters + CT_SPILL, 1);
        if (rc) return hip_fail((hipError_t)rc, "zero");
    }
    uint32_t *hw = nullptr;
    if ((rc = spill_host_word(s, group, &hw))) return rc;
    dim3 gr(wg), bl(P2G_THREADS);
    {
    PhaseTimer tk(s, TK_P2G);
    block_launcher(m.type, tile_shape(s->cfg.tile))(gr, bl, s->stream, s->g, m, pv, desc, tc, cap, scap, dt,
                                                    s->staging, s->spill, s->flags, s->counters,
                                                    s->blk_tile + (int64_t)group * s->nblocks);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "p2g_block");
    if (s->groups.size() == 1) {
        // one group: the listed particles are scattered by the block
        // compaction of the grid phase (the tiles flagged their blocks)
        s->spill_pending = true;
        s->spill_dt = dt;
        s->spill_g = s->g;          // a later scatter takes the dt this launch took
        s->spill_hw = hw;
        s->spill_seq = (uint32_t)s->tile_launches;
        return note_spills(s, group);
    }
    rc = launch_p2g_spill(s->stream, s->g, m, pv, dt, s->spill, (uint32_t)pv.n, s->counters, s->grid_mp, s->flags,
                          hw, hw + mpm_sim::SPILL_SLOTS * MAX_GROUPS, (uint32_t)s->tile_launches);
    if (rc) return hip_fail((hipError_t)rc, "p2g_spill");
    return note_spills(s, group);
}

int spill_compact(mpm_sim *s)
{
    if (!s->spill_pending) return hip_fail(hipErrorInvalidValue, "spill_compact");
    s->spill_pending = false;
    const PartView pv = view_of(s, 0);
    const int rc = launch_spill_compact(s->stream, s->spill_g, s->groups[0].mat, pv, s->spill_dt, s->spill,
                                        (uint32_t)pv.n, s->counters, s->grid_mp, s->flags, (int)s->nblocks,
                                        s->list, s->spill_hw, s->spill_hw + mpm_sim::SPILL_SLOTS * MAX_GROUPS,
                                        s->spill_seq);
    return rc ? hip_fail((hipError_t)rc, "spill_compact") : MPM_OK;
}

int flush_spill(mpm_sim *s)
{
    if (!s->spill_pending) return MPM_OK;
    s->spill_pending = false;
    const PartView pv = view_of(s, 0);
    const int rc = launch_p2g_spill(s->stream, s->spill_g, s->groups[0].mat, pv, s->spill_dt, s->spill, (uint32_t)pv.n,
                                    s->counters, s->grid_mp, s->flags, s->spill_hw,
                                    s->spill_hw + mpm_sim::SPILL_SLOTS * MAX_GROUPS, s->spill_seq);
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
    int32_t *bt = s->blk_tile + (int64_t)group * s->nblocks;
    int rc = launch_zero_u32(s->stream, s->counters + CT_SPILL, 1);
    if (rc) return hip_fail((hipError_t)rc, "zero");
    uint32_t *hw = nullptr;
    if ((rc = spill_host_word(s, group, &hw))) return rc;
    dim3 gr(wg), bl(TILE_THREADS);
    {
    PhaseTimer tk(s, TK_G2P);
    switch (m.type) {
    case MPM_JELLO: hipLaunchKernelGGL(k_g2p2g_block<MPM_JELLO>, gr, bl, 0, s->stream, s->g, m, pv, desc, tc, cap, scap, dt, dt_next, s->grid_v, s->staging, s->spill, s->flags, s->counters, sw, bt); break;
    case MPM_SNOW: hipLaunchKernelGGL(k_g2p2g_block<MPM_SNOW>, gr, bl, 0, s->stream, s->g, m, pv, desc, tc, cap, scap, dt, dt_next, s->grid_v, s->staging, s->spill, s->flags, s->counters, sw, bt); break;
    default: hipLaunchKernelGGL(k_g2p2g_block<MPM_WATER>, gr, bl, 0, s->stream, s->g, m, pv, desc, tc, cap, scap, dt, dt_next, s->grid_v, s->staging, s->spill, s->flags, s->counters, sw, bt); break;
    }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "g2p2g_block");
    rc = launch_p2g_spill(s->stream, s->g, m, pv, dt_next, s->spill, (uint32_t)pv.n, s->counters, s->grid_mp,
                          s->flags, hw, hw + mpm_sim::SPILL_SLOTS * MAX_GROUPS, (uint32_t)s->tile_launches);
    if (rc) return hip_fail((hipError_t)rc, "p2g_spill");
    return note_spills(s, group);
}

// Stand-alone sort for the probe entry point (own temporaries).
int probe_sort(const uint32_t *keys_h, int64_t n, int bits, uint32_t *sorted_h, uint32_t *perm_h)
{
    if (n <= 0) return MPM_OK;
    if (n > (1ll << 30)) return set_error(MPM_ERR_INVALID, "too many keys");
    if (bits < 1 || bits > 32) return set_error(MPM_ERR_INVALID, "bits out of range [1,32]");
    int64_t ntiles = (n + RS_TILE - 1) / RS_TILE;
    int64_t hist_len = 512 * ntiles;         // digits of up to 9 bits
    int64_t scan_words = RS_MAX_DIGITS;     // digit totals
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

// ---------------------------------------------------------------------------
// Sort that allocates nothing (mpm_sort_keys_noalloc): the ping-pong key and
// value buffers, the digit histograms and the digit totals of the radix
// sort live in one fixed device array, `scratch`, sized for
// MPM_SORT_NOALLOC_MAX_KEYS keys (about 8.9 MB, held by every GPU the
// library runs on until the process ends).  Host code never passes the
// array's name to a kernel: each device's address of it is looked up once
// (hipGetSymbolAddress with that device current) and kept per device.
// Nothing in it needs to start at zero -- the histogram kernel writes every
// entry it later reads, and so does the column-scan kernel -- but every call on one
// GPU shares it, so calls that overlap on one GPU (different streams or
// threads) race.
constexpr int64_t NA_KEYS = MPM_SORT_NOALLOC_MAX_KEYS;
constexpr int64_t NA_TILES = (NA_KEYS + RS_TILE - 1) / RS_TILE;
constexpr int64_t NA_HIST = 512 * NA_TILES;                              // digits of up to 9 bits
constexpr int64_t NA_TOTALS = RS_MAX_DIGITS;
constexpr int64_t NA_WORDS = 2 * NA_KEYS + NA_HIST + NA_TOTALS;
static __device__ __align__(256) uint32_t scratch[NA_WORDS];

constexpr int NA_MAX_DEVICES = 64;
static uint32_t *na_scratch_ptr[NA_MAX_DEVICES];
static std::once_flag na_once[NA_MAX_DEVICES];
static hipError_t na_lookup_rc[NA_MAX_DEVICES];

// This device's address of `scratch` (looked up on the first call on it).
static int noalloc_scratch(uint32_t **ptr)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= NA_MAX_DEVICES) return set_error(MPM_ERR_INVALID, "device index beyond the scratch table");
    std::call_once(na_once[dev], [dev] {
        void *p = nullptr;
        na_lookup_rc[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        na_scratch_ptr[dev] = na_lookup_rc[dev] == hipSuccess ? (uint32_t *)p : nullptr;
    });
    if (na_lookup_rc[dev] != hipSuccess) return hip_fail(na_lookup_rc[dev], "hipGetSymbolAddress(scratch)");
    *ptr = na_scratch_ptr[dev];
    return MPM_OK;
}

__global__ void k_iota(uint32_t *v, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (uint32_t)i;
}

int sort_keys_noalloc(const uint32_t *keys, int64_t n, int bits, uint32_t *sorted, uint32_t *perm, hipStream_t st)
{
    if (n < 0) return set_error(MPM_ERR_INVALID, "negative key count");
    if (n > NA_KEYS) return set_error(MPM_ERR_RANGE, "more keys than the no-allocation sort holds (MPM_SORT_NOALLOC_MAX_KEYS)");
    if (bits < 1 || bits > 32) return set_error(MPM_ERR_INVALID, "bits out of range [1,32]");
    if (n == 0) return MPM_OK;
    if (!keys || !sorted || !perm) return set_error(MPM_ERR_INVALID, "null argument");
    uint32_t *base = nullptr;
    int rc = noalloc_scratch(&base);
    if (rc) return rc;
    uint32_t *ka = base, *va = base + NA_KEYS, *hist = base + 2 * NA_KEYS, *totals = hist + NA_HIST;
    hipError_t e = hipSuccess;
    if (keys != sorted) e = hipMemcpyAsync(sorted, keys, (size_t)n * 4, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return hip_fail(e, "copy keys");
    hipLaunchKernelGGL(k_iota, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, perm, (int)n);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "iota");
    bool in_alt = false;
    rc = radix_sort(st, sorted, perm, ka, va, (int)n, bits, hist, totals, in_alt);
    if (rc) return hip_fail((hipError_t)rc, "radix_sort");
    if (in_alt) {
        if ((e = hipMemcpyAsync(sorted, ka, (size_t)n * 4, hipMemcpyDeviceToDevice, st)) != hipSuccess)
            return hip_fail(e, "copy sorted keys");
        if ((e = hipMemcpyAsync(perm, va, (size_t)n * 4, hipMemcpyDeviceToDevice, st)) != hipSuccess)
            return hip_fail(e, "copy permutation");
    }
    return MPM_OK;
}

// Batched re-sort of a one-group simulator in its current parity.
int fill_sort_slice(mpm_sim *s, SortSlice &t)
{
    if (s->groups.size() != 1) return set_error(MPM_ERR_INVALID, "a batched re-sort takes one particle group");
    int rc = ensure_fast_buffers(s);
    if (rc) return rc;
    if (!s->g2p_perm || s->g2p_perm_cap < s->cap) return set_error(MPM_ERR_INVALID, "no re-sort permutation buffer");
    const PartView pv = view_of(s, 0);
    memset(&t, 0, sizeof t);
    t.g = s->g;
    t.pv = pv;
    t.keys = s->keys;
    t.vals = s->vals;
    t.keys_alt = s->keys_alt;
    t.vals_alt = s->vals_alt;
    t.hist = s->sort_hist;
    t.scan = s->sort_scan;
    t.blk_start = s->blk_start;
    t.blk_end = s->blk_count;
    t.has = s->scan_tmp;
    t.perm = s->g2p_perm;
    t.tile_count = s->tile_counts;
    t.warm_cnt = s->warm_cnt;
    t.desc = (uint4 *)s->tile_desc;
    t.blk_tile = s->blk_tile;
    t.n = pv.n;
    t.ntiles = nblk(pv.n, RS_TILE);
    t.nblocks = (int)s->nblocks;
    t.desc_cap = (uint32_t)tile_bound(s, pv.n);
    return MPM_OK;
}

void sort_batch_plan(mpm_sim *const *sims, int ns, SortBatchPlan &pl)
{
    pl = SortBatchPlan{};
    for (int k = 0; k < ns; ++k) {
        const mpm_sim *s = sims[k];
        int bits = KEY_LOW + 1;
        while ((1ll << (bits - KEY_LOW)) < s->nblocks) ++bits;
        pl.passes = std::max(pl.passes, (bits + BS_DB - 1) / BS_DB);
        const int64_t n = s->groups[0].n;
        pl.tiles += nblk(n, RS_TILE);
        pl.runs += nblk(n, 256);
        pl.scan += nblk(s->nblocks, SCAN_TILE);
        pl.emit += nblk(s->nblocks, 256);
    }
    pl.spine = ns * ((1 << BS_DB) / 4);
}

int launch_sort_batch(hipStream_t st, const SortSlice *tab, int ns, const SortBatchPlan &pl)
{
    if (ns <= 0 || pl.tiles <= 0) return MPM_OK;
    hipLaunchKernelGGL(k_bsort_keys, dim3(pl.tiles), dim3(RS_T), 0, st, tab, ns);
    for (int pass = 0; pass < pl.passes; ++pass) {
        if (pass) hipLaunchKernelGGL(k_bsort_hist, dim3(pl.tiles), dim3(RS_T), 0, st, tab, ns, pass);
        hipLaunchKernelGGL(k_bsort_spine, dim3(pl.spine), dim3(256), 0, st, tab, ns);
        hipLaunchKernelGGL(k_bsort_scatter, dim3(pl.tiles), dim3(RS_T), 0, st, tab, ns, pass,
                           pass == pl.passes - 1 ? 1 : 0);
    }
    hipLaunchKernelGGL(k_bsort_runs, dim3(pl.runs), dim3(256), 0, st, tab, ns, pl.passes & 1);
    hipLaunchKernelGGL(k_bsort_scan_reduce, dim3(pl.scan), dim3(SCAN_T), 0, st, tab, ns);
    hipLaunchKernelGGL(k_bsort_scan_apply, dim3(pl.scan), dim3(SCAN_T), 0, st, tab, ns);
    hipLaunchKernelGGL(k_bsort_emit, dim3(pl.emit), dim3(256), 0, st, tab, ns);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? MPM_OK : hip_fail(e, "sort_batch");
}

// Batch descriptor of a one-group simulator in its current particle order.
void fill_batch_slice(mpm_sim *s, BatchSlice &b)
{
    memset(&b, 0, sizeof b);
    b.g = s->g;
    b.mat = s->groups[0].mat;
    b.mat.wlist = nullptr;          // batched scenes infer every particle's quaternion
    b.mat.wcount = nullptr;
    b.pv = view_of(s, 0);
    b.dt = (float)s->cfg.dt;
    b.sweeps = s->cfg.svd_sweeps;
    b.warm_group = -1;
    b.grid_mp = s->grid_mp;
    b.grid_v = s->grid_v;
    b.staging = s->staging;
    b.flags = s->flags;
    b.list = s->list;
    b.counters = s->counters;
    b.spill = s->spill;
    b.desc = (const uint4 *)s->tile_desc + s->group_tile_off[0];
    b.tile_count = s->tile_counts;
    b.blk_tile = s->blk_tile;
    b.desc_cap = (uint32_t)tile_bound(s, s->groups[0].n);
    b.staging_cap = (uint32_t)s->staging_cap;
    b.nblocks = s->nblocks;
    b.perm = s->g2p_perm;
    b.copy_q = s->warm_cnt ? 1 : 0;
}

// Material `mat` of the scenes [0, nslices) of tab, in chunks of P2G_CHUNK,
// tiles of `tile` particles per pass.
int launch_p2g_batch(hipStream_t st, int mat, const BatchSlice *tab, const int *types, int nslices, int wgs,
                     int tile)
{
    dim3 gr(wgs), bl(P2G_THREADS);
    for (int c0 = 0; c0 < nslices; c0 += P2G_CHUNK) {
        const int n = std::min(P2G_CHUNK, nslices - c0);
        bool any = false;
        for (int k = c0; k < c0 + n; ++k) any = any || types[k] == mat;
        if (!any) continue;
        batch_launcher(mat, tile_shape(tile))(gr, bl, st, tab + c0, n);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return (int)e;
    }
    return 0;
}

// FAST-form G2P of one group; after a deferred re-sort it writes the group
// in the new order into the other buffer (finish_g2p_move swaps them).
int launch_g2p_fast(mpm_sim *s, int group, float dt)
{
    const PartView pv = view_of(s, group);
    G2PMove mv{s->g2p_perm + pv.off, pv, s->pid, s->pid_alt, s->warm_cnt ? 1 : 0};
    mv.po.P = s->P_alt;
    mv.po.pid = s->pid_alt;
    int rc = launch_g2p(s->stream, s->g, s->groups[group].mat, pv, dt, s->grid_v, 1, s->cfg.svd_sweeps, s->counters,
                        s->cfg.threads, s->g2p_move_pending ? &mv : nullptr);
    return rc ? hip_fail((hipError_t)rc, "g2p_fast") : MPM_OK;
}

// A due FAST-form re-sort, deferred to the G2P about to run (particles that
// a P2G of this step listed for the atomic fallback are scattered first:
// the list names slots of the current order).
int sort_for_g2p(mpm_sim *s)
{
    if (!s->sorted_valid || s->g2p_move_pending || !sort_due(s, 1)) return MPM_OK;
    int rc = flush_spill(s);
    return rc ? rc : sort_all(s, true);
}

// After every group's G2P of a deferred re-sort: the written buffer becomes
// the particles.  With `moved` false (the G2P did not run) the deferred
// order is applied by the permutation kernel instead, so that the tiles and
// the particles agree again.
int finish_g2p_move(mpm_sim *s, bool moved)
{
    if (!s->g2p_move_pending) return MPM_OK;
    s->g2p_move_pending = false;
    if (!moved) {
        uint64_t plane_mask = (1ull << PL_COUNT) - 1ull;
        if (!s->tau_valid) plane_mask &= ~(((1ull << 6) - 1ull) << PL_TAU);
        if (!s->warm_cnt) plane_mask &= ~(((1ull << 4) - 1ull) << PL_Q);
        else {
            int rc = launch_zero_u32(s->stream, s->warm_cnt, 2 * MAX_GROUPS);
            if (rc) return hip_fail((hipError_t)rc, "zero");
        }
        for (size_t k = 0; k < s->groups.size(); ++k) {
            const PartView pv = view_of(s, (int)k);
            if (pv.n <= 0) continue;
            hipLaunchKernelGGL(k_permute, dim3(round_xcd((uint32_t)nblk(pv.n, 256))), dim3(256), 0, s->stream, s->P,
                               s->P_alt, s->cap, pv.off, pv.n, s->g2p_perm + pv.off, s->pid, s->pid_alt, plane_mask);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "permute");
    }
    std::swap(s->P, s->P_alt);
    std::swap(s->pid, s->pid_alt);
    return MPM_OK;
}

}  // namespace mpm

extern "C" int mpm_sort_keys_noalloc(const uint32_t *keys, int64_t n, int32_t bits, uint32_t *sorted, uint32_t *perm,
                                     void *stream)
{
    return mpm::sort_keys_noalloc(keys, n, bits, sorted, perm, (hipStream_t)stream);
}
