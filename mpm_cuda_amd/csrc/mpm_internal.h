// mpm_internal.h -- types shared by the HIP translation units of libmpm_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/mpm/mpm.h"
#include "mpm_math.h"

namespace mpm {

// Particle storage: one FP32 plane per scalar component (structure of
// arrays).  Plane p of particle i lives at P[p * cap + i].  Groups occupy
// contiguous index ranges [off, off + n) in every plane.
enum Plane : int {
    PL_X = 0,     // 3 planes: position
    PL_V = 3,     // 3: velocity
    PL_C = 6,     // 9: affine velocity C, row-major
    PL_F = 15,    // 9: (elastic) deformation gradient, row-major
    PL_S = 24,    // 1: snow J_p = det F_p, water J
    PL_TAU = 25,  // 6: Kirchhoff stress (xx,yy,zz,xy,xz,yz) for the next P2G
    PL_Q = 31,    // 4: warm-start quaternion of the SVD (written before each G2P)
    PL_COUNT = 35
};

// Grid nodes are stored block-major: 4x4x4-node blocks, block index
// (bz * nby + by) * nbx + bx, node slot ((z&3)*4 + (y&3))*4 + (x&3); one
// float4 per node.  grid_mp holds (mass, momentum) during P2G; grid_v holds
// (velocity, mass) after the grid update.
constexpr int BLK = 4;
constexpr int BLK_NODES = 64;
// FAST-form P2G arena of a block: nodes [4B-1, 4B+6] per axis (the block's
// own 4, one below and three above), staged in logical (z, y, x) order.
constexpr int ARENA = 8;
constexpr int ARENA_NODES = ARENA * ARENA * ARENA;

struct GridP {
    int nx, ny, nz;
    int nbx, nby, nbz;
    float h, inv_h;
    float ox, oy, oz;
    float gx, gy, gz;
    int pad, bc_mode;
    const float *dtp;   // step-size controller on: the step's dt is read here (device memory), else null
};

struct MatP {
    int type, water_c_mode;
    float mass, vol, mu, lam;
    float theta_c, theta_s, xi;
    float eos_k, eos_gamma;
    int warm_sweeps;    // > 0: G2P SVDs start from the stored quaternion (PL_Q) with this many sweeps
    float c2;           // squared elastic wave speed at rest: (lam + 2 mu) / rho, water k gamma / rho
    // Newton steps of the polar rotation on the snow elastic path (singular
    // values inside the clamp band): enough to bring both band edges to
    // within 1e-7 of 1
    int polar_steps = 3;
    // warm start: the G2P lists the (group-local) indices of the particles
    // that take the SVD into wlist, counted in *wcount; the next step's
    // inference runs over that list only.  Null: no list (every particle's
    // quaternion is inferred).
    uint32_t *wlist = nullptr;
    uint32_t *wcount = nullptr;
};

// Per-launch view of one group's particles.
struct PartView {
    float *P;
    uint32_t *pid;
    int64_t cap;
    int64_t off;
    int n;
};

// Device counters (one uint32 array).
enum Counter : int {
    CT_ACTIVE = 0,      // active grid blocks of the current step
    CT_NODES = 1,       // grid nodes with mass in the grid update
    CT_SPILL = 2,       // particles listed for the atomic P2G fallback
    CT_HEALTH = 3,      // MPM_H_* bits
    CT_FALLBACK = 4,    // FP64 SVD fallbacks
    CT_REJECT = 5,      // warm-start rejections
    CT_TILES = 6,       // P2G tiles (FAST / FUSED forms)
    CT_SPILLED = 7,     // spill count of the last block compaction (CT_SPILL is cleared there)
    CT_SPEED = 8,       // step-size controller: max_p |v_p| + c_p (float bits), since the last control
    CT_DEFORM = 9,      // step-size controller: max_p max(J, 1/J) (float bits)
    CT_SVD = 10,        // G2P particle updates that took the Jacobi SVD (cumulative)
    CT_WARM = 11,       // of those, started from a network quaternion (cumulative)
    CT_COUNT = 16
};

// Entry of the block -> staging slot map (blk_tile): -1 when the block has
// no tile; otherwise the slot in bits 0..23 (slots at or above the staging
// capacity are not staged; larger numbers saturate at 2^24 - 1) and in bits
// 24..29 the faces of the block's 8^3 staging arena that its last P2G wrote:
// bit 2a the low face (arena index 0) and bit 2a + 1 the high face (index 7)
// of axis a.  Arena indices 1..6 of every axis are always written.  A
// particle of the block whose base lies in the block writes indices 1..6
// only; the faces carry the particles that drifted one cell out of it since
// the last sort.
constexpr uint32_t TILE_SLOT_MASK = 0xFFFFFFu;
constexpr int TILE_FACE_SHIFT = 24;
constexpr uint32_t TILE_ALL_FACES = 0x3Fu;
__host__ __device__ inline int32_t tile_entry(uint32_t slot, uint32_t faces)
{
    return (int32_t)((slot < TILE_SLOT_MASK ? slot : TILE_SLOT_MASK) | (faces << TILE_FACE_SHIFT));
}

// Particle groups per simulator (per-group host words of the spill counts,
// warm-start list counts).
constexpr int MAX_GROUPS = MPM_MAX_GROUPS;

// Device state of the step-size controller (one per simulator).
struct DtCtl {
    float dt;           // dt of the current step, read by the step's kernels
    float pad;
    double time;        // simulated time after the last step
    double last;        // dt of the last step
};

// Memory accesses with their address space stated.  Pointers that reach
// device code through memory (batch tables, fetch structs) are generic to
// the compiler, which then emits FLAT instructions; on gfx950 a FLAT access
// also counts against the LDS wait counter, so every LDS wait in the
// surrounding code waits for it too.  Global buffers here come from
// hipMalloc, LDS ones are __shared__ arrays.  (16-byte accesses go through
// the native vector type: float4 / uint4 have no address-space-qualified
// copies.)
#define MPM_GLOBAL __attribute__((address_space(1)))
#define MPM_LDS __attribute__((address_space(3)))
typedef float mpm_f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int mpm_u32x4 __attribute__((ext_vector_type(4)));

template <class T>
__device__ __forceinline__ MPM_GLOBAL T *gp(T *p)
{
    return (MPM_GLOBAL T *)p;
}
template <class T>
__device__ __forceinline__ const MPM_GLOBAL T *gp(const T *p)
{
    return (const MPM_GLOBAL T *)p;
}
__device__ __forceinline__ float4 gload4(const float4 *p, int64_t i)
{
    const mpm_f32x4 v = ((const MPM_GLOBAL mpm_f32x4 *)p)[i];
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gstore4(float4 *p, int64_t i, float4 v)
{
    mpm_f32x4 w;
    w.x = v.x; w.y = v.y; w.z = v.z; w.w = v.w;
    ((MPM_GLOBAL mpm_f32x4 *)p)[i] = w;
}
__device__ __forceinline__ uint4 gload_u4(const uint4 *p, int64_t i)
{
    const mpm_u32x4 v = ((const MPM_GLOBAL mpm_u32x4 *)p)[i];
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float4 lload4(const float4 *p, int i)
{
    const mpm_f32x4 v = ((const MPM_LDS mpm_f32x4 *)p)[i];
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gatomic_add(float *p, float v) { (void)__builtin_amdgcn_global_atomic_fadd_f32(gp(p), v); }
__device__ __forceinline__ uint32_t gatomic_add(uint32_t *p, uint32_t v)
{
    return __hip_atomic_fetch_add(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void gatomic_or(uint32_t *p, uint32_t v)
{
    (void)__hip_atomic_fetch_or(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The step's dt: the controller's value in device memory when it is on.
__device__ __forceinline__ float step_dt(const GridP &g, float dt) { return g.dtp ? *gp(g.dtp) : dt; }

// Largest v over the wave (all 64 lanes take part; NaN loses), merged into
// *dst with an atomic max.  Non-negative floats order like their bit
// patterns as unsigned integers.  The atomic is skipped when *dst already
// holds at least v (a plain read first: once a large value is posted, most
// waves of a launch leave the single contended word alone).
// The active lanes of the wave (divergent code included) add their count of
// `pred` to *dst with one atomic.
__device__ __forceinline__ void wave_count(uint32_t *dst, bool pred)
{
    const uint64_t m = __ballot(pred);
    const int first = __ffsll((unsigned long long)__ballot(1)) - 1;
    if (m && (int)(threadIdx.x & 63) == first) (void)gatomic_add(dst, (uint32_t)__popcll(m));
}

// The active lanes of the wave append their `v` to list[*count ...) (one
// atomic per wave; the order within a wave follows the lanes).
__device__ __forceinline__ void wave_append(uint32_t *list, uint32_t *count, uint32_t v)
{
    const uint64_t m = __ballot(1);
    const int lane = (int)(threadIdx.x & 63);
    const int first = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == first) base = gatomic_add(count, (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, first);
    const uint64_t below = lane ? (m & ((~0ull) >> (64 - lane))) : 0ull;
    gp(list)[base + (uint32_t)__popcll(below)] = v;
}

__device__ __forceinline__ void wave_max_to(uint32_t *dst, float v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmaxf(v, __shfl_xor(v, d));
    if ((threadIdx.x & 63) == 0 && v > 0.f) {
        const uint32_t bits = __float_as_uint(v);
        if (bits > __hip_atomic_load(gp(dst), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            (void)__hip_atomic_fetch_max(gp(dst), bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Workgroups are dispatched to the 8 XCDs of the device round-robin
// (workgroup id k runs on XCD k % 8), and each XCD has its own L2.  Work item
// xcd_spread(k, n8) gives every XCD a contiguous eighth of [0, n8), so
// neighbouring blocks / particle ranges (which share grid nodes) meet in the
// same L2.  n8 = round_xcd(n); the map is a bijection of [0, n8).
constexpr uint32_t NXCD = 8;
__host__ __device__ inline uint32_t round_xcd(uint32_t n) { return (n + NXCD - 1) / NXCD * NXCD; }
__host__ __device__ inline uint32_t xcd_spread(uint32_t k, uint32_t n8) { return (k % NXCD) * (n8 / NXCD) + k / NXCD; }

__host__ __device__ inline int64_t node_slot(const GridP &g, int x, int y, int z)
{
    int64_t b = ((int64_t)(z >> 2) * g.nby + (y >> 2)) * g.nbx + (x >> 2);
    return b * BLK_NODES + (((z & 3) * BLK + (y & 3)) * BLK + (x & 3));
}

struct GroupInfo {
    mpm_material src;
    MatP mat;
    int64_t off = 0, n = 0;
    double mass = 0, vol = 0;
};

// One scene of a batch (a simulator with one particle group) as the batched
// kernels read it from device memory: everything a single-scene launch
// passes as arguments.
constexpr int BATCH_MAX = 256;
struct BatchSlice {
    GridP g;
    MatP mat;
    PartView pv;
    float dt;
    int sweeps;
    int warm_group;                 // warm-start weight set of the scene, -1: cold start
    int copy_q;                     // a re-sort moves the warm-start quaternions too
    const uint32_t *perm;           // the scene's re-sort order: slot j takes particle perm[j]
    uint32_t *warm_cnt;             // warm start: [0] listed particles the coming inference takes,
                                    // [1] entries the G2P appends (mat.wlist); null when cold
    float4 *grid_mp, *grid_v, *staging;
    uint32_t *flags, *list, *counters, *spill;
    const uint4 *desc;              // P2G tiles of the last sort
    const uint32_t *tile_count;
    int32_t *blk_tile;              // tile_entry per block (the P2G updates its faces)
    uint32_t desc_cap, staging_cap;
    int64_t nblocks;
};

// One scene of a batched re-sort (device table): its particles in the
// parity being sorted, its sort buffers, tile outputs and the permutation
// the batched move reads.  Keys are sorted with 8-bit digits in the batch's
// common number of passes (extra passes of a scene with fewer key bits see
// zero digits and keep the order).
struct SortSlice {
    GridP g;
    PartView pv;
    uint32_t *keys, *vals, *keys_alt, *vals_alt;
    uint32_t *hist;                 // 256 x ntiles digit counts
    uint32_t *scan;                 // radix digit totals, then the tile scan's partial sums
    uint32_t *blk_start, *blk_end, *has;
    uint32_t *perm;                 // slot j of the new order takes particle perm[j]
    uint32_t *tile_count;
    uint32_t *warm_cnt;             // warm-start list counts (cleared: the lists name old slots), or null
    uint4 *desc;
    int32_t *blk_tile;
    int n, ntiles, nblocks;
    uint32_t desc_cap;
};

// Inclusive prefix sums of per-scene work counts, held in one wave's
// registers: lane l keeps scene l + 64 j in slot j (up to BATCH_MAX scenes).
// Each wave that needs the map builds its own copy; find() turns a global
// work index into its scene with four ballots.  All 64 lanes must be active.
struct SliceScan {
    uint32_t beg[4], end[4];
    uint32_t total;
    template <class Count>
    __device__ __forceinline__ void build(int nslices, Count count)
    {
        const int lane = threadIdx.x & 63;
        uint32_t carry = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int s = j * 64 + lane;
            if (j * 64 >= nslices) {
                beg[j] = carry;
                end[j] = 0xFFFFFFFFu;
                continue;
            }
            const uint32_t c = s < nslices ? count(s) : 0u;
            uint32_t v = c;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t t = __shfl_up(v, o);
                if (lane >= o) v += t;
            }
            v += carry;
            beg[j] = v - c;
            end[j] = s < nslices ? v : 0xFFFFFFFFu;
            carry = __shfl(v, 63);
        }
        total = carry;
    }
    // scene of work index t (t < total)
    __device__ __forceinline__ uint32_t find(uint32_t t) const
    {
        uint32_t s = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) s += (uint32_t)__popcll(__ballot(end[j] <= t));
        return s;
    }
    // first work index of scene sid
    __device__ __forceinline__ uint32_t start(uint32_t sid) const
    {
        const uint32_t j = sid >> 6;
        const uint32_t v = j == 0 ? beg[0] : (j == 1 ? beg[1] : (j == 2 ? beg[2] : beg[3]));
        return __shfl(v, (int)(sid & 63u));
    }
};

}  // namespace mpm

struct mpm_sim {
    mpm_config cfg;
    mpm::GridP g;
    hipStream_t stream = nullptr;
    int64_t nblocks = 0;
    float4 *grid_mp = nullptr, *grid_v = nullptr;
    uint32_t *flags = nullptr, *list = nullptr, *counters = nullptr;
    // particles
    float *P = nullptr, *P_alt = nullptr;
    uint32_t *pid = nullptr, *pid_alt = nullptr;
    int64_t cap = 0, total = 0;
    std::vector<mpm::GroupInfo> groups;
    bool tau_valid = false;
    // FAST form: sort and tile buffers
    uint32_t *keys = nullptr, *keys_alt = nullptr, *vals = nullptr, *vals_alt = nullptr;
    uint32_t *sort_hist = nullptr, *sort_scan = nullptr;
    int64_t sort_hist_len = 0;
    uint32_t *blk_count = nullptr, *blk_start = nullptr, *tile_desc = nullptr, *scan_tmp = nullptr;
    int64_t tile_cap = 0;
    std::vector<int64_t> group_tile_off;    // per group: first tile slot
    uint32_t *tile_counts = nullptr;        // per group: tiles built by the last sort
    float4 *staging = nullptr;              // per tile: 8^3-node P2G arena written back without atomics
    int64_t staging_cap = 0;                // tiles the staging buffer holds
    int32_t *blk_tile = nullptr;            // [group][block]: tile_entry of the block (-1: no tile)
    uint32_t *spill = nullptr;              // indices of particles for the atomic P2G fallback
    bool primed = false;                    // FUSED: staging holds the scatter of the current state
    float primed_dt = 0.f;                  // FUSED: dt the primed scatter used for the force term
    bool sorted_valid = false;
    int64_t steps_since_sort = 0;
    // FAST form re-sort applied by the G2P after it: the sort leaves the
    // particle order of every group here (group-local sources at the group's
    // offset) and the G2P writes the updated particles in that order into the
    // other buffer
    uint32_t *g2p_perm = nullptr;
    int64_t g2p_perm_cap = 0;
    bool g2p_move_pending = false;
    // spill counts of the tile P2G launches, written to pinned host memory
    // by the spill kernel with the launch number after them; a re-sort is
    // brought forward when they grow.  Two slots: the count of the launch two
    // steps back is read (usually written by then; waiting for its number
    // keeps at most two steps queued ahead of the device, and no event
    // packet sits in the stream).
    static constexpr int SPILL_SLOTS = 2;
    uint32_t *spill_host = nullptr;         // [slot][MAX_GROUPS] counts, then [slot][MAX_GROUPS] launch numbers
    int64_t spill_launch[SPILL_SLOTS] = {-1, -1};   // tile launch number held by the slot
    int64_t tile_launches = 0;
    bool counters_clean = false;            // CT_SPILL cleared by the last grid phase (FAST form)
    // one-group tile P2G: its spill scatter waits for the grid phase
    bool spill_pending = false;
    float spill_dt = 0.f;
    mpm::GridP spill_g{};                      // the grid arguments of that tile launch (its dt source)
    uint32_t *spill_hw = nullptr;
    uint32_t spill_seq = 0;
    int64_t sorted_at_launch = 0;           // tile_launches when the last sort ran
    int64_t sort_interval = 0;              // steps between re-sorts (learned from the spill counts)
    int64_t sort_safe = 0;                  // longest interval since the last spill-forced re-sort (0: sort_every)
    int64_t sort_clean = 0;                 // re-sorts in a row that the interval, not spills, called for
    int last_phase = -1;
    // warm-start MLP weights (device), a host copy and a version that
    // changes with every mpm_set_warm_start_weights
    float *mlp_w = nullptr;
    int64_t mlp_count = 0;
    // warm start by list: warm_list holds, per group at the group's offset,
    // the particles whose last G2P took the SVD; warm_cnt[2 k + p] counts
    // group k's entries.  The inference of a step reads the entries counted
    // in parity warm_par; the G2P after it (stream order) rewrites the list
    // counting in parity warm_par ^ 1.
    uint32_t *warm_list = nullptr;
    uint32_t *warm_cnt = nullptr;
    int64_t warm_list_cap = 0;
    int warm_par = 0;
    std::vector<float> mlp_host;
    int64_t mlp_version = 0;
    // step-size controller (cfg.dt_mode != MPM_DT_FIXED)
    mpm::DtCtl *dtc = nullptr;
    bool dt_valid = false;                  // counters hold the maxima of the current state
    double last_dt = 0;                     // fixed mode: dt of the last step
    // phase timing (HIP events on `stream`): ids 0-3 phases, 4-7 kernels
    static constexpr int TIMED_IDS = 8;
    double tacc_ms[TIMED_IDS] = {};
    int64_t tacc_n[TIMED_IDS] = {};
    bool timing = false;
    struct Timed { int phase; hipEvent_t a, b; };
    std::vector<Timed> timed;
    std::vector<hipEvent_t> event_pool;
    // host bookkeeping
    int64_t steps = 0;
    int64_t sorts = 0;
    double time = 0;
    uint32_t health_host = 0;
};

namespace mpm {

// Timed ids: phases (sort, P2G, grid, G2P) and the main kernel of each.
enum { TK_P2G = 4, TK_GRID = 5, TK_G2P = 6 };

hipEvent_t take_event(mpm_sim *s);

// Records a HIP event pair around its scope on the simulator's stream when
// timing is on.
struct PhaseTimer {
    mpm_sim *s;
    int id;
    hipEvent_t a = nullptr;
    PhaseTimer(mpm_sim *s_, int id_) : s(s_), id(id_)
    {
        if (s->timing && (a = take_event(s))) (void)hipEventRecord(a, s->stream);
    }
    ~PhaseTimer()
    {
        if (!a) return;
        hipEvent_t b = take_event(s);
        if (b && hipEventRecord(b, s->stream) == hipSuccess) s->timed.push_back({id, a, b});
        else { s->event_pool.push_back(a); if (b) s->event_pool.push_back(b); }
    }
};


int set_error(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);
}  // namespace mpm

#define MPM_HIP(call)                                                     \
    do {                                                                  \
        hipError_t e_ = (call);                                           \
        if (e_ != hipSuccess) return mpm::hip_fail(e_, #call);            \
    } while (0)
