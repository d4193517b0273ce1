// mpm_launch.h -- host-side launchers of the step kernels.
#pragma once

#include "mpm_internal.h"

namespace mpm {

int launch_zero_u32(hipStream_t st, uint32_t *p, int n);
int launch_p2g_atomic(hipStream_t st, const GridP &g, const MatP &m, const PartView &pv, float dt,
                      int sweeps, int tau_stored, float4 *grid, uint32_t *flags, uint32_t *counters,
                      int threads);
int launch_compact(hipStream_t st, uint32_t *flags, int nblocks, uint32_t *list, uint32_t *counters);
int launch_count_nodes(hipStream_t st, const uint32_t *list, uint32_t *counters, const float4 *gv);
int launch_grid_update(hipStream_t st, const GridP &g, float dt, const uint32_t *list,
                       uint32_t *counters, float4 *gmp, float4 *gv, int nblocks, const float4 *staging,
                       const int32_t *blk_tile, int ngroups, uint32_t staging_cap, int clear_spill = 0);
// A G2P that applies a re-sort: slot j of po receives particle perm[j] of
// the view, its id (pid -> pid_out) and, with copy_q, its quaternion.
struct G2PMove {
    const uint32_t *perm;
    PartView po;
    const uint32_t *pid;
    uint32_t *pid_out;
    int copy_q;
};
int launch_g2p(hipStream_t st, const GridP &g, const MatP &m, const PartView &pv, float dt,
               const float4 *gv, int write_tau, int sweeps, uint32_t *counters, int threads,
               const G2PMove *mv = nullptr);
int launch_compute_tau(hipStream_t st, const MatP &m, const PartView &pv, int sweeps, uint32_t *counters);
int launch_p2g_spill(hipStream_t st, const GridP &g, const MatP &m, const PartView &pv, float dt,
                     const uint32_t *list, uint32_t list_cap, uint32_t *counters, float4 *grid, uint32_t *flags,
                     uint32_t *host_count, uint32_t *host_seq = nullptr, uint32_t seq = 0,
                     int clear_active = 0);
int launch_spill_compact(hipStream_t st, const GridP &g, const MatP &m, const PartView &pv, float dt,
                         const uint32_t *spill, uint32_t spill_cap, uint32_t *counters, float4 *grid,
                         uint32_t *flags, int nblocks, uint32_t *list, uint32_t *host_count, uint32_t *host_seq,
                         uint32_t seq);
int launch_dt_scan(hipStream_t st, const MatP &m, const PartView &pv, uint32_t *counters);
int launch_dt_control(hipStream_t st, DtCtl *c, uint32_t *counters, float max_dt, float cfl_h, int mode);

// FAST form (mpm_fast.hip)
int fast_prepare(mpm_sim *s);                 // (re)sort particles and build P2G tiles when due
int launch_p2g_tiles(mpm_sim *s, int group, float dt);
int spill_compact(mpm_sim *s);                // grid phase after a one-group tile P2G: spill scatter + compaction
int flush_spill(mpm_sim *s);                  // spill scatter of a tile P2G whose grid phase did not run
int launch_g2p_fast(mpm_sim *s, int group, float dt);
int launch_g2p2g(mpm_sim *s, int group, float dt, float dt_next);   // FUSED form
int sort_all(mpm_sim *s, bool defer = false);  // sort every group now and rebuild the tiles
int sort_for_g2p(mpm_sim *s);                 // a due re-sort, applied by the coming FAST G2P
int finish_g2p_move(mpm_sim *s, bool moved);  // after that G2P (or instead of it): the particles in the new order
int fast_ready(mpm_sim *s);                   // sorted once, stored stress current
PartView view_of(mpm_sim *s, int g);

// Batched stages over a BatchSlice table in device memory (mpm_step.hip,
// mpm_fast.hip); host side in mpm_batch.hip.
void fill_batch_slice(mpm_sim *s, BatchSlice &b);
// Batched re-sort (mpm_fast.hip): launch sizes of a set of scenes, one
// scene's table entry in its current parity, and the launches.
struct SortBatchPlan {
    int passes = 0, tiles = 0, spine = 0, runs = 0, scan = 0, emit = 0;
};
void sort_batch_plan(mpm_sim *const *sims, int ns, SortBatchPlan &pl);
int fill_sort_slice(mpm_sim *s, SortSlice &t);
int launch_sort_batch(hipStream_t st, const SortSlice *tab, int ns, const SortBatchPlan &pl);
int launch_batch_zero(hipStream_t st, const BatchSlice *tab, int nslices);
// Warm-start particle lists of a simulator (mpm_api.hip).
int ensure_warm_lists(mpm_sim *s);
int launch_p2g_batch(hipStream_t st, int mat, const BatchSlice *tab, const int *types, int nslices, int wgs,
                     int tile);
// FAST P2G tile shapes (particles per pass over a tile): 0 = the default.
int tile_shape(int tile);
bool tile_shape_valid(int tile);
int launch_spill_batch(hipStream_t st, const BatchSlice *tab, int nslices, int wgs);
int launch_compact_batch(hipStream_t st, const BatchSlice *tab, int nslices, int wgs);
int launch_permute_batch(hipStream_t st, const BatchSlice *src, const BatchSlice *dst, int nslices, int wgs);
int launch_grid_batch(hipStream_t st, const BatchSlice *tab, int nslices, int wgs);
int launch_g2p_batch(hipStream_t st, int mat, const BatchSlice *tab, int nslices, int wgs, int threads);
int launch_mlp_infer(hipStream_t st, const float *params, const PartView &pv);
int launch_mlp_infer_batch(hipStream_t st, const float *params, const BatchSlice *tab, int nslices, int group,
                           int64_t chunks);

// Scene sampling and files (scene.cpp)
int sample_object(const mpm_config &cfg, const mpm_object &obj, std::vector<float> &xyz);
int write_bgeo(const char *path, const float *xyz, int64_t n);

}  // namespace mpm
