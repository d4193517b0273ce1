// mpm_batch.hip -- many independent scenes advanced together.
//
// No counterpart in the reference (it steps one MPMSimulator per process).
// Each scene is an mpm_sim with one particle group in the FAST form.  The
// batch keeps a table of the scenes' parameters (BatchSlice) in device memory
// and advances every scene by one step with one launch per stage for the
// whole batch -- counters, atomic spill fallback, active-block compaction,
// grid update -- and one per stage and material for P2G tiles (per 64
// scenes) and G2P, each a persistent launch over the scenes of that
// material, plus one warm-start inference launch per set of scenes with
// equal network weights.  That sequence is recorded once into a hipGraph and replayed.
// Every `sort_every` steps all scenes are re-sorted by a second graph with
// one branch per scene, on the scenes' own streams.  Sorting moves particle
// storage between two buffers, so there is one step graph and one sort graph
// per buffer parity.  Launch widths come from mpm_batch_autotune, which times
// the candidates of each stage on live steps; results do not depend on them
// (every tile, block and particle is processed by exactly one workgroup).
#include "mpm_launch.h"

#include <string.h>

#include <algorithm>
#include <vector>

using namespace mpm;

namespace {

constexpr int SPILL_WGS = 512;
constexpr int NCAND = 4;
const int P2G_CAND[NCAND] = {1024, 2048, 4096, 8192};   // one-wave workgroups
const int GRID_CAND[NCAND] = {1024, 2048, 4096, 8192};
const int G2P_CAND[NCAND] = {64, 128, 192, 256};
constexpr int NTILE = 3;
const int TILE_CAND[NTILE] = {256, 384, 512};            // P2G particles per pass over a tile

// What the captured graphs depend on, per scene.
struct SceneSig {
    float *P0, *P1;
    uint32_t *pid0, *pid1;
    float *mlp_w;
    int64_t mlp_version;
    int warm;
    double dt;
    int sweeps;
    bool operator==(const SceneSig &o) const
    {
        return P0 == o.P0 && P1 == o.P1 && pid0 == o.pid0 && pid1 == o.pid1 && mlp_w == o.mlp_w &&
               mlp_version == o.mlp_version && warm == o.warm && dt == o.dt && sweeps == o.sweeps;
    }
};

}  // namespace

struct mpm_batch {
    std::vector<mpm_sim *> sims;
    hipStream_t stream = nullptr;
    mpm_batch_shape shape{};
    // particle buffers of parity 0 / 1 (P, P_alt of each scene at creation)
    std::vector<float *> P0, P1;
    std::vector<uint32_t *> pid0, pid1;
    BatchSlice *tab[2] = {};
    std::vector<SceneSig> sig;          // signature the tables and graphs were built for
    bool built = false;
    hipGraphExec_t step_exec[2] = {}, sort_exec[2] = {};
    int captured_p2g = 0, captured_grid = 0, captured_g2p = 0, captured_tile = 0;
    int parity = 0;
    int64_t steps_since_sort = 0;
    int64_t total = 0;
    int compact_wgs = 0;
    std::vector<int> types;             // material of each scene
    // scenes are kept ordered by material: material m holds table entries
    // [mat_begin[m], mat_begin[m] + mat_count[m]); its block compaction
    // takes mat_compact_wgs[m] workgroups
    int mat_begin[3] = {0, 0, 0}, mat_count[3] = {0, 0, 0}, mat_compact_wgs[3] = {0, 0, 0};
    int permute_wgs = 0;                // workgroups of the batched re-sort move
    SortSlice *stab[2] = {};            // batched re-sort tables of both parities
    SortBatchPlan splan;
    // warm-start sets present among each material's scenes, with their
    // 64-particle chunks there
    std::vector<std::pair<int, int64_t>> mat_sets[3];
    // warm-start weight sets: scenes with equal weights share one inference launch
    std::vector<const float *> warm_params; // device weights of each set
    std::vector<int64_t> warm_chunks;       // 64-particle chunks of each set
    std::vector<int> warm_group;            // per scene, -1: cold
    hipEvent_t ev_fork = nullptr, ev_done = nullptr;
    std::vector<hipEvent_t> ev_join;
    std::vector<hipEvent_t> tev;        // step timing events
};

namespace {

bool warm_on(const mpm_sim *s)
{
    return s->cfg.warm_start == MPM_WARM_MLP && s->mlp_w && s->mlp_count == mpm_mlp_param_count() &&
           s->groups[0].mat.type != MPM_WATER;
}

SceneSig sig_of(const mpm_batch *b, size_t k)
{
    const mpm_sim *s = b->sims[k];
    return SceneSig{b->P0[k], b->P1[k], b->pid0[k], b->pid1[k], s->mlp_w, s->mlp_version, warm_on(s) ? 1 : 0,
                    s->cfg.dt, s->cfg.svd_sweeps};
}

int parity_of(const mpm_batch *b, size_t k)
{
    const mpm_sim *s = b->sims[k];
    if (s->P == b->P0[k] && s->P_alt == b->P1[k]) return 0;
    if (s->P == b->P1[k] && s->P_alt == b->P0[k]) return 1;
    return -1;
}

// Point the scene's host view at the buffers of parity p.
void set_parity(mpm_batch *b, size_t k, int p)
{
    mpm_sim *s = b->sims[k];
    s->P = p ? b->P1[k] : b->P0[k];
    s->P_alt = p ? b->P0[k] : b->P1[k];
    s->pid = p ? b->pid1[k] : b->pid0[k];
    s->pid_alt = p ? b->pid0[k] : b->pid1[k];
}

// G2P workgroups of the scenes of material `mat` (0: none of them).
int g2p_wgs(const mpm_batch *b, int mat, int threads)
{
    uint32_t w = 0;
    for (mpm_sim *s : b->sims)
        if (s->groups[0].mat.type == mat) w += (uint32_t)((s->groups[0].n + threads - 1) / threads);
    return (int)round_xcd(w);
}


int check(int rc, const char *what) { return rc ? hip_fail((hipError_t)rc, what) : MPM_OK; }

// The launches of every material present (`launch(stream, material)`), as
// parallel branches: forked from the batch stream onto the stream of each
// material's first scene and joined back (the materials' scenes are
// disjoint, so one launch's tail overlaps the others' work).
template <class L>
int fork_materials(mpm_batch *b, L launch)
{
    MPM_HIP(hipEventRecord(b->ev_fork, b->stream));
    for (int m = MPM_JELLO; m <= MPM_WATER; ++m) {
        size_t k = 0;
        while (k < b->types.size() && b->types[k] != m) ++k;
        if (k == b->types.size()) continue;
        hipStream_t ms = b->sims[k]->stream;
        MPM_HIP(hipStreamWaitEvent(ms, b->ev_fork, 0));
        if (int rc = launch(ms, m)) return rc;
        MPM_HIP(hipEventRecord(b->ev_join[k], ms));
        MPM_HIP(hipStreamWaitEvent(b->stream, b->ev_join[k], 0));
    }
    return MPM_OK;
}

// The step of one material's scenes (table range [s0, s0 + cnt)) on stream
// ms: counter reset, P2G, spill scatter, block compaction, grid update, the
// warm-start inference of each weight set among them, G2P.
int enqueue_material(mpm_batch *b, hipStream_t ms, int m, const BatchSlice *tab, int s0, int cnt, int p2g, int grid,
                     int g2p_threads, int tile)
{
    const BatchSlice *t = tab + s0;
    int rc;
    if ((rc = check(launch_batch_zero(ms, t, cnt), "batch_zero"))) return rc;
    if ((rc = check(launch_p2g_batch(ms, m, t, b->types.data() + s0, cnt, p2g, tile), "p2g_batch"))) return rc;
    if ((rc = check(launch_spill_batch(ms, t, cnt, SPILL_WGS), "spill_batch"))) return rc;
    if ((rc = check(launch_compact_batch(ms, t, cnt, b->mat_compact_wgs[m]), "compact_batch"))) return rc;
    if ((rc = check(launch_grid_batch(ms, t, cnt, grid), "grid_batch"))) return rc;
    for (const auto &ws : b->mat_sets[m])
        if ((rc = check(launch_mlp_infer_batch(ms, b->warm_params[ws.first], t, cnt, ws.first, ws.second),
                        "mlp_infer_batch")))
            return rc;
    const int wgs = g2p_wgs(b, m, g2p_threads);
    if (wgs && (rc = check(launch_g2p_batch(ms, m, t, cnt, wgs, g2p_threads), "g2p_batch"))) return rc;
    return MPM_OK;
}

// One step of every scene (parity p tables).  Without `ev` each material's
// scenes run their whole step as one branch beside the other materials'.
// With `ev` (6 events, the autotuner's stage timing) the stages follow one
// another on the batch stream, only the materials' P2G and G2P launches side
// by side, and the P2G, grid-update and G2P stages are bracketed.
int enqueue_step(mpm_batch *b, int p, int p2g, int grid, int g2p_threads, int tile, hipEvent_t *ev)
{
    hipStream_t st = b->stream;
    const BatchSlice *tab = b->tab[p];
    const int n = (int)b->sims.size();
    int rc;
    if (!ev)
        return fork_materials(b, [&](hipStream_t ms, int m) {
            return enqueue_material(b, ms, m, tab, b->mat_begin[m], b->mat_count[m], p2g, grid, g2p_threads, tile);
        });
    if ((rc = check(launch_batch_zero(st, tab, n), "batch_zero"))) return rc;
    if (ev) MPM_HIP(hipEventRecord(ev[0], st));
    rc = fork_materials(b, [&](hipStream_t ms, int m) {
        return check(launch_p2g_batch(ms, m, tab, b->types.data(), n, p2g, tile), "p2g_batch");
    });
    if (rc) return rc;
    if (ev) MPM_HIP(hipEventRecord(ev[1], st));
    if ((rc = check(launch_spill_batch(st, tab, n, SPILL_WGS), "spill_batch"))) return rc;
    if ((rc = check(launch_compact_batch(st, tab, n, b->compact_wgs), "compact_batch"))) return rc;
    if (ev) MPM_HIP(hipEventRecord(ev[2], st));
    if ((rc = check(launch_grid_batch(st, tab, n, grid), "grid_batch"))) return rc;
    if (ev) MPM_HIP(hipEventRecord(ev[3], st));
    for (size_t w = 0; w < b->warm_params.size(); ++w)
        if ((rc = check(launch_mlp_infer_batch(st, b->warm_params[w], tab, n, (int)w, b->warm_chunks[w]),
                        "mlp_infer_batch")))
            return rc;
    if (ev) MPM_HIP(hipEventRecord(ev[4], st));
    rc = fork_materials(b, [&](hipStream_t ms, int m) {
        const int wgs = g2p_wgs(b, m, g2p_threads);
        return wgs ? check(launch_g2p_batch(ms, m, tab, n, wgs, g2p_threads), "g2p_batch") : MPM_OK;
    });
    if (rc) return rc;
    if (ev) MPM_HIP(hipEventRecord(ev[5], st));
    return MPM_OK;
}

// Re-sort every scene from parity p into parity 1 - p on the batch stream:
// keys, order and tiles of all scenes with one launch per sort stage, then
// one launch moves every scene's particles into the parity 1 - p buffers in
// its new order (the warm-start lists, which name old slots, are dropped).
// The caller points the host views at parity 1 - p.
int enqueue_sort(mpm_batch *b, int p)
{
    const int n = (int)b->sims.size();
    int rc = launch_sort_batch(b->stream, b->stab[p], n, b->splan);
    if (rc) return rc;
    return check(launch_permute_batch(b->stream, b->tab[p], b->tab[p ^ 1], n, b->permute_wgs), "permute_batch");
}

void drop_graphs(mpm_batch *b)
{
    for (int p = 0; p < 2; ++p) {
        if (b->step_exec[p]) (void)hipGraphExecDestroy(b->step_exec[p]);
        if (b->sort_exec[p]) (void)hipGraphExecDestroy(b->sort_exec[p]);
        b->step_exec[p] = b->sort_exec[p] = nullptr;
    }
}

// Scene bookkeeping kept by sort_all, saved around captures and replays.
struct SortBook {
    bool sorted_valid;
    int64_t steps_since_sort, sorted_at_launch, sorts;
};

SortBook book_of(const mpm_sim *s) { return SortBook{s->sorted_valid, s->steps_since_sort, s->sorted_at_launch, s->sorts}; }
void set_book(mpm_sim *s, const SortBook &k)
{
    s->sorted_valid = k.sorted_valid;
    s->steps_since_sort = k.steps_since_sort;
    s->sorted_at_launch = k.sorted_at_launch;
    s->sorts = k.sorts;
}

int capture(mpm_batch *b, hipGraphExec_t *out, bool sort, int p)
{
    std::vector<SortBook> book;
    for (mpm_sim *s : b->sims) book.push_back(book_of(s));
    MPM_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
    int rc = sort ? enqueue_sort(b, p)
                  : enqueue_step(b, p, b->shape.p2g_workgroups, b->shape.grid_workgroups, b->shape.g2p_threads,
                                 b->shape.p2g_tile, nullptr);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &graph);
    for (size_t k = 0; k < b->sims.size(); ++k) set_book(b->sims[k], book[k]);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    return MPM_OK;
}

// Device tables for both parities; graphs when the shape asks for them.
// Warm scenes grouped by identical weights.
void group_warm(mpm_batch *b)
{
    b->warm_params.clear();
    b->warm_chunks.clear();
    b->warm_group.assign(b->sims.size(), -1);
    std::vector<size_t> rep;                        // first scene of each set
    for (size_t k = 0; k < b->sims.size(); ++k) {
        mpm_sim *s = b->sims[k];
        if (!warm_on(s)) continue;
        int g = -1;
        for (size_t r = 0; r < rep.size() && g < 0; ++r)
            if (b->sims[rep[r]]->mlp_host == s->mlp_host) g = (int)r;
        if (g < 0) {
            g = (int)rep.size();
            rep.push_back(k);
            b->warm_params.push_back(s->mlp_w);
            b->warm_chunks.push_back(0);
        }
        b->warm_group[k] = g;
        b->warm_chunks[g] += (s->groups[0].n + 63) / 64;
    }
    for (int m = 0; m < 3; ++m) {
        b->mat_sets[m].clear();
        for (int k = b->mat_begin[m]; k < b->mat_begin[m] + b->mat_count[m]; ++k) {
            const int g = b->warm_group[k];
            if (g < 0) continue;
            size_t j = 0;
            while (j < b->mat_sets[m].size() && b->mat_sets[m][j].first != g) ++j;
            if (j == b->mat_sets[m].size()) b->mat_sets[m].push_back({g, 0});
            b->mat_sets[m][j].second += (b->sims[k]->groups[0].n + 63) / 64;
        }
    }
}

int build(mpm_batch *b)
{
    const size_t n = b->sims.size();
    MPM_HIP(hipStreamSynchronize(b->stream));       // no replay still reads the tables
    group_warm(b);
    for (mpm_sim *s : b->sims) {
        int rc = warm_on(s) ? ensure_warm_lists(s) : MPM_OK;
        if (rc) return rc;
        // the re-sort order the batched move reads (allocated here, never
        // inside a captured sort)
        if (s->g2p_perm_cap < s->cap) {
            if (s->g2p_perm) (void)hipFree(s->g2p_perm);
            s->g2p_perm = nullptr;
            s->g2p_perm_cap = 0;
            MPM_HIP(hipMalloc(&s->g2p_perm, (size_t)s->cap * sizeof(uint32_t)));
            s->g2p_perm_cap = s->cap;
        }
    }
    for (int p = 0; p < 2; ++p) {
        std::vector<BatchSlice> h(n);
        for (size_t k = 0; k < n; ++k) {
            mpm_sim *s = b->sims[k];
            float *P = s->P, *Pa = s->P_alt;
            uint32_t *id = s->pid, *ida = s->pid_alt;
            set_parity(b, k, p);
            fill_batch_slice(s, h[k]);
            h[k].mat.warm_sweeps = warm_on(s) ? std::max(1, s->cfg.warm_sweeps) : 0;
            h[k].warm_group = b->warm_group[k];
            if (warm_on(s)) {
                // the G2P lists the particles that take the SVD; the next
                // step's inference runs over them (k_batch_zero hands them over)
                h[k].mat.wlist = s->warm_list;
                h[k].mat.wcount = s->warm_cnt + 1;
                h[k].warm_cnt = s->warm_cnt;
            }
            s->P = P; s->P_alt = Pa; s->pid = id; s->pid_alt = ida;
        }
        MPM_HIP(hipMemcpy(b->tab[p], h.data(), n * sizeof(BatchSlice), hipMemcpyHostToDevice));
        std::vector<SortSlice> hs(n);
        for (size_t k = 0; k < n; ++k) {
            mpm_sim *s = b->sims[k];
            float *P = s->P, *Pa = s->P_alt;
            uint32_t *id = s->pid, *ida = s->pid_alt;
            set_parity(b, k, p);
            int rc = fill_sort_slice(s, hs[k]);
            s->P = P; s->P_alt = Pa; s->pid = id; s->pid_alt = ida;
            if (rc) return rc;
        }
        MPM_HIP(hipMemcpy(b->stab[p], hs.data(), n * sizeof(SortSlice), hipMemcpyHostToDevice));
    }
    b->sig.clear();
    for (size_t k = 0; k < n; ++k) b->sig.push_back(sig_of(b, k));
    drop_graphs(b);
    b->built = true;
    if (!b->shape.use_graph) return MPM_OK;
    std::vector<int> par(n);
    for (size_t k = 0; k < n; ++k) par[k] = parity_of(b, k);
    int rc = MPM_OK;
    for (int p = 0; p < 2 && !rc; ++p) {
        rc = capture(b, &b->sort_exec[p], true, p);
        if (!rc) rc = capture(b, &b->step_exec[p], false, p);
    }
    for (size_t k = 0; k < n; ++k) set_parity(b, k, par[k]);
    if (rc) { drop_graphs(b); b->built = false; return rc; }
    b->captured_p2g = b->shape.p2g_workgroups;
    b->captured_grid = b->shape.grid_workgroups;
    b->captured_g2p = b->shape.g2p_threads;
    b->captured_tile = b->shape.p2g_tile;
    return MPM_OK;
}

int validate_shape(const mpm_batch_shape &s)
{
    if (s.p2g_workgroups < 1 || s.p2g_workgroups > (1 << 20)) return set_error(MPM_ERR_INVALID, "p2g_workgroups out of range");
    if (s.grid_workgroups < 1 || s.grid_workgroups > (1 << 20)) return set_error(MPM_ERR_INVALID, "grid_workgroups out of range");
    if (s.g2p_threads < 64 || s.g2p_threads > 256 || s.g2p_threads % 64)
        return set_error(MPM_ERR_INVALID, "g2p_threads must be 64, 128, 192 or 256");
    if (s.sort_every < 1) return set_error(MPM_ERR_INVALID, "sort_every must be >= 1");
    if (s.p2g_tile != 0 && !tile_shape_valid(s.p2g_tile)) return set_error(MPM_ERR_INVALID, "p2g_tile must be 256, 384 or 512");
    return MPM_OK;
}

// Before a run: every scene in one buffer parity with a valid stress, the
// tables and graphs current, and the batch stream after the scenes' streams.
int prepare(mpm_batch *b)
{
    const size_t n = b->sims.size();
    for (size_t k = 0; k < n; ++k) {
        mpm_sim *s = b->sims[k];
        s->counters_clean = false;     // batch steps reset the counters themselves
        if (int rc = flush_spill(s)) return rc;
        if (parity_of(b, k) < 0) return set_error(MPM_ERR_INVALID, "a scene's particle storage changed since mpm_batch_create");
        if (s->groups.size() != 1) return set_error(MPM_ERR_INVALID, "a scene of the batch gained a particle group");
        int rc = fast_ready(s);
        if (rc) return rc;
    }
    // scenes stepped on their own may have re-sorted an odd number of times
    int ones = 0;
    for (size_t k = 0; k < n; ++k) ones += parity_of(b, k);
    b->parity = 2 * ones > (int)n ? 1 : 0;
    for (size_t k = 0; k < n; ++k)
        if (parity_of(b, k) != b->parity) {
            int rc = sort_all(b->sims[k]);
            if (rc) return rc;
        }
    bool same = b->built && b->sig.size() == n;
    for (size_t k = 0; same && k < n; ++k) same = sig_of(b, k) == b->sig[k];
    if (same && b->shape.use_graph)
        same = b->step_exec[0] && b->captured_p2g == b->shape.p2g_workgroups &&
               b->captured_grid == b->shape.grid_workgroups && b->captured_g2p == b->shape.g2p_threads &&
               b->captured_tile == b->shape.p2g_tile;
    if (!same) {
        for (mpm_sim *s : b->sims) MPM_HIP(hipStreamSynchronize(s->stream));
        int rc = build(b);
        if (rc) return rc;
    }
    b->steps_since_sort = 0;
    for (mpm_sim *s : b->sims) b->steps_since_sort = std::max(b->steps_since_sort, s->steps_since_sort);
    for (size_t k = 0; k < n; ++k) {
        MPM_HIP(hipEventRecord(b->ev_join[k], b->sims[k]->stream));
        MPM_HIP(hipStreamWaitEvent(b->stream, b->ev_join[k], 0));
    }
    return MPM_OK;
}

// Scene streams continue after the batch's work.
int release(mpm_batch *b)
{
    MPM_HIP(hipEventRecord(b->ev_done, b->stream));
    for (mpm_sim *s : b->sims) MPM_HIP(hipStreamWaitEvent(s->stream, b->ev_done, 0));
    return MPM_OK;
}

void after_sort(mpm_batch *b)
{
    b->parity ^= 1;
    for (size_t k = 0; k < b->sims.size(); ++k) {
        mpm_sim *s = b->sims[k];
        set_parity(b, k, b->parity);
        s->sorted_valid = true;
        s->steps_since_sort = 0;
        s->sorted_at_launch = s->tile_launches;
        s->sorts += 1;
    }
    b->steps_since_sort = 0;
}

void after_step(mpm_batch *b)
{
    for (mpm_sim *s : b->sims) {
        s->steps += 1;
        s->time += s->cfg.dt;
        s->last_dt = s->cfg.dt;
        s->steps_since_sort += 1;
        s->tau_valid = true;
        s->last_phase = MPM_PHASE_G2P;
    }
    b->steps_since_sort += 1;
}

// Sort when due, then one step; graphs or direct launches.
int one_step(mpm_batch *b, int p2g, int grid, int g2p_threads, int tile, hipEvent_t *ev, bool graphs)
{
    if (b->steps_since_sort >= b->shape.sort_every) {
        if (graphs) {
            MPM_HIP(hipGraphLaunch(b->sort_exec[b->parity], b->stream));
        } else {
            std::vector<SortBook> book;
            for (mpm_sim *s : b->sims) book.push_back(book_of(s));
            int rc = enqueue_sort(b, b->parity);
            for (size_t k = 0; k < b->sims.size(); ++k) set_book(b->sims[k], book[k]);
            if (rc) return rc;
        }
        after_sort(b);
    }
    if (graphs) {
        MPM_HIP(hipGraphLaunch(b->step_exec[b->parity], b->stream));
    } else {
        int rc = enqueue_step(b, b->parity, p2g, grid, g2p_threads, tile, ev);
        if (rc) return rc;
    }
    after_step(b);
    return MPM_OK;
}

hipEvent_t *timing_events(mpm_batch *b, size_t need)
{
    while (b->tev.size() < need) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        b->tev.push_back(e);
    }
    return b->tev.data();
}

}  // namespace

extern "C" {

void mpm_batch_shape_default(mpm_batch_shape *s)
{
    if (!s) return;
    memset(s, 0, sizeof *s);
    s->p2g_workgroups = 4096;
    s->grid_workgroups = 4096;
    s->g2p_threads = 256;
    s->sort_every = 24;
    s->use_graph = 1;
    s->p2g_tile = 384;
}

int mpm_batch_create(mpm_sim *const *sims, int32_t n, const mpm_batch_shape *shape, mpm_batch **out)
{
    if (!sims || !out) return set_error(MPM_ERR_INVALID, "null argument");
    *out = nullptr;
    if (n < 1 || n > BATCH_MAX) return set_error(MPM_ERR_INVALID, "a batch holds 1 to 256 scenes");
    mpm_batch_shape sh;
    mpm_batch_shape_default(&sh);
    if (shape) sh = *shape;
    int rc = validate_shape(sh);
    if (rc) return rc;
    for (int k = 0; k < n; ++k) {
        mpm_sim *s = sims[k];
        if (!s) return set_error(MPM_ERR_INVALID, "scene is null");
        if (s->cfg.form != MPM_FORM_FAST) return set_error(MPM_ERR_INVALID, "batched scenes use the FAST form");
        if (s->cfg.dt_mode != MPM_DT_FIXED)
            return set_error(MPM_ERR_INVALID, "batched scenes step with their fixed dt (dt_mode FIXED)");
        if (s->groups.size() != 1 || s->groups[0].n <= 0)
            return set_error(MPM_ERR_INVALID, "a batched scene holds exactly one non-empty particle group");
        for (int j = 0; j < k; ++j)
            if (sims[j] == s) return set_error(MPM_ERR_INVALID, "a scene appears twice in the batch");
    }
    mpm_batch *b = new mpm_batch();
    b->sims.assign(sims, sims + n);
    // ordered by material (stable): each material's scenes form one range of
    // the tables, so a step can run the materials side by side
    std::stable_sort(b->sims.begin(), b->sims.end(), [](const mpm_sim *x, const mpm_sim *y) {
        return x->groups[0].mat.type < y->groups[0].mat.type;
    });
    b->shape = sh;
    auto fail = [&](int code) { mpm_batch_destroy(b); return code; };
    for (mpm_sim *s : b->sims) {
        if ((rc = fast_ready(s))) return fail(rc);         // buffers, first sort, stress
        b->total += s->groups[0].n;
    }
    for (mpm_sim *s : b->sims)
        if (hipStreamSynchronize(s->stream) != hipSuccess) return fail(hip_fail(hipErrorUnknown, "scene sync"));
    for (mpm_sim *s : b->sims) b->types.push_back(s->groups[0].mat.type);
    for (mpm_sim *s : b->sims) {
        b->P0.push_back(s->P); b->P1.push_back(s->P_alt);
        b->pid0.push_back(s->pid); b->pid1.push_back(s->pid_alt);
        b->compact_wgs += (int)((s->nblocks + 255) / 256);
    }
    uint32_t pblocks = 0;
    for (mpm_sim *s : b->sims) pblocks += (uint32_t)((s->groups[0].n + 255) / 256);
    b->permute_wgs = (int)round_xcd(pblocks);
    for (int k = 0; k < n; ++k) {
        const int m = b->types[k];
        if (b->mat_count[m] == 0) b->mat_begin[m] = k;
        b->mat_count[m] += 1;
        b->mat_compact_wgs[m] += (int)((b->sims[k]->nblocks + 255) / 256);
    }
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) return fail(hip_fail(e, "stream"));
    if ((e = hipMalloc(&b->tab[0], n * sizeof(BatchSlice))) != hipSuccess) return fail(hip_fail(e, "hipMalloc"));
    if ((e = hipMalloc(&b->tab[1], n * sizeof(BatchSlice))) != hipSuccess) return fail(hip_fail(e, "hipMalloc"));
    for (int p = 0; p < 2; ++p)
        if ((e = hipMalloc(&b->stab[p], n * sizeof(SortSlice))) != hipSuccess) return fail(hip_fail(e, "hipMalloc"));
    sort_batch_plan(b->sims.data(), n, b->splan);
    if ((e = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming)) != hipSuccess) return fail(hip_fail(e, "event"));
    if ((e = hipEventCreateWithFlags(&b->ev_done, hipEventDisableTiming)) != hipSuccess) return fail(hip_fail(e, "event"));
    b->ev_join.assign(n, nullptr);
    for (auto &ev : b->ev_join)
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) return fail(hip_fail(e, "event"));
    *out = b;
    return MPM_OK;
}

int mpm_batch_get_shape(const mpm_batch *b, mpm_batch_shape *out)
{
    if (!b || !out) return set_error(MPM_ERR_INVALID, "null argument");
    *out = b->shape;
    return MPM_OK;
}

int mpm_batch_set_shape(mpm_batch *b, const mpm_batch_shape *shape)
{
    if (!b || !shape) return set_error(MPM_ERR_INVALID, "null argument");
    int rc = validate_shape(*shape);
    if (rc) return rc;
    b->shape = *shape;
    return MPM_OK;
}

int64_t mpm_batch_particles(const mpm_batch *b) { return b ? b->total : 0; }
int32_t mpm_batch_scenes(const mpm_batch *b) { return b ? (int32_t)b->sims.size() : 0; }

int mpm_batch_run(mpm_batch *b, int32_t nsteps, float *step_ms)
{
    if (!b) return set_error(MPM_ERR_INVALID, "batch is null");
    if (nsteps < 0) return set_error(MPM_ERR_INVALID, "nsteps must be >= 0");
    if (nsteps == 0) return MPM_OK;
    int rc = prepare(b);
    if (rc) return rc;
    hipEvent_t *ev = step_ms ? timing_events(b, (size_t)nsteps + 1) : nullptr;
    if (step_ms && !ev) return set_error(MPM_ERR_HIP, "hipEventCreate failed");
    const bool graphs = b->shape.use_graph != 0;
    for (int32_t k = 0; k < nsteps; ++k) {
        if (ev) MPM_HIP(hipEventRecord(ev[k], b->stream));
        rc = one_step(b, b->shape.p2g_workgroups, b->shape.grid_workgroups, b->shape.g2p_threads, b->shape.p2g_tile,
                      nullptr, graphs);
        if (rc) return rc;
    }
    if (ev) MPM_HIP(hipEventRecord(ev[nsteps], b->stream));
    if ((rc = release(b))) return rc;
    if (ev) {
        MPM_HIP(hipEventSynchronize(ev[nsteps]));
        for (int32_t k = 0; k < nsteps; ++k) MPM_HIP(hipEventElapsedTime(&step_ms[k], ev[k], ev[k + 1]));
    }
    return MPM_OK;
}

// Times each stage's launch-width candidates on live steps (rounds of
// NCAND steps, candidate c in step c of every round) and keeps the fastest
// median per stage; then, with those widths, the P2G tile shapes (rounds of
// NTILE steps) and keeps the fastest.  Advances every scene by
// reps * (NCAND + NTILE) steps.  ms_out (optional, 3 x NCAND + NTILE):
// median milliseconds of P2G, grid update and G2P per width candidate, then
// of P2G per tile shape.
int mpm_batch_autotune(mpm_batch *b, int32_t reps, float *ms_out)
{
    if (!b) return set_error(MPM_ERR_INVALID, "batch is null");
    if (reps < 1 || reps > 1000) return set_error(MPM_ERR_INVALID, "reps out of range [1,1000]");
    int rc = prepare(b);
    if (rc) return rc;
    const size_t per = 6;
    hipEvent_t *ev = timing_events(b, (size_t)reps * (NCAND + NTILE) * per);
    if (!ev) return set_error(MPM_ERR_HIP, "hipEventCreate failed");
    const int tile0 = tile_shape(b->shape.p2g_tile);
    for (int32_t r = 0; r < reps; ++r)
        for (int c = 0; c < NCAND; ++c) {
            hipEvent_t *e = ev + ((size_t)r * NCAND + c) * per;
            rc = one_step(b, P2G_CAND[c], GRID_CAND[c], G2P_CAND[c], tile0, e, false);
            if (rc) return rc;
        }
    MPM_HIP(hipStreamSynchronize(b->stream));
    // median of the span [e[k0], e[k0 + 1]] over the rounds
    auto median = [&](size_t first, size_t stride, int k0, float &out) -> int {
        std::vector<float> t;
        for (int32_t r = 0; r < reps; ++r) {
            hipEvent_t *e = ev + (first + (size_t)r * stride) * per;
            float v = 0;
            MPM_HIP(hipEventElapsedTime(&v, e[k0], e[k0 + 1]));
            t.push_back(v);
        }
        std::sort(t.begin(), t.end());
        out = t[t.size() / 2];
        return MPM_OK;
    };
    float med[3][NCAND];
    for (int c = 0; c < NCAND; ++c)
        for (int j = 0; j < 3; ++j)
            if ((rc = median((size_t)c, NCAND, 2 * j, med[j][c]))) return rc;
    int best[3] = {0, 0, 0};
    for (int j = 0; j < 3; ++j)
        for (int c = 1; c < NCAND; ++c)
            if (med[j][c] < med[j][best[j]]) best[j] = c;
    b->shape.p2g_workgroups = P2G_CAND[best[0]];
    b->shape.grid_workgroups = GRID_CAND[best[1]];
    b->shape.g2p_threads = G2P_CAND[best[2]];
    // tile shapes with the chosen widths
    const size_t t0 = (size_t)reps * NCAND;
    for (int32_t r = 0; r < reps; ++r)
        for (int c = 0; c < NTILE; ++c) {
            hipEvent_t *e = ev + (t0 + (size_t)r * NTILE + c) * per;
            rc = one_step(b, b->shape.p2g_workgroups, b->shape.grid_workgroups, b->shape.g2p_threads, TILE_CAND[c], e,
                          false);
            if (rc) return rc;
        }
    if ((rc = release(b))) return rc;
    MPM_HIP(hipStreamSynchronize(b->stream));
    float tmed[NTILE];
    int tbest = 0;
    for (int c = 0; c < NTILE; ++c) {
        if ((rc = median(t0 + (size_t)c, NTILE, 0, tmed[c]))) return rc;
        if (tmed[c] < tmed[tbest]) tbest = c;
    }
    b->shape.p2g_tile = TILE_CAND[tbest];
    if (ms_out) {
        for (int j = 0; j < 3; ++j)
            for (int c = 0; c < NCAND; ++c) ms_out[j * NCAND + c] = med[j][c];
        for (int c = 0; c < NTILE; ++c) ms_out[3 * NCAND + c] = tmed[c];
    }
    return MPM_OK;
}

int mpm_batch_synchronize(mpm_batch *b)
{
    if (!b) return set_error(MPM_ERR_INVALID, "batch is null");
    MPM_HIP(hipStreamSynchronize(b->stream));
    return MPM_OK;
}

void mpm_batch_destroy(mpm_batch *b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graphs(b);
    for (int p = 0; p < 2; ++p) {
        if (b->tab[p]) (void)hipFree(b->tab[p]);
        if (b->stab[p]) (void)hipFree(b->stab[p]);
    }
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->ev_done) (void)hipEventDestroy(b->ev_done);
    for (hipEvent_t e : b->ev_join)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->tev)
        if (e) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

}  // extern "C"
