// mpm_api.hip -- C ABI of libmpm_hip (declared in include/mpm/mpm.h).
//
// Host orchestration of the step: MPMSimulator (reference
// src/simulator/MPMSimulator.cpp, MPMCUDA.cu:793-925) restated over one HIP
// stream per simulator, with status codes instead of exit(1).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "mpm_launch.h"

namespace mpm {

static thread_local std::string g_err;

int set_error(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return set_error(MPM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

static int check_launch(int rc, const char *what)
{
    if (rc != 0) return hip_fail((hipError_t)rc, what);
    return MPM_OK;
}

static int width_of(int field)
{
    switch (field) {
    case MPM_P_POS: case MPM_P_VEL: return 3;
    case MPM_P_C: case MPM_P_F: return 9;
    case MPM_P_JP: return 1;
    case MPM_P_TAU: return 6;
    default: return -1;
    }
}

static int plane_of(int field)
{
    switch (field) {
    case MPM_P_POS: return PL_X;
    case MPM_P_VEL: return PL_V;
    case MPM_P_C: return PL_C;
    case MPM_P_F: return PL_F;
    case MPM_P_JP: return PL_S;
    case MPM_P_TAU: return PL_TAU;
    default: return -1;
    }
}

PartView view_of(mpm_sim *s, int g)
{
    PartView pv;
    pv.P = s->P;
    pv.pid = s->pid;
    pv.cap = s->cap;
    pv.off = s->groups[g].off;
    pv.n = (int)s->groups[g].n;
    return pv;
}

// Grow particle storage to hold `need` particles (planes keep their data).
static int ensure_capacity(mpm_sim *s, int64_t need)
{
    if (need <= s->cap) return MPM_OK;
    int64_t ncap = std::max<int64_t>(need, s->cap + s->cap / 2);
    ncap = (ncap + 255) / 256 * 256;
    float *P = nullptr, *Pa = nullptr;
    uint32_t *pid = nullptr, *pida = nullptr;
    MPM_HIP(hipMalloc(&P, (size_t)ncap * PL_COUNT * sizeof(float)));
    MPM_HIP(hipMalloc(&Pa, (size_t)ncap * PL_COUNT * sizeof(float)));
    MPM_HIP(hipMalloc(&pid, (size_t)ncap * sizeof(uint32_t)));
    MPM_HIP(hipMalloc(&pida, (size_t)ncap * sizeof(uint32_t)));
    MPM_HIP(hipMemsetAsync(P, 0, (size_t)ncap * PL_COUNT * sizeof(float), s->stream));
    if (s->P && s->total > 0) {
        for (int p = 0; p < PL_COUNT; ++p)
            MPM_HIP(hipMemcpyAsync(P + (size_t)p * ncap, s->P + (size_t)p * s->cap,
                                   (size_t)s->total * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
        MPM_HIP(hipMemcpyAsync(pid, s->pid, (size_t)s->total * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                               s->stream));
    }
    MPM_HIP(hipStreamSynchronize(s->stream));
    if (s->P) { (void)hipFree(s->P); (void)hipFree(s->P_alt); (void)hipFree(s->pid); (void)hipFree(s->pid_alt); }
    s->P = P; s->P_alt = Pa; s->pid = pid; s->pid_alt = pida;
    s->cap = ncap;
    return MPM_OK;
}

// Newton steps R <- (R + R^-T) / 2 from R = F that bring every singular
// value of the band [lo, hi] to within 1e-7 of 1 (the iteration maps a
// singular value s to (s + 1/s) / 2; one FP32 ulp at 1 is 1.2e-7).  The snow
// clamp band (0.975, 1.0075) takes two.
static int polar_steps_for_band(double lo, double hi)
{
    int steps = 1;
    for (double s0 : {lo, hi}) {
        double s = s0;
        int k = 0;
        while (std::fabs(s - 1.0) > 1e-7 && k < 8) { s = 0.5 * (s + 1.0 / s); ++k; }
        steps = std::max(steps, k);
    }
    return steps;
}

static MatP make_matp(const mpm_material &m, double mass, double vol)
{
    MatP p;
    p.type = m.type;
    p.water_c_mode = m.water_c_mode;
    p.mass = (float)mass;
    p.vol = (float)vol;
    p.mu = (float)(m.E / (2.0 * (1.0 + m.nu)));                          // Elasiticity.cuh:42
    p.lam = (float)(m.E * m.nu / ((1.0 + m.nu) * (1.0 - 2.0 * m.nu)));   // Elasiticity.cuh:43
    p.theta_c = (float)m.theta_c;
    p.theta_s = (float)m.theta_s;
    p.xi = (float)m.xi;
    p.eos_k = (float)m.eos_k;
    p.eos_gamma = (float)m.eos_gamma;
    p.warm_sweeps = 0;
    p.polar_steps = polar_steps_for_band(1.0 - m.theta_c, 1.0 + m.theta_s);
    // squared wave speed at rest for the step-size controller: P waves
    // (lam + 2 mu) / rho, water's bulk modulus k gamma over rho
    const double rho = mass / vol;
    const double mu = m.E / (2.0 * (1.0 + m.nu)), lam = m.E * m.nu / ((1.0 + m.nu) * (1.0 - 2.0 * m.nu));
    p.c2 = (float)((m.type == MPM_WATER ? m.eos_k * m.eos_gamma : lam + 2.0 * mu) / rho);
    return p;
}

static int validate_material(const mpm_material *m)
{
    if (!m) return set_error(MPM_ERR_INVALID, "material is null");
    if (m->type < MPM_JELLO || m->type > MPM_WATER)
        return set_error(MPM_ERR_INVALID, "unknown material type");
    if (!(m->E >= 0)) return set_error(MPM_ERR_INVALID, "Youngs Modulus is out of range.");   // Elasiticity.cuh:34-37
    if (!(m->nu >= 0 && m->nu < 0.5)) return set_error(MPM_ERR_INVALID, "Poison Ratio is out of range."); // :38-41
    if (!(m->rho > 0)) return set_error(MPM_ERR_INVALID, "density must be positive");
    return MPM_OK;
}

static int run_phase(mpm_sim *s, int phase, float dt);
static int fused_step(mpm_sim *s, float dt);

}  // namespace mpm

using namespace mpm;

namespace mpm {
int points_in_mesh(const char *path, const double *pts, int64_t n, uint8_t *inside);
}

extern "C" {

const char *mpm_last_error(void) { return g_err.c_str(); }

void mpm_config_default(mpm_config *c)
{
    if (!c) return;
    memset(c, 0, sizeof *c);
    c->h = 0.04;                 // main.cpp:161
    c->nx = c->ny = c->nz = 64;
    c->dt = 1.0 / 5000.0;        // main.cpp:161
    c->gravity[1] = -9.8;        // MPMSimulator.cpp:186
    c->bc_pad = 10;              // MPMCUDA.cu:845-850
    c->bc_mode = MPM_BC_AXIS;
    c->ppc = 10;                 // main.cpp:161
    c->volume_mode = MPM_VOLUME_REF;
    c->seed = 20261202ull;
    c->form = MPM_FORM_FAST;
    c->sort_every = 24;
    c->svd_sweeps = 4;
    c->warm_start = MPM_WARM_NONE;
    c->warm_sweeps = 1;
    c->tile = 0;
    c->threads = 256;
    c->dt_mode = MPM_DT_FIXED;   // adp_dt = max_dt (MPMSimulator.cpp:184-185)
    c->cfl = 0.5f;
}

int mpm_material_init(mpm_material *m, double E, double nu, double rho, int type)
{
    if (!m) return set_error(MPM_ERR_INVALID, "material is null");
    memset(m, 0, sizeof *m);
    m->type = type;
    m->water_c_mode = MPM_WATER_C_APIC;
    m->E = E; m->nu = nu; m->rho = rho;
    m->theta_c = 8e-2; m->theta_s = 2e-2;   // MPMCUDA.cu:658-660
    m->xi = 10.0;                           // MPMCUDA.cu:344
    m->eos_k = 2.51; m->eos_gamma = 7.0;    // MPMCUDA.cu:353
    return validate_material(m);
}

int mpm_create(const mpm_config *cfg, mpm_sim **out)
{
    if (!cfg || !out) return set_error(MPM_ERR_INVALID, "null argument");
    *out = nullptr;
    if (!(cfg->h > 0)) return set_error(MPM_ERR_INVALID, "grid spacing must be positive");
    if (cfg->nx < 4 || cfg->ny < 4 || cfg->nz < 4) return set_error(MPM_ERR_INVALID, "grid needs >= 4 nodes per axis");
    if ((int64_t)cfg->nx * cfg->ny * cfg->nz > (1ll << 31))
        return set_error(MPM_ERR_INVALID, "grid too large");
    if (cfg->form != MPM_FORM_PLAIN && cfg->form != MPM_FORM_FAST && cfg->form != MPM_FORM_FUSED)
        return set_error(MPM_ERR_INVALID, "unknown form");
    if (!tile_shape_valid(cfg->tile)) return set_error(MPM_ERR_INVALID, "tile must be 0, 256, 384 or 512");
    if (cfg->bc_mode != MPM_BC_AXIS && cfg->bc_mode != MPM_BC_STICKY) return set_error(MPM_ERR_INVALID, "unknown bc_mode");
    if (cfg->svd_sweeps < 1 || cfg->svd_sweeps > 16) return set_error(MPM_ERR_INVALID, "svd_sweeps out of range [1,16]");
    if (cfg->dt_mode < MPM_DT_FIXED || cfg->dt_mode > MPM_DT_CFL_DEFORM) return set_error(MPM_ERR_INVALID, "unknown dt_mode");
    if (cfg->dt_mode != MPM_DT_FIXED && cfg->form == MPM_FORM_FUSED)
        return set_error(MPM_ERR_INVALID, "the FUSED form scatters the next step before its dt is known; "
                                          "use dt_mode FIXED or the FAST form");
    if (cfg->dt_mode != MPM_DT_FIXED && !(cfg->cfl > 0.0f && cfg->cfl <= 1.0f))
        return set_error(MPM_ERR_INVALID, "cfl must be in (0, 1]");
    int threads = cfg->threads ? cfg->threads : 256;
    if (threads != 64 && threads != 128 && threads != 256 && threads != 512)
        return set_error(MPM_ERR_INVALID, "threads must be 64, 128, 256 or 512");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(MPM_ERR_HIP, "no HIP device");
    mpm_sim *s = new mpm_sim();
    s->cfg = *cfg;
    s->cfg.threads = threads;
    if (s->cfg.sort_every < 1) s->cfg.sort_every = 1;
    GridP &g = s->g;
    g.nx = cfg->nx; g.ny = cfg->ny; g.nz = cfg->nz;
    g.nbx = (g.nx + BLK - 1) / BLK; g.nby = (g.ny + BLK - 1) / BLK; g.nbz = (g.nz + BLK - 1) / BLK;
    g.h = (float)cfg->h;
    g.inv_h = (float)(1.0 / cfg->h);
    g.ox = (float)cfg->origin[0]; g.oy = (float)cfg->origin[1]; g.oz = (float)cfg->origin[2];
    g.gx = (float)cfg->gravity[0]; g.gy = (float)cfg->gravity[1]; g.gz = (float)cfg->gravity[2];
    g.pad = cfg->bc_pad; g.bc_mode = cfg->bc_mode;
    g.dtp = nullptr;
    s->nblocks = (int64_t)g.nbx * g.nby * g.nbz;
    int rc = MPM_OK;
    auto fail = [&](hipError_t e, const char *w) { rc = hip_fail(e, w); };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking)) != hipSuccess) fail(e, "hipStreamCreate");
    size_t gbytes = (size_t)s->nblocks * BLK_NODES * sizeof(float4);
    if (!rc && (e = hipMalloc(&s->grid_mp, gbytes)) != hipSuccess) fail(e, "hipMalloc grid_mp");
    if (!rc && (e = hipMalloc(&s->grid_v, gbytes)) != hipSuccess) fail(e, "hipMalloc grid_v");
    if (!rc && (e = hipMalloc(&s->flags, (size_t)s->nblocks * 4)) != hipSuccess) fail(e, "hipMalloc flags");
    if (!rc && (e = hipMalloc(&s->list, (size_t)s->nblocks * 4)) != hipSuccess) fail(e, "hipMalloc list");
    if (!rc && (e = hipMalloc(&s->counters, CT_COUNT * 4)) != hipSuccess) fail(e, "hipMalloc counters");
    if (!rc && (e = hipMemsetAsync(s->grid_mp, 0, gbytes, s->stream)) != hipSuccess) fail(e, "memset");
    if (!rc && (e = hipMemsetAsync(s->grid_v, 0, gbytes, s->stream)) != hipSuccess) fail(e, "memset");
    if (!rc && (e = hipMemsetAsync(s->flags, 0, (size_t)s->nblocks * 4, s->stream)) != hipSuccess) fail(e, "memset");
    if (!rc && (e = hipMemsetAsync(s->counters, 0, CT_COUNT * 4, s->stream)) != hipSuccess) fail(e, "memset");
    if (!rc && cfg->dt_mode != MPM_DT_FIXED) {
        if ((e = hipMalloc(&s->dtc, sizeof(DtCtl))) != hipSuccess) fail(e, "hipMalloc dt state");
        else if ((e = hipMemsetAsync(s->dtc, 0, sizeof(DtCtl), s->stream)) != hipSuccess) fail(e, "memset");
        else g.dtp = &s->dtc->dt;
    }
    if (!rc && (e = hipStreamSynchronize(s->stream)) != hipSuccess) fail(e, "sync");
    if (rc) { mpm_destroy(s); return rc; }
    *out = s;
    return MPM_OK;
}

int mpm_add_particles(mpm_sim *s, const mpm_material *mat, int64_t n, const float *xyz, const float *vel,
                      const double v0[3], double particle_volume, int32_t *group)
{
    if (!s) return set_error(MPM_ERR_INVALID, "sim is null");
    int rc = validate_material(mat);
    if (rc) return rc;
    if (n <= 0 || !xyz) return set_error(MPM_ERR_INVALID, "no particles");
    if ((rc = flush_spill(s))) return rc;     // a P2G phase's spill scatter reads the state it saw
    // particle offsets inside a plane are 32-bit byte offsets on the device
    if (s->total + n > MPM_MAX_PARTICLES) return set_error(MPM_ERR_INVALID, "too many particles");
    if ((int64_t)s->groups.size() >= MAX_GROUPS) return set_error(MPM_ERR_INVALID, "at most 256 particle groups");
    // every particle's stencil must lie inside the grid (MPMSimulator.cpp:210-222)
    const GridP &g = s->g;
    const float o[3] = {g.ox, g.oy, g.oz};
    const int nn[3] = {g.nx, g.ny, g.nz};
    for (int64_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            float x = xyz[3 * i + a];
            float hi = o[a] + g.h * (float)(nn[a] - 1);
            if (!(x >= o[a] && x <= hi))
                return set_error(MPM_ERR_RANGE, "particle outside the grid [origin, origin+(N-1)h]");
        }
    rc = ensure_capacity(s, s->total + n);
    if (rc) return rc;
    GroupInfo gi;
    gi.src = *mat;
    gi.off = s->total;
    gi.n = n;
    double h3 = s->cfg.h * s->cfg.h * s->cfg.h;
    gi.vol = particle_volume > 0 ? particle_volume : h3;   // V_p = occ h^3 / N_p = h^3 (MPMSimulator.cpp:263)
    gi.mass = mat->rho * gi.vol;                            // MPMSimulator.cpp:264
    gi.mat = make_matp(*mat, gi.mass, gi.vol);
    // host staging in plane order
    std::vector<float> host((size_t)n * PL_COUNT, 0.0f);
    auto plane = [&](int p) { return host.data() + (size_t)p * n; };
    for (int64_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            plane(PL_X + a)[i] = xyz[3 * i + a];
            plane(PL_V + a)[i] = vel ? vel[3 * i + a] : (float)(v0 ? v0[a] : 0.0);
        }
        plane(PL_F + 0)[i] = 1.0f; plane(PL_F + 4)[i] = 1.0f; plane(PL_F + 8)[i] = 1.0f;   // F = I (:20-26)
        plane(PL_S)[i] = 1.0f;                                                             // det Fp = 1, J = 1
    }
    for (int p = 0; p < PL_COUNT; ++p)
        MPM_HIP(hipMemcpyAsync(s->P + (size_t)p * s->cap + gi.off, plane(p), (size_t)n * sizeof(float),
                               hipMemcpyHostToDevice, s->stream));
    std::vector<uint32_t> ids((size_t)n);
    for (int64_t i = 0; i < n; ++i) ids[i] = (uint32_t)i;
    MPM_HIP(hipMemcpyAsync(s->pid + gi.off, ids.data(), (size_t)n * 4, hipMemcpyHostToDevice, s->stream));
    MPM_HIP(hipStreamSynchronize(s->stream));
    s->groups.push_back(gi);
    s->total += n;
    s->tau_valid = false;
    s->sorted_valid = false;
    s->primed = false;
    s->dt_valid = false;
    if (group) *group = (int32_t)(s->groups.size() - 1);
    return MPM_OK;
}

int mpm_add_object(mpm_sim *s, const mpm_object *obj, int32_t *group)
{
    if (!s || !obj) return set_error(MPM_ERR_INVALID, "null argument");
    std::vector<float> xyz;
    int rc = sample_object(s->cfg, *obj, xyz);
    if (rc) return rc;
    int64_t n = (int64_t)xyz.size() / 3;
    if (n == 0) return set_error(MPM_ERR_RANGE, "object produced no particles");
    double h3 = s->cfg.h * s->cfg.h * s->cfg.h;
    double vol = s->cfg.volume_mode == MPM_VOLUME_PPC ? h3 / std::max(1, s->cfg.ppc) : h3;
    return mpm_add_particles(s, &obj->material, n, xyz.data(), nullptr, obj->init_vel, vol, group);
}

int mpm_step_phase(mpm_sim *s, int32_t phase, double dt)
{
    if (!s) return set_error(MPM_ERR_INVALID, "sim is null");
    if (phase < MPM_PHASE_P2G || phase > MPM_PHASE_G2P) return set_error(MPM_ERR_INVALID, "unknown phase");
    float fdt = (float)(dt > 0 ? dt : s->cfg.dt);
    s->primed = false;
    // single phases always take the given dt; the controller re-reduces its
    // maxima at the next full step
    s->dt_valid = false;
    const float *dtp = s->g.dtp;
    s->g.dtp = nullptr;
    int rc = run_phase(s, phase, fdt);
    s->g.dtp = dtp;
    return rc;
}

// Step-size control at the start of a step: the maxima of the current state
// (reduced by the previous G2P, or scanned now) give the step's dt on the
// device.
static int dt_control(mpm_sim *s, double max_dt)
{
    int rc = MPM_OK;
    if (!s->dt_valid) {
        rc = check_launch(launch_zero_u32(s->stream, s->counters + CT_SPEED, 2), "zero");
        for (size_t k = 0; k < s->groups.size() && !rc; ++k)
            rc = check_launch(launch_dt_scan(s->stream, s->groups[k].mat, view_of(s, (int)k), s->counters), "dt_scan");
        if (rc) return rc;
    }
    rc = check_launch(launch_dt_control(s->stream, s->dtc, s->counters, (float)max_dt,
                                        (float)(s->cfg.cfl * s->cfg.h), s->cfg.dt_mode),
                      "dt_control");
    if (!rc) s->dt_valid = true;        // this step's G2P reduces the maxima of the state it writes
    return rc;
}

int mpm_step(mpm_sim *s, double dt)
{
    if (!s) return set_error(MPM_ERR_INVALID, "sim is null");
    double d = dt > 0 ? dt : s->cfg.dt;
    float fdt = (float)d;
    if (s->dtc) {
        int rc = dt_control(s, d);
        if (rc) return rc;
    }
    if (s->cfg.form == MPM_FORM_FUSED) {
        int rc = fused_step(s, fdt);
        if (rc) return rc;
    } else {
        for (int ph = MPM_PHASE_P2G; ph <= MPM_PHASE_G2P; ++ph) {
            int rc = run_phase(s, ph, fdt);
            if (rc) return rc;
        }
    }
    s->steps += 1;          // ++current_frame (MPMCUDA.cu:923)
    if (!s->dtc) {
        s->time += d;       // current_time += adp_dt (MPMCUDA.cu:924); with a controller the device adds it
        s->last_dt = d;
    }
    return MPM_OK;
}

int mpm_run(mpm_sim *s, int32_t nsteps, double dt)
{
    for (int k = 0; k < nsteps; ++k) {
        int rc = mpm_step(s, dt);
        if (rc) return rc;
    }
    return MPM_OK;
}

int64_t mpm_total_particles(const mpm_sim *s) { return s ? s->total : 0; }
int32_t mpm_group_count(const mpm_sim *s) { return s ? (int32_t)s->groups.size() : 0; }
int64_t mpm_group_size(const mpm_sim *s, int32_t g)
{
    if (!s || g < 0 || g >= (int32_t)s->groups.size()) return -1;
    return s->groups[g].n;
}

int mpm_synchronize(mpm_sim *s)
{
    if (!s) return set_error(MPM_ERR_INVALID, "sim is null");
    MPM_HIP(hipStreamSynchronize(s->stream));
    return MPM_OK;
}

int mpm_read_particles(mpm_sim *s, int32_t group, int32_t field, float *out)
{
    if (!s || !out) return set_error(MPM_ERR_INVALID, "null argument");
    if (group < 0 || group >= (int32_t)s->groups.size()) return set_error(MPM_ERR_INVALID, "bad group");
    int w = width_of(field), p0 = plane_of(field);
    if (w < 0) return set_error(MPM_ERR_INVALID, "unknown particle field");
    const GroupInfo &gi = s->groups[group];
    if (field == MPM_P_TAU && !s->tau_valid) {
        // stress of the current state, as the next P2G will use it
        for (size_t k = 0; k < s->groups.size(); ++k) {
            int rc = check_launch(launch_compute_tau(s->stream, s->groups[k].mat, view_of(s, (int)k),
                                                     s->cfg.svd_sweeps, s->counters), "compute_tau");
            if (rc) return rc;
        }
        s->tau_valid = true;
    }
    int64_t n = gi.n;
    std::vector<float> buf((size_t)n * w);
    std::vector<uint32_t> ids((size_t)n);
    for (int c = 0; c < w; ++c)
        MPM_HIP(hipMemcpyAsync(buf.data() + (size_t)c * n, s->P + (size_t)(p0 + c) * s->cap + gi.off,
                               (size_t)n * 4, hipMemcpyDeviceToHost, s->stream));
    MPM_HIP(hipMemcpyAsync(ids.data(), s->pid + gi.off, (size_t)n * 4, hipMemcpyDeviceToHost, s->stream));
    MPM_HIP(hipStreamSynchronize(s->stream));
    for (int64_t j = 0; j < n; ++j) {
        uint32_t i = ids[j];
        if (i >= (uint32_t)n) return set_error(MPM_ERR_STATE, "particle id table corrupted");
        for (int c = 0; c < w; ++c) out[(size_t)i * w + c] = buf[(size_t)c * n + j];
    }
    return MPM_OK;
}

int mpm_write_particles(mpm_sim *s, int32_t group, int32_t field, const float *in)
{
    if (!s || !in) return set_error(MPM_ERR_INVALID, "null argument");
    if (group < 0 || group >= (int32_t)s->groups.size()) return set_error(MPM_ERR_INVALID, "bad group");
    if (field == MPM_P_TAU) return set_error(MPM_ERR_INVALID, "TAU is derived state and read-only");
    int w = width_of(field), p0 = plane_of(field);
    if (w < 0) return set_error(MPM_ERR_INVALID, "unknown particle field");
    if (int rc = flush_spill(s)) return rc;   // a P2G phase's spill scatter reads the state it saw
    const GroupInfo &gi = s->groups[group];
    int64_t n = gi.n;
    if (field == MPM_P_POS) {
        const GridP &g = s->g;
        const float o[3] = {g.ox, g.oy, g.oz};
        const int nn[3] = {g.nx, g.ny, g.nz};
        for (int64_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) {
                float x = in[3 * i + a];
                if (!(x >= o[a] && x <= o[a] + g.h * (float)(nn[a] - 1)))
                    return set_error(MPM_ERR_RANGE, "particle outside the grid");
            }
    }
    std::vector<uint32_t> ids((size_t)n);
    MPM_HIP(hipMemcpyAsync(ids.data(), s->pid + gi.off, (size_t)n * 4, hipMemcpyDeviceToHost, s->stream));
    MPM_HIP(hipStreamSynchronize(s->stream));
    std::vector<float> buf((size_t)n * w);
    for (int64_t j = 0; j < n; ++j) {
        uint32_t i = ids[j];
        if (i >= (uint32_t)n) return set_error(MPM_ERR_STATE, "particle id table corrupted");
        for (int c = 0; c < w; ++c) buf[(size_t)c * n + j] = in[(size_t)i * w + c];
    }
    for (int c = 0; c < w; ++c)
        MPM_HIP(hipMemcpyAsync(s->P + (size_t)(p0 + c) * s->cap + gi.off, buf.data() + (size_t)c * n,
                               (size_t)n * 4, hipMemcpyHostToDevice, s->stream));
    MPM_HIP(hipStreamSynchronize(s->stream));
    if (field == MPM_P_F || field == MPM_P_JP) s->tau_valid = false;
    if (field == MPM_P_POS) s->sorted_valid = false;
    s->primed = false;
    s->dt_valid = false;
    return MPM_OK;
}

int mpm_get_positions(mpm_sim *s, float *out)
{
    if (!s || !out) return set_error(MPM_ERR_INVALID, "null argument");
    int64_t off = 0;
    for (size_t g = 0; g < s->groups.size(); ++g) {
        int rc = mpm_read_particles(s, (int32_t)g, MPM_P_POS, out + 3 * off);
        if (rc) return rc;
        off += s->groups[g].n;
    }
    return MPM_OK;
}

int mpm_read_grid(mpm_sim *s, int32_t field, float *out)
{
    if (!s || !out) return set_error(MPM_ERR_INVALID, "null argument");
    if (field < MPM_G_MASS || field > MPM_G_VELOCITY) return set_error(MPM_ERR_INVALID, "unknown grid field");
    if (int rc = flush_spill(s)) return rc;   // after a P2G phase: its spilled particles' share
    const GridP &g = s->g;
    size_t nslots = (size_t)s->nblocks * BLK_NODES;
    std::vector<float4> mp(nslots), gv(nslots);
    std::vector<uint32_t> cnt(CT_COUNT);
    MPM_HIP(hipMemcpyAsync(mp.data(), s->grid_mp, nslots * sizeof(float4), hipMemcpyDeviceToHost, s->stream));
    MPM_HIP(hipMemcpyAsync(gv.data(), s->grid_v, nslots * sizeof(float4), hipMemcpyDeviceToHost, s->stream));
    MPM_HIP(hipMemcpyAsync(cnt.data(), s->counters, CT_COUNT * 4, hipMemcpyDeviceToHost, s->stream));
    MPM_HIP(hipStreamSynchronize(s->stream));
    uint32_t nact = cnt[CT_ACTIVE];
    std::vector<uint32_t> lst(nact);
    if (nact) {
        MPM_HIP(hipMemcpy(lst.data(), s->list, (size_t)nact * 4, hipMemcpyDeviceToHost));
    }
    std::vector<char> active((size_t)s->nblocks, 0);
    for (uint32_t k = 0; k < nact; ++k) if (lst[k] < (uint32_t)s->nblocks) active[lst[k]] = 1;
    // grid_mp holds (m, p) between P2G and the grid update; grid_v after it
    const bool after_p2g = s->last_phase == MPM_PHASE_P2G;
    if (after_p2g && s->cfg.form != MPM_FORM_PLAIN && s->blk_tile) {
        // FAST form: add the staged block arenas of the last P2G
        std::vector<uint32_t> tcount(s->groups.size());
        MPM_HIP(hipMemcpy(tcount.data(), s->tile_counts, tcount.size() * 4, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < s->groups.size(); ++k) {
            uint32_t nt = std::min<uint32_t>(tcount[k], (uint32_t)std::min<int64_t>(s->nblocks, s->groups[k].n));
            if (!nt) continue;
            std::vector<uint4> desc(nt);
            MPM_HIP(hipMemcpy(desc.data(), (const uint4 *)s->tile_desc + s->group_tile_off[k], nt * sizeof(uint4),
                              hipMemcpyDeviceToHost));
            std::vector<int32_t> entry((size_t)s->nblocks);
            MPM_HIP(hipMemcpy(entry.data(), s->blk_tile + (int64_t)k * s->nblocks, entry.size() * sizeof(int32_t),
                              hipMemcpyDeviceToHost));
            std::vector<float4> arena(ARENA_NODES);
            for (uint32_t t = 0; t < nt; ++t) {
                if (desc[t].w >= (uint32_t)s->staging_cap) continue;
                MPM_HIP(hipMemcpy(arena.data(), s->staging + (int64_t)desc[t].w * ARENA_NODES,
                                  ARENA_NODES * sizeof(float4), hipMemcpyDeviceToHost));
                const int b = (int)desc[t].z;
                const int bx = b % g.nbx, by = (b / g.nbx) % g.nby, bz = b / (g.nbx * g.nby);
                // arena faces the P2G wrote (the rest of the slot's face nodes is stale)
                const uint32_t f = (uint32_t)entry[b] >> TILE_FACE_SHIFT;
                auto face_ok = [&](int i, int a) { return (i != 0 || (f >> (2 * a)) & 1u) && (i != 7 || (f >> (2 * a + 1)) & 1u); };
                for (int q = 0; q < ARENA_NODES; ++q) {
                    const int x = bx * BLK - 1 + (q & 7), y = by * BLK - 1 + ((q >> 3) & 7), z = bz * BLK - 1 + (q >> 6);
                    if (x < 0 || y < 0 || z < 0 || x >= g.nx || y >= g.ny || z >= g.nz) continue;
                    if (!face_ok(q & 7, 0) || !face_ok((q >> 3) & 7, 1) || !face_ok(q >> 6, 2)) continue;
                    float4 &dst = mp[node_slot(g, x, y, z)];
                    dst.x += arena[q].x; dst.y += arena[q].y; dst.z += arena[q].z; dst.w += arena[q].w;
                }
            }
        }
    }
    int w = field == MPM_G_MASS ? 1 : 3;
    for (int z = 0; z < g.nz; ++z)
        for (int y = 0; y < g.ny; ++y)
            for (int x = 0; x < g.nx; ++x) {
                int64_t slot = node_slot(g, x, y, z);
                int64_t blk = slot / BLK_NODES;
                size_t o = ((size_t)z * g.ny + y) * g.nx + x;
                float4 a = mp[slot];
                float4 b = active[blk] ? gv[slot] : make_float4(0, 0, 0, 0);
                if (field == MPM_G_MASS) out[o] = after_p2g ? a.x : b.w;
                else if (field == MPM_G_MOMENTUM) { out[3 * o] = a.y; out[3 * o + 1] = a.z; out[3 * o + 2] = a.w; }
                else { out[3 * o] = b.x; out[3 * o + 1] = b.y; out[3 * o + 2] = b.z; }
            }
    (void)w;
    return MPM_OK;
}

int mpm_health(mpm_sim *s, uint32_t *bits)
{
    if (!s || !bits) return set_error(MPM_ERR_INVALID, "null argument");
    uint32_t v = 0;
    MPM_HIP(hipMemcpyAsync(&v, s->counters + CT_HEALTH, 4, hipMemcpyDeviceToHost, s->stream));
    MPM_HIP(hipStreamSynchronize(s->stream));
    int rc = check_launch(launch_zero_u32(s->stream, s->counters + CT_HEALTH, 1), "zero");
    if (rc) return rc;
    MPM_HIP(hipStreamSynchronize(s->stream));
    *bits = v;
    return MPM_OK;
}

int mpm_stats(mpm_sim *s, double st[8])
{
    if (!s || !st) return set_error(MPM_ERR_INVALID, "null argument");
    uint32_t c[CT_COUNT];
    MPM_HIP(hipMemcpyAsync(c, s->counters, sizeof c, hipMemcpyDeviceToHost, s->stream));
    MPM_HIP(hipStreamSynchronize(s->stream));
    st[0] = (double)s->steps;
    st[1] = c[CT_ACTIVE];
    st[2] = c[CT_FALLBACK];
    st[3] = c[CT_REJECT];
    st[4] = (double)s->sorts;
    st[5] = mpm_current_time(s);
    st[6] = c[CT_TILES];
    st[7] = s->cfg.form == MPM_FORM_FUSED ? c[CT_SPILL] : c[CT_SPILLED];
    return MPM_OK;
}

int mpm_svd_stats(mpm_sim *s, double out[4])
{
    if (!s || !out) return set_error(MPM_ERR_INVALID, "null argument");
    uint32_t c[CT_COUNT];
    MPM_HIP(hipMemcpyAsync(c, s->counters, sizeof c, hipMemcpyDeviceToHost, s->stream));
    MPM_HIP(hipStreamSynchronize(s->stream));
    out[0] = c[CT_SVD];
    out[1] = c[CT_WARM];
    out[2] = c[CT_REJECT];
    out[3] = c[CT_FALLBACK];
    return MPM_OK;
}

int mpm_set_timing(mpm_sim *s, int32_t enable)
{
    if (!s) return set_error(MPM_ERR_INVALID, "sim is null");
    s->timing = enable != 0;
    return MPM_OK;
}

// Elapsed times of the recorded event pairs into the per-id accumulators.
static int drain_timed(mpm_sim *s)
{
    MPM_HIP(hipStreamSynchronize(s->stream));
    for (auto &t : s->timed) {
        float e = 0.f;
        MPM_HIP(hipEventElapsedTime(&e, t.a, t.b));
        if (t.phase >= 0 && t.phase < mpm_sim::TIMED_IDS) { s->tacc_ms[t.phase] += e; s->tacc_n[t.phase] += 1; }
        s->event_pool.push_back(t.a);
        s->event_pool.push_back(t.b);
    }
    s->timed.clear();
    return MPM_OK;
}

int mpm_phase_times(mpm_sim *s, double ms[4], int64_t counts[4], int64_t *stats_nodes)
{
    if (!s || !ms || !counts) return set_error(MPM_ERR_INVALID, "null argument");
    int rc = drain_timed(s);
    if (rc) return rc;
    for (int k = 0; k < 4; ++k) {
        ms[k] = s->tacc_ms[k]; counts[k] = s->tacc_n[k];
        s->tacc_ms[k] = 0.0; s->tacc_n[k] = 0;
    }
    if (stats_nodes) {
        uint32_t c[CT_COUNT];
        MPM_HIP(hipMemcpy(c, s->counters, sizeof c, hipMemcpyDeviceToHost));
        *stats_nodes = c[CT_NODES];
    }
    return MPM_OK;
}

int mpm_kernel_times(mpm_sim *s, double ms[4], int64_t counts[4])
{
    if (!s || !ms || !counts) return set_error(MPM_ERR_INVALID, "null argument");
    int rc = drain_timed(s);
    if (rc) return rc;
    for (int k = 0; k < 4; ++k) {
        ms[k] = s->tacc_ms[4 + k]; counts[k] = s->tacc_n[4 + k];
        s->tacc_ms[4 + k] = 0.0; s->tacc_n[4 + k] = 0;
    }
    return MPM_OK;
}

// With a controller the clock lives on the device; reading it synchronizes
// the stream (the handle's host copy is refreshed).
static bool read_dt_state(const mpm_sim *s, DtCtl *out)
{
    if (!s->dtc) return false;
    if (hipMemcpyAsync(out, s->dtc, sizeof *out, hipMemcpyDeviceToHost, s->stream) != hipSuccess) return false;
    return hipStreamSynchronize(s->stream) == hipSuccess;
}

double mpm_current_time(const mpm_sim *s)
{
    if (!s) return 0.0;
    DtCtl c;
    if (read_dt_state(s, &c)) const_cast<mpm_sim *>(s)->time = c.time;
    return s->time;
}

double mpm_last_dt(const mpm_sim *s)
{
    if (!s) return 0.0;
    DtCtl c;
    if (read_dt_state(s, &c)) const_cast<mpm_sim *>(s)->last_dt = c.last;
    return s->last_dt;
}
int64_t mpm_current_frame(const mpm_sim *s) { return s ? s->steps : 0; }

int mpm_write_bgeo(const char *path, const float *xyz, int64_t n)
{
    if (!path || (n > 0 && !xyz) || n < 0) return set_error(MPM_ERR_INVALID, "bad argument");
    return write_bgeo(path, xyz, n);
}

extern "C" int mpm_mlp_param_count(void);

int mpm_probe_point_in_mesh(const char *obj_path, const double *pts, int64_t n, uint8_t *inside)
{
    if (!obj_path || n < 0 || (n > 0 && (!pts || !inside))) return set_error(MPM_ERR_INVALID, "bad argument");
    return mpm::points_in_mesh(obj_path, pts, n, inside);
}

int mpm_set_warm_start_weights(mpm_sim *s, const float *w, int64_t count)
{
    if (!s || !w || count <= 0) return set_error(MPM_ERR_INVALID, "bad argument");
    if (count != mpm_mlp_param_count()) return set_error(MPM_ERR_INVALID, "warm-start weights: wrong parameter count");
    if (s->mlp_w) { (void)hipFree(s->mlp_w); s->mlp_w = nullptr; }
    MPM_HIP(hipMalloc(&s->mlp_w, (size_t)count * 4));
    MPM_HIP(hipMemcpy(s->mlp_w, w, (size_t)count * 4, hipMemcpyHostToDevice));
    s->mlp_count = count;
    s->mlp_host.assign(w, w + count);
    s->mlp_version += 1;
    return MPM_OK;
}

void mpm_destroy(mpm_sim *s)
{
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    void *ptrs[] = {s->grid_mp, s->grid_v, s->flags, s->list, s->counters, s->P, s->P_alt, s->pid,
                    s->pid_alt, s->keys, s->keys_alt, s->vals, s->vals_alt, s->sort_hist, s->sort_scan,
                    s->blk_count, s->blk_start, s->tile_desc, s->scan_tmp, s->tile_counts, s->mlp_w,
                    s->staging, s->blk_tile, s->spill, s->dtc, s->warm_list, s->warm_cnt, s->g2p_perm};
    for (void *p : ptrs) if (p) (void)hipFree(p);
    if (s->spill_host) (void)hipHostFree(s->spill_host);
    for (auto &t : s->timed) { (void)hipEventDestroy(t.a); (void)hipEventDestroy(t.b); }
    for (hipEvent_t e : s->event_pool) (void)hipEventDestroy(e);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

}  // extern "C"

namespace mpm {

// Event pair around a phase when timing is on.

hipEvent_t take_event(mpm_sim *s)
{
    hipEvent_t e = nullptr;
    if (!s->event_pool.empty()) {
        e = s->event_pool.back();
        s->event_pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        e = nullptr;
    }
    return e;
}

int launch_mlp_infer_list(hipStream_t st, const float *params, const PartView &pv, const uint32_t *list,
                          const uint32_t *count, uint32_t *next_count);

// The particle list of the warm start (and its per-group counts, cleared),
// sized for the particle capacity.
int ensure_warm_lists(mpm_sim *s)
{
    if (s->warm_list && s->warm_list_cap >= s->cap) return MPM_OK;
    if (s->warm_list) (void)hipFree(s->warm_list);
    s->warm_list = nullptr;
    s->warm_list_cap = 0;
    MPM_HIP(hipMalloc(&s->warm_list, (size_t)s->cap * sizeof(uint32_t)));
    if (!s->warm_cnt) MPM_HIP(hipMalloc(&s->warm_cnt, 2 * MAX_GROUPS * sizeof(uint32_t)));
    s->warm_list_cap = s->cap;
    return check_launch(launch_zero_u32(s->stream, s->warm_cnt, 2 * MAX_GROUPS), "zero");
}

// Warm start of the G2P SVDs: quaternions from the network for the particles
// whose last G2P took the SVD (snow and jello groups), the list the coming
// G2P fills, and the sweep count for the G2P kernels.  A particle entering
// the SVD path for the first time has no fresh quaternion and is listed for
// the next step; between steps its planes keep the last one inferred.
static int prepare_warm(mpm_sim *s)
{
    const bool on = s->cfg.warm_start == MPM_WARM_MLP && s->mlp_w && s->mlp_count == mpm_mlp_param_count();
    int rc = on ? ensure_warm_lists(s) : MPM_OK;
    if (rc) return rc;
    const int p = s->warm_par;
    for (size_t k = 0; k < s->groups.size(); ++k) {
        MatP &m = s->groups[k].mat;
        m.warm_sweeps = (on && m.type != MPM_WATER) ? std::max(1, s->cfg.warm_sweeps) : 0;
        m.wlist = nullptr;
        m.wcount = nullptr;
        if (m.warm_sweeps > 0) {
            const int64_t off = s->groups[k].off;
            uint32_t *cnt = s->warm_cnt + 2 * k;
            rc = check_launch(launch_mlp_infer_list(s->stream, s->mlp_w, view_of(s, (int)k), s->warm_list + off,
                                                    cnt + p, cnt + (p ^ 1)),
                              "mlp_infer_list");
            if (rc) return rc;
            m.wlist = s->warm_list + off;
            m.wcount = cnt + (p ^ 1);
        }
    }
    if (on) s->warm_par ^= 1;
    return MPM_OK;
}

// One phase of MPMSimulator::step (MPMCUDA.cu:793-925).
static int run_phase(mpm_sim *s, int phase, float dt)
{
    const bool fast = s->cfg.form != MPM_FORM_PLAIN;
    const int thr = s->cfg.threads;
    int rc = MPM_OK;
    s->last_phase = phase;
    if (phase == MPM_PHASE_P2G) {
        // FAST form: the last block compaction cleared the spill counter and
        // this P2G's spill kernels clear the active-block and node counters
        const bool clean = s->cfg.form == MPM_FORM_FAST && s->counters_clean && s->total > 0;
        s->counters_clean = false;
        if ((rc = flush_spill(s))) return rc;
        if (!clean) {
            rc = check_launch(launch_zero_u32(s->stream, s->counters + CT_ACTIVE, 3), "zero");
            if (rc) return rc;
        }
        if (fast) {
            {
                PhaseTimer t(s, 0);
                rc = fast_prepare(s);
            }
            if (rc) return rc;
            PhaseTimer t(s, 1);
            for (size_t k = 0; k < s->groups.size() && !rc; ++k) rc = launch_p2g_tiles(s, (int)k, dt);
        } else {
            PhaseTimer t(s, 1), tk(s, TK_P2G);
            for (size_t k = 0; k < s->groups.size() && !rc; ++k)
                rc = check_launch(launch_p2g_atomic(s->stream, s->g, s->groups[k].mat, view_of(s, (int)k), dt,
                                                    s->cfg.svd_sweeps, 0, s->grid_mp, s->flags, s->counters, thr),
                                  "p2g");
        }
        return rc;
    }
    if (phase == MPM_PHASE_GRID) {
        {
            PhaseTimer t(s, 2);
            const bool fused = s->spill_pending;
            rc = fused ? spill_compact(s)
                       : check_launch(launch_compact(s->stream, s->flags, (int)s->nblocks, s->list, s->counters),
                                      "compact");
            if (rc) return rc;
            s->counters_clean = true;
            const bool staged = fast && s->blk_tile;
            PhaseTimer tk(s, TK_GRID);
            rc = check_launch(launch_grid_update(s->stream, s->g, dt, s->list, s->counters, s->grid_mp, s->grid_v,
                                                 (int)s->nblocks, staged ? s->staging : nullptr,
                                                 staged ? s->blk_tile : nullptr, (int)s->groups.size(),
                                                 (uint32_t)s->staging_cap, fused ? 1 : 0),
                              "grid_update");
            if (rc) return rc;
        }
        if (s->timing) rc = check_launch(launch_count_nodes(s->stream, s->list, s->counters, s->grid_v), "count");
        return rc;
    }
    // G2P; a due re-sort of the FAST form is applied by this G2P
    if (fast) {
        PhaseTimer t(s, 0);
        rc = sort_for_g2p(s);
        if (rc) return rc;
    }
    rc = prepare_warm(s);
    if (rc) {
        (void)finish_g2p_move(s, false);
        return rc;
    }
    {
        PhaseTimer t(s, 3), tk(s, TK_G2P);
        for (size_t k = 0; k < s->groups.size() && !rc; ++k) {
            if (fast) rc = launch_g2p_fast(s, (int)k, dt);
            else
                rc = check_launch(launch_g2p(s->stream, s->g, s->groups[k].mat, view_of(s, (int)k), dt, s->grid_v,
                                             0, s->cfg.svd_sweeps, s->counters, thr),
                                  "g2p");
        }
    }
    if (rc) return rc;          // a failed launch leaves the device state undefined either way
    if ((rc = finish_g2p_move(s, true))) return rc;
    s->tau_valid = fast;
    if (fast) s->steps_since_sort += 1;
    return rc;
}

// FUSED form step: grid update of the primed scatter, periodic re-sort, then
// G2P of this step and P2G of the next in one pass (MPMCUDA.cu:793-925 with
// the P2G of step n+1 moved into the G2P kernel of step n).  The first step,
// and any step after the state was rewritten or dt changed, primes the
// scatter with the FAST P2G.
int sort_if_due(mpm_sim *s);

static int fused_step(mpm_sim *s, float dt)
{
    int rc = MPM_OK;
    if (!s->primed || s->primed_dt != dt) {
        rc = run_phase(s, MPM_PHASE_P2G, dt);
        if (rc) return rc;
    } else {
        rc = check_launch(launch_zero_u32(s->stream, s->counters + CT_ACTIVE, 2), "zero");
        if (rc) return rc;
    }
    rc = run_phase(s, MPM_PHASE_GRID, dt);
    if (rc) return rc;
    {
        PhaseTimer t(s, 0);
        rc = sort_if_due(s);
    }
    if (rc) return rc;
    s->last_phase = MPM_PHASE_G2P;
    rc = prepare_warm(s);
    if (rc) return rc;
    {
        PhaseTimer t(s, 3);
        for (size_t k = 0; k < s->groups.size() && !rc; ++k) rc = launch_g2p2g(s, (int)k, dt, dt);
    }
    if (rc) return rc;
    s->primed = true;
    s->primed_dt = dt;
    s->tau_valid = false;           // the stress stayed in registers
    s->steps_since_sort += 1;
    return MPM_OK;
}

}  // namespace mpm
