"""The step-size controller on the device (dt_mode CFL / CFL_DEFORM) against
the oracle's restatement (oracle.controller_dt, pinned by closed forms in
test_oracle_dt.py): the dt the device picks from its own FP32 state, every
step, for every material in the PLAIN and FAST forms; rollouts whose dt
sequence comes from the device against the FP64 oracle stepped with the
same sequence (so the kernels must use the controller's dt for the force,
gravity and advection); the clock; the rescan after the host rewrites the
state; and single phases, which keep the dt they are given."""
import numpy as np
import pytest

import oracle as O
from mpm_cuda_amd import capi
from tests.scenes import Pair, rel

pytestmark = pytest.mark.gpu

MATS = [capi.JELLO, capi.SNOW, capi.WATER]
MAT_ID = {capi.JELLO: "jello", capi.SNOW: "snow", capi.WATER: "water"}
FORM_ID = {capi.FORM_PLAIN: "plain", capi.FORM_FAST: "fast"}
MODE_ID = {capi.DT_CFL: "cfl", capi.DT_CFL_DEFORM: "cfl_deform"}
CFL = 0.5
MAX_DT = 4e-4


def _cfg(mode):
    return {"dt_mode": mode, "cfl": CFL}


def device_group(p):
    """The device's current FP32 state as an oracle group (same material)."""
    g = p.gpu
    n = p.n
    m = p.orc.groups[0].mat
    S = g.read(0, capi.P_JP)[:, 0].astype(np.float64)
    kw = {}
    if p.type == capi.SNOW:
        kw["Fp"] = np.cbrt(S)[:, None, None] * np.eye(3)
    if p.type == capi.WATER:
        kw["J"] = S
    return O.Group(g.read(0, capi.P_POS).astype(np.float64), g.read(0, capi.P_VEL).astype(np.float64),
                   type=p.type, mass=m.mass, vol=m.vol, mu=m.mu, lam=m.lam, xi=m.xi, eos_k=m.eos_k,
                   eos_gamma=m.eos_gamma, Fe=g.read(0, capi.P_F).astype(np.float64).reshape(n, 3, 3), **kw)


def test_fixed_mode_keeps_the_given_dt(gpu):
    p = Pair(type_=capi.JELLO, form=capi.FORM_FAST, seed=3)
    try:
        for _ in range(3):
            p.gpu.step(p.dt)
        assert p.gpu.last_dt == p.dt
        assert abs(p.gpu.current_time - 3 * p.dt) < 1e-15
    finally:
        p.close()


@pytest.mark.parametrize("form", [capi.FORM_PLAIN, capi.FORM_FAST], ids=["plain", "fast"])
def test_uniform_motion_closed_form(gpu, form):
    """Jello translating at (0.3, -1, 0.5) m/s, F = I, no gravity: every step
    takes cfl h / (|v| + c_P) with c_P the P-wave speed."""
    p = Pair(type_=capi.JELLO, form=form, seed=4, randomize=False, gravity=(0.0, 0.0, 0.0), dt=1e-2,
             extra_cfg=_cfg(capi.DT_CFL))
    try:
        E, nu, rho = 5e4, 0.3, 1e3
        c = np.sqrt(E * (1 - nu) / ((1 + nu) * (1 - 2 * nu)) / rho)
        want = CFL * p.h / (np.linalg.norm([0.3, -1.0, 0.5]) + c)
        dts = []
        for _ in range(5):
            p.gpu.step(1e-2)
            dts.append(p.gpu.last_dt)
        print("dt", dts, "closed form", want)
        assert np.all(np.abs(np.array(dts) / want - 1) < 2e-6), dts
        assert abs(p.gpu.current_time - sum(dts)) < 1e-12 * sum(dts)
        assert p.gpu.current_frame == 5
    finally:
        p.close()


CASES = [(m, f, d) for m in MATS for f in (capi.FORM_PLAIN, capi.FORM_FAST) for d in (capi.DT_CFL, capi.DT_CFL_DEFORM)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{MAT_ID[c[0]]}-{FORM_ID[c[1]]}-{MODE_ID[c[2]]}")
def test_dt_follows_the_state(gpu, case):
    """Before every step the oracle's controller, applied to the device's own
    FP32 state, predicts the dt the device then takes (first step: the scan
    kernel; later steps: the maxima the previous G2P reduced)."""
    mat, form, mode = case
    p = Pair(type_=mat, form=form, seed=31, dt=MAX_DT, extra_cfg=_cfg(mode))
    try:
        for k in range(6):
            want = O.controller_dt([device_group(p)], p.h, MAX_DT, CFL, mode)
            p.gpu.step(MAX_DT)
            got = p.gpu.last_dt
            print(MAT_ID[mat], FORM_ID[form], MODE_ID[mode], k, f"dt {got:.6e} want {want:.6e}")
            assert abs(got / want - 1) < 5e-6, (k, got, want)
            assert got <= MAX_DT
        assert p.gpu.health() & ~capi.H_ARENA_SPILL == 0
    finally:
        p.close()


@pytest.mark.parametrize("mat", MATS, ids=[MAT_ID[m] for m in MATS])
def test_adaptive_rollout_matches_oracle(gpu, mat):
    """30 FAST steps under CFL_DEFORM: the oracle, stepped with the device's
    dt sequence, agrees to the fixed-dt rollout bars, and its own controller
    on its FP64 state picks the same dt."""
    p = Pair(type_=mat, form=capi.FORM_FAST, seed=5, dt=MAX_DT, extra_cfg=_cfg(capi.DT_CFL_DEFORM))
    try:
        t = 0.0
        for k in range(30):
            want = O.controller_dt(p.orc.groups, p.h, MAX_DT, CFL, capi.DT_CFL_DEFORM)
            p.gpu.step(MAX_DT)
            dt = p.gpu.last_dt
            assert abs(dt / want - 1) < 1e-4, (k, dt, want)
            assert p.orc.step(dt) == 0
            t += dt
        g, o = p.gpu_particles(), p.orc_particles()
        ex, ev = rel(g["x"], o["x"]), rel(g["v"], o["v"])
        eF = rel(g["F"], o["F"]) if mat != capi.WATER else 0.0
        eS = float(np.abs(g["S"] - o["S"]).max()) if mat != capi.JELLO else 0.0
        print(MAT_ID[mat], f"t {t:.4e} x {ex:.2e} v {ev:.2e} F {eF:.2e} S {eS:.2e}")
        assert ex < 5e-6 and ev < 1e-4 and eF < 1e-5 and eS < 1e-5
        assert abs(p.gpu.current_time - t) < 1e-12 * t
    finally:
        p.close()


def test_rewritten_state_is_rescanned(gpu):
    """After the host writes velocities, the next step's dt comes from the
    new state, not from the maxima of the previous G2P."""
    p = Pair(type_=capi.SNOW, form=capi.FORM_FAST, seed=6, dt=MAX_DT, extra_cfg=_cfg(capi.DT_CFL))
    try:
        p.gpu.step(MAX_DT)
        p.gpu.step(MAX_DT)
        v = p.gpu.read(0, capi.P_VEL)
        v[7] = (40.0, 0.0, 0.0)
        p.gpu.write(0, capi.P_VEL, v)
        want = O.controller_dt([device_group(p)], p.h, MAX_DT, CFL, capi.DT_CFL)
        p.gpu.step(MAX_DT)
        got = p.gpu.last_dt
        print("dt after rewrite", got, "want", want)
        assert abs(got / want - 1) < 5e-6
    finally:
        p.close()


def test_phases_take_the_given_dt(gpu):
    """mpm_step_phase ignores the controller: the three phases at dt = 1e-5
    match the oracle's phases at 1e-5; the next full step rescans."""
    p = Pair(type_=capi.JELLO, form=capi.FORM_FAST, seed=8, dt=MAX_DT, extra_cfg=_cfg(capi.DT_CFL))
    try:
        dt = 1e-5
        p.gpu.step(MAX_DT)
        dt0 = p.gpu.last_dt
        assert p.orc.step(dt0) == 0
        for ph in (capi.PHASE_P2G, capi.PHASE_GRID, capi.PHASE_G2P):
            p.gpu.phase(ph, dt)
        p.orc.clear_grid()
        assert p.orc.p2g(dt) == 0
        p.orc.grid_update(dt)
        assert p.orc.g2p(dt) == 0
        g, o = p.gpu_particles(), p.orc_particles()
        ex, ev = rel(g["x"], o["x"]), rel(g["v"], o["v"])
        print(f"after phases x {ex:.2e} v {ev:.2e}")
        assert ex < 1e-6 and ev < 1e-4
        assert abs(p.gpu.current_time - dt0) < 1e-15        # phases do not advance the clock
        want = O.controller_dt([device_group(p)], p.h, MAX_DT, CFL, capi.DT_CFL)
        p.gpu.step(MAX_DT)
        assert abs(p.gpu.last_dt / want - 1) < 5e-6
    finally:
        p.close()


def test_batch_refuses_a_controlled_scene(gpu):
    a = Pair(type_=capi.SNOW, form=capi.FORM_FAST, seed=9)
    b = Pair(type_=capi.SNOW, form=capi.FORM_FAST, seed=10, extra_cfg=_cfg(capi.DT_CFL))
    try:
        with pytest.raises(capi.MpmError, match="dt_mode"):
            capi.Batch([a.gpu, b.gpu])
    finally:
        a.close()
        b.close()


def test_deferred_spill_scatter_takes_the_phase_dt(gpu):
    """A one-group FAST P2G leaves its spilled particles' atomic scatter to
    the next grid phase, or to a grid read / particle write before it.  With
    the controller on, that scatter must use the dt the P2G phase was given
    (phases ignore the controller), not the controller's device dt: the grid
    read right after each P2G phase equals the plain form's at that dt, on a
    fresh simulator (controller never run, device dt 0) and after full
    controlled steps, on steps where particles spilled."""
    kw = dict(type_=capi.JELLO, seed=17, randomize=False, dt=1e-4, gravity=(0.0, 0.0, 0.0))
    a = Pair(form=capi.FORM_PLAIN, **kw)
    b = Pair(form=capi.FORM_FAST, sort_every=1000, extra_cfg=_cfg(capi.DT_CFL), **kw)
    try:
        rng = np.random.default_rng(18)
        v = np.tile(np.array([[20.0, 5.0, -3.0]], np.float32), (a.n, 1))
        C = (30.0 * rng.normal(size=(a.n, 9))).astype(np.float32)
        F = (np.eye(3).reshape(1, 9) + 0.05 * rng.normal(size=(a.n, 9))).astype(np.float32)
        for p in (a, b):
            p.gpu.write(0, capi.P_VEL, v)
            p.gpu.write(0, capi.P_C, C)
            p.gpu.write(0, capi.P_F, F)
        spilled = 0
        for k in range(24):
            if k == 12:              # full controlled steps in between: the device dt is now set
                b.gpu.step(MAX_DT)
                a.gpu.step(b.gpu.last_dt)
            dt = 1e-4 if k % 2 else 5e-5
            for p in (a, b):
                p.gpu.phase(capi.PHASE_P2G, dt)
            for f in (capi.G_MASS, capi.G_MOMENTUM):
                e = rel(b.gpu.grid(f), a.gpu.grid(f))
                assert e < 1e-5, (k, f, e)
            for p in (a, b):
                p.gpu.phase(capi.PHASE_GRID, dt)
                p.gpu.phase(capi.PHASE_G2P, dt)
            spilled += b.gpu.stats()["spilled"] > 0
        print("steps with spills:", spilled)
        assert spilled >= 1
    finally:
        a.close()
        b.close()
