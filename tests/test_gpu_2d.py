"""The device 2D snow step (mpm2d_*, the reference's MPMSnow.py substep)
against the FP64 2D oracle (oracle/mpm_oracle2d.c) on the same FP32 initial
state: the grid and every particle field after one step, rollouts, graph
replays against plain launches, and the full 2000-step workload."""
import numpy as np
import pytest

import oracle as O
from mpm_cuda_amd import capi, workloads
from tests.scenes import rel

pytestmark = pytest.mark.gpu


def _state(seed, n=9000, moving=True):
    rng = np.random.default_rng(seed)
    x = np.concatenate([rng.random((n // 3, 2)) * 0.2 + [0.3 + 0.1 * g, 0.05 + 0.32 * g] for g in range(3)])
    v = (2.0 * rng.normal(size=x.shape)) if moving else np.zeros_like(x)
    return x.astype(np.float32), v.astype(np.float32)


def _pair(x, v, **kw):
    g = capi.Sim2D(**kw)
    g.add_particles(x, v)
    c = g.cfg
    o = O.Sim2D(x.astype(np.float64), v.astype(np.float64), n_grid=c.n_grid, wall=c.wall, E=c.E, nu=c.nu,
                rho=c.rho, xi=c.xi, theta_c=c.theta_c, theta_s=c.theta_s, gravity=tuple(c.gravity))
    return g, o


def _fields(g, o):
    gf = {"x": g.read(capi.P2_POS), "v": g.read(capi.P2_VEL), "C": g.read(capi.P2_C), "F": g.read(capi.P2_F),
          "J_p": g.read(capi.P2_JP)[:, 0]}
    of = {"x": o.x, "v": o.v, "C": o.C.reshape(-1, 4), "F": o.Fe.reshape(-1, 4), "J_p": np.linalg.det(o.Fp)}
    return gf, of


def test_one_step_grid_and_particles(gpu):
    x, v = _state(1)
    g, o = _pair(x, v, graph_steps=0)
    try:
        g.run(1)
        assert o.step(1e-4) == 0
        m, gv = g.grid()
        assert rel(m, o.grid_m) < 1e-6
        # mass-weighted grid velocity (empty nodes carry none)
        assert rel(m[:, None] * gv, o.grid_m[:, None] * o.grid_v.reshape(-1, 2)) < 1e-5
        gf, of = _fields(g, o)
        errs = {k: rel(gf[k], of[k]) for k in gf}
        print("2D one step", {k: f"{e:.1e}" for k, e in errs.items()})
        for k, e in errs.items():
            assert e < 1e-5, (k, e)
        assert g.health() == 0
    finally:
        g.close()


@pytest.mark.parametrize("moving", [False, True], ids=["rest", "random_v"])
def test_rollout_matches_oracle(gpu, moving):
    """200 steps (the squares fall and the lowest lands on the floor band):
    x within 5e-6, v within 1e-4, F within 1e-5 normwise; J_p within 1e-5
    normwise from rest and 5e-5 with random 2 m/s particle velocities, where
    most particles sit on the plastic bound every step and J_p is a product
    of 200 FP32 ratios of singular values.  (The largest single J_p
    difference is printed, not bounded: a singular value on the clamp bound
    may be clamped in FP32 and not in FP64.)"""
    x, v = _state(2, moving=moving)
    g, o = _pair(x, v)
    try:
        g.run(200)
        for _ in range(200):
            assert o.step(1e-4) == 0
        gf, of = _fields(g, o)
        ex, ev, eF = rel(gf["x"], of["x"]), rel(gf["v"], of["v"]), rel(gf["F"], of["F"])
        eJ, mJ = rel(gf["J_p"], of["J_p"]), float(np.abs(gf["J_p"] - of["J_p"]).max())
        print("2D rollout", moving, f"x {ex:.1e} v {ev:.1e} F {eF:.1e} J_p {eJ:.1e} (max {mJ:.1e})")
        assert ex < 5e-6 and ev < 1e-4 and eF < 1e-5 and eJ < (5e-5 if moving else 1e-5)
        assert g.health() == 0
    finally:
        g.close()


def test_graph_replay_matches_plain_launches(gpu):
    """Replays of the captured 50-step graph, with reads between runs, equal
    the same steps launched one by one, to the parity bar: the P2G float
    atomics make any two runs differ at the rounding level."""
    x, v = _state(3, moving=False)
    a = capi.Sim2D(graph_steps=50)
    b = capi.Sim2D(graph_steps=0)
    try:
        a.add_particles(x, v)
        b.add_particles(x, v)
        for n in (120, 50, 73):
            a.run(n)
            b.run(n)
            for f in (capi.P2_POS, capi.P2_VEL, capi.P2_F):
                assert rel(a.read(f), b.read(f)) < 1e-5
        assert a.steps == b.steps == 243
    finally:
        a.close()
        b.close()


def test_full_workload_2000_steps(gpu):
    """The 9,000-particle workload for its 2000 steps: finite, inside the
    domain, positive J_p, and the grid mass equals the particle mass."""
    sim, info = workloads.snow2d_squares()
    try:
        sim.run(info["steps"])
        x = sim.read(capi.P2_POS)
        J = sim.read(capi.P2_JP)[:, 0]
        assert sim.health() == 0
        assert np.all(np.isfinite(x)) and x.min() > 0.0 and x.max() < 1.0
        assert np.all(J > 0)
        m, _ = sim.grid()
        p_mass = (0.5 / 128) ** 2
        assert abs(m.sum() - len(x) * p_mass) < 1e-4 * len(x) * p_mass
        assert x[:, 1].min() < 0.1            # the squares have fallen to the floor band
    finally:
        sim.close()


def test_errors(gpu):
    with pytest.raises(capi.MpmError, match="wall"):
        capi.Sim2D(n_grid=16, wall=7)
    s = capi.Sim2D()
    try:
        with pytest.raises(capi.MpmError, match="no particles"):
            s.run(1)
        with pytest.raises(capi.MpmError, match="outside"):
            s.add_particles(np.array([[0.5, 0.999]], np.float32))
        s.add_particles(np.array([[0.5, 0.5]], np.float32))
        with pytest.raises(capi.MpmError, match="already"):
            s.add_particles(np.array([[0.5, 0.5]], np.float32))
    finally:
        s.close()
