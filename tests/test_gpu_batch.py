"""Batched scenes (mpm_batch_*): one launch per stage for all scenes,
replayed from hipGraphs, against the same scenes stepped one by one, against
the FP64 oracle, and across graph replays, sorts of both buffer parities,
single steps in between and the autotuner."""
import numpy as np
import pytest

from mpm_cuda_amd import capi
from mpm_cuda_amd import warm_start
from tests.scenes import Pair, rel

pytestmark = pytest.mark.gpu

# uneven scenes: materials, grid sizes (48^3 .. 80^3, one not a multiple of
# the 4^3 block) and time steps differ
SPECS = [
    dict(type_=capi.JELLO, n_grid=64, seed=31),
    dict(type_=capi.SNOW, n_grid=48, seed=32, box=((0.3, 0.3, 0.3), (0.6, 0.55, 0.6))),
    dict(type_=capi.WATER, n_grid=80, seed=33, dt=5e-5, box=((0.2, 0.2, 0.2), (0.45, 0.4, 0.45))),
    dict(type_=capi.SNOW, n_grid=64, seed=34, box=((0.35, 0.3, 0.3), (0.5, 0.5, 0.6))),
    dict(type_=capi.JELLO, n_grid=42, seed=35, box=((0.3, 0.3, 0.3), (0.6, 0.6, 0.6))),
]


def _pairs(sort_every=6, specs=SPECS, **kw):
    return [Pair(form=capi.FORM_FAST, sort_every=sort_every, **{**s, **kw}) for s in specs]


def _close(ps):
    for p in ps:
        p.close()


def _compare(pa, pb, tol=1e-5):
    ga, gb = pa.gpu_particles(), pb.gpu_particles()
    keys = ("x", "v", "C") + (("F",) if pa.type != capi.WATER else ("S",))
    errs = {k: rel(ga[k], gb[k]) for k in keys}
    for k, e in errs.items():
        assert e < tol, (pa.type, k, e)
    return errs


def test_batch_matches_single_scene_stepping(gpu):
    A, B = _pairs(), _pairs()
    try:
        with capi.Batch([p.gpu for p in A], capi.batch_shape(sort_every=6)) as b:
            assert b.scenes == len(A) and b.particles == sum(p.n for p in A)
            ms = b.run(20, timed=True)
            assert ms.shape == (20,) and np.all(ms > 0)
            for p in B:
                for _ in range(20):
                    p.gpu.step(p.dt)
            for pa, pb in zip(A, B):
                print(pa.type, {k: f"{v:.1e}" for k, v in _compare(pa, pb).items()})
                st = pa.gpu.stats()
                assert st["sorts"] >= 4, st          # first sort + 3 batch re-sorts
                assert pa.gpu.current_frame == 20
                assert abs(pa.gpu.current_time - 20 * pa.dt) < 1e-12
                assert pa.gpu.health() & ~capi.H_ARENA_SPILL == 0
    finally:
        _close(A)
        _close(B)


def test_batch_rollout_matches_oracle(gpu):
    """30 batched steps of a jello, a snow and a water scene against the FP64
    oracle, with the bars of the single-scene rollout test (x 5e-6, v 1e-4,
    F 1e-5 of ||F||, J_p / J 1e-5)."""
    P = _pairs(sort_every=5, specs=SPECS[:3])
    try:
        with capi.Batch([p.gpu for p in P], capi.batch_shape(sort_every=5)) as b:
            b.run(30)
            for p in P:
                for _ in range(30):
                    assert p.orc.step(p.dt) == 0
            for p in P:
                g, o = p.gpu_particles(), p.orc_particles()
                ex, ev = rel(g["x"], o["x"]), rel(g["v"], o["v"])
                eF = rel(g["F"], o["F"]) if p.type != capi.WATER else 0.0
                eS = float(np.abs(g["S"] - o["S"]).max()) if p.type != capi.JELLO else 0.0
                print(p.type, f"x {ex:.2e} v {ev:.2e} F {eF:.2e} S {eS:.2e}")
                assert ex < 5e-6 and ev < 1e-4 and eF < 1e-5 and eS < 1e-5
    finally:
        _close(P)


def test_graph_replays_match_direct_launches(gpu):
    """Four runs of 7 steps (re-sorts every 5 steps land in both buffer
    parities and across run boundaries), reading every scene between runs:
    each replay of the captured graphs equals the same stages launched
    directly."""
    A, B = _pairs(sort_every=5), _pairs(sort_every=5)
    try:
        ba = capi.Batch([p.gpu for p in A], capi.batch_shape(sort_every=5, use_graph=1))
        bb = capi.Batch([p.gpu for p in B], capi.batch_shape(sort_every=5, use_graph=0))
        try:
            for r in range(4):
                ba.run(7)
                bb.run(7)
                for pa, pb in zip(A, B):
                    _compare(pa, pb, tol=1e-6)
        finally:
            ba.close()
            bb.close()
    finally:
        _close(A)
        _close(B)


def test_batch_interleaved_with_single_steps(gpu):
    """Batch steps, then scene 0 on its own (its re-sort flips its buffer
    parity), then batch steps again."""
    A, B = _pairs(), _pairs()
    try:
        with capi.Batch([p.gpu for p in A], capi.batch_shape(sort_every=6)) as b:
            b.run(5)
            for _ in range(4):
                A[0].gpu.step(A[0].dt)
            b.run(5)
        for k, p in enumerate(B):
            for _ in range(14 if k == 0 else 10):
                p.gpu.step(p.dt)
        for pa, pb in zip(A, B):
            _compare(pa, pb)
        assert A[0].gpu.current_frame == 14 and A[1].gpu.current_frame == 10
    finally:
        _close(A)
        _close(B)


def test_autotune_keeps_results(gpu):
    A, B = _pairs(), _pairs()
    try:
        with capi.Batch([p.gpu for p in A]) as b:
            ms = b.autotune(reps=2)                  # 14 steps
            print("autotune medians (ms)", ms)
            assert set(ms) == {"p2g", "grid", "g2p", "tile"}
            for st, cand in ms.items():
                assert len(cand) == len(capi.BATCH_CANDIDATES[st]) and all(v > 0 for v in cand.values())
            sh = b.shape
            assert sh.p2g_workgroups in capi.BATCH_CANDIDATES["p2g"]
            assert sh.grid_workgroups in capi.BATCH_CANDIDATES["grid"]
            assert sh.g2p_threads in capi.BATCH_CANDIDATES["g2p"]
            assert sh.p2g_workgroups == min(ms["p2g"], key=ms["p2g"].get)
            assert sh.p2g_tile == min(ms["tile"], key=ms["tile"].get)
            b.run(6)
        for p in B:
            for _ in range(20):
                p.gpu.step(p.dt)
        for pa, pb in zip(A, B):
            _compare(pa, pb)
    finally:
        _close(A)
        _close(B)


def _out_of_band(p, seed):
    """A third of the particles get F outside the band of the cheap stress
    path, so the G2P takes the (warm-started) SVD: snow singular values in
    [0.93, 1.03] (clamp band 0.975-1.0075), jello in [1.8, 2.4]."""
    rng = np.random.default_rng(seed)
    lo, hi = (0.93, 1.03) if p.type == capi.SNOW else (1.8, 2.4)
    F = np.tile(np.eye(3), (p.n, 1, 1))
    pick = rng.random(p.n) < 1 / 3
    m = int(pick.sum())
    U, _ = np.linalg.qr(rng.normal(size=(m, 3, 3)))
    V, _ = np.linalg.qr(rng.normal(size=(m, 3, 3)))
    U *= np.sign(np.linalg.det(U))[:, None, None]
    V *= np.sign(np.linalg.det(V))[:, None, None]
    F[pick] = U * rng.uniform(lo, hi, (m, 1, 3)) @ V.transpose(0, 2, 1)
    p.gpu.write(0, capi.P_F, F.reshape(p.n, 9).astype(np.float32))


def test_batch_warm_start(gpu):
    """Snow and jello scenes with the learned warm start (two sweeps, cold
    fallback) in a batch match the same scenes with the warm start stepped
    one by one.  Two weight sets (the trained one and a perturbed copy) give
    two batched inference launches.  A third of each scene's particles take
    the SVD: the network runs for the particles the step before listed, so
    after the first step most SVD-path updates start from a quaternion, in
    the batch as in the scenes stepped alone."""
    w = warm_start.load_weights()
    w2 = (w + 0.01 * np.random.default_rng(40).normal(size=w.size)).astype(np.float32)
    specs = [SPECS[1], SPECS[3], SPECS[0]]
    extra = {"warm_start": capi.WARM_MLP, "warm_sweeps": 2}
    A = _pairs(specs=specs, extra_cfg=extra)
    B = _pairs(specs=specs, extra_cfg=extra)
    try:
        for k, p in enumerate(A + B):
            p.gpu.set_warm_start_weights(w2 if k % len(specs) == 1 else w)
            _out_of_band(p, 41 + k % len(specs))
        with capi.Batch([p.gpu for p in A], capi.batch_shape(sort_every=6)) as b:
            b.run(10)
        for p in B:
            for _ in range(10):
                p.gpu.step(p.dt)
        for pa, pb in zip(A, B):
            _compare(pa, pb)
            sa, sb = pa.gpu.svd_stats(), pb.gpu.svd_stats()
            print(pa.type, "batched", sa, "alone", sb)
            assert sa["svd"] >= 0.1 * 10 * pa.n and sb["svd"] >= 0.1 * 10 * pb.n
            assert sa["warm"] >= 0.5 * sa["svd"] and sb["warm"] >= 0.5 * sb["svd"]
            assert sa["rejected"] <= sa["warm"] and sb["rejected"] <= sb["warm"]
    finally:
        _close(A)
        _close(B)


def test_batch_of_70_scenes(gpu):
    """More scenes than one batched P2G launch maps (64): the P2G stage runs
    in chunks; every scene still matches its twin stepped on its own."""
    mats = (capi.JELLO, capi.SNOW, capi.WATER)
    specs = [dict(type_=mats[k % 3], n_grid=24 + 4 * (k % 3), seed=100 + k, ppc=2,
                  box=((0.35, 0.35, 0.35), (0.6, 0.55 + 0.01 * (k % 5), 0.6)), dt=5e-5 if k % 3 == 2 else 1e-4)
             for k in range(70)]
    A, B = _pairs(specs=specs), _pairs(specs=specs)
    try:
        with capi.Batch([p.gpu for p in A], capi.batch_shape(sort_every=4)) as b:
            assert b.scenes == 70
            b.run(9)
        for p in B:
            for _ in range(9):
                p.gpu.step(p.dt)
        for pa, pb in zip(A, B):
            _compare(pa, pb)
            assert pa.gpu.stats()["tiles"] > 0
    finally:
        _close(A)
        _close(B)


def test_batch_rejects_bad_scenes(gpu):
    plain = Pair(form=capi.FORM_PLAIN, **SPECS[0])
    fast = Pair(form=capi.FORM_FAST, **SPECS[0])
    two = Pair(form=capi.FORM_FAST, **SPECS[4])
    try:
        two.gpu.add_particles(capi.material(5e4, 0.3, 1e3, capi.JELLO), np.full((10, 3), 0.5, np.float32))
        for sims, what in (([plain.gpu], "FAST"), ([two.gpu], "one"), ([fast.gpu, fast.gpu], "twice"), ([], "1 to 256")):
            with pytest.raises(capi.MpmError, match=what):
                capi.Batch(sims)
        with pytest.raises(capi.MpmError, match="g2p_threads"):
            capi.Batch([fast.gpu], capi.batch_shape(g2p_threads=96))
        with pytest.raises(capi.MpmError, match="sort_every"):
            capi.Batch([fast.gpu], capi.batch_shape(sort_every=0))
    finally:
        plain.close()
        fast.close()
        two.close()
