"""The learned SVD warm start on the device: forward pass and gradients
against the CPU oracle (oracle/mlp.py), training, what the trained network
gains over the cold start, and the simulator running with it."""
import numpy as np
import pytest

from mpm_cuda_amd import capi
from oracle import mlp as M
from tests.scenes import Pair, rel

pytestmark = pytest.mark.gpu

# largest share of warm-started SVDs whose start the acceptance test rejects
# (measured on the MI355X: jello 0.66, snow 0.41 with the network below)
WARM_REJECT_BAR = {capi.JELLO: 0.8, capi.SNOW: 0.55}


def _deformations(n, seed, spread=0.3):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(n, 3, 3)))
    Q = Q * np.sign(np.linalg.det(Q))[:, None, None]
    s = np.exp(spread * rng.normal(size=(n, 3)))
    Q2, _ = np.linalg.qr(rng.normal(size=(n, 3, 3)))
    Q2 = Q2 * np.sign(np.linalg.det(Q2))[:, None, None]
    return (Q * s[:, None, :] @ Q2).reshape(n, 9).astype(np.float32)


def _perturbed(p, seed, scale=0.05):
    return (p + scale * np.random.default_rng(seed).normal(size=p.size)).astype(np.float32)


def test_forward_matches_oracle(gpu):
    p = _perturbed(M.init_params(1), 2)
    F = _deformations(1000, 3)
    q = capi.mlp_forward(p, F)
    qo, _ = M.forward(p, F, bf16_points=True)
    # FP32 (device) vs FP64 (oracle) accumulation between the same BF16 roundings
    assert rel(q, qo) < 5e-3, rel(q, qo)
    assert q.shape == (1000, 4) and np.all(np.isfinite(q))


def test_gradient_matches_oracle(gpu):
    p = _perturbed(M.init_params(4), 5)
    F = _deformations(3000, 6)
    L, g = capi.mlp_grad(p, F)
    Lo, go = M.grad(p, F, bf16_points=True)
    assert abs(L - Lo) < 2e-3 * Lo, (L, Lo)
    err = rel(g, go)
    print("gradient rel error", err)
    assert err < 2e-2, err
    # every block of the parameter vector carries gradient
    for name, (a, b) in {"W1": (0, 576), "W2": (640, 4736), "W3": (4800, 5056)}.items():
        assert rel(g[a:b], go[a:b]) < 5e-2, name


def test_training_reduces_loss(gpu):
    F = _deformations(20000, 7)
    p0 = M.init_params(8)
    L0 = M.loss(p0, F[:2000])
    p, hist = capi.mlp_train(F, p0, steps=1000, batch=2048, lr=2e-3, seed=1)
    L1 = M.loss(p, F[:2000])
    print("loss", L0, "->", L1, "history", hist[0], hist[-1])
    assert np.all(np.isfinite(hist))
    # eigenvectors of arbitrary F (order and sign ambiguous) are a hard
    # target for a 64-wide network; the loss must still fall clearly
    assert L1 < 0.6 * L0, (L0, L1)


def test_trained_warm_start_gain(gpu):
    """On held-out deformation gradients the trained network makes two
    Jacobi sweeps reach the 1e-6 acceptance bound (largest cosine between
    the columns of F V) much more often than the cold start does.  (One
    sweep reaches it for almost no matrix either way: a BF16 network's
    rotation is good to ~1e-2, and a sweep squares the error.)"""
    F = _deformations(40000, 9, spread=0.25)
    Ftr, Fte = F[:30000], F[30000:]
    p, _ = capi.mlp_train(Ftr, M.init_params(10), steps=1500, batch=4096, lr=2e-3, seed=2)
    q = capi.mlp_forward(p, Fte)
    off_w = capi.probe_svd_residual(Fte, q, sweeps=1)
    off_c = capi.probe_svd_residual(Fte, None, sweeps=1)
    acc_w, acc_c = float(np.mean(off_w <= 1e-6)), float(np.mean(off_c <= 1e-6))
    off_w2 = capi.probe_svd_residual(Fte, q, sweeps=2)
    off_c2 = capi.probe_svd_residual(Fte, None, sweeps=2)
    acc_w2, acc_c2 = float(np.mean(off_w2 <= 1e-6)), float(np.mean(off_c2 <= 1e-6))
    print(f"accepted after 1 sweep: warm {acc_w:.3f} cold {acc_c:.3f}; after 2: warm {acc_w2:.3f} cold {acc_c2:.3f}")
    assert acc_w2 > acc_c2 + 0.1, (acc_w2, acc_c2)
    assert acc_w >= acc_c
    # the median residual after one sweep is smaller with the warm start
    assert np.median(off_w) < np.median(off_c)


def _out_of_band(n, seed, frac, lo, hi):
    """F = U diag(s) V^T with s uniform in [lo, hi] for a fraction `frac` of
    n particles (the rest identity), FP32 row-major."""
    rng = np.random.default_rng(seed)
    F = np.tile(np.eye(3), (n, 1, 1))
    pick = rng.random(n) < frac
    m = int(pick.sum())
    U, _ = np.linalg.qr(rng.normal(size=(m, 3, 3)))
    V, _ = np.linalg.qr(rng.normal(size=(m, 3, 3)))
    U *= np.sign(np.linalg.det(U))[:, None, None]
    V *= np.sign(np.linalg.det(V))[:, None, None]
    F[pick] = U * rng.uniform(lo, hi, (m, 1, 3)) @ V.transpose(0, 2, 1)
    return F.reshape(n, 9).astype(np.float32)


def _svd_calls(sim, mat, n):
    """Particles of the step just taken whose stress or plasticity took the
    (warm-started) SVD: snow particles whose J_p changed (the clamp), jello
    particles whose new F has a singular value outside (0.5, 2) (outside
    the Newton polar path)."""
    if mat == capi.SNOW:
        return None
    F = sim.read(0, capi.P_F).reshape(n, 3, 3).astype(np.float64)
    s = np.linalg.svd(F, compute_uv=False)
    return int(np.count_nonzero((s.min(1) <= 0.5) | (s.max(1) >= 2.0)))


@pytest.mark.parametrize("mat", [capi.JELLO, capi.SNOW], ids=["jello", "snow"])
def test_simulation_with_warm_start(gpu, mat):
    """FAST form with the warm start (two sweeps, cold fallback) against the
    cold start, with a third of the particles deformed so that the SVD path
    runs: snow trial F outside the clamp band (singular values in
    [0.93, 1.03] against [0.975, 1.0075]), jello F outside (0.5, 2) (the
    stress takes the SVD instead of the Newton polar rotation).  Fields
    match the cold run within the FP32 parity bars.  The device counters
    agree with the SVD-path particles counted on the host; the network is
    run for the particles whose previous G2P took the SVD, so every
    SVD-path update after the first step starts from a quaternion, and at
    most the stated share of those starts is rejected."""
    F = _deformations(20000, 11, spread=0.05)
    p, _ = capi.mlp_train(F, M.init_params(12), steps=300, batch=2048, lr=2e-3, seed=3)
    a = Pair(type_=mat, form=capi.FORM_FAST, seed=21)
    b = Pair(type_=mat, form=capi.FORM_FAST, seed=21, extra_cfg={"warm_start": capi.WARM_MLP, "warm_sweeps": 2})
    try:
        n = b.n
        F0 = _out_of_band(n, 22, 1 / 3, 0.93, 1.03) if mat == capi.SNOW else _out_of_band(n, 22, 1 / 3, 0.3, 0.45)
        if mat == capi.JELLO:
            big = np.random.default_rng(23).random(n) < 0.5         # half of them stretched instead
            F0 = F0.reshape(n, 3, 3)
            s = np.linalg.svd(F0.astype(np.float64), compute_uv=False)
            sel = big & (s.max(1) < 0.5)
            F0[sel] *= 6.0                                          # [0.3, 0.45] -> [1.8, 2.7]
            F0 = F0.reshape(n, 9).astype(np.float32)
        for q in (a, b):
            q.gpu.write(0, capi.P_F, F0)
        b.gpu.set_warm_start_weights(p)
        calls, first = 0, 0
        jp = b.gpu.read(0, capi.P_JP)[:, 0].copy()
        for k in range(5):
            a.gpu.step(a.dt)
            b.gpu.step(b.dt)
            if mat == capi.SNOW:
                j = b.gpu.read(0, capi.P_JP)[:, 0]
                c = int(np.count_nonzero(j != jp))
                jp = j.copy()
            else:
                c = _svd_calls(b.gpu, mat, n)
            calls += c
            if k == 0:
                first = b.gpu.svd_stats()["svd"]
        ga, gb = a.gpu_particles(), b.gpu_particles()
        for k in ("x", "v"):
            assert rel(gb[k], ga[k]) < 1e-5, k
        assert rel(gb["F"] - np.eye(3), ga["F"] - np.eye(3)) < 1e-4
        st = b.gpu.svd_stats()
        cold = a.gpu.svd_stats()
        frac = st["rejected"] / max(st["warm"], 1)
        print(f"SVD-path updates {st['svd']} (host count {calls}, cold run {cold['svd']}) of {5 * n} "
              f"particle-steps, warm-started {st['warm']}, rejected {st['rejected']} ({frac:.3f})")
        assert calls >= 0.1 * 5 * n
        # host and device agree on the SVD path (a particle at rounding
        # distance of the band edge may take it without a clamp)
        assert calls <= st["svd"] <= 1.01 * calls + 10, (calls, st["svd"])
        # the cold run takes the same path up to particles at rounding
        # distance of the band edge (the two runs differ by rounding)
        assert cold["warm"] == 0 and cold["rejected"] == 0
        assert abs(cold["svd"] - st["svd"]) <= 0.01 * st["svd"] + 10, (cold["svd"], st["svd"])
        # the first step has no quaternion yet; later SVD-path updates of
        # particles listed by the step before start from the network's
        assert st["warm"] >= 0.8 * (st["svd"] - first), (st, first)
        assert frac <= WARM_REJECT_BAR[mat], frac
        assert b.gpu.health() & ~capi.H_ARENA_SPILL == 0
    finally:
        a.close()
        b.close()
