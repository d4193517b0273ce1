"""Pins of the 2D FP64 oracle (oracle/mpm_oracle2d.c, the restatement of the
reference's MPMSnow.py substep): spline identities, the 2x2 SVD against
LAPACK, the stress against closed forms and finite differences of its
energy, and invariants of whole steps (conservation, affine-field
reproduction, plasticity).  The Taichi script itself cannot run here
(taichi is not installed), and the reference ships no fixtures."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

L = O.lib()


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def test_spline_identities_and_derivative():
    h = 1.0 / 128
    rng = np.random.default_rng(1)
    for _ in range(200):
        xp = rng.uniform(10 * h, 20 * h, 2)
        base = (xp / h - 0.5).astype(int)
        sw, sx, g, M = 0.0, np.zeros(2), np.zeros(2), np.zeros((2, 2))
        for i in range(3):
            for j in range(3):
                xi = (base + [i, j]) * h
                u = (xp - xi) / h
                w = L.orc2_N(u[0]) * L.orc2_N(u[1])
                gw = np.array([L.orc2_dN(u[0]) * L.orc2_N(u[1]), L.orc2_N(u[0]) * L.orc2_dN(u[1])]) / h
                sw += w
                sx += w * (xi - xp)
                g += gw
                M += np.outer(xi - xp, gw)
        assert abs(sw - 1) < 1e-14 and np.abs(sx).max() < 1e-16
        assert np.abs(g).max() < 1e-10
        # d/dx_p of sum w (x_i - x_p) = 0 gives sum (x_i - x_p) grad w^T = I
        assert np.abs(M - np.eye(2)).max() < 1e-12, M
    for u in np.linspace(-1.45, 1.45, 59):
        if min(abs(abs(u) - 0.5), abs(abs(u) - 1.5)) < 1e-3:
            continue
        e = 1e-6
        fd = (L.orc2_N(u + e) - L.orc2_N(u - e)) / (2 * e)
        assert abs(fd - L.orc2_dN(u)) < 1e-8, u
    assert L.orc2_N(1.5) == 0.0 and L.orc2_dN(-1.5) == 0.0 and L.orc2_dN(0.5) == -1.0


def _svd(F):
    U, s, V = np.zeros(4), np.zeros(2), np.zeros(4)
    F = np.ascontiguousarray(F, np.float64).reshape(4)
    L.orc2_svd(_p(F), _p(U), _p(s), _p(V))
    return U.reshape(2, 2), s, V.reshape(2, 2)


@pytest.mark.parametrize("kind", ["random", "near_identity", "rotation", "diagonal", "repeated", "inverted"])
def test_svd2_against_lapack(kind):
    rng = np.random.default_rng({"random": 2, "near_identity": 3, "rotation": 4, "diagonal": 5, "repeated": 6,
                                 "inverted": 7}[kind])
    for _ in range(100):
        if kind == "random":
            F = rng.normal(size=(2, 2))
            if np.linalg.det(F) < 0:
                F[0] *= -1
        elif kind == "near_identity":
            F = np.eye(2) + 1e-4 * rng.normal(size=(2, 2))
        elif kind == "rotation":
            t = rng.uniform(-np.pi, np.pi)
            F = np.array([[np.cos(t), -np.sin(t)], [np.sin(t), np.cos(t)]])
        elif kind == "diagonal":
            F = np.diag(rng.uniform(0.5, 2, 2))
        elif kind == "repeated":
            t = rng.uniform(-np.pi, np.pi)
            F = rng.uniform(0.5, 2) * np.array([[np.cos(t), -np.sin(t)], [np.sin(t), np.cos(t)]])
        else:
            F = rng.normal(size=(2, 2))
            if np.linalg.det(F) > 0:
                F[0] *= -1
        U, s, V = _svd(F)
        assert np.abs(U @ np.diag(s) @ V.T - F).max() < 1e-12 * max(1.0, np.abs(F).max())
        assert abs(np.linalg.det(U) - 1) < 1e-12 and abs(np.linalg.det(V) - 1) < 1e-12
        assert np.abs(U.T @ U - np.eye(2)).max() < 1e-12
        ref = np.linalg.svd(F, compute_uv=False)
        assert np.allclose(np.sort(np.abs(s))[::-1], ref, rtol=1e-12, atol=1e-14)
        assert s[0] >= s[1]
        if np.linalg.det(F) < 0:
            assert s[1] < 0            # the sign of det F lands on the smaller value
    U, s, V = _svd(np.zeros((2, 2)))
    assert np.all(s == 0)


def _tau(F, mu, lam):
    t = np.zeros(4)
    F = np.ascontiguousarray(F, np.float64).reshape(4)
    L.orc2_fcr_tau(_p(F), mu, lam, _p(t))
    return t.reshape(2, 2)


def test_fcr_stress_closed_forms_and_energy_derivative():
    mu, lam = O.lame(1e3, 0.2)
    t = 0.7
    R = np.array([[np.cos(t), -np.sin(t)], [np.sin(t), np.cos(t)]])
    assert np.abs(_tau(R, mu, lam)).max() < 1e-10
    for s in (0.9, 1.0, 1.1):
        expect = (2 * mu * (s - 1) * s + lam * (s * s - 1) * s * s) * np.eye(2)
        assert np.abs(_tau(s * np.eye(2), mu, lam) - expect).max() < 1e-9
    # tau = dPsi/dF F^T with Psi = mu |F - R|^2 + lam/2 (J - 1)^2
    rng = np.random.default_rng(8)

    def psi(F):
        U, s, V = _svd(F)
        return mu * np.sum((F - U @ V.T) ** 2) + 0.5 * lam * (np.linalg.det(F) - 1) ** 2

    for _ in range(20):
        F = np.eye(2) + 0.2 * rng.normal(size=(2, 2))
        P = np.zeros((2, 2))
        e = 1e-6
        for a in range(2):
            for b in range(2):
                d = np.zeros((2, 2)); d[a, b] = e
                P[a, b] = (psi(F + d) - psi(F - d)) / (2 * e)
        assert np.abs(P @ F.T - _tau(F, mu, lam)).max() < 1e-5 * max(1.0, np.abs(P).max())


def _squares(seed, n=400):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.4, 0.6, (n, 2))


def test_step_conserves_mass_and_momentum():
    x = _squares(9)
    rng = np.random.default_rng(10)
    v = rng.normal(size=x.shape)
    Fe = np.eye(2) + 0.05 * rng.normal(size=(len(x), 2, 2))
    s = O.Sim2D(x, v, gravity=(0.0, 0.0), Fe=Fe, C=5 * rng.normal(size=(len(x), 2, 2)))
    m_p = s.params.p_mass
    assert s.step(1e-4) == 0
    m = s.grid_m
    gv = s.grid_v.reshape(-1, 2)
    assert abs(m.sum() - len(x) * m_p) < 1e-13 * len(x) * m_p
    # sum_i m_i v_i = sum_p m_p v_p: the C term vanishes (sum w (x_i - x_p) = 0)
    # and so does the stress term (sum grad w = 0); no node is in a wall band
    assert np.abs((m[:, None] * gv).sum(0) - m_p * v.sum(0)).max() < 1e-12 * np.abs(m_p * v).sum()


def test_step_reproduces_affine_velocity():
    """With no stress (F = I), no gravity and v_p = A x_p + b, C_p = A, the
    grid velocity is the affine field and G2P returns v_p and C_p = A
    exactly (the quadratic-spline APIC transfer is exact for affine fields),
    and grad v = A enters F."""
    x = _squares(11)
    A = np.array([[0.3, -1.2], [0.7, 0.1]])
    b = np.array([0.5, -0.25])
    v = x @ A.T + b
    C = np.broadcast_to(A, (len(x), 2, 2))
    s = O.Sim2D(x, v, gravity=(0.0, 0.0), C=C, theta_c=0.5, theta_s=0.5)
    x0 = s.x.copy()
    dt = 1e-5
    assert s.step(dt) == 0
    m = s.grid_m
    gv = s.grid_v.reshape(-1, 2)
    n = s.n_grid
    idx = np.nonzero(m > 0)[0]
    xi = np.stack([idx % n, idx // n], 1) / n
    assert np.abs(gv[idx] - (xi @ A.T + b)).max() < 1e-12
    assert np.abs(s.v - (x0 @ A.T + b)).max() < 1e-12
    assert np.abs(s.C - A).max() < 1e-10
    assert np.abs(s.x - (x0 + dt * s.v)).max() < 1e-15
    assert np.abs(s.Fe - (np.eye(2) + dt * A)).max() < 1e-12


def test_plasticity_invariants():
    rng = np.random.default_rng(12)
    x = _squares(13)
    Fe = np.eye(2) + 0.2 * rng.normal(size=(len(x), 2, 2))
    Fe[np.linalg.det(Fe) < 0, 0, :] *= -1          # det F_e > 0
    Fp = np.eye(2) * rng.uniform(0.9, 1.1, (len(x), 1, 1))
    s = O.Sim2D(x, 0.0, gravity=(0.0, 0.0), Fe=Fe, Fp=Fp, E=1e-3)
    prod0 = np.einsum("nab,nbc->nac", Fe, Fp)
    assert s.step(1e-7) == 0
    sv = np.linalg.svd(s.Fe, compute_uv=False)
    assert sv.min() >= 1 - 2.5e-2 - 1e-12 and sv.max() <= 1 + 7.5e-3 + 1e-12
    # F_e F_p is the elastic-plastic product of the trial state (dt -> 0: Fe_trial ~ Fe)
    prod = np.einsum("nab,nbc->nac", s.Fe, s.Fp)
    assert np.abs(prod - prod0).max() < 1e-4
    # inside the band nothing is clamped: F_e = F_trial, F_p unchanged
    Fe2 = np.eye(2) + 1e-3 * rng.normal(size=(len(x), 2, 2))
    s2 = O.Sim2D(x, 0.0, gravity=(0.0, 0.0), Fe=Fe2, E=1e-3)
    s2.step(1e-9)
    assert np.abs(s2.Fe - Fe2).max() < 1e-8
    assert np.abs(s2.Fp - np.eye(2)).max() < 1e-12


@pytest.mark.parametrize("Jp", [0.85, 1.0, 1.08])
def test_p2g_force_carries_the_hardening(Jp):
    """P2G of MPMSnow.py:86-103 with v = C = 0 and no gravity: node momenta
    m_i v_i are dt f_i, f_i = -V h tau grad w_i with h = exp(xi (1 - det F_p))
    (:87).  For F_e = s R the 2D fixed-corotated stress is isotropic,
    tau = [2 mu (s - 1) s + lam (s^2 - 1) s^2] I, and
    sum_i grad w_i (x_i - x_p)^T = I, so the first moment of the momenta is
    -dt V exp(xi (1 - J_p)) [2 mu (s-1) s + lam (s^2-1) s^2] I.  Lame
    constants from (E, nu) by the textbook formulas."""
    E, nu, xi, s = 1e3, 0.2, 10.0, 0.96
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    Fp = np.array([[Jp, 0.3], [0.0, 1.0]])                     # det = Jp
    x = np.array([[0.5031, 0.4917]])
    sim = O.Sim2D(x, 0.0, gravity=(0.0, 0.0), E=E, nu=nu, xi=xi, Fe=(s * R)[None], Fp=Fp[None],
                  theta_c=0.5, theta_s=0.5)
    dt = 1e-5
    assert sim.step(dt) == 0
    n = sim.n_grid
    m = sim.grid_m
    mom = m[:, None] * sim.grid_v.reshape(-1, 2)
    idx = np.nonzero(m > 0)[0]
    xi_ = np.stack([idx % n, idx // n], 1) / n
    M = (mom[idx][:, :, None] * (xi_ - x[0])[:, None, :]).sum(0)
    tau_iso = 2.0 * mu * (s - 1.0) * s + lam * (s * s - 1.0) * s * s
    want = -dt * sim.params.p_vol * np.exp(xi * (1.0 - Jp)) * tau_iso
    assert np.abs(mom[idx].sum(0)).max() < 1e-12 * abs(want) * n
    assert np.allclose(M, want * np.eye(2), rtol=1e-9, atol=1e-12 * abs(want))
    assert want > 0                                   # compressed: pushes outward


def test_walls_and_gravity():
    x = np.array([[0.5, 0.5], [3.2 / 128, 0.5]])
    s = O.Sim2D(x, 0.0, gravity=(0.0, -50.0))
    s.step(1e-4)
    n = s.n_grid
    m = s.grid_m
    gv = s.grid_v.reshape(-1, 2)
    for k in np.nonzero(m > 0)[0]:
        i, j = k % n, k // n
        if i <= 5 or i >= n - 5 or j <= 5 or j >= n - 5:
            assert np.all(gv[k] == 0.0)
        else:
            assert abs(gv[k][1] - (-50.0 * 1e-4)) < 1e-12 and abs(gv[k][0]) < 1e-15
