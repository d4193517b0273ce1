"""Pins of the FP64 oracle (CPU only).

Each check is chosen so that a plausible restatement mistake -- a dropped
factor, a wrong sign, a transposed operand, a wrong interval end -- fails:
LAPACK golden vectors for the SVD / polar / stress, closed forms and a
finite-difference derivative of the fixed-corotated energy, the quadratic
B-spline identities, exact reproduction of affine velocity fields by
G2P (pins 4/h^2 and grad w), conservation in P2G, and the invariants of the
snow return mapping.
"""
import numpy as np
import pytest

import oracle as O

FIX = np.load("tests/golden/svd_fixtures.npz")


def test_svd_matches_lapack_fixtures():
    F, S, R, kind = FIX["F"], FIX["S"], FIX["R"], FIX["kind"]
    for i in range(len(F)):
        U, s, V = O.svd3(F[i])
        scale = max(abs(S[i][0]), 1e-300)
        assert np.abs(s - S[i]).max() <= 1e-12 * scale + 1e-300, (i, kind[i])
        if kind[i] == "zero":
            continue
        assert abs(np.linalg.det(U) - 1) < 1e-12 and abs(np.linalg.det(V) - 1) < 1e-12
        assert np.abs(U.T @ U - np.eye(3)).max() < 1e-12
        assert np.abs(U @ np.diag(s) @ V.T - F[i]).max() < 1e-12 * scale
        assert s[0] >= s[1] >= abs(s[2])
        cond = abs(S[i][2]) / scale
        if kind[i] not in ("repeated", "isotropic", "identity", "diagonal") and cond > 1e-7:
            assert np.abs(U @ V.T - R[i]).max() < 1e-9 / cond, (i, kind[i])


def test_fcr_stress_matches_fixtures():
    mu, lam = float(FIX["mu"]), float(FIX["lam"])
    for i, F in enumerate(FIX["F"]):
        if FIX["kind"][i] in ("zero",):
            continue
        t = O.fcr_tau(F, mu, lam)
        ref = FIX["tau"][i]
        assert np.abs(t - ref).max() <= 1e-8 * (np.abs(ref).max() + mu), (i, FIX["kind"][i])


@pytest.mark.parametrize("s", [0.7, 0.95, 1.0, 1.05, 1.3])
def test_fcr_closed_forms(s):
    mu, lam = O.lame(1e4, 0.3)
    t = O.fcr_tau(s * np.eye(3), mu, lam)
    want = (2 * mu * (s - 1) * s + lam * s ** 3 * (s ** 3 - 1)) * np.eye(3)
    assert np.allclose(t, want, rtol=1e-12, atol=1e-9)
    rng = np.random.default_rng(1)
    q = rng.normal(size=(3, 3))
    Rm, _ = np.linalg.qr(q)
    if np.linalg.det(Rm) < 0:
        Rm[:, 0] *= -1
    assert np.abs(O.fcr_tau(Rm, mu, lam)).max() < 1e-9


def test_fcr_stress_is_energy_derivative():
    """tau = dPsi/dF F^T with Psi = mu ||F - R(F)||^2 + lam/2 (J - 1)^2 (central differences)."""
    mu, lam = O.lame(5e4, 0.3)

    def psi(F):
        U, s, V = O.svd3(F)
        R = U @ V.T
        J = np.linalg.det(F)
        return mu * np.sum((F - R) ** 2) + 0.5 * lam * (J - 1) ** 2

    rng = np.random.default_rng(4)
    for _ in range(10):
        F = np.eye(3) + 0.2 * rng.normal(size=(3, 3))
        P = np.zeros((3, 3))
        eps = 1e-6
        for a in range(3):
            for b in range(3):
                d = np.zeros((3, 3)); d[a, b] = eps
                P[a, b] = (psi(F + d) - psi(F - d)) / (2 * eps)
        tau = O.fcr_tau(F, mu, lam)
        assert np.allclose(tau, P @ F.T, rtol=1e-5, atol=1e-3 * mu * 1e-3)
        assert np.allclose(tau, tau.T, atol=1e-8 * mu)


def _stencil(xp, h):
    b = [O.base_ref(xp[a], 0.0, h) for a in range(3)]
    for ox in range(3):
        for oy in range(3):
            for oz in range(3):
                xi = np.array([(b[0] + ox) * h, (b[1] + oy) * h, (b[2] + oz) * h])
                yield xi


def test_bspline_identities():
    rng = np.random.default_rng(2)
    h = 1.0 / 64
    for _ in range(200):
        xp = rng.uniform(0.2, 0.8, 3)
        sw = 0.0; sd = np.zeros(3); sg = np.zeros(3); sdg = np.zeros((3, 3))
        for xi in _stencil(xp, h):
            w = O.weight(xp, xi, h)
            g = O.weight_grad(xp, xi, h)
            sw += w; sd += w * (xi - xp); sg += g; sdg += np.outer(xi - xp, g)
        assert abs(sw - 1) < 1e-13
        assert np.abs(sd).max() < 1e-15
        assert np.abs(sg).max() < 1e-9
        # d/dx_p of sum w (x_i - x_p) = 0 gives sum (x_i - x_p) grad w^T = I
        assert np.abs(sdg - np.eye(3)).max() < 1e-12


def test_bspline_derivative_matches_finite_difference():
    for x in np.linspace(-1.7, 1.7, 341):
        if min(abs(abs(x) - 0.5), abs(abs(x) - 1.5)) < 1e-3:
            continue
        fd = (O.N(x + 1e-7) - O.N(x - 1e-7)) / 2e-7
        assert abs(O.dN(x) - fd) < 1e-6, x
    # values at the interval ends of MPMCUDA.cu:267-290
    assert O.N(0.5) == 0.5 and O.N(-0.5) == 0.5 and O.N(1.5) == 0.0 and O.N(0.0) == 0.75
    assert O.dN(0.5) == -1.0 and O.dN(-0.5) == 1.0 and O.dN(1.5) == 0.0 and O.dN(-1.5) == 0.0


def test_base_index_forms_agree_off_threshold():
    rng = np.random.default_rng(5)
    h = 1.0 / 192
    x = rng.random(100000).astype(np.float32)
    ref = np.array([O.base_ref(float(v), 0.0, h) for v in x])
    f32 = np.array([O.base_f32(v, 0.0, 192.0) for v in x])
    d = ref != f32
    assert d.mean() < 1e-4
    assert np.all(np.abs(ref[d] - f32[d]) == 1)
    # away from thresholds they must agree
    frac = (x.astype(np.float64) / h - 0.5) % 1.0
    far = (frac > 1e-3) & (frac < 1 - 1e-3)
    assert np.array_equal(ref[far], f32[far])


def _affine_grid(sim, v0, A):
    nz, ny, nx = sim.nz, sim.ny, sim.nx
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    X = np.stack([x, y, z], -1) * sim.h
    V = v0 + X @ A.T
    sim.p[:] = V.reshape(-1)
    sim.m[:] = 1.0


@pytest.mark.parametrize("kind", [O.JELLO, O.WATER])
def test_g2p_reproduces_affine_field(kind):
    n_grid, h = 32, 1.0 / 32
    rng = np.random.default_rng(6)
    x = rng.uniform(0.3, 0.7, (100, 3))
    A = rng.normal(size=(3, 3))
    v0 = rng.normal(size=3)
    g = O.Group(x, np.zeros(3), type=kind, mass=1.0, vol=h ** 3, E=1e3, nu=0.3)
    sim = O.Sim(n_grid, n_grid, n_grid, h, [g])
    _affine_grid(sim, v0, A)
    dt = 1e-3
    x0 = g.x.copy()
    J0 = g.J.copy()
    assert sim.g2p(dt) == 0
    vexp = v0 + x0 @ A.T
    assert np.abs(g.v - vexp).max() < 1e-12
    assert np.abs(g.C - A).max() < 1e-10
    assert np.abs(g.x - (x0 + dt * vexp)).max() < 1e-14
    if kind == O.JELLO:
        assert np.abs(g.Fe - (np.eye(3) + dt * A)).max() < 1e-12
    else:
        assert np.abs(g.J - (1 + dt * np.trace(A)) * J0).max() < 1e-12


def test_p2g_conserves_mass_and_momentum():
    n_grid, h = 32, 1.0 / 32
    rng = np.random.default_rng(8)
    x = rng.uniform(0.3, 0.7, (500, 3))
    v = rng.normal(size=(500, 3))
    C = rng.normal(size=(500, 3, 3))
    g = O.Group(x, v, type=O.WATER, mass=0.5, vol=h ** 3, eos_k=1e4, C=C)   # J = 1 -> zero pressure
    sim = O.Sim(n_grid, n_grid, n_grid, h, [g])
    sim.clear_grid()
    assert sim.p2g(1e-4) == 0
    assert abs(sim.m.sum() - 0.5 * 500) < 1e-10
    P = sim.p.reshape(-1, 3).sum(0)
    assert np.abs(P - 0.5 * v.sum(0)).max() < 1e-10


def test_p2g_force_is_minus_volume_stress_gradient():
    """With v = C = 0 the node momenta are dt * (-V tau grad w); their sum is
    zero and their first moment is -dt V tau (sum grad w (x_i - x_p)^T = I)."""
    n_grid, h = 32, 1.0 / 32
    x = np.array([[0.51, 0.47, 0.53]])
    F = np.eye(3) + np.array([[0.1, 0.02, 0.0], [0.0, -0.05, 0.01], [0.03, 0.0, 0.07]])
    g = O.Group(x, np.zeros(3), type=O.JELLO, mass=1.0, vol=h ** 3, E=5e4, nu=0.3, Fe=F[None])
    sim = O.Sim(n_grid, n_grid, n_grid, h, [g])
    sim.clear_grid()
    dt = 1e-4
    sim.p2g(dt)
    tau = O.fcr_tau(F, g.mat.mu, g.mat.lam)
    p = sim.grid_vec()
    assert np.abs(p.reshape(-1, 3).sum(0)).max() < 1e-12
    z, y, xx = np.nonzero(sim.grid_mass())
    M = np.zeros((3, 3))
    for a, b, c in zip(z, y, xx):
        xi = np.array([c, b, a]) * h
        M += np.outer(p[a, b, c], xi - x[0])
    assert np.allclose(M, -dt * h ** 3 * tau, rtol=1e-9, atol=1e-18)


def _first_moment(sim, xp):
    """sum_i p_i (x_i - x_p)^T over the nodes with mass, and sum_i p_i."""
    h = sim.h
    p = sim.grid_vec()
    z, y, xx = np.nonzero(sim.grid_mass())
    M = np.zeros((3, 3))
    for a, b, c in zip(z, y, xx):
        M += np.outer(p[a, b, c], np.array([c, b, a]) * h - xp)
    return M, p.reshape(-1, 3).sum(0)


def _rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


@pytest.mark.parametrize("Jp", [0.8, 1.0, 1.1])
def test_p2g_snow_force_carries_the_hardening(Jp):
    """Snow P2G (MPMCUDA.cu:342-349): with v = C = 0 the node momenta are
    dt f_i, f_i = -V h tau grad w_i, h = exp(xi (1 - det F_p)).  For
    F_e = s R (R a rotation) the fixed-corotated stress is isotropic,
    tau = [2 mu (s - 1) s + lam s^3 (s^3 - 1)] I, and
    sum_i grad w_i (x_i - x_p)^T = I, so the first moment is
    -dt V exp(xi (1 - J_p)) [2 mu (s-1) s + lam s^3 (s^3-1)] I.
    Lame constants from (E, nu) by the textbook formulas; F_p is a
    rotated non-diagonal factor with det J_p.  A flipped or inverted
    hardening exponent changes the result by exp(+-2 xi (1 - J_p))."""
    rng = np.random.default_rng(21)
    n_grid, h = 32, 1.0 / 32
    E, nu, xi, s = 1.4e5, 0.2, 10.0, 0.97
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    x = np.array([[0.503, 0.491, 0.517]])
    R = _rotation(rng)
    Q = _rotation(rng)
    Fp = Q @ np.diag([Jp, 1.0, 1.0]) @ Q.T @ _rotation(rng)   # det = Jp, not symmetric
    assert abs(np.linalg.det(Fp) - Jp) < 1e-12
    g = O.Group(x, np.zeros(3), type=O.SNOW, mass=1.0, vol=h ** 3, E=E, nu=nu, xi=xi,
                Fe=(s * R)[None], Fp=Fp[None])
    sim = O.Sim(n_grid, n_grid, n_grid, h, [g])
    sim.clear_grid()
    dt = 1e-4
    assert sim.p2g(dt) == 0
    M, P = _first_moment(sim, x[0])
    tau_iso = 2.0 * mu * (s - 1.0) * s + lam * s ** 3 * (s ** 3 - 1.0)
    want = -dt * h ** 3 * np.exp(xi * (1.0 - Jp)) * tau_iso
    assert np.abs(P).max() < 1e-12 * abs(want) / h
    assert np.allclose(M, want * np.eye(3), rtol=1e-9, atol=1e-12 * abs(want))
    # compressed snow (s < 1) pushes its nodes outward: positive first moment
    assert want > 0


@pytest.mark.parametrize("J,pk", [(0.5, 127.0), (1.25, 0.2097152 - 1.0), (1.0, 0.0)])
def test_p2g_water_pressure_force(J, pk):
    """Water P2G (MPMCUDA.cu:351-354, 412-414): p = k (J^-gamma - 1),
    f_i = +V p J grad w_i, so with v = C = 0 the first moment of the node
    momenta is +dt V p J I.  gamma = 7: J = 1/2 gives J^-7 - 1 = 127 and
    J = 5/4 gives 0.8^7 - 1 = 0.2097152 - 1 (both exact).  Compressed water
    (J < 1) pushes outward (positive moment), stretched water pulls in."""
    n_grid, h = 32, 1.0 / 32
    k, dt = 1e4, 1e-4
    x = np.array([[0.488, 0.509, 0.521]])
    g = O.Group(x, np.zeros(3), type=O.WATER, mass=1.0, vol=h ** 3, eos_k=k, eos_gamma=7.0, J=np.array([J]))
    sim = O.Sim(n_grid, n_grid, n_grid, h, [g])
    sim.clear_grid()
    assert sim.p2g(dt) == 0
    M, P = _first_moment(sim, x[0])
    want = dt * h ** 3 * k * pk * J
    scale = dt * h ** 3 * k
    assert np.abs(P).max() < 1e-12 * scale / h
    assert np.allclose(M, want * np.eye(3), rtol=1e-9, atol=1e-12 * scale)
    assert np.sign(M[0, 0]) == np.sign(1.0 - J)


def test_grid_update_gravity_and_walls():
    n, h = 24, 1.0 / 24
    sim = O.Sim(n, n, n, h, [], bc_pad=3, bc_mode=O.BC_AXIS)
    sim.m[:] = 2.0
    sim.p[:] = 4.0
    sim.m[0] = 1e-17    # below DBL_EPSILON: zeroed
    dt = 1e-3
    sim.grid_update(dt)
    v = sim.grid_vec()
    idx = np.arange(n)
    wall = (idx <= 3) | (idx >= n - 3)
    vx = np.where(wall[None, None, :], 0.0, 2.0)
    vy = np.where(wall[None, :, None], 0.0, 2.0 - 9.8 * dt)
    vz = np.where(wall[:, None, None], 0.0, 2.0)
    exp = np.stack(np.broadcast_arrays(vx, vy, vz), -1)
    exp[0, 0, 0] = 0.0
    assert np.allclose(v, exp, atol=1e-15)
    assert sim.grid_mass()[0, 0, 0] == 0.0
    sticky = O.Sim(n, n, n, h, [], bc_pad=3, bc_mode=O.BC_STICKY)
    sticky.m[:] = 1.0; sticky.p[:] = 1.0
    sticky.grid_update(dt)
    vs = sticky.grid_vec()
    anywall = wall[:, None, None] | wall[None, :, None] | wall[None, None, :]
    assert np.all(vs[anywall] == 0.0)


def test_snow_return_mapping_invariants():
    rng = np.random.default_rng(9)
    tc, ts = 2.5e-2, 7.5e-3
    for _ in range(100):
        Ftr = np.eye(3) + 0.05 * rng.normal(size=(3, 3))
        Fp = np.eye(3) * rng.uniform(0.9, 1.1)
        Fe, Fp2 = O.snow_plastic(Ftr, Fp, tc, ts)
        s = np.linalg.svd(Fe, compute_uv=False)
        assert s.max() <= 1 + ts + 1e-12 and s.min() >= 1 - tc - 1e-12
        assert np.allclose(Fe @ Fp2, Ftr @ Fp, atol=1e-12)
        st = np.linalg.svd(Ftr, compute_uv=False)
        ratio = np.prod(st / np.clip(st, 1 - tc, 1 + ts))
        assert abs(np.linalg.det(Fp2) - np.linalg.det(Fp) * ratio) < 1e-12
    Fin = np.diag([1.001, 0.999, 1.0])
    Fe, Fp2 = O.snow_plastic(Fin, np.eye(3), tc, ts)
    assert np.allclose(Fe, Fin, atol=1e-14) and np.allclose(Fp2, np.eye(3), atol=1e-14)


def test_uniform_translation_step():
    """A rigid translation without gravity or walls keeps every particle's
    velocity, leaves F = I and moves x by dt v (P2G -> update -> G2P)."""
    n_grid, h = 32, 1.0 / 32
    rng = np.random.default_rng(10)
    x = rng.uniform(0.4, 0.6, (300, 3))
    vel = np.array([0.5, -1.0, 0.25])
    g = O.Group(x, vel, type=O.JELLO, mass=1e-3, vol=h ** 3, E=1e4, nu=0.3)
    sim = O.Sim(n_grid, n_grid, n_grid, h, [g], gravity=(0, 0, 0), bc_pad=2)
    x0 = g.x.copy()
    dt = 1e-4
    assert sim.step(dt) == 0
    assert np.abs(g.v - vel).max() < 1e-12
    assert np.abs(g.C).max() < 1e-9
    assert np.abs(g.Fe - np.eye(3)).max() < 1e-12
    assert np.abs(g.x - (x0 + dt * vel)).max() < 1e-15
