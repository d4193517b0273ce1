"""Pins of the CPU restatement of the warm-start network (oracle/mlp.py):
finite differences, closed forms and invariances; no device needed."""
import numpy as np

from oracle import mlp as M


def _rand_F(n, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(n, 3, 3)))
    s = np.exp(0.3 * rng.normal(size=(n, 3)))
    Q2, _ = np.linalg.qr(rng.normal(size=(n, 3, 3)))
    return (Q * s[:, None, :] @ Q2).reshape(n, 9)


def test_bf16_rounding():
    # bf16 keeps 7 fraction bits: the spacing at 1 is 2^-7
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -9, 1.0 + 2 ** -8 + 2 ** -10, 1.0 + 3 * 2 ** -8,
                  -3.14159, 0.0], np.float32)
    r = M.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                         # halfway: to even
    assert r[2] == 1.0 + 2 ** -7               # below halfway
    assert r[3] == 1.0 + 2 ** -7               # above halfway
    assert r[4] == 1.0 + 2 ** -6               # halfway between odd 1+2^-7 and even 1+2^-6
    assert abs(r[5] + 3.140625) < 1e-12
    assert r[6] == 0.0


def test_loss_vanishes_on_eigenvectors_and_is_symmetric():
    F = _rand_F(50, 1)
    x = M.net_input(F)
    A = np.einsum("nra,nrb->nab", x.reshape(-1, 3, 3), x.reshape(-1, 3, 3))
    _, V = np.linalg.eigh(A)
    V = V * np.sign(np.linalg.det(V))[:, None, None]      # proper rotations
    # quaternion of V (w >= 0 branch is enough for a check)
    q = np.empty((50, 4))
    for k in range(50):
        R = V[k]
        t = np.trace(R)
        w = np.sqrt(max(0.0, 1 + t)) / 2
        if w > 1e-3:
            q[k] = [w, (R[2, 1] - R[1, 2]) / (4 * w), (R[0, 2] - R[2, 0]) / (4 * w), (R[1, 0] - R[0, 1]) / (4 * w)]
        else:
            i = np.argmax(np.diag(R))
            j, l = (i + 1) % 3, (i + 2) % 3
            s = np.sqrt(max(0.0, 1 + R[i, i] - R[j, j] - R[l, l])) * 2
            v = np.zeros(4)
            v[1 + i] = s / 4
            v[0] = (R[l, j] - R[j, l]) / s
            v[1 + j] = (R[j, i] + R[i, j]) / s
            v[1 + l] = (R[l, i] + R[i, l]) / s
            q[k] = v
    assert np.abs(M.quat_to_mat(q) - V).max() < 1e-9
    assert M.loss_per_sample(x, q).max() < 1e-20
    q2 = np.random.default_rng(2).normal(size=(50, 4))
    l1 = M.loss_per_sample(x, q2)
    assert np.allclose(l1, M.loss_per_sample(x, -3.0 * q2))      # sign and scale of q
    assert np.allclose(l1, M.loss_per_sample(7.0 * x, q2))       # scale of F
    assert (l1 > 1e-4).mean() > 0.9


def test_dloss_dq_finite_differences():
    F = _rand_F(20, 3)
    x = M.net_input(F)
    q = np.random.default_rng(4).normal(size=(20, 4))
    g = M.dloss_dq(x, q)
    eps = 1e-6
    for k in range(4):
        e = np.zeros(4)
        e[k] = eps
        fd = (M.loss_per_sample(x, q + e) - M.loss_per_sample(x, q - e)) / (2 * eps)
        assert np.abs(fd - g[:, k]).max() < 1e-7 * max(1.0, np.abs(g).max()), k


def test_parameter_gradient_finite_differences():
    F = _rand_F(64, 5)
    p = M.init_params(7).astype(np.float64)
    rng = np.random.default_rng(8)
    p = p + 0.05 * rng.normal(size=p.size)          # leave the symmetric start
    L, g = M.grad(p, F, bf16_points=False)
    assert abs(L - M.loss(p, F)) < 1e-15
    idx = np.concatenate([rng.integers(0, M.N_PARAMS, 40), [M.N_PARAMS - 1, M.N_PARAMS - 4, 9 * 64]])
    eps = 1e-6
    for k in idx:
        e = np.zeros(M.N_PARAMS)
        e[k] = eps
        fd = (M.loss(p + e, F) - M.loss(p - e, F)) / (2 * eps)
        assert abs(fd - g[k]) < 1e-6 * max(1e-3, np.abs(g).max()), (k, fd, g[k])


def test_pack_unpack_roundtrip():
    p = M.init_params(3)
    assert p.size == M.N_PARAMS == 5060
    assert np.array_equal(M.pack(M.unpack(p)), p)
