"""CPU restatement of the learned SVD warm start (test infrastructure only).

The device network (``mpm_cuda_amd/csrc/mpm_mlp.hip``) has no counterpart in
the reference, whose SVD always starts from the identity (``svd3_cuda.h``);
this module is the checker for its forward pass, loss and gradients:

* ``forward`` / ``grad`` with ``bf16=True`` round the inputs, weights,
  hidden activations and back-propagated deltas to bfloat16 at the points
  where the device kernels store them in BF16 for the MFMA units, and
  accumulate in FP64 (the device accumulates in FP32);
* with ``bf16=False`` everything is FP64, which the tests pin against
  central finite differences of ``loss``.

Parameter vector (FP32): W1[9][64] b1[64] W2[64][64] b2[64] W3[64][4] b3[4];
a layer is y_j = sum_i W[i][j] x_i + b_j; hidden layers use ReLU.  Input:
F / |F|_F (row-major).  Loss per sample, with A = F^T F / |F|_F^2 and
R = R(q / |q|) (q = (w, x, y, z)):  sum_{i != j} (R^T A R)_ij^2 / |A|_F^2.
"""
from __future__ import annotations

import numpy as np

N_IN, N_H, N_OUT = 9, 64, 4
SHAPES = [("W1", (N_IN, N_H)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)), ("W3", (N_H, N_OUT)),
          ("b3", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)       # 5060


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def unpack(params):
    p = np.asarray(params, np.float64).reshape(-1)
    assert p.size == N_PARAMS, p.size
    out, k = {}, 0
    for name, shape in SHAPES:
        size = int(np.prod(shape))
        out[name] = p[k:k + size].reshape(shape)
        k += size
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).reshape(-1) for name, _ in SHAPES]).astype(np.float32)


def init_params(seed=0):
    """He-normal weights, zero biases; b3 = (1, 0, 0, 0) (the identity rotation)."""
    rng = np.random.default_rng(seed)
    d = {"W1": rng.normal(0, np.sqrt(2.0 / N_IN), (N_IN, N_H)), "b1": np.zeros(N_H),
         "W2": rng.normal(0, np.sqrt(2.0 / N_H), (N_H, N_H)), "b2": np.zeros(N_H),
         "W3": rng.normal(0, 0.1 * np.sqrt(1.0 / N_H), (N_H, N_OUT)), "b3": np.array([1.0, 0.0, 0.0, 0.0])}
    return pack(d)


def net_input(F):
    F = np.asarray(F, np.float64).reshape(-1, 9)
    n = np.sqrt((F * F).sum(1, keepdims=True))
    return np.where(n > 0, F / np.where(n > 0, n, 1.0), 0.0)


def quat_to_mat(q):
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.empty((q.shape[0], 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - w * z); R[:, 0, 2] = 2 * (x * z + w * y)
    R[:, 1, 0] = 2 * (x * y + w * z); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - w * x)
    R[:, 2, 0] = 2 * (x * z - w * y); R[:, 2, 1] = 2 * (y * z + w * x); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def forward(params, F, bf16_points=True):
    """q[n][4] and the activations (x, h1, h2) used by grad."""
    P = unpack(params)
    rnd = bf16 if bf16_points else (lambda a: np.asarray(a, np.float64))
    x = rnd(net_input(F).astype(np.float32) if bf16_points else net_input(F))
    W1, W2, W3 = rnd(P["W1"]), rnd(P["W2"]), rnd(P["W3"])
    h1 = rnd(np.maximum(x @ W1 + P["b1"], 0.0))
    h2 = rnd(np.maximum(h1 @ W2 + P["b2"], 0.0))
    q = h2 @ W3 + P["b3"]
    return q, (x, h1, h2, W1, W2, W3)


def loss_per_sample(x, q):
    """Loss of each sample (x: network inputs F / |F|)."""
    F = x.reshape(-1, 3, 3)
    A = np.einsum("nra,nrb->nab", F, F)
    nA = (A * A).sum((1, 2))
    qh = q / np.maximum(np.linalg.norm(q, axis=1, keepdims=True), 1e-12)
    R = quat_to_mat(qh)
    B = np.einsum("nki,nkl,nlj->nij", R, A, R)
    off = B - np.einsum("nii->ni", B)[:, :, None] * np.eye(3)
    return np.where(nA > 0, (off * off).sum((1, 2)) / np.where(nA > 0, nA, 1.0), 0.0)


def dloss_dq(x, q):
    """d loss / d q per sample (analytic; pinned by finite differences in the tests)."""
    F = x.reshape(-1, 3, 3)
    A = np.einsum("nra,nrb->nab", F, F)
    nA = (A * A).sum((1, 2))
    inv = np.where(nA > 0, 1.0 / np.where(nA > 0, nA, 1.0), 0.0)
    qn = np.maximum(np.linalg.norm(q, axis=1, keepdims=True), 1e-12)
    qh = q / qn
    R = quat_to_mat(qh)
    B = np.einsum("nki,nkl,nlj->nij", R, A, R)
    GB = 2.0 * (B - np.einsum("nii->ni", B)[:, :, None] * np.eye(3)) * inv[:, None, None]
    dR = 2.0 * A @ R @ GB
    w, X, Y, Z = qh[:, 0], qh[:, 1], qh[:, 2], qh[:, 3]
    zero = np.zeros_like(w)
    # dR/dq_k for q = (w, x, y, z)
    dRdq = [
        2 * np.stack([zero, -Z, Y, Z, zero, -X, -Y, X, zero], 1),
        2 * np.stack([zero, Y, Z, Y, -2 * X, -w, Z, w, -2 * X], 1),
        2 * np.stack([-2 * Y, X, w, X, zero, Z, -w, Z, -2 * Y], 1),
        2 * np.stack([-2 * Z, -w, X, w, -2 * Z, Y, X, Y, zero], 1),
    ]
    g = np.stack([(dR.reshape(-1, 9) * d).sum(1) for d in dRdq], 1)
    return (g - qh * (qh * g).sum(1, keepdims=True)) / qn


def loss(params, F, bf16_points=False):
    q, (x, *_rest) = forward(params, F, bf16_points)
    return float(loss_per_sample(x, q).mean())


def grad(params, F, bf16_points=True):
    """(mean loss, gradient[N_PARAMS]) with the device's rounding points."""
    q, (x, h1, h2, W1, W2, W3) = forward(params, F, bf16_points)
    rnd = bf16 if bf16_points else (lambda a: np.asarray(a, np.float64))
    n = x.shape[0]
    L = loss_per_sample(x, q)
    d3 = rnd(dloss_dq(x, q) / n)
    gW3, gb3 = h2.T @ d3, d3.sum(0)
    d2 = rnd(np.where(h2 > 0, d3 @ W3.T, 0.0))
    gW2, gb2 = h1.T @ d2, d2.sum(0)
    d1 = rnd(np.where(h1 > 0, d2 @ W2.T, 0.0))
    gW1, gb1 = x.T @ d1, d1.sum(0)
    g = pack({"W1": gW1, "b1": gb1, "W2": gW2, "b2": gb2, "W3": gW3, "b3": gb3}).astype(np.float64)
    return float(L.mean()), g
