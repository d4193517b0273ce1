"""Training of the learned SVD warm start on deformation gradients the
simulator itself produces.

    python -m mpm_cuda_amd.warm_start            # collect, train on the GPU, save

Samples: the deformation gradients of the 3D jello drop (262,144 particles,
128^3 grid, onto a sticky floor) and of the 1M-particle snow collision, read
every few steps of their rollouts.  Training runs on the device through
``capi.mlp_train`` (the project's own BF16 MFMA forward/backward kernels and
Adam).  The result is written to ``warm_start_weights.npy`` beside this
file, which ``load_weights`` reads.
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "warm_start_weights.npy")
N_PARAMS = 5060


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases, output bias = the identity quaternion."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, np.sqrt(2.0 / 9), (9, 64))
    W2 = rng.normal(0, np.sqrt(2.0 / 64), (64, 64))
    W3 = rng.normal(0, 0.1 * np.sqrt(1.0 / 64), (64, 4))
    p = np.concatenate([W1.ravel(), np.zeros(64), W2.ravel(), np.zeros(64), W3.ravel(), [1.0, 0.0, 0.0, 0.0]])
    return p.astype(np.float32)


def load_weights(path: str = WEIGHTS) -> np.ndarray:
    w = np.load(path, allow_pickle=False).astype(np.float32).reshape(-1)
    if w.size != N_PARAMS:
        raise ValueError(f"{path}: {w.size} parameters, expected {N_PARAMS}")
    return w


def rollout_samples(workload: str, steps, every: int, per_read: int, seed: int = 0) -> np.ndarray:
    """F of `per_read` random particles every `every` steps up to `steps`."""
    from . import capi, workloads
    sim, info = workloads.WORKLOADS[workload](form=capi.FORM_FAST)
    rng = np.random.default_rng(seed)
    dt = info["cfg"].dt
    out = []
    done = 0
    for s in range(every, steps + 1, every):
        sim.run(s - done, dt)
        done = s
        F = sim.read(0, capi.P_F)
        ok = np.all(np.isfinite(F), axis=1)
        F = F[ok]
        k = rng.choice(F.shape[0], size=min(per_read, F.shape[0]), replace=False)
        out.append(F[k])
    sim.close()
    return np.concatenate(out).astype(np.float32)


def collect(seed: int = 0) -> np.ndarray:
    """Training set: jello drop (config 2) and snow collision (config 3)."""
    jello = rollout_samples("jello_cube", 500, 25, 20000, seed)
    snow = rollout_samples("snow_spheres", 200, 10, 20000, seed + 1)
    return np.concatenate([jello, snow])


def train(F: np.ndarray, steps: int = 3000, batch: int = 8192, lr: float = 2e-3, seed: int = 0, init=None):
    from . import capi
    p0 = init_params(seed) if init is None else np.asarray(init, np.float32)
    return capi.mlp_train(F, p0, steps, batch, lr, seed)


def main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--steps", type=int, default=3000)
    a = ap.parse_args(argv)
    t0 = time.time()
    F = collect()
    print(f"collected {F.shape[0]} deformation gradients in {time.time() - t0:.1f} s", flush=True)
    t1 = time.time()
    p, hist = train(F, steps=a.steps)
    print(f"trained {len(hist)} steps in {time.time() - t1:.1f} s; loss {hist[0]:.3e} -> {hist[-10:].mean():.3e}",
          flush=True)
    np.save(a.out, p)
    np.save(os.path.splitext(a.out)[0] + "_loss.npy", hist)
    print(f"wrote {a.out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
