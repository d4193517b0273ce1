#!/usr/bin/env python3
"""Throughput benchmark of the explicit MPM step on one MI355X.

Default workload: 3D snow, two jittered spheres (r = 0.15, +-3 m/s), exactly
1,000,000 particles on a 192^3 grid (dx = 1/192), dt = 1e-4, FP32 -- the
1M-particle snow configuration the metric is quoted on.  A "step" is one
full explicit step (P2G -> grid update -> G2P, plus the periodic re-sort of
the FAST form).  Particle state is resident in HBM before timing starts.

Prints one JSON line: particle-steps/s, ms per step, the roofline of the
dominant kernel (algorithmic bytes / average launch time from HIP events on
the simulator's stream, against 8 TB/s) and the CPU baseline (the FP64
oracle, single thread, on a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per particle of each kernel: every persistent FP32
# particle field counted once per read pass and once per write pass; grid
# nodes with mass counted at 16 B (P2G flush), 12 B (G2P gather).
# FAST form: P2G reads {x, v, C, tau(6)}; G2P reads {x, F, J_p} and writes
# {x, v, C, F, J_p, tau}.  PLAIN form: P2G reads {x, v, C, F, J_p}; G2P reads
# {x, F, J_p} and writes {x, v, C, F, J_p}.
PART_BYTES = {
    # (material, form): (p2g_read, g2p_read, g2p_write)
    ("snow", "fast"): (84, 52, 124), ("snow", "plain"): (100, 52, 100),
    ("jello", "fast"): (84, 48, 120), ("jello", "plain"): (96, 48, 96),
    ("water", "fast"): (64, 16, 64), ("water", "plain"): (64, 16, 64),
}
NODE_BYTES = {"p2g": 16, "g2p": 12, "grid": 48}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="snow_spheres", choices=["snow_spheres", "jello_cube", "water_dam"])
    ap.add_argument("--form", default="fast", choices=["fast", "plain"])
    ap.add_argument("--sort-every", type=int, default=0)
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-variants", action="store_true", help="skip the plain-form comparison run")
    ap.add_argument("--cpu-steps", type=int, default=4)
    return ap.parse_args()


def run_gpu(workload, form, steps, warmup, sort_every=0, tile=0, timing=True):
    from mpm_cuda_amd import capi, workloads
    kw = {}
    if sort_every:
        kw["sort_every"] = sort_every
    if tile:
        kw["tile"] = tile
    f = capi.FORM_FAST if form == "fast" else capi.FORM_PLAIN
    sim, info = workloads.WORKLOADS[workload](form=f, **kw)
    n = sim.total
    dt = info["cfg"].dt
    sim.run(warmup, dt)
    sim.sync()
    sim.phase_times()          # drop warm-up records
    sim.health()
    sim.set_timing(timing)
    t0 = time.perf_counter()
    sim.run(steps, dt)
    sim.sync()
    t1 = time.perf_counter()
    ph = sim.phase_times()
    health = sim.health()
    stats = sim.stats()
    sim.close()
    return {"n": n, "seconds": t1 - t0, "phases": ph, "health": health, "stats": stats, "info": info}


def roofline(res, material, form, steps):
    n = res["n"]
    ph = res["phases"]
    nodes = max(ph["nodes"], 1)
    p2g_r, g2p_r, g2p_w = PART_BYTES[(material, form)]
    bytes_per = {
        "p2g": p2g_r * n + NODE_BYTES["p2g"] * nodes,
        "g2p": (g2p_r + g2p_w) * n + NODE_BYTES["g2p"] * nodes,
    }
    avg = {k: ph[k][0] / max(ph[k][1], 1) for k in ("sort", "p2g", "grid", "g2p")}   # ms per launch
    dom = max(("p2g", "g2p"), key=lambda k: avg[k])
    achieved = bytes_per[dom] / (avg[dom] * 1e-3) / 1e9
    return {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "avg_launch_ms": {k: round(v, 4) for k, v in avg.items()},
            "bytes_per_launch": bytes_per[dom], "nodes_with_mass": ph["nodes"],
            "other": {k: round(bytes_per[k] / (avg[k] * 1e-3) / 1e9, 1) for k in bytes_per if avg[k] > 0}}


def cpu_baseline(workload, steps):
    """FP64 oracle, one host thread, on the same initial state."""
    import oracle as O
    from mpm_cuda_amd import workloads
    if workload != "snow_spheres":
        return None
    cfg, mat, x, v = workloads.snow_spheres_state()
    h = cfg.h
    mu, lam = O.lame(mat.E, mat.nu)
    g = O.Group(x.astype(np.float64), v.astype(np.float64), type=O.SNOW, mass=mat.rho * h ** 3, vol=h ** 3, mu=mu,
                lam=lam, theta_c=mat.theta_c, theta_s=mat.theta_s, xi=mat.xi)
    sim = O.Sim(cfg.nx, cfg.ny, cfg.nz, h, [g], gravity=tuple(cfg.gravity), bc_pad=cfg.bc_pad, bc_mode=cfg.bc_mode)
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step(cfg.dt)
    t1 = time.perf_counter()
    n = len(x)
    return {"value": round(n * steps / (t1 - t0), 1), "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"FP64 oracle (oracle/mpm_oracle.c), all {n:,} particles of the same workload, {steps} steps, "
                      f"dense FP64 grid as in the reference; {t1 - t0:.1f} s"}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    material = {"snow_spheres": "snow", "jello_cube": "jello", "water_dam": "water"}[a.workload]
    res = run_gpu(a.workload, a.form, a.steps, a.warmup, a.sort_every, a.tile)
    n = res["n"]
    value = n * a.steps / res["seconds"]
    out = {
        "metric": "particle-steps/s (1M-particle 3D snow)" if a.workload == "snow_spheres" else "particle-steps/s",
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(res["seconds"] / a.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": res["info"]["desc"], "particles": n, "form": a.form, "warm_start": "none",
                   "sort_every": int(res["info"]["cfg"].sort_every), "tile": int(res["info"]["cfg"].tile)},
        "roofline": roofline(res, material, a.form, a.steps),
        "health_bits": res["health"],
        "fp64_svd_fallbacks": res["stats"]["fp64_fallbacks"],
    }
    if not a.no_variants and a.form == "fast":
        ksteps = max(10, a.steps // 4)
        pl = run_gpu(a.workload, "plain", ksteps, min(a.warmup, 5))
        out["variants"] = {"plain": {"value": round(pl["n"] * ksteps / pl["seconds"], 1),
                                     "ms_per_step": round(pl["seconds"] / ksteps * 1e3, 4),
                                     "roofline": roofline(pl, material, "plain", ksteps)}}
        out["variants"]["fast_over_plain"] = round(value / out["variants"]["plain"]["value"], 3)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(a.workload, a.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
