#!/usr/bin/env python3
"""Throughput benchmark of the explicit MPM step on one MI355X.

Default workload: 3D snow, two jittered spheres (r = 0.15, +-3 m/s), exactly
1,000,000 particles on a 192^3 grid (dx = 1/192), dt = 1e-4, FP32 -- the
1M-particle snow configuration the metric is quoted on.  A "step" is one
full explicit step (P2G -> grid update -> G2P, plus the re-sorts of the FAST
form).  Particle state is resident in HBM before timing starts.  The timed
steps end at step 200, inside the collision (the spheres' stencils meet at
~150): an untimed pre-roll of up to 175 steps precedes the warm-up, so a
short window (e.g. 5 warm-up + 20 timed steps) times plastic flow; the
default 20 + 180 steps cover the approach and the collision.  Past step ~210
this explicit scheme is unstable at dt = 1e-4 (CFL of the hardened snow), in
the FP64 oracle as on the device; the step-size controller's leg times the
collision beyond it.

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

# Algorithmic bytes of each kernel (per particle, per grid node with mass):
# every persistent FP32 particle field of the material counted once per read
# pass and once per write pass, whatever the form keeps in between, and per
# grid node with mass 16 B for the P2G write-back and 12 B for the G2P
# velocity gather:
#   P2G reads {x, v, C, F_e, J_p} (snow) / {x, v, C, F} (jello) / {x, v, C, J} (water);
#   G2P reads {x, F_e, J_p} / {x, F} / {x, J} and writes the P2G set back.
# The FAST form's P2G reads the stress its G2P stored instead of F (and J_p);
# the figures stay those of the state the step advances.  FUSED: one kernel
# does both.
ALG_BYTES = {
    "snow": {"p2g": (100, 16), "g2p": (152, 12)},
    "jello": {"p2g": (96, 16), "g2p": (144, 12)},
    "water": {"p2g": (64, 16), "g2p": (80, 12)},
}
KERNEL_BYTES = {}
for _mat, _b in ALG_BYTES.items():
    for _form in ("plain", "fast"):
        KERNEL_BYTES[(_mat, _form)] = dict(_b)
    KERNEL_BYTES[(_mat, "fused")] = {"g2p": (_b["p2g"][0] + _b["g2p"][0], _b["p2g"][1] + _b["g2p"][1])}
KERNEL_NAME = {("plain", "p2g"): "k_p2g_atomic", ("plain", "g2p"): "k_g2p", ("fast", "p2g"): "k_p2g_block",
               ("fast", "g2p"): "k_g2p", ("fused", "g2p"): "k_g2p2g_block", ("plain", "grid"): "k_grid_update",
               ("fast", "grid"): "k_grid_update", ("fused", "grid"): "k_grid_update"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 20 + 180 steps end at step 200: the spheres touch at ~145 and the
    # collision runs stably at dt = 1e-4 until ~210, when the hardened snow's
    # wave speed passes the CFL limit (the FP64 oracle diverges at the same
    # step as the device)
    ap.add_argument("--steps", type=int, default=180)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--preroll", type=int, default=-1,
                    help="untimed steps before the warm-up (default: the snow collision window -- the timed "
                         "steps end at step 200, at most 175 steps of pre-roll; 0 for the other workloads)")
    ap.add_argument("--workload", default="snow_spheres", choices=["snow_spheres", "jello_cube", "water_dam"])
    ap.add_argument("--form", default="fast", choices=["fused", "fast", "plain"])
    ap.add_argument("--sort-every", type=int, default=0)
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-variants", action="store_true", help="skip the plain-form comparison run")
    ap.add_argument("--cpu-steps", type=int, default=20,
                    help="oracle steps for the CPU baseline and the FP32-vs-FP64 error")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batched workload")
    ap.add_argument("--no-dt-control", action="store_true",
                    help="skip the long snow run with the fixed dt and with the CFL step-size controller")
    ap.add_argument("--no-2d", action="store_true", help="skip the 2D snow workload")
    ap.add_argument("--batch-only", action="store_true", help="run only the batched workload")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB three times (HBM counter calibration)")
    return ap.parse_args()


def calibrate_copy(mib):
    """A plain copy of known size: under `rocprofv3 --pmc FETCH_SIZE` (or
    WRITE_SIZE) its per-dispatch counter reads mib * 1024 KiB when the
    counter is exact, which gives the scale applied to the MPM kernels."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "copy_kib": n * 4 // 1024, "launches": 3}), flush=True)


def _make_sim(workload, form, sort_every=0, tile=0, warm_sweeps=0):
    from mpm_cuda_amd import capi, workloads
    kw = {}
    if warm_sweeps:
        kw["warm_start"] = capi.WARM_MLP
        kw["warm_sweeps"] = warm_sweeps
    if sort_every:
        kw["sort_every"] = sort_every
    if tile:
        kw["tile"] = tile
    f = {"fused": capi.FORM_FUSED, "fast": capi.FORM_FAST, "plain": capi.FORM_PLAIN}[form]
    sim, info = workloads.WORKLOADS[workload](form=f, **kw)
    if warm_sweeps:
        from mpm_cuda_amd import warm_start
        sim.set_warm_start_weights(warm_start.load_weights())
    return sim, info


COLLISION_END = 200      # the snow spheres meet at ~150; fixed dt = 1e-4 holds until ~210


def preroll_steps(workload, steps, warmup, preroll=-1):
    """Untimed steps before the warm-up: by default the 1M snow's timed steps
    end at step 200, inside the collision (at most 175 steps of pre-roll)."""
    if preroll >= 0:
        return preroll
    if workload != "snow_spheres":
        return 0
    return max(0, min(175, COLLISION_END - warmup - steps))


def run_gpu(workload, form, steps, warmup, sort_every=0, tile=0, warm_sweeps=0, preroll=0):
    """Throughput run (no per-kernel events in the timed region), then the
    same pre-roll + warm-up + timed steps again on a fresh simulator with HIP
    events around each phase and each phase's main kernel (the events cost
    host time, so they stay out of the throughput run).  The device's SVD
    counters give the particle-steps of the timed window whose G2P took the
    Jacobi SVD (snow: trial F outside the clamp band, i.e. plastic flow),
    how many of those started from the network's quaternion and how many
    of those starts were rejected; J_p at the end gives the share of
    particles compressed plastically (J_p < 1 - 1e-3)."""
    from mpm_cuda_amd import capi
    sim, info = _make_sim(workload, form, sort_every, tile, warm_sweeps)
    n = sim.total
    dt = info["cfg"].dt
    sim.run(preroll + warmup, dt)
    sim.sync()
    sim.health()
    st0 = sim.stats()
    sv0 = sim.svd_stats()
    t0 = time.perf_counter()
    sim.run(steps, dt)
    sim.sync()
    t1 = time.perf_counter()
    health = sim.health()
    stats = sim.stats()
    sv1 = sim.svd_stats()
    stats["sorts_timed"] = stats["sorts"] - st0["sorts"]
    svd = {k: (sv1[k] - sv0[k]) % (1 << 32) for k in sv1}      # 32-bit device counters
    plastic = float(np.mean(sim.read(0, capi.P_JP)[:, 0] < 1 - 1e-3)) if info["mat"].type == capi.SNOW else 0.0
    sim.close()
    # timing pass: identical steps, events on during the timed region
    sim, _ = _make_sim(workload, form, sort_every, tile, warm_sweeps)
    sim.run(preroll + warmup, dt)
    sim.sync()
    sim.phase_times()
    sim.kernel_times()
    sim.set_timing(True)
    sim.run(steps, dt)
    sim.sync()
    ph = sim.phase_times()
    kt = sim.kernel_times()
    sim.close()
    return {"n": n, "seconds": t1 - t0, "phases": ph, "kernels": kt, "health": health, "stats": stats, "info": info,
            "svd": svd, "svd_fraction": svd["svd"] / (n * steps), "plastic_fraction": plastic}


def run_batch(steps, warmup, warm=False, autotune_reps=2, seq_steps=5):
    """The batched workload: 64 scenes (log-uniform 10k-500k particles, mixed
    jello / snow / water, 64^3-160^3 grids) advanced together, one launch per
    stage replayed from hipGraphs, launch widths from the autotuner.  With
    `warm`, snow scenes use the learned warm start.  Also times every scene
    stepped on its own (the unbatched path) for comparison."""
    from mpm_cuda_amd import capi, workloads, warm_start
    kw = {"warm_start": capi.WARM_MLP, "warm_sweeps": 2} if warm else {}
    scenes = workloads.batch_scenes(**kw)
    sims = [s for s, _ in scenes]
    if warm:
        w = warm_start.load_weights()
        for s, info in scenes:
            if info["type"] == capi.SNOW:
                s.set_warm_start_weights(w)
    n = sum(s.total for s in sims)
    b = capi.Batch(sims)
    tune = b.autotune(reps=autotune_reps)
    b.run(warmup)
    b.sync()
    st0 = [s.svd_stats() for s in sims]
    t0 = time.perf_counter()
    b.run(steps)
    b.sync()
    t1 = time.perf_counter()
    ms = b.run(max(20, steps // 2), timed=True)
    sh = b.shape
    b.close()
    st1 = [s.svd_stats() for s in sims]
    health = 0
    for s in sims:
        health |= s.health()
    snow = [k for k, (_, info) in enumerate(scenes) if info["type"] == capi.SNOW]
    warm_calls = sum(st1[k]["warm"] - st0[k]["warm"] for k in snow)
    rejects = sum(st1[k]["rejected"] - st0[k]["rejected"] for k in snow)
    # each scene on its own (mpm_run on its stream), same state
    seq = []
    for s in sims:
        s.run(1)
        s.sync()
        a0 = time.perf_counter()
        s.run(seq_steps)
        s.sync()
        seq.append((time.perf_counter() - a0) / seq_steps * 1e3)
    for s in sims:
        s.close()
    seq = np.array(seq)
    counts = np.array([info["n"] for _, info in scenes])
    return {
        "workload": f"{len(sims)} scenes, log-uniform {counts.min():,}-{counts.max():,} particles "
                    f"({n:,} total), jello/snow/water, grids 64^3-160^3, FP32",
        "particles": n, "scenes": len(sims),
        "value": round(n * steps / (t1 - t0), 1), "unit": "particle-steps/s",
        "ms_per_step": round((t1 - t0) / steps * 1e3, 4),
        "step_ms_p50": round(float(np.percentile(ms, 50)), 4), "step_ms_p99": round(float(np.percentile(ms, 99)), 4),
        "shape": {"p2g_workgroups": sh.p2g_workgroups, "grid_workgroups": sh.grid_workgroups,
                  "g2p_threads": sh.g2p_threads, "sort_every": sh.sort_every, "graph": bool(sh.use_graph)},
        "autotune_ms": {k: {str(c): round(v, 4) for c, v in d.items()} for k, d in tune.items()},
        "warm_start": "snow scenes, 2 sweeps" if warm else "none",
        "warm_svd_calls": int(warm_calls) if warm else None,
        "warm_rejected_fraction": round(rejects / warm_calls, 4) if warm and warm_calls else None,
        "health_bits": health,
        "unbatched": {"value": round(n / (seq.sum() * 1e-3), 1), "scene_step_ms_p50": round(float(np.percentile(seq, 50)), 4),
                      "scene_step_ms_p99": round(float(np.percentile(seq, 99)), 4),
                      "sum_ms_per_step": round(float(seq.sum()), 4)},
    }


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the same
    counters on a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_snow1m.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d["kernels"][kernel]["hbm_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def roofline(res, material, form, steps, workload):
    """Achieved = algorithmic bytes of the dominant kernel per launch / its
    average launch time (HIP events around that launch alone)."""
    n = res["n"]
    ph = res["phases"]
    kt = res["kernels"]
    nodes = max(ph["nodes"], 1)
    model = KERNEL_BYTES[(material, form)]
    avg = {k: kt[k][0] / max(kt[k][1], 1) for k in ("p2g", "grid", "g2p")}          # ms per launch
    bytes_per = {k: pb * n + nb * nodes for k, (pb, nb) in model.items()}
    dom = max(model, key=lambda k: avg[k])
    achieved = bytes_per[dom] / (avg[dom] * 1e-3) / 1e9
    kname = KERNEL_NAME[(form, dom)]
    traffic = measured_traffic(kname) if workload == "snow_spheres" else None
    return {"kernel": kname, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": round(traffic) if traffic else None,
            "kernel_avg_ms": {KERNEL_NAME.get((form, k), k): round(v, 4) for k, v in avg.items() if v > 0},
            "phase_avg_ms": {k: round(ph[k][0] / max(ph[k][1], 1), 4) for k in ("sort", "p2g", "grid", "g2p")},
            "bytes_per_launch": bytes_per[dom], "nodes_with_mass": ph["nodes"],
            "per_kernel_GBs": {KERNEL_NAME[(form, k)]: round(bytes_per[k] / (avg[k] * 1e-3) / 1e9, 1)
                               for k in bytes_per if avg[k] > 0}}


def cpu_baseline(workload, steps):
    """The FP64 oracle on the host cores: OpenMP over all cores available to
    this process (at most 16, the host share of one GPU), and one thread."""
    import oracle as O
    from mpm_cuda_amd import workloads
    if workload != "snow_spheres":
        return None
    cfg, mat, x, v = workloads.snow_spheres_state()
    h = cfg.h
    mu, lam = O.lame(mat.E, mat.nu)
    n = len(x)
    cores = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count() or 1))

    def rate(threads, nsteps):
        O.set_threads(threads)
        g = O.Group(x.astype(np.float64), v.astype(np.float64), type=O.SNOW, mass=mat.rho * h ** 3, vol=h ** 3,
                    mu=mu, lam=lam, theta_c=mat.theta_c, theta_s=mat.theta_s, xi=mat.xi)
        sim = O.Sim(cfg.nx, cfg.ny, cfg.nz, h, [g], gravity=tuple(cfg.gravity), bc_pad=cfg.bc_pad,
                    bc_mode=cfg.bc_mode)
        t0 = time.perf_counter()
        for _ in range(nsteps):
            sim.step(cfg.dt)
        dt_s = time.perf_counter() - t0
        return n * nsteps / dt_s, dt_s, sim

    multi, tm, orc = rate(cores, max(steps, 2))
    single, ts, _ = rate(1, 1)
    O.set_threads(1)
    return {"value": round(multi, 1), "unit": "particle-steps/s", "cores": cores, "kind": "port",
            "sample": f"FP64 oracle (oracle/mpm_oracle.c, OpenMP, atomic P2G node sums) on all {n:,} particles of "
                      f"the same workload, {max(steps, 2)} steps in {tm:.1f} s on {cores} threads; one thread: "
                      f"{single:.0f} particle-steps/s (1 step, {ts:.1f} s); dense FP64 grid as in the reference",
            "single_thread_value": round(single, 1)}, orc


def run_dt_control(steps=600, every=20, max_dt=1e-4, cfl=0.5):
    """The 1M-particle snow collision run past the point where the fixed dt
    of its configuration turns unstable (the hardened snow's wave speed
    outgrows h / dt): fixed dt until the health check fails, then the CFL
    controller (dt_mode CFL) for `steps` steps.  Reports the first failing
    check of the fixed run, and for the controlled run the simulated time,
    health, the dt range and the throughput."""
    from mpm_cuda_amd import capi, workloads
    out = {}
    for name, mode in (("fixed", capi.DT_FIXED), ("cfl", capi.DT_CFL)):
        sim, info = workloads.snow_spheres(form=capi.FORM_FAST, dt_mode=mode, cfl=cfl)
        n = sim.total
        dts, bad, k = [], None, 0
        t0 = time.perf_counter()
        while k < steps and bad is None:
            sim.run(every, max_dt)
            k += every
            if sim.health() & ~capi.H_ARENA_SPILL:
                bad = k
            dts.append(sim.last_dt)
        t1 = time.perf_counter()
        out[name] = {"steps": k, "failed_by_step": bad, "sim_time": round(sim.current_time, 6),
                     "dt_min": min(dts), "dt_max": max(dts)}
        if bad is None:
            out[name]["value"] = round(n * k / (t1 - t0), 1)
            out[name]["unit"] = "particle-steps/s (health checks every %d steps included)" % every
        sim.close()
    # the collision under the controller, timed without health checks:
    # 175 untimed steps, 20 warm-up, then `window` timed steps (past the
    # step where the fixed dt fails), then the health check
    pre, wu, window = 175, 20, 300
    sim, info = workloads.snow_spheres(form=capi.FORM_FAST, dt_mode=capi.DT_CFL, cfl=cfl)
    sim.run(pre + wu, max_dt)
    dt0 = sim.last_dt
    sim.sync()
    t0 = time.perf_counter()
    sim.run(window, max_dt)
    sim.sync()
    t1 = time.perf_counter()
    out["cfl_collision"] = {"timed_steps": [pre + wu, pre + wu + window], "value": round(sim.total * window / (t1 - t0), 1),
                            "unit": "particle-steps/s", "ms_per_step": round((t1 - t0) / window * 1e3, 4),
                            "dt_first": dt0, "dt_last": sim.last_dt, "sim_time": round(sim.current_time, 6),
                            "health_bits": sim.health() & ~capi.H_ARENA_SPILL}
    sim.close()
    return out


def run_2d(oracle_steps=400):
    """The reference's 2D snow step (MPMSnow.py) on its 9,000-particle
    workload for its 2000 steps, timed on the device from the initial state
    (hipGraph replays of 50 steps; the capture is inside the timed region),
    and the FP64 2D oracle on one host thread over the first `oracle_steps`
    steps of the same workload."""
    import oracle as O
    from mpm_cuda_amd import capi, workloads
    warm, info = workloads.snow2d_squares()          # loads the kernels
    warm.run(60)
    warm.sync()
    warm.close()
    sim, info = workloads.snow2d_squares()
    n, steps = sim.n, info["steps"]
    t0 = time.perf_counter()
    sim.run(steps)
    sim.sync()
    t1 = time.perf_counter()
    health = sim.health()
    sim.close()
    c = info["cfg"]
    o = O.Sim2D(info["x"].astype(np.float64), 0.0, n_grid=c.n_grid, wall=c.wall, E=c.E, nu=c.nu, rho=c.rho,
                xi=c.xi, theta_c=c.theta_c, theta_s=c.theta_s, gravity=tuple(c.gravity))
    a0 = time.perf_counter()
    for _ in range(oracle_steps):
        o.step(c.dt)
    ta = time.perf_counter() - a0
    return {"workload": info["desc"], "particles": n, "steps": steps, "value": round(n * steps / (t1 - t0), 1),
            "unit": "particle-steps/s", "ms_per_step": round((t1 - t0) / steps * 1e3, 5),
            "seconds": round(t1 - t0, 4), "health_bits": health,
            "cpu_baseline": {"value": round(n * oracle_steps / ta, 1), "unit": "particle-steps/s", "cores": 1,
                             "kind": "port", "sample": f"FP64 2D oracle (oracle/mpm_oracle2d.c), first "
                                                       f"{oracle_steps} steps of the same workload, {ta:.1f} s"}}


def _field_errors(g, ref, h):
    out = {}
    vscale = float(np.abs(ref["v"]).max()) / h          # scale of a velocity gradient
    for k, b in ref.items():
        a = np.asarray(g[k], np.float64).reshape(b.shape)
        nb = np.linalg.norm(b)
        e = {"rel_l2": float(f"{np.linalg.norm(a - b) / nb:.3e}") if nb > 0 else None,
             "max_abs": float(f"{np.abs(a - b).max():.3e}")}
        if k == "C":
            e["max_abs_over_vmax_per_h"] = float(f"{np.abs(a - b).max() / max(vscale, 1e-300):.3e}")
        out[k] = e
    return out


def _gpu_fields(sim):
    from mpm_cuda_amd import capi
    return {"x": sim.read(0, capi.P_POS), "v": sim.read(0, capi.P_VEL), "C": sim.read(0, capi.P_C),
            "F": sim.read(0, capi.P_F), "J_p": sim.read(0, capi.P_JP)[:, 0]}


def _orc_fields(orc):
    o = orc.groups[0]
    return {"x": o.x, "v": o.v, "C": o.C.reshape(-1, 9), "F": o.Fe.reshape(-1, 9), "J_p": np.linalg.det(o.Fp)}


def fp32_vs_fp64(workload, orc, steps, restart_at=150):
    """Per-field error of the device FP32 FAST form against the FP64 oracle:
    (a) after `steps` steps from the same initial state (the oracle run of the
    CPU baseline); (b) in the collision: the device state after `restart_at`
    steps is copied into the oracle (J_p as F_p = J_p^(1/3) I, the only use
    of F_p) and both advance `steps` more steps.  Relative L2 error
    ||a - b|| / ||b|| and largest absolute difference per field; C also
    against the velocity-gradient scale max|v| / h (C is ~0 before impact)."""
    import oracle as O
    sim, info = _make_sim(workload, "fast")
    cfg, mat = info["cfg"], info["mat"]
    h = cfg.h
    sim.run(steps, cfg.dt)
    sim.sync()
    out = {"from_start": {"steps": steps, **_field_errors(_gpu_fields(sim), _orc_fields(orc), h)}}
    sim.run(restart_at - steps, cfg.dt)
    sim.sync()
    g0 = _gpu_fields(sim)
    mu, lam = O.lame(mat.E, mat.nu)
    n = g0["x"].shape[0]
    grp = O.Group(g0["x"].astype(np.float64), g0["v"].astype(np.float64), type=O.SNOW, mass=mat.rho * h ** 3,
                  vol=h ** 3, mu=mu, lam=lam, theta_c=mat.theta_c, theta_s=mat.theta_s, xi=mat.xi,
                  C=g0["C"].astype(np.float64).reshape(n, 3, 3), Fe=g0["F"].astype(np.float64).reshape(n, 3, 3),
                  Fp=np.cbrt(g0["J_p"].astype(np.float64))[:, None, None] * np.eye(3))
    o2 = O.Sim(cfg.nx, cfg.ny, cfg.nz, h, [grp], gravity=tuple(cfg.gravity), bc_pad=cfg.bc_pad, bc_mode=cfg.bc_mode)
    O.set_threads(max(1, min(16, len(os.sched_getaffinity(0)))))
    for _ in range(steps):
        o2.step(cfg.dt)
    O.set_threads(1)
    sim.run(steps, cfg.dt)
    sim.sync()
    out["in_collision"] = {"start_step": restart_at, "steps": steps, **_field_errors(_gpu_fields(sim), _orc_fields(o2), h)}
    out["health_bits"] = sim.health()
    sim.close()
    return out


def main():
    a = parse()
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    if a.batch_only:
        print(json.dumps({"batched": run_batch(a.steps, a.warmup),
                          "batched_warm": run_batch(a.steps, a.warmup, warm=True)}), flush=True)
        return
    material = {"snow_spheres": "snow", "jello_cube": "jello", "water_dam": "water"}[a.workload]
    pre = preroll_steps(a.workload, a.steps, a.warmup, a.preroll)
    res = run_gpu(a.workload, a.form, a.steps, a.warmup, a.sort_every, a.tile, preroll=pre)
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
                   "sort_every": int(res["info"]["cfg"].sort_every), "tile": int(res["info"]["cfg"].tile),
                   "timed_steps": [pre + a.warmup, pre + a.warmup + a.steps]},
        "roofline": roofline(res, material, a.form, a.steps, a.workload),
        "health_bits": res["health"],
        "fp64_svd_fallbacks": res["stats"]["fp64_fallbacks"],
        "sorts_in_timed_steps": int(res["stats"]["sorts_timed"]),
    }
    # particle-steps of the timed window whose G2P took the Jacobi SVD (snow:
    # trial F outside the clamp band, plastic flow) instead of the elastic
    # path, and the particles compressed plastically by its end
    out["svd_path_fraction"] = round(res["svd_fraction"], 5)
    if material == "snow":
        out["plastic_fraction"] = round(res["plastic_fraction"], 5)
    # the achievable bandwidth beside the published peak
    from mpm_cuda_amd import capi
    cp = round(capi.probe_copy_bandwidth(1 << 30, 10), 1)
    out["roofline"]["copy_GBs"] = cp
    out["roofline"]["frac_of_copy"] = round(out["roofline"]["achieved"] / cp, 4)
    if not a.no_variants:
        ksteps = max(10, a.steps // 4)
        out["variants"] = {}
        kwu = min(a.warmup, 5)
        kpre = preroll_steps(a.workload, ksteps, kwu, a.preroll)
        for form in ("fast", "fused", "plain"):
            if form == a.form:
                continue
            r = run_gpu(a.workload, form, ksteps, kwu, preroll=kpre)
            v = r["n"] * ksteps / r["seconds"]
            out["variants"][form] = {"value": round(v, 1), "ms_per_step": round(r["seconds"] / ksteps * 1e3, 4),
                                     "roofline": roofline(r, material, form, ksteps, a.workload)}
        # learned warm start of the G2P SVD: network quaternion + 2 Jacobi
        # sweeps, cold 4-sweep SVD for every particle above the 1e-6 bound
        # (the network runs for the particles whose previous G2P took the SVD)
        r = run_gpu(a.workload, a.form, ksteps, kwu, warm_sweeps=2, preroll=kpre)
        rc = run_gpu(a.workload, a.form, ksteps, kwu, preroll=kpre)
        v = r["n"] * ksteps / r["seconds"]
        sv = r["svd"]
        out["variants"]["warm_start"] = {
            "value": round(v, 1), "ms_per_step": round(r["seconds"] / ksteps * 1e3, 4), "form": a.form,
            "warm_sweeps": 2, "timed_steps": [kpre + kwu, kpre + kwu + ksteps],
            "svd_path_fraction": round(r["svd_fraction"], 5),
            "svd_calls": sv["svd"], "warm_svd_calls": sv["warm"], "warm_rejects": sv["rejected"],
            "rejected_fraction": round(sv["rejected"] / sv["warm"], 4) if sv["warm"] else None,
            "kernel_avg_ms": roofline(r, material, a.form, ksteps, a.workload)["kernel_avg_ms"],
            "cold_same_window": {"value": round(rc["n"] * ksteps / rc["seconds"], 1),
                                 "ms_per_step": round(rc["seconds"] / ksteps * 1e3, 4),
                                 "kernel_avg_ms": roofline(rc, material, a.form, ksteps, a.workload)["kernel_avg_ms"]}}
        if "plain" in out["variants"]:
            out["variants"]["speedup_over_plain"] = round(value / out["variants"]["plain"]["value"], 3)
    if not a.no_2d:
        out["snow2d"] = run_2d()
    if not a.no_dt_control and a.workload == "snow_spheres":
        out["dt_control"] = run_dt_control()
    if not a.no_batch:
        out["batched"] = run_batch(a.steps, a.warmup)
        out["batched_warm"] = run_batch(a.steps, a.warmup, warm=True)
    if not a.no_cpu_baseline:
        cb = cpu_baseline(a.workload, a.cpu_steps)
        if cb is not None:
            out["cpu_baseline"], orc = cb
            out["fp32_vs_fp64"] = fp32_vs_fp64(a.workload, orc, max(a.cpu_steps, 2))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
