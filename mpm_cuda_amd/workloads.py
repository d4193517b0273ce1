"""Named synthetic workloads (seeded, FP32 initial states).

* ``snow_spheres``  -- 3D snow, 192^3 grid (dx = 1/192), two jittered spheres
  r = 0.15 at (0.3, 0.5, 0.5) and (0.7, 0.5, 0.5) moving at +-3 m/s along x,
  5 samples per cell, trimmed to 500,000 particles each (1,000,000 total);
  E = 1.4e5, nu = 0.2, rho = 400, theta_c = 2.5e-2, theta_s = 7.5e-3, xi = 10,
  dt = 1e-4.
* ``jello_cube``    -- 3D jello, 128^3 grid (dx = 1/128), a 0.25^3 cube
  (32^3 cells' worth at 8 samples per cell; its bottom at y = 0.1 is not on
  a cell face, and 262,146 of the jittered samples fall inside), v0 = (0,-2,0),
  E = 5e4, nu = 0.3, rho = 1e3, sticky walls, dt = 1e-4.
* ``water_dam``     -- 3D water dam break, 256^3 grid (dx = 1/256), a
  0.5 x 0.4 x 0.5 block in the lower-left corner (inside the 10-cell wall
  band), 2,000,000 particles, k = 1e4, gamma = 7, C = grad v, dt = 5e-5.

* ``batch_scenes``  -- 64 independent scenes for the batched form: particle
  counts log-uniform in [10,000, 500,000] (seeded), materials cycling jello /
  snow / water, grids of 64^3 to 160^3 nodes (dx = 1/N, 3-cell walls).
  Jello: a cube dropped at 2 m/s onto a sticky floor (E = 5e4, nu = 0.3);
  snow: a ball thrown at (2, -3, 0) m/s (the snow material above); water: a
  dam-break block in the lower corner (k = 1e4, gamma = 7, dt = 5e-5).  8
  samples per cell, trimmed to the drawn count.

* ``snow2d_squares`` -- the reference's 2D snow step (MPMSnow.py): 128^2 grid
  (dx = 1/128), three 0.2-side squares of 3,000 seeded uniform-random particles
  each, lower-left corners at (0.3 + 0.1 g, 0.05 + 0.32 g), starting at rest
  (the script's (0, -6.4) m/s start is unstable with E = 1e3 at dt = 1e-4: the
  FP64 oracle blows up at step ~630), E = 1e3, nu = 0.2, rho = 1,
  theta_c = 2.5e-2, theta_s = 7.5e-3, xi = 10, gravity -50, dt = 1e-4, 2000
  steps, FP32.

Each builder returns ``(sim, info)`` where ``info`` holds the arrays needed
to rebuild the same state elsewhere (positions, velocities, material).
"""
from __future__ import annotations

import numpy as np

from . import capi

SEED0 = 20261202


def _sphere(cfg, center, radius, mat, max_particles):
    o = capi.MpmObject()
    o.shape = capi.SHAPE_SPHERE
    o.obj_path = b""
    o.shape_param = (capi.C.c_double * 6)(*center, radius, 0.0, 0.0)
    o.translation = (capi.C.c_float * 3)(0, 0, 0)
    o.scale = (capi.C.c_float * 3)(1, 1, 1)
    o.rotation_axis = (capi.C.c_float * 3)(1, 0, 0)
    o.rotation_deg = 0.0
    o.material = mat
    o.max_particles = max_particles
    return capi.sample_object(cfg, o)


def _box(cfg, lo, hi, mat, max_particles=0):
    o = capi.MpmObject()
    o.shape = capi.SHAPE_BOX
    o.obj_path = b""
    o.shape_param = (capi.C.c_double * 6)(*lo, *hi)
    o.translation = (capi.C.c_float * 3)(0, 0, 0)
    o.scale = (capi.C.c_float * 3)(1, 1, 1)
    o.rotation_axis = (capi.C.c_float * 3)(1, 0, 0)
    o.rotation_deg = 0.0
    o.material = mat
    o.max_particles = max_particles
    return capi.sample_object(cfg, o)


def snow_spheres_state(n_grid=192, per_sphere=500_000, ppc=5, seed=SEED0 + 3):
    cfg = capi.default_config(h=1.0 / n_grid, nx=n_grid, ny=n_grid, nz=n_grid, dt=1e-4, ppc=ppc, seed=seed)
    mat = capi.material(1.4e5, 0.2, 400.0, capi.SNOW, theta_c=2.5e-2, theta_s=7.5e-3, xi=10.0)
    a = _sphere(cfg, (0.3, 0.5, 0.5), 0.15, mat, per_sphere)
    b = _sphere(cfg, (0.7, 0.5, 0.5), 0.15, mat, per_sphere)
    x = np.concatenate([a, b])
    v = np.zeros_like(x)
    v[: len(a), 0] = 3.0
    v[len(a):, 0] = -3.0
    return cfg, mat, x, v


def snow_spheres(form=capi.FORM_FAST, n_grid=192, per_sphere=500_000, **cfg_kw):
    cfg, mat, x, v = snow_spheres_state(n_grid, per_sphere)
    cfg.form = form
    for k, val in cfg_kw.items():
        setattr(cfg, k, val)
    sim = capi.Sim(cfg)
    sim.add_particles(mat, x, vel=v, volume=cfg.h ** 3)
    return sim, {"cfg": cfg, "mat": mat, "x": x, "v": v, "name": "snow_spheres",
                 "desc": f"3D snow, two jittered spheres r=0.15 at +-3 m/s, {len(x):,} particles, "
                         f"{n_grid}^3 grid, dt=1e-4, FP32"}


def jello_cube(form=capi.FORM_FAST, n_grid=128, **cfg_kw):
    cfg = capi.default_config(h=1.0 / n_grid, nx=n_grid, ny=n_grid, nz=n_grid, dt=1e-4, ppc=8,
                              seed=SEED0 + 2, bc_mode=capi.BC_STICKY, form=form)
    for k, val in cfg_kw.items():
        setattr(cfg, k, val)
    mat = capi.material(5e4, 0.3, 1e3, capi.JELLO)
    side = 32 / n_grid
    lo = (0.5 - side / 2, 0.1, 0.5 - side / 2)
    hi = (lo[0] + side, lo[1] + side, lo[2] + side)
    x = _box(cfg, lo, hi, mat)
    sim = capi.Sim(cfg)
    sim.add_particles(mat, x, v0=(0.0, -2.0, 0.0), volume=cfg.h ** 3)
    v = np.tile(np.array([[0.0, -2.0, 0.0]], np.float32), (len(x), 1))
    return sim, {"cfg": cfg, "mat": mat, "x": x, "v": v, "name": "jello_cube",
                 "desc": f"3D jello cube, {len(x):,} particles, {n_grid}^3 grid, dt=1e-4, FP32"}


def water_dam(form=capi.FORM_FAST, n_grid=256, n_particles=2_000_000, **cfg_kw):
    cfg = capi.default_config(h=1.0 / n_grid, nx=n_grid, ny=n_grid, nz=n_grid, dt=5e-5, ppc=2,
                              seed=SEED0 + 4, form=form)
    for k, val in cfg_kw.items():
        setattr(cfg, k, val)
    mat = capi.material(1e4, 0.2, 1e3, capi.WATER, eos_k=1e4, eos_gamma=7.0,
                        water_c_mode=capi.WATER_C_GRADV)
    pad = 11.0 / n_grid
    x = _box(cfg, (pad, pad, pad), (pad + 0.5, pad + 0.4, pad + 0.5), mat, max_particles=n_particles)
    sim = capi.Sim(cfg)
    sim.add_particles(mat, x, volume=cfg.h ** 3)
    return sim, {"cfg": cfg, "mat": mat, "x": x, "v": np.zeros_like(x), "name": "water_dam",
                 "desc": f"3D water dam break, {len(x):,} particles, {n_grid}^3 grid, dt=5e-5, FP32"}


BATCH_GRIDS = (64, 80, 96, 112, 128, 144, 160)


def batch_plan(n_scenes=64, seed=SEED0 + 5, n_min=10_000, n_max=500_000):
    """(material, grid, particles) of every scene of the batched workload."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(n_min), np.log(n_max), n_scenes)).astype(np.int64)
    plan = []
    for k, n in enumerate(counts):
        side = (n / 8.0) ** (1.0 / 3.0)                  # cells of a cube holding n at 8 per cell
        fits = [N for N in BATCH_GRIDS if N - 8 >= 1.9 * side]
        plan.append(((capi.JELLO, capi.SNOW, capi.WATER)[k % 3], int(rng.choice(fits)), int(n)))
    return plan


def batch_scene(type_, n_grid, n_particles, seed, form=capi.FORM_FAST, **cfg_kw):
    h = 1.0 / n_grid
    pad = 3
    dt = 5e-5 if type_ == capi.WATER else 1e-4
    cfg = capi.default_config(h=h, nx=n_grid, ny=n_grid, nz=n_grid, dt=dt, ppc=8, seed=seed, bc_pad=pad,
                              bc_mode=capi.BC_STICKY if type_ == capi.JELLO else capi.BC_AXIS, form=form)
    for k, val in cfg_kw.items():
        setattr(cfg, k, val)
    side = 1.12 * (n_particles / 8.0) ** (1.0 / 3.0) * h    # sampled a little larger, then trimmed
    lo_wall = (pad + 2) * h
    v = (0.0, 0.0, 0.0)
    if type_ == capi.JELLO:
        mat = capi.material(5e4, 0.3, 1e3, capi.JELLO)
        lo = (0.5 - side / 2, lo_wall + 0.1 * (1 - side), 0.5 - side / 2)
        x = _box(cfg, lo, tuple(a + side for a in lo), mat, max_particles=n_particles)
        v = (0.0, -2.0, 0.0)
    elif type_ == capi.SNOW:
        mat = capi.material(1.4e5, 0.2, 400.0, capi.SNOW, theta_c=2.5e-2, theta_s=7.5e-3, xi=10.0)
        r = side * (3.0 / (4.0 * np.pi)) ** (1.0 / 3.0) * 1.05
        x = _sphere(cfg, (0.45, 0.55, 0.5), r, mat, n_particles)
        v = (2.0, -3.0, 0.0)
    else:
        mat = capi.material(1e4, 0.2, 1e3, capi.WATER, eos_k=1e4, eos_gamma=7.0, water_c_mode=capi.WATER_C_GRADV)
        # 0.5 : 0.4 : 0.5 block of the same volume
        a = side / (0.1 ** (1.0 / 3.0))
        lo = (lo_wall, lo_wall, lo_wall)
        x = _box(cfg, lo, (lo_wall + 0.5 * a, lo_wall + 0.4 * a, lo_wall + 0.5 * a), mat, max_particles=n_particles)
    sim = capi.Sim(cfg)
    sim.add_particles(mat, x, v0=v, volume=cfg.h ** 3)
    return sim, {"cfg": cfg, "mat": mat, "n": len(x), "type": type_, "grid": n_grid, "dt": dt}


def batch_scenes(n_scenes=64, seed=SEED0 + 5, form=capi.FORM_FAST, **cfg_kw):
    """The batched workload: a list of (sim, info)."""
    return [batch_scene(t, N, n, seed + 17 * k, form=form, **cfg_kw)
            for k, (t, N, n) in enumerate(batch_plan(n_scenes, seed))]


def snow2d_state(seed=SEED0 + 6, per_square=3000):
    """Positions [9000][2] (FP32) of the three squares."""
    rng = np.random.default_rng(seed)
    sq = [rng.random((per_square, 2)) * 0.2 + np.array([0.3 + 0.1 * g, 0.05 + 0.32 * g]) for g in range(3)]
    return np.concatenate(sq).astype(np.float32)


def snow2d_squares(**cfg_kw):
    x = snow2d_state()
    sim = capi.Sim2D(**cfg_kw)
    sim.add_particles(x)
    return sim, {"x": x, "cfg": sim.cfg, "name": "snow2d_squares", "steps": 2000,
                 "desc": f"2D snow, three 0.2 squares, {len(x):,} particles, 128^2 grid, dt=1e-4, FP32"}


WORKLOADS = {"snow_spheres": snow_spheres, "jello_cube": jello_cube, "water_dam": water_dam}
