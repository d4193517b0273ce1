"""Named synthetic workloads (seeded, FP32 initial states).

* ``snow_spheres``  -- 3D snow, 192^3 grid (dx = 1/192), two jittered spheres
  r = 0.15 at (0.3, 0.5, 0.5) and (0.7, 0.5, 0.5) moving at +-3 m/s along x,
  5 samples per cell, trimmed to 500,000 particles each (1,000,000 total);
  E = 1.4e5, nu = 0.2, rho = 400, theta_c = 2.5e-2, theta_s = 7.5e-3, xi = 10,
  dt = 1e-4.
* ``jello_cube``    -- 3D jello, 128^3 grid (dx = 1/128), a 0.25^3 cube of
  32^3 cells x 8 samples = 262,144 particles, bottom at y = 0.1, v0 = (0,-2,0),
  E = 5e4, nu = 0.3, rho = 1e3, sticky walls, dt = 1e-4.
* ``water_dam``     -- 3D water dam break, 256^3 grid (dx = 1/256), a
  0.5 x 0.4 x 0.5 block in the lower-left corner (inside the 10-cell wall
  band), 2,000,000 particles, k = 1e4, gamma = 7, C = grad v, dt = 5e-5.

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


WORKLOADS = {"snow_spheres": snow_spheres, "jello_cube": jello_cube, "water_dam": water_dam}
