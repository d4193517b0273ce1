"""Shared scene builders for the parity tests: identical FP32 initial states
for the HIP library (through the C ABI) and the FP64 oracle."""
from __future__ import annotations

import numpy as np

import oracle as O
from mpm_cuda_amd import capi

MAT_PARAMS = {
    # type: (E, nu, rho)
    capi.JELLO: (5e4, 0.3, 1e3),
    capi.SNOW: (1.4e5, 0.2, 400.0),
    capi.WATER: (1e4, 0.2, 1e3),
}


def jittered_box(lo, hi, h, ppc, seed):
    """ppc uniform samples in each cell of the box [lo, hi) (cell aligned), FP32."""
    rng = np.random.default_rng(seed)
    lo_i = np.floor(np.asarray(lo) / h).astype(int)
    hi_i = np.floor(np.asarray(hi) / h).astype(int)
    cells = np.stack(np.meshgrid(*[np.arange(a, b) for a, b in zip(lo_i, hi_i)], indexing="ij"), -1).reshape(-1, 3)
    pts = (cells[:, None, :] + rng.random((cells.shape[0], ppc, 3))) * h
    return pts.reshape(-1, 3).astype(np.float32)


def random_state(n, seed, type_, *, v_scale=1.0, c_scale=20.0, f_scale=0.05):
    """Random particle state near the rest state (FP32)."""
    rng = np.random.default_rng(seed + 1)
    v = (v_scale * rng.normal(size=(n, 3))).astype(np.float32)
    C = (c_scale * rng.normal(size=(n, 3, 3))).astype(np.float32)
    F = (np.eye(3) + f_scale * rng.normal(size=(n, 3, 3))).astype(np.float32)
    S = (1.0 + 0.02 * rng.normal(size=n)).astype(np.float32)   # J_p (snow) / J (water)
    if type_ == capi.SNOW:
        # keep F inside the plastic band so the state is reachable
        U, s, Vt = np.linalg.svd(F.astype(np.float64))
        s = np.clip(s, 1 - 2.5e-2, 1 + 7.5e-3)
        F = (U * s[:, None, :] @ Vt).astype(np.float32)
    return v, C, F, S


class Pair:
    """The same scene in the HIP library and in the oracle."""

    def __init__(self, *, n_grid=64, type_=capi.JELLO, form=capi.FORM_PLAIN, ppc=4, box=((0.3, 0.3, 0.3), (0.55, 0.55, 0.55)),
                 seed=1, randomize=True, dt=1e-4, bc_pad=3, bc_mode=capi.BC_AXIS, gravity=(0.0, -9.8, 0.0),
                 theta_c=2.5e-2, theta_s=7.5e-3, eos_k=1e4, eos_gamma=7.0, sort_every=4, water_c_mode=0,
                 sweeps=4, tile=0, extra_cfg=None):
        h = 1.0 / n_grid
        self.h, self.dt = h, dt
        x = jittered_box(box[0], box[1], h, ppc, seed)
        n = x.shape[0]
        self.n = n
        E, nu, rho = MAT_PARAMS[type_]
        cfg = capi.default_config(h=h, nx=n_grid, ny=n_grid, nz=n_grid, dt=dt, bc_pad=bc_pad, bc_mode=bc_mode,
                                  gravity=gravity, form=form, sort_every=sort_every, svd_sweeps=sweeps, tile=tile,
                                  **(extra_cfg or {}))
        self.gpu = capi.Sim(cfg)
        mat = capi.material(E, nu, rho, type_, theta_c=theta_c, theta_s=theta_s, eos_k=eos_k, eos_gamma=eos_gamma,
                            water_c_mode=water_c_mode)
        vol = h ** 3
        mass = rho * vol
        if randomize:
            v, C, F, S = random_state(n, seed, type_)
        else:
            v = np.tile(np.array([[0.3, -1.0, 0.5]], np.float32), (n, 1))
            C = np.zeros((n, 3, 3), np.float32)
            F = np.tile(np.eye(3, dtype=np.float32), (n, 1, 1))
            S = np.ones(n, np.float32)
        self.gpu.add_particles(mat, x, vel=v, volume=vol)
        self.gpu.write(0, capi.P_C, C.reshape(n, 9))
        self.gpu.write(0, capi.P_F, F.reshape(n, 9))
        self.gpu.write(0, capi.P_JP, S.reshape(n, 1))
        x64 = x.astype(np.float64)
        Fp = None
        J = None
        if type_ == capi.SNOW:
            Fp = np.cbrt(S.astype(np.float64))[:, None, None] * np.eye(3)
        if type_ == capi.WATER:
            J = S.astype(np.float64)
        grp = O.Group(x64, v.astype(np.float64), type=type_, mass=mass, vol=vol, E=E, nu=nu, theta_c=theta_c,
                      theta_s=theta_s, xi=10.0, eos_k=eos_k, eos_gamma=eos_gamma, water_c_mode=water_c_mode,
                      C=C.astype(np.float64), Fe=F.astype(np.float64), Fp=Fp, J=J)
        self.orc = O.Sim(n_grid, n_grid, n_grid, h, [grp], gravity=gravity, bc_pad=bc_pad, bc_mode=bc_mode)
        self.type = type_

    def close(self):
        self.gpu.close()

    # particle fields in a common layout
    def gpu_particles(self):
        g = self.gpu
        return {"x": g.read(0, capi.P_POS), "v": g.read(0, capi.P_VEL), "C": g.read(0, capi.P_C).reshape(-1, 3, 3),
                "F": g.read(0, capi.P_F).reshape(-1, 3, 3), "S": g.read(0, capi.P_JP)[:, 0]}

    def orc_particles(self):
        o = self.orc.groups[0]
        S = np.linalg.det(o.Fp) if self.type == capi.SNOW else o.J
        return {"x": o.x, "v": o.v, "C": o.C, "F": o.Fe, "S": S}


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def device_fields(sim, group):
    """One group's device state (FP32) in a common layout."""
    return {"x": sim.read(group, capi.P_POS), "v": sim.read(group, capi.P_VEL),
            "C": sim.read(group, capi.P_C).reshape(-1, 3, 3), "F": sim.read(group, capi.P_F).reshape(-1, 3, 3),
            "S": sim.read(group, capi.P_JP)[:, 0]}


def oracle_group(mat, vol, f):
    """An oracle group holding a device group's state `f` (J_p enters the
    oracle as F_p = J_p^(1/3) I: only det F_p is ever used)."""
    S = f["S"].astype(np.float64)
    kw = {}
    if mat.type == capi.SNOW:
        kw["Fp"] = np.cbrt(S)[:, None, None] * np.eye(3)
    if mat.type == capi.WATER:
        kw["J"] = S
    return O.Group(f["x"].astype(np.float64), f["v"].astype(np.float64), type=mat.type, mass=mat.rho * vol, vol=vol,
                   E=mat.E, nu=mat.nu, theta_c=mat.theta_c, theta_s=mat.theta_s, xi=mat.xi, eos_k=mat.eos_k,
                   eos_gamma=mat.eos_gamma, water_c_mode=mat.water_c_mode, C=f["C"].astype(np.float64),
                   Fe=f["F"].astype(np.float64), **kw)


def oracle_from_device(sim, cfg, mats, vol=None):
    """The FP64 oracle holding the device's current state, group by group
    (mats: one material per group, in group order)."""
    vol = cfg.h ** 3 if vol is None else vol
    groups = [oracle_group(m, vol, device_fields(sim, k)) for k, m in enumerate(mats)]
    return O.Sim(cfg.nx, cfg.ny, cfg.nz, cfg.h, groups, gravity=tuple(cfg.gravity), bc_pad=cfg.bc_pad,
                 bc_mode=cfg.bc_mode)


def oracle_fields(o: "O.Group"):
    S = np.linalg.det(o.Fp) if o.mat.type == capi.SNOW else o.J
    return {"x": o.x, "v": o.v, "C": o.C, "F": o.Fe, "S": S}


def host_threads(cap=16):
    import os
    try:
        return max(1, min(cap, len(os.sched_getaffinity(0))))
    except AttributeError:
        return max(1, min(cap, os.cpu_count() or 1))
