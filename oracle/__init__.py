"""ctypes front end of the FP64 CPU oracle (test infrastructure only).

The oracle restates JerryYan97/MPM_CUDA ``src/simulator/MPMCUDA.cu`` (P2G
:310-433, VelUpdate :435-472, InterpolateAndMove :474-741, step :793-925) in C;
see ``mpm_oracle.c``.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module, always as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

JELLO, SNOW, WATER = 0, 1, 2
BC_AXIS, BC_STICKY = 0, 1
WATER_C_APIC, WATER_C_GRADV = 0, 1
FLAG_STENCIL_OOB, FLAG_WEIGHT_SUM, FLAG_DET_NONPOS, FLAG_LEFT_DOMAIN, FLAG_NONFINITE = 1, 2, 4, 8, 16

_dp = C.POINTER(C.c_double)


class OrcGrid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("reserved", C.c_int32),
                ("h", C.c_double), ("origin", C.c_double * 3), ("m", _dp), ("p", _dp)]


class OrcMaterial(C.Structure):
    _fields_ = [("type", C.c_int32), ("water_c_mode", C.c_int32),
                ("mass", C.c_double), ("vol", C.c_double), ("mu", C.c_double), ("lam", C.c_double),
                ("theta_c", C.c_double), ("theta_s", C.c_double), ("xi", C.c_double),
                ("eos_k", C.c_double), ("eos_gamma", C.c_double)]


class OrcGroup(C.Structure):
    _fields_ = [("n", C.c_int64), ("mat", OrcMaterial),
                ("x", _dp), ("v", _dp), ("C", _dp), ("Fe", _dp), ("Fp", _dp), ("J", _dp),
                ("dgdiff", _dp)]


class OrcParams(C.Structure):
    _fields_ = [("gravity", C.c_double * 3), ("bc_pad", C.c_int32), ("bc_mode", C.c_int32)]


_d2 = C.c_double * 2


class Orc2Params(C.Structure):
    _fields_ = [("n_grid", C.c_int32), ("wall", C.c_int32), ("dx", C.c_double), ("p_vol", C.c_double),
                ("p_mass", C.c_double), ("mu0", C.c_double), ("lam0", C.c_double), ("xi", C.c_double),
                ("theta_c", C.c_double), ("theta_s", C.c_double), ("gravity", C.c_double * 2)]


class Orc2Particles(C.Structure):
    _fields_ = [("n", C.c_int64), ("x", _dp), ("v", _dp), ("C", _dp), ("Fe", _dp), ("Fp", _dp)]


REF_MESH_QUERY = os.path.join(_HERE, "_ref", "libmesh_query.so")
_REF_SRC = "/root/reference/thirdparties/mesh_query0.1"


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile, and the reference's
    mesh_query into _ref/ when the reference sources are present."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    if os.path.isdir(_REF_SRC) and (force or not os.path.exists(REF_MESH_QUERY)):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "ref"], check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def set_threads(n: int) -> int:
    """OpenMP threads of the oracle's P2G / grid / G2P loops (1 = sequential)."""
    return int(lib().orc_set_threads(int(n)))


def mesh_query():
    """ctypes handle of the reference's mesh_query (oracle/_ref), or None."""
    if not os.path.exists(REF_MESH_QUERY):
        return None
    L = C.CDLL(REF_MESH_QUERY)
    L.construct_mesh_object.restype = C.c_void_p
    L.construct_mesh_object.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p]
    L.destroy_mesh_object.restype = None
    L.destroy_mesh_object.argtypes = [C.c_void_p]
    L.point_inside_mesh.restype = C.c_bool
    L.point_inside_mesh.argtypes = [C.c_void_p, C.c_void_p]
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        d3 = C.c_double * 3
        L.orc_N.restype = C.c_double; L.orc_N.argtypes = [C.c_double]
        L.orc_dN.restype = C.c_double; L.orc_dN.argtypes = [C.c_double]
        L.orc_weight.restype = C.c_double; L.orc_weight.argtypes = [_dp, _dp, C.c_double]
        L.orc_weight_grad.restype = None; L.orc_weight_grad.argtypes = [_dp, _dp, C.c_double, _dp]
        L.orc_base_ref.restype = C.c_int; L.orc_base_ref.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_base_f32.restype = C.c_int; L.orc_base_f32.argtypes = [C.c_float, C.c_float, C.c_float]
        L.orc_set_threads.restype = C.c_int; L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads(1)          # sequential and deterministic unless asked (set_threads)
        L.orc_block_key.restype = C.c_uint32
        L.orc_block_key.argtypes = [C.c_int] * 5
        L.orc_det3.restype = C.c_double; L.orc_det3.argtypes = [_dp]
        L.orc_inv3.restype = C.c_int; L.orc_inv3.argtypes = [_dp, _dp]
        L.orc_svd3.restype = None; L.orc_svd3.argtypes = [_dp, _dp, _dp, _dp]
        L.orc_fcr_tau.restype = None; L.orc_fcr_tau.argtypes = [_dp, C.c_double, C.c_double, _dp]
        L.orc_snow_plastic.restype = None
        L.orc_snow_plastic.argtypes = [_dp, _dp, C.c_double, C.c_double, _dp, _dp]
        L.orc_clear_grid.restype = None; L.orc_clear_grid.argtypes = [C.POINTER(OrcGrid)]
        L.orc_p2g.restype = C.c_int; L.orc_p2g.argtypes = [C.POINTER(OrcGrid), C.POINTER(OrcGroup), C.c_double]
        L.orc_grid_update.restype = C.c_int
        L.orc_grid_update.argtypes = [C.POINTER(OrcGrid), C.POINTER(OrcParams), C.c_double]
        L.orc_g2p.restype = C.c_int; L.orc_g2p.argtypes = [C.POINTER(OrcGrid), C.POINTER(OrcGroup), C.c_double]
        L.orc_step.restype = C.c_int
        L.orc_step.argtypes = [C.POINTER(OrcGrid), C.POINTER(OrcGroup), C.c_int, C.POINTER(OrcParams), C.c_double]
        L.orc2_N.restype = C.c_double; L.orc2_N.argtypes = [C.c_double]
        L.orc2_dN.restype = C.c_double; L.orc2_dN.argtypes = [C.c_double]
        L.orc2_svd.restype = None; L.orc2_svd.argtypes = [_dp, _dp, _dp, _dp]
        L.orc2_fcr_tau.restype = None; L.orc2_fcr_tau.argtypes = [_dp, C.c_double, C.c_double, _dp]
        L.orc2_step.restype = C.c_int
        L.orc2_step.argtypes = [_dp, _dp, C.POINTER(Orc2Particles), C.POINTER(Orc2Params), C.c_double]
        del d3
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


# ---------------------------------------------------------------- scalar API
def N(x: float) -> float:
    return lib().orc_N(float(x))


def dN(x: float) -> float:
    return lib().orc_dN(float(x))


def weight(xp, xi, h):
    xp = np.ascontiguousarray(xp, np.float64); xi = np.ascontiguousarray(xi, np.float64)
    return lib().orc_weight(_ptr(xp), _ptr(xi), float(h))


def weight_grad(xp, xi, h):
    xp = np.ascontiguousarray(xp, np.float64); xi = np.ascontiguousarray(xi, np.float64)
    g = np.zeros(3)
    lib().orc_weight_grad(_ptr(xp), _ptr(xi), float(h), _ptr(g))
    return g


def base_ref(x, o, h) -> int:
    return lib().orc_base_ref(float(x), float(o), float(h))


def base_f32(x, o, inv_h) -> int:
    return lib().orc_base_f32(float(np.float32(x)), float(np.float32(o)), float(np.float32(inv_h)))


def det3(A) -> float:
    A = np.ascontiguousarray(A, np.float64).reshape(9)
    return lib().orc_det3(_ptr(A))


def svd3(F):
    F = np.ascontiguousarray(F, np.float64).reshape(9)
    U, s, V = np.zeros(9), np.zeros(3), np.zeros(9)
    lib().orc_svd3(_ptr(F), _ptr(U), _ptr(s), _ptr(V))
    return U.reshape(3, 3), s, V.reshape(3, 3)


def fcr_tau(F, mu, lam):
    F = np.ascontiguousarray(F, np.float64).reshape(9)
    t = np.zeros(9)
    lib().orc_fcr_tau(_ptr(F), float(mu), float(lam), _ptr(t))
    return t.reshape(3, 3)


def snow_plastic(Ftr, Fp, theta_c, theta_s):
    Ftr = np.ascontiguousarray(Ftr, np.float64).reshape(9)
    Fp = np.ascontiguousarray(Fp, np.float64).reshape(9)
    Fe_new, Fp_new = np.zeros(9), np.zeros(9)
    lib().orc_snow_plastic(_ptr(Ftr), _ptr(Fp), float(theta_c), float(theta_s), _ptr(Fe_new), _ptr(Fp_new))
    return Fe_new.reshape(3, 3), Fp_new.reshape(3, 3)


# ---------------------------------------------------------------- simulator
def lame(E: float, nu: float):
    """(E, nu) -> (mu, lambda), Elasiticity.cuh:42-43."""
    return E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))


class Group:
    """One particle group in the reference's AoS FP64 layout (MPMSimulator.cuh:30-52)."""

    def __init__(self, x, v, *, type, mass, vol, E=0.0, nu=0.0, mu=None, lam=None,
                 theta_c=8e-2, theta_s=2e-2, xi=10.0, eos_k=2.51, eos_gamma=7.0,
                 water_c_mode=WATER_C_APIC, C=None, Fe=None, Fp=None, J=None):
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        n = x.shape[0]
        self.n = n
        self.x = x.copy()
        self.v = np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64), (n, 3))).copy()
        eye = np.broadcast_to(np.eye(3), (n, 3, 3))
        self.C = np.zeros((n, 3, 3)) if C is None else np.ascontiguousarray(C, np.float64).reshape(n, 3, 3).copy()
        self.Fe = eye.copy() if Fe is None else np.ascontiguousarray(Fe, np.float64).reshape(n, 3, 3).copy()
        self.Fp = eye.copy() if Fp is None else np.ascontiguousarray(Fp, np.float64).reshape(n, 3, 3).copy()
        self.J = np.ones(n) if J is None else np.ascontiguousarray(J, np.float64).reshape(n).copy()
        self.dgdiff = np.zeros(n)
        if mu is None or lam is None:
            mu, lam = lame(E, nu)
        self.mat = OrcMaterial(type, water_c_mode, mass, vol, mu, lam, theta_c, theta_s, xi, eos_k, eos_gamma)

    def struct(self) -> OrcGroup:
        return OrcGroup(self.n, self.mat, _ptr(self.x), _ptr(self.v), _ptr(self.C), _ptr(self.Fe),
                        _ptr(self.Fp), _ptr(self.J), _ptr(self.dgdiff))


class Sim:
    """FP64 reference-semantics simulator over a dense grid (reference node order)."""

    def __init__(self, nx, ny, nz, h, groups, *, origin=(0.0, 0.0, 0.0), gravity=(0.0, -9.8, 0.0),
                 bc_pad=10, bc_mode=BC_AXIS):
        self.nx, self.ny, self.nz, self.h = nx, ny, nz, float(h)
        self.m = np.zeros(nx * ny * nz)
        self.p = np.zeros(nx * ny * nz * 3)
        self.grid = OrcGrid(nx, ny, nz, 0, self.h, (C.c_double * 3)(*origin), _ptr(self.m), _ptr(self.p))
        self.params = OrcParams((C.c_double * 3)(*gravity), bc_pad, bc_mode)
        self.groups = list(groups)

    def _gs(self):
        arr = (OrcGroup * len(self.groups))(*[g.struct() for g in self.groups])
        return arr

    def clear_grid(self):
        lib().orc_clear_grid(C.byref(self.grid))

    def p2g(self, dt):
        f = 0
        for g in self.groups:
            s = g.struct()
            f |= lib().orc_p2g(C.byref(self.grid), C.byref(s), float(dt))
        return f

    def grid_update(self, dt):
        return lib().orc_grid_update(C.byref(self.grid), C.byref(self.params), float(dt))

    def g2p(self, dt):
        f = 0
        for g in self.groups:
            s = g.struct()
            f |= lib().orc_g2p(C.byref(self.grid), C.byref(s), float(dt))
        return f

    def step(self, dt):
        arr = self._gs()
        return lib().orc_step(C.byref(self.grid), arr, len(self.groups), C.byref(self.params), float(dt))

    def grid_mass(self):
        return self.m.reshape(self.nz, self.ny, self.nx)

    def grid_vec(self):
        return self.p.reshape(self.nz, self.ny, self.nx, 3)


DT_FIXED, DT_CFL, DT_CFL_DEFORM = 0, 1, 2


def controller_measures(g: Group):
    """Per-particle inputs of the step-size controller for one group's FP64
    state: |v| + c (c the elastic wave speed at the particle's deformation:
    P waves sqrt((lam + 2 mu) / rho), hardened by exp(xi (1 - det F_p)) for
    snow; water sqrt(k gamma J^(1 - gamma) / rho)) and the deformation
    measure max(J, 1/J) of the reference's unused time-step factor
    (MPMCUDA.cu:613-619 jello det F, :700-708 snow det F_p; never set for
    water, :710-713, so 0 here)."""
    m = g.mat
    rho = m.mass / m.vol
    if m.type == WATER:
        J = g.J
        c2 = m.eos_k * m.eos_gamma / rho * np.power(np.where(J > 0, J, np.nan), 1.0 - m.eos_gamma)
        d = np.zeros(g.n)
    else:
        J = np.linalg.det(g.Fp if m.type == SNOW else g.Fe)
        c2 = (m.lam + 2.0 * m.mu) / rho * np.ones(g.n)
        if m.type == SNOW:
            c2 = c2 * np.exp(m.xi * (1.0 - J))
        d = np.where(J > 0, np.maximum(J, 1.0 / np.where(J > 0, J, 1.0)), 1e30)
    c = np.sqrt(np.maximum(np.nan_to_num(c2, nan=0.0), 0.0))
    return np.linalg.norm(g.v, axis=1) + c, d


def controller_dt(groups, h: float, max_dt: float, cfl: float, mode: int) -> float:
    """dt the controller gives the next step: min(max_dt, cfl h / max(|v|+c))
    and, with DT_CFL_DEFORM, at most max_dt / clamp(max(J, 1/J)^3, 1, 10)
    (the reference's factor min(10, max^3), MPMCUDA.cu:915-920, which it
    computes but never applies)."""
    if mode == DT_FIXED:
        return float(max_dt)
    speed, deform = 0.0, 0.0
    for g in groups:
        s, d = controller_measures(g)
        if g.n:
            speed = max(speed, float(s.max()))
            deform = max(deform, float(d.max()))
    dt = float(max_dt)
    if speed > 0:
        dt = min(dt, cfl * h / speed)
    if mode == DT_CFL_DEFORM:
        dt = min(dt, max_dt / min(10.0, max(1.0, deform ** 3)))
    return dt


class Sim2D:
    """FP64 2D snow of MPMSnow.py:77-150 on an n x n grid (index j * n + i).
    Particles in AoS FP64 (x, v: [n][2]; C, Fe, Fp: [n][2][2]); the script's
    constants are keyword arguments (defaults: the 9,000-particle snow
    workload -- E = 1e3, nu = 0.2, rho = 1, theta_c = 2.5e-2,
    theta_s = 7.5e-3, xi = 10, gravity (0, -50), 5-node sticky walls)."""

    def __init__(self, x, v, *, n_grid=128, wall=5, E=1e3, nu=0.2, rho=1.0, xi=10.0, theta_c=2.5e-2,
                 theta_s=7.5e-3, gravity=(0.0, -50.0), C=None, Fe=None, Fp=None):
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 2)
        n = x.shape[0]
        self.n = n
        self.x = x.copy()
        self.v = np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64), (n, 2))).copy()
        eye = np.broadcast_to(np.eye(2), (n, 2, 2))
        self.C = np.zeros((n, 2, 2)) if C is None else np.ascontiguousarray(C, np.float64).reshape(n, 2, 2).copy()
        self.Fe = eye.copy() if Fe is None else np.ascontiguousarray(Fe, np.float64).reshape(n, 2, 2).copy()
        self.Fp = eye.copy() if Fp is None else np.ascontiguousarray(Fp, np.float64).reshape(n, 2, 2).copy()
        dx = 1.0 / n_grid
        p_vol = (0.5 * dx) ** 2
        mu, lam = lame(E, nu)
        self.params = Orc2Params(n_grid, wall, dx, p_vol, p_vol * rho, mu, lam, xi, theta_c, theta_s,
                                 _d2(*gravity))
        self.n_grid = n_grid
        self.grid_m = np.zeros(n_grid * n_grid)
        self.grid_v = np.zeros(n_grid * n_grid * 2)

    def step(self, dt):
        ps = Orc2Particles(self.n, _ptr(self.x), _ptr(self.v), _ptr(self.C), _ptr(self.Fe), _ptr(self.Fp))
        return lib().orc2_step(_ptr(self.grid_m), _ptr(self.grid_v), C.byref(ps), C.byref(self.params), float(dt))
