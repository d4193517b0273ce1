"""ctypes binding of the C ABI in include/mpm/mpm.h (libmpm_hip.so).

The library is loaded from this package directory only; when it is missing
every call raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpm_hip.so")

# enums (mpm.h)
MPM_OK = 0
JELLO, SNOW, WATER = 0, 1, 2
BC_AXIS, BC_STICKY = 0, 1
WATER_C_APIC, WATER_C_GRADV = 0, 1
FORM_PLAIN, FORM_FAST, FORM_FUSED = 0, 1, 2
WARM_NONE, WARM_MLP = 0, 1
DT_FIXED, DT_CFL, DT_CFL_DEFORM = 0, 1, 2
VOLUME_REF, VOLUME_PPC = 0, 1
SHAPE_MESH, SHAPE_BOX, SHAPE_SPHERE = 0, 1, 2
P_POS, P_VEL, P_C, P_F, P_JP, P_TAU = 0, 1, 2, 3, 4, 5
P_WIDTH = {P_POS: 3, P_VEL: 3, P_C: 9, P_F: 9, P_JP: 1, P_TAU: 6}
G_MASS, G_MOMENTUM, G_VELOCITY = 0, 1, 2
PHASE_P2G, PHASE_GRID, PHASE_G2P = 0, 1, 2
H_STENCIL_OOB, H_WEIGHT_SUM, H_DET_NONPOS, H_LEFT_DOMAIN, H_NONFINITE, H_ARENA_SPILL = 1, 2, 4, 8, 16, 32


class MpmConfig(C.Structure):
    _fields_ = [("h", C.c_double), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("origin", C.c_double * 3), ("dt", C.c_double), ("gravity", C.c_double * 3),
                ("bc_pad", C.c_int32), ("bc_mode", C.c_int32), ("ppc", C.c_int32), ("volume_mode", C.c_int32),
                ("seed", C.c_uint64), ("form", C.c_int32), ("sort_every", C.c_int32),
                ("svd_sweeps", C.c_int32), ("warm_start", C.c_int32), ("warm_sweeps", C.c_int32),
                ("tile", C.c_int32), ("threads", C.c_int32), ("dt_mode", C.c_int32), ("cfl", C.c_float),
                ("reserved", C.c_int32 * 3)]


class MpmMaterial(C.Structure):
    _fields_ = [("type", C.c_int32), ("water_c_mode", C.c_int32), ("E", C.c_double), ("nu", C.c_double),
                ("rho", C.c_double), ("theta_c", C.c_double), ("theta_s", C.c_double), ("xi", C.c_double),
                ("eos_k", C.c_double), ("eos_gamma", C.c_double)]


class MpmObject(C.Structure):
    _fields_ = [("shape", C.c_int32), ("reserved0", C.c_int32), ("obj_path", C.c_char_p),
                ("shape_param", C.c_double * 6), ("init_vel", C.c_double * 3), ("translation", C.c_float * 3),
                ("scale", C.c_float * 3), ("rotation_axis", C.c_float * 3), ("rotation_deg", C.c_float),
                ("material", MpmMaterial), ("max_particles", C.c_int64)]


class MpmBatchShape(C.Structure):
    _fields_ = [("p2g_workgroups", C.c_int32), ("grid_workgroups", C.c_int32), ("g2p_threads", C.c_int32),
                ("sort_every", C.c_int32), ("use_graph", C.c_int32), ("p2g_tile", C.c_int32)]


class Mpm2dConfig(C.Structure):
    _fields_ = [("n_grid", C.c_int32), ("wall", C.c_int32), ("dt", C.c_double), ("E", C.c_double),
                ("nu", C.c_double), ("rho", C.c_double), ("theta_c", C.c_double), ("theta_s", C.c_double),
                ("xi", C.c_double), ("gravity", C.c_double * 2), ("graph_steps", C.c_int32),
                ("reserved", C.c_int32 * 3)]


class MpmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mpm error {code}: {msg}")
        self.code = code


_lib = None

_SIGS = {
    "mpm_config_default": (None, [C.POINTER(MpmConfig)]),
    "mpm_material_init": (C.c_int, [C.POINTER(MpmMaterial), C.c_double, C.c_double, C.c_double, C.c_int]),
    "mpm_create": (C.c_int, [C.POINTER(MpmConfig), C.POINTER(C.c_void_p)]),
    "mpm_add_object": (C.c_int, [C.c_void_p, C.POINTER(MpmObject), C.POINTER(C.c_int32)]),
    "mpm_add_particles": (C.c_int, [C.c_void_p, C.POINTER(MpmMaterial), C.c_int64, C.c_void_p, C.c_void_p,
                                    C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_int32)]),
    "mpm_step": (C.c_int, [C.c_void_p, C.c_double]),
    "mpm_run": (C.c_int, [C.c_void_p, C.c_int32, C.c_double]),
    "mpm_step_phase": (C.c_int, [C.c_void_p, C.c_int32, C.c_double]),
    "mpm_total_particles": (C.c_int64, [C.c_void_p]),
    "mpm_group_count": (C.c_int32, [C.c_void_p]),
    "mpm_group_size": (C.c_int64, [C.c_void_p, C.c_int32]),
    "mpm_read_particles": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]),
    "mpm_write_particles": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]),
    "mpm_get_positions": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mpm_read_grid": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "mpm_health": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "mpm_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "mpm_svd_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "mpm_synchronize": (C.c_int, [C.c_void_p]),
    "mpm_set_timing": (C.c_int, [C.c_void_p, C.c_int32]),
    "mpm_phase_times": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "mpm_mlp_param_count": (C.c_int, []),
    "mpm_mlp_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "mpm_mlp_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.POINTER(C.c_float)]),
    "mpm_mlp_train": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_float, C.c_uint64, C.c_void_p,
                                C.c_void_p]),
    "mpm_probe_svd_residual": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p]),
    "mpm_probe_point_in_mesh": (C.c_int, [C.c_char_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "mpm_kernel_times": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "mpm_current_time": (C.c_double, [C.c_void_p]),
    "mpm_current_frame": (C.c_int64, [C.c_void_p]),
    "mpm_last_dt": (C.c_double, [C.c_void_p]),
    "mpm_write_bgeo": (C.c_int, [C.c_char_p, C.c_void_p, C.c_int64]),
    "mpm_set_warm_start_weights": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    "mpm_destroy": (None, [C.c_void_p]),
    "mpm_last_error": (C.c_char_p, []),
    "mpm_sample_object": (C.c_int, [C.POINTER(MpmConfig), C.POINTER(MpmObject), C.c_void_p, C.POINTER(C.c_int64)]),
    "mpm_probe_base_index": (C.c_int, [C.c_void_p, C.c_int64, C.c_float, C.c_float, C.c_void_p]),
    "mpm_probe_svd": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mpm_probe_sort": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p]),
    "mpm_sort_keys_noalloc": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mpm_probe_copy_bandwidth": (C.c_int, [C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
    "mpm_probe_soa_bandwidth": (C.c_int, [C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_double)]),
    "mpm_probe_lds_scatter": (C.c_int, [C.c_int32, C.c_int32, C.POINTER(C.c_double)]),
    "mpm_probe_stress": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_double, C.c_double, C.c_double,
                                   C.c_int32, C.c_void_p]),
    "mpm_batch_shape_default": (None, [C.POINTER(MpmBatchShape)]),
    "mpm_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(MpmBatchShape),
                                   C.POINTER(C.c_void_p)]),
    "mpm_batch_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "mpm_batch_get_shape": (C.c_int, [C.c_void_p, C.POINTER(MpmBatchShape)]),
    "mpm_batch_set_shape": (C.c_int, [C.c_void_p, C.POINTER(MpmBatchShape)]),
    "mpm_batch_run": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "mpm_batch_synchronize": (C.c_int, [C.c_void_p]),
    "mpm_batch_particles": (C.c_int64, [C.c_void_p]),
    "mpm_batch_scenes": (C.c_int32, [C.c_void_p]),
    "mpm_batch_destroy": (None, [C.c_void_p]),
    "mpm2d_config_default": (None, [C.POINTER(Mpm2dConfig)]),
    "mpm2d_create": (C.c_int, [C.POINTER(Mpm2dConfig), C.POINTER(C.c_void_p)]),
    "mpm2d_add_particles": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "mpm2d_run": (C.c_int, [C.c_void_p, C.c_int32]),
    "mpm2d_read": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "mpm2d_read_grid": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "mpm2d_health": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "mpm2d_particles": (C.c_int64, [C.c_void_p]),
    "mpm2d_steps": (C.c_int64, [C.c_void_p]),
    "mpm2d_synchronize": (C.c_int, [C.c_void_p]),
    "mpm2d_destroy": (None, [C.c_void_p]),
}

EXPORTED = tuple(_SIGS)


def lib():
    """Load libmpm_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m mpm_cuda_amd.build` "
                              "(the MPM path has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != MPM_OK:
        raise MpmError(rc, lib().mpm_last_error().decode(errors="replace"))


def default_config(**kw) -> MpmConfig:
    c = MpmConfig()
    lib().mpm_config_default(C.byref(c))
    for k, v in kw.items():
        if k in ("origin", "gravity"):
            setattr(c, k, (C.c_double * 3)(*v))
        else:
            setattr(c, k, v)
    return c


def material(E: float, nu: float, rho: float, type: int, **kw) -> MpmMaterial:
    m = MpmMaterial()
    check(lib().mpm_material_init(C.byref(m), float(E), float(nu), float(rho), int(type)))
    for k, v in kw.items():
        setattr(m, k, v)
    return m


def _f32(a, width=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a


class Sim:
    """Thin owner of an ``mpm_sim*`` handle."""

    def __init__(self, cfg: MpmConfig):
        self._h = C.c_void_p()
        check(lib().mpm_create(C.byref(cfg), C.byref(self._h)))
        self.cfg = cfg

    def close(self):
        if self._h:
            lib().mpm_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def add_particles(self, mat: MpmMaterial, xyz, vel=None, v0=(0.0, 0.0, 0.0), volume=0.0) -> int:
        xyz = _f32(xyz).reshape(-1, 3)
        vel_p = None
        if vel is not None:
            vel = _f32(vel).reshape(-1, 3)
            vel_p = vel.ctypes.data_as(C.c_void_p)
        g = C.c_int32()
        v0a = (C.c_double * 3)(*v0)
        check(lib().mpm_add_particles(self._h, C.byref(mat), xyz.shape[0], xyz.ctypes.data_as(C.c_void_p), vel_p,
                                      v0a, float(volume), C.byref(g)))
        return g.value

    def add_object(self, obj: MpmObject) -> int:
        g = C.c_int32()
        check(lib().mpm_add_object(self._h, C.byref(obj), C.byref(g)))
        return g.value

    def step(self, dt: float = 0.0):
        check(lib().mpm_step(self._h, float(dt)))

    def run(self, nsteps: int, dt: float = 0.0):
        check(lib().mpm_run(self._h, int(nsteps), float(dt)))

    def phase(self, ph: int, dt: float = 0.0):
        check(lib().mpm_step_phase(self._h, int(ph), float(dt)))

    def sync(self):
        check(lib().mpm_synchronize(self._h))

    @property
    def total(self) -> int:
        return lib().mpm_total_particles(self._h)

    def group_count(self) -> int:
        return lib().mpm_group_count(self._h)

    def group_size(self, g: int) -> int:
        return lib().mpm_group_size(self._h, g)

    def read(self, group: int, field: int) -> np.ndarray:
        n = self.group_size(group)
        w = P_WIDTH[field]
        out = np.empty((n, w), np.float32)
        check(lib().mpm_read_particles(self._h, group, field, out.ctypes.data_as(C.c_void_p)))
        return out

    def write(self, group: int, field: int, data) -> None:
        n = self.group_size(group)
        w = P_WIDTH[field]
        arr = _f32(data).reshape(n, w)
        check(lib().mpm_write_particles(self._h, group, field, arr.ctypes.data_as(C.c_void_p)))

    def positions(self) -> np.ndarray:
        out = np.empty((self.total, 3), np.float32)
        check(lib().mpm_get_positions(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def grid(self, field: int) -> np.ndarray:
        c = self.cfg
        w = 1 if field == G_MASS else 3
        out = np.empty((c.nz, c.ny, c.nx, w), np.float32)
        check(lib().mpm_read_grid(self._h, field, out.ctypes.data_as(C.c_void_p)))
        return out[..., 0] if w == 1 else out

    def health(self) -> int:
        b = C.c_uint32()
        check(lib().mpm_health(self._h, C.byref(b)))
        return b.value

    def stats(self) -> dict:
        s = (C.c_double * 8)()
        check(lib().mpm_stats(self._h, s))
        return {"steps": s[0], "active_blocks": s[1], "fp64_fallbacks": s[2], "warm_rejects": s[3],
                "sorts": s[4], "time": s[5], "tiles": s[6], "spilled": s[7]}

    def svd_stats(self) -> dict:
        """Cumulative G2P SVD counts: SVD-path updates, warm-started ones,
        rejected warm results, FP64 fallbacks."""
        s = (C.c_double * 4)()
        check(lib().mpm_svd_stats(self._h, s))
        return {"svd": int(s[0]), "warm": int(s[1]), "rejected": int(s[2]), "fp64_fallbacks": int(s[3])}

    def set_timing(self, on: bool = True) -> None:
        check(lib().mpm_set_timing(self._h, 1 if on else 0))

    def phase_times(self) -> dict:
        """{phase: (total_ms, launches)} since the previous call, plus 'nodes'."""
        ms = (C.c_double * 4)()
        cnt = (C.c_int64 * 4)()
        nodes = C.c_int64()
        check(lib().mpm_phase_times(self._h, ms, cnt, C.byref(nodes)))
        names = ("sort", "p2g", "grid", "g2p")
        out = {names[k]: (ms[k], cnt[k]) for k in range(4)}
        out["nodes"] = nodes.value
        return out

    def kernel_times(self) -> dict:
        """{kernel: (total_ms, launches)} of the main kernel of each phase
        (P2G, grid update, G2P) since the previous call."""
        ms = (C.c_double * 4)()
        cnt = (C.c_int64 * 4)()
        check(lib().mpm_kernel_times(self._h, ms, cnt))
        names = ("p2g", "grid", "g2p")
        return {names[k]: (ms[k], cnt[k]) for k in range(3)}

    def set_warm_start_weights(self, w) -> None:
        w = _f32(w).reshape(-1)
        check(lib().mpm_set_warm_start_weights(self._h, w.ctypes.data_as(C.c_void_p), w.size))

    @property
    def current_time(self) -> float:
        return lib().mpm_current_time(self._h)

    @property
    def current_frame(self) -> int:
        return lib().mpm_current_frame(self._h)

    @property
    def last_dt(self) -> float:
        """dt of the last step (adp_dt); synchronizes with a dt controller."""
        return lib().mpm_last_dt(self._h)


BATCH_STAGES = ("p2g", "grid", "g2p")
BATCH_CANDIDATES = {"p2g": (1024, 2048, 4096, 8192), "grid": (1024, 2048, 4096, 8192), "g2p": (64, 128, 192, 256),
                    "tile": (256, 384, 512)}


def batch_shape(**kw) -> MpmBatchShape:
    sh = MpmBatchShape()
    lib().mpm_batch_shape_default(C.byref(sh))
    for k, v in kw.items():
        setattr(sh, k, v)
    return sh


class Batch:
    """Owner of an ``mpm_batch*``: independent one-group FAST-form scenes
    advanced together (the scenes stay owned by their ``Sim`` objects and
    must outlive the batch)."""

    def __init__(self, sims, shape: MpmBatchShape | None = None):
        self.sims = list(sims)
        arr = (C.c_void_p * len(self.sims))(*[s.handle for s in self.sims])
        self._h = C.c_void_p()
        check(lib().mpm_batch_create(arr, len(self.sims), C.byref(shape) if shape is not None else None,
                                     C.byref(self._h)))

    def close(self):
        if self._h:
            lib().mpm_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def autotune(self, reps: int = 3) -> dict:
        """Median ms per stage and candidate (launch widths of P2G, grid update
        and G2P, then P2G per tile shape); keeps the fastest of each."""
        ms = np.zeros(15, np.float32)
        check(lib().mpm_batch_autotune(self._h, int(reps), ms.ctypes.data_as(C.c_void_p)))
        out = {st: dict(zip(BATCH_CANDIDATES[st], ms[4 * i:4 * i + 4].tolist())) for i, st in enumerate(BATCH_STAGES)}
        out["tile"] = dict(zip(BATCH_CANDIDATES["tile"], ms[12:15].tolist()))
        return out

    @property
    def shape(self) -> MpmBatchShape:
        sh = MpmBatchShape()
        check(lib().mpm_batch_get_shape(self._h, C.byref(sh)))
        return sh

    @shape.setter
    def shape(self, sh: MpmBatchShape) -> None:
        check(lib().mpm_batch_set_shape(self._h, C.byref(sh)))

    def run(self, nsteps: int, timed: bool = False):
        """Advance every scene by nsteps; with timed, the device ms of each step."""
        if not timed:
            check(lib().mpm_batch_run(self._h, int(nsteps), None))
            return None
        ms = np.zeros(max(int(nsteps), 1), np.float32)
        check(lib().mpm_batch_run(self._h, int(nsteps), ms.ctypes.data_as(C.c_void_p)))
        return ms[:int(nsteps)]

    def sync(self):
        check(lib().mpm_batch_synchronize(self._h))

    @property
    def particles(self) -> int:
        return lib().mpm_batch_particles(self._h)

    @property
    def scenes(self) -> int:
        return lib().mpm_batch_scenes(self._h)


P2_POS, P2_VEL, P2_C, P2_F, P2_JP = 0, 1, 2, 3, 4
_P2_WIDTH = {P2_POS: 2, P2_VEL: 2, P2_C: 4, P2_F: 4, P2_JP: 1}


def config2d(**kw) -> Mpm2dConfig:
    c = Mpm2dConfig()
    lib().mpm2d_config_default(C.byref(c))
    for k, v in kw.items():
        if k == "gravity":
            c.gravity = (C.c_double * 2)(*v)
        else:
            setattr(c, k, v)
    return c


class Sim2D:
    """Owner of an ``mpm2d_sim*``: the reference's 2D snow step (MPMSnow.py)."""

    def __init__(self, cfg: Mpm2dConfig | None = None, **kw):
        self.cfg = cfg if cfg is not None else config2d(**kw)
        self._h = C.c_void_p()
        check(lib().mpm2d_create(C.byref(self.cfg), C.byref(self._h)))

    def close(self):
        if self._h:
            lib().mpm2d_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add_particles(self, xy, vel=None) -> None:
        xy = _f32(xy).reshape(-1, 2)
        vp = None
        if vel is not None:
            vel = _f32(np.broadcast_to(np.asarray(vel, np.float32), xy.shape))
            vp = vel.ctypes.data_as(C.c_void_p)
        check(lib().mpm2d_add_particles(self._h, xy.shape[0], xy.ctypes.data_as(C.c_void_p), vp))

    def run(self, nsteps: int) -> None:
        check(lib().mpm2d_run(self._h, int(nsteps)))

    def sync(self) -> None:
        check(lib().mpm2d_synchronize(self._h))

    @property
    def n(self) -> int:
        return lib().mpm2d_particles(self._h)

    @property
    def steps(self) -> int:
        return lib().mpm2d_steps(self._h)

    def read(self, field: int) -> np.ndarray:
        w = _P2_WIDTH[field]
        out = np.empty((self.n, w), np.float32)
        check(lib().mpm2d_read(self._h, int(field), out.ctypes.data_as(C.c_void_p)))
        return out

    def grid(self):
        """(mass [n*n], velocity [n*n][2]) after the last grid update; node (i, j) at j * n + i."""
        nn = self.cfg.n_grid * self.cfg.n_grid
        m = np.empty(nn, np.float32)
        v = np.empty((nn, 2), np.float32)
        check(lib().mpm2d_read_grid(self._h, m.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p)))
        return m, v

    def health(self) -> int:
        b = C.c_uint32()
        check(lib().mpm2d_health(self._h, C.byref(b)))
        return b.value


def sample_object(cfg: MpmConfig, obj: MpmObject) -> np.ndarray:
    """Host-only particle sampling of one object (what mpm_add_object adds)."""
    n = C.c_int64(0)
    check(lib().mpm_sample_object(C.byref(cfg), C.byref(obj), None, C.byref(n)))
    out = np.empty((n.value, 3), np.float32)
    check(lib().mpm_sample_object(C.byref(cfg), C.byref(obj), out.ctypes.data_as(C.c_void_p), C.byref(n)))
    return out


def probe_base_index(x, origin: float, inv_h: float) -> np.ndarray:
    x = _f32(x).reshape(-1)
    out = np.empty(x.size, np.int32)
    check(lib().mpm_probe_base_index(x.ctypes.data_as(C.c_void_p), x.size, float(origin), float(inv_h),
                                     out.ctypes.data_as(C.c_void_p)))
    return out


def probe_svd(F, sweeps: int = 4):
    F = _f32(F).reshape(-1, 9)
    n = F.shape[0]
    U, s, V = np.empty((n, 9), np.float32), np.empty((n, 3), np.float32), np.empty((n, 9), np.float32)
    check(lib().mpm_probe_svd(F.ctypes.data_as(C.c_void_p), n, int(sweeps), U.ctypes.data_as(C.c_void_p),
                              s.ctypes.data_as(C.c_void_p), V.ctypes.data_as(C.c_void_p)))
    return U.reshape(n, 3, 3), s, V.reshape(n, 3, 3)


def probe_stress(F, type: int, mu: float, lam: float, xi: float = 10.0, Jp=None, sweeps: int = 4):
    """Kirchhoff stress of the step kernels, (n, 6) as xx, yy, zz, xy, xz, yz."""
    F = _f32(F).reshape(-1, 9)
    n = F.shape[0]
    tau = np.empty((n, 6), np.float32)
    jp = None
    if Jp is not None:
        Jp = _f32(Jp).reshape(n)
        jp = Jp.ctypes.data_as(C.c_void_p)
    check(lib().mpm_probe_stress(F.ctypes.data_as(C.c_void_p), jp, n, int(type), float(mu), float(lam), float(xi),
                                 int(sweeps), tau.ctypes.data_as(C.c_void_p)))
    return tau


def probe_svd_residual(F, q=None, sweeps: int = 1):
    """Column orthogonality residual of F V after `sweeps` Jacobi sweeps from
    V0 = R(q) (identity when q is None)."""
    F = _f32(F).reshape(-1, 9)
    n = F.shape[0]
    off = np.empty(n, np.float32)
    qp = None
    if q is not None:
        q = _f32(q).reshape(n, 4)
        qp = q.ctypes.data_as(C.c_void_p)
    check(lib().mpm_probe_svd_residual(F.ctypes.data_as(C.c_void_p), qp, n, int(sweeps), off.ctypes.data_as(C.c_void_p)))
    return off


def probe_point_in_mesh(obj_path: str, pts) -> np.ndarray:
    """The sampler's inside test for points in the OBJ file's coordinates (host only)."""
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    out = np.empty(pts.shape[0], np.uint8)
    check(lib().mpm_probe_point_in_mesh(obj_path.encode(), pts.ctypes.data_as(C.c_void_p), pts.shape[0],
                                        out.ctypes.data_as(C.c_void_p)))
    return out.astype(bool)


def mlp_param_count() -> int:
    return int(lib().mpm_mlp_param_count())


def mlp_forward(params, F):
    """Warm-start network output q[n][4] (device, BF16 MFMA)."""
    params = _f32(params).reshape(-1)
    F = _f32(F).reshape(-1, 9)
    n = F.shape[0]
    q = np.empty((n, 4), np.float32)
    check(lib().mpm_mlp_forward(params.ctypes.data_as(C.c_void_p), F.ctypes.data_as(C.c_void_p), n,
                                q.ctypes.data_as(C.c_void_p)))
    return q


def mlp_grad(params, F):
    """(mean loss, gradient) of the warm-start loss over F (device)."""
    params = _f32(params).reshape(-1)
    F = _f32(F).reshape(-1, 9)
    g = np.empty(params.size, np.float32)
    loss = C.c_float()
    check(lib().mpm_mlp_grad(params.ctypes.data_as(C.c_void_p), F.ctypes.data_as(C.c_void_p), F.shape[0],
                             g.ctypes.data_as(C.c_void_p), C.byref(loss)))
    return float(loss.value), g


def mlp_train(F, params, steps: int, batch: int = 4096, lr: float = 1e-3, seed: int = 0):
    """Adam on the device; returns (trained params, per-step mean loss)."""
    p = np.array(_f32(params).reshape(-1), copy=True)
    F = _f32(F).reshape(-1, 9)
    hist = np.empty(max(steps, 1), np.float32)
    check(lib().mpm_mlp_train(F.ctypes.data_as(C.c_void_p), F.shape[0], int(steps), int(batch), float(lr), int(seed),
                              p.ctypes.data_as(C.c_void_p), hist.ctypes.data_as(C.c_void_p)))
    return p, hist[:steps]


def probe_copy_bandwidth(bytes_: int = 1 << 30, reps: int = 10) -> float:
    """GB/s of a streaming device copy (read + write), median of reps."""
    g = C.c_double()
    check(lib().mpm_probe_copy_bandwidth(int(bytes_), int(reps), C.byref(g)))
    return g.value


def probe_soa_bandwidth(n: int = 1_000_000, planes_in: int = 13, planes_out: int = 31, reps: int = 10) -> float:
    """GB/s (read + write) of a structure-of-arrays pass: n elements,
    planes_in FP32 planes read, planes_out written, median of reps."""
    g = C.c_double()
    check(lib().mpm_probe_soa_bandwidth(int(n), int(planes_in), int(planes_out), int(reps), C.byref(g)))
    return g.value


LDS_PROBE_MODES = {0: "float atomics, distinct cells", 1: "float atomics, hashed cells",
                   2: "float4 read-modify-write, distinct cells", 3: "float4 read-modify-write, hashed cells"}


def probe_lds_scatter(mode: int, reps: int = 256) -> float:
    """ms of one mpm_probe_lds_scatter launch (8,192 one-wave workgroups,
    reps x 64 particles each, 27 nodes x 4 floats per particle)."""
    ms = C.c_double()
    check(lib().mpm_probe_lds_scatter(int(mode), int(reps), C.byref(ms)))
    return ms.value


def probe_sort(keys, bits: int):
    keys = np.ascontiguousarray(keys, np.uint32).reshape(-1)
    n = keys.size
    srt, perm = np.empty(n, np.uint32), np.empty(n, np.uint32)
    check(lib().mpm_probe_sort(keys.ctypes.data_as(C.c_void_p), n, int(bits), srt.ctypes.data_as(C.c_void_p),
                               perm.ctypes.data_as(C.c_void_p)))
    return srt, perm


SORT_NOALLOC_MAX_KEYS = 1 << 20


def sort_keys_noalloc(keys_ptr: int, n: int, bits: int, sorted_ptr: int, perm_ptr: int, stream: int = 0) -> None:
    """mpm_sort_keys_noalloc on device pointers (queued on `stream`)."""
    check(lib().mpm_sort_keys_noalloc(C.c_void_p(keys_ptr), int(n), int(bits), C.c_void_p(sorted_ptr),
                                      C.c_void_p(perm_ptr), C.c_void_p(stream)))


def write_bgeo(path: str, xyz) -> None:
    xyz = _f32(xyz).reshape(-1, 3)
    check(lib().mpm_write_bgeo(path.encode(), xyz.ctypes.data_as(C.c_void_p), xyz.shape[0]))
