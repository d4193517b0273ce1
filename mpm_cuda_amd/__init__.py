"""MI355X-native explicit MLS-MPM simulator.

Python mirror of the reference simulator's interface (JerryYan97/MPM_CUDA):

* ``Material(E, nu, rho, type)``               -- Elasiticity.cuh:14-46
* ``ObjInitInfo``                               -- MPMSimulator.cuh:19-27
* ``MPMSimulator(gap, dt, nx, ny, nz, ppc, infos)``
  with ``step()``, ``totalParticlesNum()``, ``getGLParticlesPos()``,
  ``adp_dt``, ``current_time``, ``current_frame`` -- MPMSimulator.cuh:66-102
* ``OutputBego(path, xyz)``                     -- FilesIO.h:10

Everything runs through the C ABI of ``libmpm_hip.so`` (``capi``); errors
raise ``capi.MpmError`` where the reference prints and exits.
"""
from __future__ import annotations

import enum
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import capi
from .capi import MpmError  # noqa: F401


class MaterialType(enum.IntEnum):
    """Elasiticity.cuh:10-12 (SAND is declared there but never simulated)."""
    JELLO = 0
    SNOW = 1
    WATER = 2
    SAND = 3


class Material:
    """Material(E, nu, rho, type): Lame constants mu = E/(2(1+nu)),
    lambda = E nu/((1+nu)(1-2nu)) (Elasiticity.cuh:42-43); the range checks
    of :34-41 raise MpmError instead of exiting."""

    def __init__(self, iYM: float = 0.0, iPR: float = 0.0, iDensity: float = 1.0,
                 iType: MaterialType = MaterialType.JELLO, **extra):
        if int(iType) == MaterialType.SAND:
            raise MpmError(1, "SAND is not simulated (no constitutive model in the reference)")
        self.mType = MaterialType(int(iType))
        self._m = capi.material(iYM, iPR, iDensity, int(iType), **extra)
        self.mYoungsModulus, self.mPoissonRatio, self.mDensity = iYM, iPR, iDensity
        self.mMu = iYM / (2 * (1 + iPR))
        self.mLambda = iYM * iPR / ((1 + iPR) * (1 - 2 * iPR))

    @property
    def c_struct(self) -> capi.MpmMaterial:
        return self._m


@dataclass
class ObjInitInfo:
    """MPMSimulator.cuh:19-27.  ``objPath`` names a Wavefront OBJ; instead a
    procedural ``shape`` ("box" with ``shape_param`` = local min/max, or
    "sphere" with centre/radius) may be given."""
    initVel: Sequence[float] = (0.0, 0.0, 0.0)
    initTranslation: Sequence[float] = (0.0, 0.0, 0.0)
    initRotationAxis: Sequence[float] = (1.0, 0.0, 0.0)
    initScale: Sequence[float] = (1.0, 1.0, 1.0)
    initRotationDegree: float = 0.0
    objPath: str = ""
    mMaterial: Optional[Material] = None
    shape: str = "mesh"
    shape_param: Sequence[float] = (0.0, 0.0, 0.0, 1.0, 1.0, 1.0)
    max_particles: int = 0

    def c_struct(self) -> capi.MpmObject:
        o = capi.MpmObject()
        o.shape = {"mesh": capi.SHAPE_MESH, "box": capi.SHAPE_BOX, "sphere": capi.SHAPE_SPHERE}[self.shape]
        self._path = self.objPath.encode()
        o.obj_path = self._path
        sp = list(self.shape_param) + [0.0] * (6 - len(self.shape_param))
        o.shape_param = (capi.C.c_double * 6)(*sp)
        o.init_vel = (capi.C.c_double * 3)(*self.initVel)
        o.translation = (capi.C.c_float * 3)(*self.initTranslation)
        o.scale = (capi.C.c_float * 3)(*self.initScale)
        o.rotation_axis = (capi.C.c_float * 3)(*self.initRotationAxis)
        o.rotation_deg = float(self.initRotationDegree)
        if self.mMaterial is None:
            raise MpmError(1, "ObjInitInfo.mMaterial is not set")
        o.material = self.mMaterial.c_struct
        o.max_particles = int(self.max_particles)
        return o


class MPMSimulator:
    """MPMSimulator(gap, dt, nx, ny, nz, ppc, objInitInfoVec), MPMSimulator.cuh:92-96.

    Extra keyword arguments are fields of ``mpm_config`` (form, sort_every,
    bc_mode, gravity, ...)."""

    def __init__(self, gap: float, dt: float, nodeNumDimX: int, nodeNumDimY: int, nodeNumDimZ: int,
                 particleNumPerCell: int, objInitInfoVec: List[ObjInitInfo], **cfg):
        c = capi.default_config(h=float(gap), dt=float(dt), nx=int(nodeNumDimX), ny=int(nodeNumDimY),
                                nz=int(nodeNumDimZ), ppc=int(particleNumPerCell), **cfg)
        self._sim = capi.Sim(c)
        self.max_dt = float(dt)
        self.adp_dt = float(dt)       # MPMSimulator.cpp:184-185 (the reference never changes it)
        self._adaptive = int(cfg.get("dt_mode", capi.DT_FIXED)) != capi.DT_FIXED
        for info in objInitInfoVec:
            self._sim.add_object(info.c_struct())

    @property
    def sim(self) -> capi.Sim:
        return self._sim

    @property
    def current_time(self) -> float:
        return self._sim.current_time

    @property
    def current_frame(self) -> int:
        return self._sim.current_frame

    def step(self) -> None:
        """One step of adp_dt (MPMCUDA.cu:793-925).  With a dt controller
        (``dt_mode``) the step takes at most max_dt and adp_dt reports the dt
        the controller chose."""
        if self._adaptive:
            self._sim.step(self.max_dt)
            self.adp_dt = self._sim.last_dt
        else:
            self._sim.step(self.adp_dt)

    def totalParticlesNum(self) -> int:
        return self._sim.total

    def getGLParticlesPos(self) -> np.ndarray:
        """Positions of all groups, flattened xyz (MPMSimulator.cpp:362-371)."""
        return self._sim.positions().reshape(-1)

    def close(self) -> None:
        self._sim.close()


def OutputBego(fileNamePath: str, outPosData) -> None:
    """Write a BGEO particle file with one "position" attribute (FilesIO.cpp:6-23)."""
    capi.write_bgeo(fileNamePath, np.asarray(outPosData, np.float32).reshape(-1, 3))


__all__ = ["Material", "MaterialType", "ObjInitInfo", "MPMSimulator", "OutputBego", "MpmError", "capi"]
