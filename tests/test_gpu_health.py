"""The health bits (MPM_H_*) on the paths that set them, in every form: the
reference stops on these conditions with asserts and printf (MPMCUDA.cu:41-42
singular F, :722-739 a particle out of bounds); here a step flags them and
goes on (stencils are clamped to the grid, so nothing is read or written
outside it), and mpm_health reports and clears the bits.

Each case injects one bad particle through the C ABI and takes one step:
* an inverted deformation gradient (det F < 0): MPM_H_DET_NONPOS
  (jello det F_new, snow J_p after the clamp);
* a non-finite velocity: its P2G spreads it into the grid, the G2P of its
  neighbours reads it back: MPM_H_NONFINITE;
* a particle advected out of the grid (no wall band, 1000 m/s towards the
  x = 0 face): MPM_H_LEFT_DOMAIN.
"""
import numpy as np
import pytest

from mpm_cuda_amd import capi
from tests.scenes import MAT_PARAMS, Pair

pytestmark = pytest.mark.gpu

FORMS = {"plain": capi.FORM_PLAIN, "fast": capi.FORM_FAST, "fused": capi.FORM_FUSED}


@pytest.mark.parametrize("form", list(FORMS))
@pytest.mark.parametrize("type_", [capi.JELLO, capi.SNOW], ids=["jello", "snow"])
def test_inverted_F_sets_det_nonpos(gpu, form, type_):
    p = Pair(type_=type_, form=FORMS[form], seed=61, randomize=False, n_grid=32,
             box=((0.3, 0.3, 0.3), (0.5, 0.5, 0.5)))
    try:
        p.gpu.step(p.dt)
        assert p.gpu.health() & ~capi.H_ARENA_SPILL == 0          # clean before the injection
        F = p.gpu.read(0, capi.P_F)
        F[7] = np.diag([-1.0, 1.0, 1.0]).reshape(9)
        p.gpu.write(0, capi.P_F, F)
        p.gpu.step(p.dt)
        bits = p.gpu.health()
        print(form, type_, f"health {bits:#x}")
        assert bits & capi.H_DET_NONPOS
        assert p.gpu.health() == 0                                # read and cleared
    finally:
        p.close()


@pytest.mark.parametrize("form", list(FORMS))
def test_nonfinite_velocity_sets_nonfinite(gpu, form):
    p = Pair(type_=capi.WATER, form=FORMS[form], seed=62, randomize=False, n_grid=32,
             box=((0.3, 0.3, 0.3), (0.5, 0.5, 0.5)))
    try:
        p.gpu.step(p.dt)
        assert p.gpu.health() & ~capi.H_ARENA_SPILL == 0
        v = p.gpu.read(0, capi.P_VEL)
        v[11] = (np.nan, 0.0, 0.0)
        p.gpu.write(0, capi.P_VEL, v)
        p.gpu.step(p.dt)
        bits = p.gpu.health()
        print(form, f"health {bits:#x}")
        assert bits & capi.H_NONFINITE
    finally:
        p.close()


@pytest.mark.parametrize("form", list(FORMS))
def test_particle_leaving_the_grid_sets_left_domain(gpu, form):
    n_grid = 32
    h = 1.0 / n_grid
    cfg = capi.default_config(h=h, nx=n_grid, ny=n_grid, nz=n_grid, dt=1e-4, bc_pad=0, gravity=(0.0, 0.0, 0.0),
                              form=FORMS[form])
    sim = capi.Sim(cfg)
    try:
        E, nu, rho = MAT_PARAMS[capi.JELLO]
        x = np.array([[0.6 * h, 0.5, 0.5], [0.5, 0.5, 0.5]], np.float32)
        v = np.array([[-1000.0, 0.0, 0.0], [0.0, 0.0, 0.0]], np.float32)
        sim.add_particles(capi.material(E, nu, rho, capi.JELLO), x, vel=v, volume=h ** 3)
        sim.step(1e-4)
        bits = sim.health()
        xn = sim.read(0, capi.P_POS)
        print(form, f"health {bits:#x}", "x of the fast particle", xn[0])
        assert xn[0, 0] < 0.0
        assert bits & capi.H_LEFT_DOMAIN
        assert np.allclose(xn[1], x[1])                           # the other particle stays put
    finally:
        sim.close()
