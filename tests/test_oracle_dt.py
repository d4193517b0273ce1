"""Pins of the oracle's step-size controller (oracle.controller_dt) against
closed forms: the P-wave speed from (E, nu, rho), the snow hardening
factor, water's sound speed from a finite difference of its equation of
state, and the reference's deformation factor min(10, max(J, 1/J)^3)
(MPMCUDA.cu:613-619, 700-708, 915-920).  Also the C ABI's checks of the
controller settings (no GPU needed: they run before any device call)."""
import math

import numpy as np
import pytest

import oracle as O
from mpm_cuda_amd import capi

H = 1.0 / 64
CFL = 0.5


def _group(type_, n=5, *, v=0.0, E=5e4, nu=0.3, rho=1e3, Fe=None, Fp=None, J=None, eos_k=1e4, eos_gamma=7.0):
    vol = H ** 3
    x = np.full((n, 3), 0.5)
    return O.Group(x, v, type=type_, mass=rho * vol, vol=vol, E=E, nu=nu, eos_k=eos_k, eos_gamma=eos_gamma,
                   Fe=Fe, Fp=Fp, J=J)


def p_wave_speed(E, nu, rho):
    # P-wave modulus M = E (1 - nu) / ((1 + nu) (1 - 2 nu))
    return math.sqrt(E * (1 - nu) / ((1 + nu) * (1 - 2 * nu)) / rho)


def test_at_rest_is_p_wave_cfl():
    g = _group(O.JELLO)
    c = p_wave_speed(5e4, 0.3, 1e3)
    assert abs(O.controller_dt([g], H, 1.0, CFL, O.DT_CFL) - CFL * H / c) < 1e-15
    # fixed mode and a max_dt below the CFL bound both return max_dt
    assert O.controller_dt([g], H, 1.0, CFL, O.DT_FIXED) == 1.0
    assert O.controller_dt([g], H, 1e-6, CFL, O.DT_CFL) == 1e-6


def test_speed_adds_particle_velocity():
    v = np.zeros((5, 3))
    v[2] = (3.0, 4.0, 0.0)
    g = _group(O.JELLO, v=v)
    c = p_wave_speed(5e4, 0.3, 1e3)
    assert abs(O.controller_dt([g], H, 1.0, CFL, O.DT_CFL) - CFL * H / (5.0 + c)) < 1e-15


def test_snow_hardening_raises_wave_speed():
    Jp = 0.9
    Fp = np.broadcast_to(np.eye(3) * Jp ** (1 / 3), (5, 3, 3))
    g = _group(O.SNOW, E=1.4e5, nu=0.2, rho=400.0, Fp=Fp)
    c = p_wave_speed(1.4e5, 0.2, 400.0) * math.exp(0.5 * 10.0 * (1 - Jp))
    assert abs(O.controller_dt([g], H, 1.0, CFL, O.DT_CFL) / (CFL * H / c) - 1) < 1e-12


def test_water_sound_speed_from_eos():
    k, gam, rho0, J = 1e4, 7.0, 1e3, 0.8
    g = _group(O.WATER, rho=rho0, J=np.full(5, J), eos_k=k, eos_gamma=gam)
    # c^2 = dp/drho at rho = rho0 / J, with p(J) = k (J^-gamma - 1) (MPMCUDA.cu:353)
    p = lambda j: k * (j ** -gam - 1.0)
    e = 1e-6
    dp_dJ = (p(J + e) - p(J - e)) / (2 * e)
    drho_dJ = -rho0 / J ** 2
    c = math.sqrt(dp_dJ / drho_dJ)
    assert abs(O.controller_dt([g], H, 1.0, CFL, O.DT_CFL) / (CFL * H / c) - 1) < 1e-7


@pytest.mark.parametrize("J,factor", [(1.0, 1.0), (1.5, 3.375), (0.5, 8.0), (0.2, 10.0)])
def test_deformation_factor(J, factor):
    Fe = np.broadcast_to(np.diag([J, 1.0, 1.0]), (5, 3, 3))
    g = _group(O.JELLO, Fe=Fe)
    max_dt = 1e-5                       # far below the CFL bound
    assert abs(O.controller_dt([g], H, max_dt, CFL, O.DT_CFL_DEFORM) - max_dt / factor) < 1e-18
    assert O.controller_dt([g], H, max_dt, CFL, O.DT_CFL) == max_dt


def test_water_does_not_enter_the_deformation_factor():
    g = _group(O.WATER, J=np.full(5, 0.5))
    assert O.controller_dt([g], H, 1e-5, CFL, O.DT_CFL_DEFORM) == 1e-5
    jel = _group(O.JELLO, Fe=np.broadcast_to(np.diag([0.5, 1.0, 1.0]), (5, 3, 3)))
    assert O.controller_dt([g, jel], H, 1e-5, CFL, O.DT_CFL_DEFORM) == 1e-5 / 8.0


@pytest.mark.parametrize("kw,msg", [
    ({"dt_mode": capi.DT_CFL, "form": capi.FORM_FUSED}, "FUSED"),
    ({"dt_mode": capi.DT_CFL, "cfl": 0.0}, "cfl"),
    ({"dt_mode": capi.DT_CFL_DEFORM, "cfl": 1.5}, "cfl"),
    ({"dt_mode": 7}, "dt_mode"),
])
def test_controller_settings_are_checked(kw, msg):
    cfg = capi.default_config(h=H, nx=64, ny=64, nz=64, **kw)
    with pytest.raises(capi.MpmError, match=msg):
        capi.Sim(cfg)


def test_default_config_is_the_reference_clock():
    cfg = capi.default_config()
    assert cfg.dt_mode == capi.DT_FIXED and abs(cfg.cfl - 0.5) < 1e-7
