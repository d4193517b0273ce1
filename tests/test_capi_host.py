"""C ABI checks that need no GPU: the library loads, exports every entry point
declared in include/mpm/mpm.h, and its host-only paths (material checks,
sampling, BGEO output, error reporting) behave as documented."""
import os
import re
import struct

import numpy as np
import pytest

import mpm_cuda_amd as M
from mpm_cuda_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "mpm", "mpm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mpm(?:2d)?_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(capi.EXPORTED)


def test_material_range_checks():
    m = capi.material(1e4, 0.3, 1e3, capi.JELLO)
    assert abs(m.E - 1e4) < 1e-9 and m.xi == 10.0 and m.theta_c == 8e-2 and m.eos_k == 2.51
    with pytest.raises(capi.MpmError, match="Youngs"):
        capi.material(-1.0, 0.3, 1.0, capi.JELLO)
    with pytest.raises(capi.MpmError, match="Poison"):
        capi.material(1e4, 0.6, 1.0, capi.JELLO)
    with pytest.raises(capi.MpmError):
        capi.material(1e4, 0.3, 1.0, 7)
    mt = M.Material(1e4, 0.4, 0.8, M.MaterialType.JELLO)
    assert abs(mt.mMu - 1e4 / 2.8) < 1e-9
    assert abs(mt.mLambda - 1e4 * 0.4 / (1.4 * 0.2)) < 1e-6
    with pytest.raises(capi.MpmError):
        M.Material(1e4, 0.3, 1.0, M.MaterialType.SAND)


def _obj(shape, params, ppc_mat=capi.JELLO, translation=(0, 0, 0), scale=(1, 1, 1), axis=(1, 0, 0), deg=0.0,
         max_particles=0, path=b""):
    o = capi.MpmObject()
    o.shape = shape
    o.obj_path = path
    o.shape_param = (capi.C.c_double * 6)(*(list(params) + [0.0] * (6 - len(params))))
    o.translation = (capi.C.c_float * 3)(*translation)
    o.scale = (capi.C.c_float * 3)(*scale)
    o.rotation_axis = (capi.C.c_float * 3)(*axis)
    o.rotation_deg = deg
    o.material = capi.material(1e4, 0.3, 1e3, ppc_mat)
    o.max_particles = max_particles
    return o


def test_sample_box_and_sphere():
    cfg = capi.default_config(h=1 / 64, nx=64, ny=64, nz=64, ppc=8)
    box = capi.sample_object(cfg, _obj(capi.SHAPE_BOX, (0.25, 0.25, 0.25, 0.5, 0.5, 0.5)))
    assert box.dtype == np.float32
    assert np.all(box >= 0.25) and np.all(box <= 0.5)
    assert abs(len(box) - 16 ** 3 * 8) < 0.02 * 16 ** 3 * 8
    sph = capi.sample_object(cfg, _obj(capi.SHAPE_SPHERE, (0.5, 0.5, 0.5, 0.15)))
    expect = 4 / 3 * np.pi * 0.15 ** 3 * 64 ** 3 * 8
    assert abs(len(sph) - expect) < 0.02 * expect
    assert np.all(np.linalg.norm(sph - 0.5, axis=1) <= 0.15 + 1e-6)
    # deterministic for a fixed seed, different for another
    again = capi.sample_object(cfg, _obj(capi.SHAPE_SPHERE, (0.5, 0.5, 0.5, 0.15)))
    assert np.array_equal(sph, again)
    cfg2 = capi.default_config(h=1 / 64, nx=64, ny=64, nz=64, ppc=8, seed=99)
    other = capi.sample_object(cfg2, _obj(capi.SHAPE_SPHERE, (0.5, 0.5, 0.5, 0.15)))
    assert not np.array_equal(sph[:100], other[:100])
    # trimmed subset keeps the order and exact count
    sub = capi.sample_object(cfg, _obj(capi.SHAPE_SPHERE, (0.5, 0.5, 0.5, 0.15), max_particles=1000))
    assert sub.shape == (1000, 3)


def test_sample_transform_and_range_errors():
    cfg = capi.default_config(h=1 / 64, nx=64, ny=64, nz=64, ppc=4)
    # unit box scaled by 0.2 and translated: M = T S R (model.cpp:202-206)
    pts = capi.sample_object(cfg, _obj(capi.SHAPE_BOX, (0, 0, 0, 1, 1, 1), translation=(0.3, 0.4, 0.5),
                                       scale=(0.2, 0.1, 0.2)))
    assert np.all(pts[:, 0] >= 0.3 - 1e-6) and np.all(pts[:, 0] <= 0.5 + 1e-6)
    assert np.all(pts[:, 1] >= 0.4 - 1e-6) and np.all(pts[:, 1] <= 0.5 + 1e-6)
    # 90 degrees about z maps local +x to world +y
    rot = capi.sample_object(cfg, _obj(capi.SHAPE_BOX, (0, 0, 0, 1, 0.5, 0.5), translation=(0.5, 0.3, 0.3),
                                       scale=(0.2, 0.2, 0.2), axis=(0, 0, 1), deg=90.0))
    assert rot[:, 1].max() - rot[:, 1].min() > 0.15 and rot[:, 0].max() - rot[:, 0].min() < 0.11
    with pytest.raises(capi.MpmError, match="lower bound"):
        capi.sample_object(cfg, _obj(capi.SHAPE_SPHERE, (0.05, 0.5, 0.5, 0.1)))
    with pytest.raises(capi.MpmError, match="upper bound"):
        capi.sample_object(cfg, _obj(capi.SHAPE_SPHERE, (0.95, 0.5, 0.5, 0.1)))


def test_sample_obj_mesh(tmp_path):
    """A closed OBJ cube (quads) samples the same set as the procedural box."""
    p = tmp_path / "cube.obj"
    v = [(x, y, z) for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (-0.5, 0.5)]
    faces = [(1, 3, 4, 2), (5, 6, 8, 7), (1, 2, 6, 5), (3, 7, 8, 4), (1, 5, 7, 3), (2, 4, 8, 6)]
    p.write_text("\n".join([f"v {a} {b} {c}" for a, b, c in v] + [f"f {' '.join(map(str, f))}" for f in faces]) + "\n")
    cfg = capi.default_config(h=1 / 64, nx=64, ny=64, nz=64, ppc=4)
    mesh = capi.sample_object(cfg, _obj(capi.SHAPE_MESH, (), translation=(0.5, 0.5, 0.5), scale=(0.3, 0.3, 0.3),
                                        path=str(p).encode()))
    box = capi.sample_object(cfg, _obj(capi.SHAPE_BOX, (-0.5, -0.5, -0.5, 0.5, 0.5, 0.5), translation=(0.5, 0.5, 0.5),
                                       scale=(0.3, 0.3, 0.3)))
    assert abs(len(mesh) - len(box)) <= 0.002 * len(box)
    with pytest.raises(capi.MpmError, match="cannot open"):
        capi.sample_object(cfg, _obj(capi.SHAPE_MESH, (), path=b"/nonexistent.obj"))


def read_bgeo(path):
    b = open(path, "rb").read()
    magic, ver_c, ver = struct.unpack(">IcI", b[:9])
    counts = struct.unpack(">8I", b[9:41])
    n = counts[0]
    pts = np.frombuffer(b[41:41 + 16 * n], dtype=">f4").reshape(n, 4)
    return magic, ver_c, ver, counts, pts, b[41 + 16 * n:]


def test_bgeo_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    xyz = rng.random((1234, 3)).astype(np.float32)
    path = str(tmp_path / "Frame000000.bgeo")
    M.OutputBego(path, xyz.reshape(-1))
    magic, vc, ver, counts, pts, tail = read_bgeo(path)
    assert magic == int.from_bytes(b"Bgeo", "big") and vc == b"V" and ver == 5
    assert counts == (1234, 0, 0, 0, 0, 0, 0, 0)
    assert np.array_equal(pts[:, :3].astype(np.float32), xyz) and np.all(pts[:, 3] == 1.0)
    assert tail == b"\x00\xff"
    empty = str(tmp_path / "empty.bgeo")
    capi.write_bgeo(empty, np.zeros((0, 3), np.float32))
    assert read_bgeo(empty)[3][0] == 0
    with pytest.raises(capi.MpmError):
        capi.write_bgeo(str(tmp_path / "no_dir" / "x.bgeo"), xyz)


def test_config_tile_shape_checked_before_device():
    """The FAST tile shape is one of 256, 384, 512 (0 = default); anything else
    is refused before a device is touched."""
    for bad in (1, 128, 400, 1024, -384):
        with pytest.raises(capi.MpmError, match="tile"):
            capi.Sim(capi.default_config(nx=32, ny=32, nz=32, h=1 / 32, tile=bad))


def test_bandwidth_probe_arguments_checked_before_device():
    """The bandwidth probes refuse bad sizes before touching a device."""
    for args in ((0, 13, 31), (1 << 29, 13, 31), (1000, 0, 31), (1000, 13, 65)):
        with pytest.raises(capi.MpmError, match="range"):
            capi.probe_soa_bandwidth(*args, reps=5)
    with pytest.raises(capi.MpmError, match="reps"):
        capi.probe_soa_bandwidth(1000, 13, 31, reps=0)
    with pytest.raises(capi.MpmError, match="multiple of 16"):
        capi.probe_copy_bandwidth(1000, 5)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_create_without_gpu_reports_error():
    with pytest.raises(capi.MpmError, match="no HIP device"):
        capi.Sim(capi.default_config())
