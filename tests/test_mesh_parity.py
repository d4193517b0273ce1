"""The sampler's point-in-mesh test against the reference's own routine.

The reference decides which per-cell samples belong to an object with
mesh_query0.1's point_inside_mesh (MPMSimulator.cpp:251; exact predicates).
oracle/Makefile compiles that code from the reference's sources into
oracle/_ref (it is not committed and the reference is not present on every
machine, hence the skip).  Meshes here are generated (the reference ships no
.obj files): a subdivided icosphere, a torus, and a non-convex extruded L.
Random points are in general position; degenerate ones (rays through the
projected mesh's vertices and edges, points on faces) exercise the exact
orientation signs and the half-open tie-break rules, which must give the
reference's decision on every point."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
from mpm_cuda_amd import capi

MQ = oracle.mesh_query()
pytestmark = pytest.mark.skipif(MQ is None, reason="reference mesh_query not built (oracle/_ref needs /root/reference)")


def icosphere(levels=2, r=1.0):
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    v = [np.array(p, float) / np.linalg.norm(p) for p in v]
    for _ in range(levels):
        mid = {}
        nf = []

        def m(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                p = v[a] + v[b]
                v.append(p / np.linalg.norm(p))
                mid[k] = len(v) - 1
            return mid[k]
        for a, b, c in f:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return np.array(v) * r, np.array(f)


def torus(R=1.0, r=0.35, nu=24, nv=12):
    v, f = [], []
    for i in range(nu):
        for j in range(nv):
            a, b = 2 * np.pi * i / nu, 2 * np.pi * j / nv
            v.append(((R + r * np.cos(b)) * np.cos(a), (R + r * np.cos(b)) * np.sin(a), r * np.sin(b)))
    for i in range(nu):
        for j in range(nv):
            p, q = i * nv + j, ((i + 1) % nu) * nv + j
            p2, q2 = i * nv + (j + 1) % nv, ((i + 1) % nu) * nv + (j + 1) % nv
            f += [(p, q, q2), (p, q2, p2)]
    return np.array(v), np.array(f)


def extruded_L(h=0.6):
    poly = [(0, 0), (2, 0), (2, 0.7), (0.8, 0.7), (0.8, 2), (0, 2)]       # counter-clockwise
    n = len(poly)
    v = [(x, y, 0.0) for x, y in poly] + [(x, y, h) for x, y in poly]
    f = []
    # caps: fan from the reflex-free vertex 0 is not valid for an L; use explicit triangles
    caps = [(0, 1, 2), (0, 2, 3), (0, 3, 4), (0, 4, 5)]
    for a, b, c in caps:
        f.append((a, c, b))                 # bottom faces down
        f.append((a + n, b + n, c + n))     # top faces up
    for i in range(n):
        j = (i + 1) % n
        f += [(i, j, j + n), (i, j + n, i + n)]
    return np.array(v, float), np.array(f)


def write_obj(path, v, f):
    with open(path, "w") as fh:
        for p in v:
            fh.write("v %.17g %.17g %.17g\n" % tuple(p))
        for t in f:
            fh.write("f %d %d %d\n" % tuple(int(k) + 1 for k in t))


def reference_inside(v, f, pts):
    vv = np.ascontiguousarray(v, np.float64)
    ff = np.ascontiguousarray(f, np.int32)
    mesh = MQ.construct_mesh_object(len(vv), vv.ctypes.data_as(C.c_void_p), len(ff), ff.ctypes.data_as(C.c_void_p))
    try:
        out = np.empty(len(pts), bool)
        for i, p in enumerate(np.ascontiguousarray(pts, np.float64)):
            out[i] = MQ.point_inside_mesh(p.ctypes.data_as(C.c_void_p), mesh)
        return out
    finally:
        MQ.destroy_mesh_object(mesh)


@pytest.mark.parametrize("shape", ["icosphere", "torus", "L"])
def test_point_in_mesh_matches_reference(tmp_path, shape):
    v, f = {"icosphere": icosphere, "torus": torus, "L": extruded_L}[shape]()
    path = os.path.join(tmp_path, shape + ".obj")
    write_obj(path, v, f)
    rng = np.random.default_rng(5)
    lo, hi = v.min(0), v.max(0)
    pad = 0.1 * (hi - lo)
    pts = rng.uniform(lo - pad, hi + pad, size=(6000, 3))
    ours = capi.probe_point_in_mesh(path, pts)
    ref = reference_inside(v, f, pts)
    assert 0.05 < ref.mean() < 0.95, ref.mean()       # a real mix of inside and outside points
    mism = np.nonzero(ours != ref)[0]
    assert mism.size == 0, (shape, mism.size, pts[mism[:5]])


def degenerate_points(v, f, zs):
    """x-y positions on the projected mesh's vertices, on its edges (a
    quarter, half and three quarters along, rounded to doubles: on the edge
    or within an ulp of it) and at triangle centroids, each at heights zs."""
    xy = [v[:, :2]]
    for a, b in ((0, 1), (1, 2), (2, 0)):
        pa, pb = v[f[:, a], :2], v[f[:, b], :2]
        for t in (0.25, 0.5, 0.75):
            xy.append(pa + t * (pb - pa))
    xy.append(v[f].mean(1)[:, :2])
    xy = np.unique(np.concatenate(xy), axis=0)
    return np.array([(x, y, z) for x, y in xy for z in zs])


@pytest.mark.parametrize("shape", ["cube", "boat", "L", "icosphere"])
def test_point_in_mesh_degenerate_rays_match_reference(tmp_path, shape):
    from tests.meshes import boat_hull, unit_cube
    v, f = {"cube": unit_cube, "boat": boat_hull, "L": extruded_L, "icosphere": lambda: icosphere(1)}[shape]()
    path = os.path.join(tmp_path, shape + ".obj")
    write_obj(path, v, f)
    lo, hi = v.min(0)[2], v.max(0)[2]
    zs = sorted(set([lo - 0.5, lo, 0.5 * (lo + hi), hi, hi + 0.5] + list(np.unique(v[:, 2]))))
    pts = degenerate_points(v, f, zs)
    ours = capi.probe_point_in_mesh(path, pts)
    ref = reference_inside(v, f, pts)
    assert 0.02 < ref.mean() < 0.98, ref.mean()
    mism = np.nonzero(ours != ref)[0]
    assert mism.size == 0, (shape, mism.size, len(pts), pts[mism[:5]], ref[mism[:5]])
