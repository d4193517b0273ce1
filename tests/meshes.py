"""Closed triangle meshes written as OBJ files for the tests (the reference
ships no .obj models: its models/*.obj are not in the repository)."""
from __future__ import annotations

import numpy as np


def write_obj(path, v, f):
    with open(path, "w") as fh:
        for p in v:
            fh.write("v %.17g %.17g %.17g\n" % tuple(p))
        for t in f:
            fh.write("f %d %d %d\n" % tuple(int(k) + 1 for k in t))


def unit_cube(half=0.5):
    """The axis-aligned cube [-half, half]^3, 12 triangles."""
    v = np.array([[x, y, z] for z in (-half, half) for y in (-half, half) for x in (-half, half)], float)
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    f = []
    for a, b, c, d in quads:
        f += [(a, b, c), (a, c, d)]
    return v, np.array(f)


def extrude(poly, caps, depth):
    """A prism: polygon `poly` ((a, b) vertices in order) with cap triangles
    `caps` (indices into poly) extruded over [0, depth] along the third
    coordinate."""
    n = len(poly)
    v = [(a, b, 0.0) for a, b in poly] + [(a, b, depth) for a, b in poly]
    f = []
    for a, b, c in caps:
        f.append((a, c, b))
        f.append((a + n, b + n, c + n))
    for i in range(n):
        j = (i + 1) % n
        f += [(i, j, j + n), (i, j + n, i + n)]
    return np.array(v, float), np.array(f)


def boat_hull(length=2.0, beam=1.0, height=0.6, wall=0.2, floor=0.2):
    """An open-topped hull: the U cross-section (beam x height, walls and
    floor `wall` / `floor` thick) extruded over `length` along x, centred at
    the origin (x along the keel, y up, z across)."""
    hb, hh = beam / 2, height / 2
    # U polygon in (z, y), counter-clockwise; caps as three quads
    poly = [(-hb, -hh), (hb, -hh), (hb, hh), (hb - wall, hh), (hb - wall, -hh + floor), (-hb + wall, -hh + floor),
            (-hb + wall, hh), (-hb, hh)]
    caps = [(1, 2, 3), (1, 3, 4), (0, 1, 4), (0, 4, 5), (0, 5, 6), (0, 6, 7)]
    pv, f = extrude(poly, caps, length)
    v = np.stack([pv[:, 2] - length / 2, pv[:, 1], pv[:, 0]], 1)
    return v, f
