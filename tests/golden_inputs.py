"""Seeded inputs of the builder fixtures (shared by tests/golden/make_golden.py and the tests).

Every input is a pure function of its seed: random triangles from the library's MWC64X
generator (itself checked against the reference's), icospheres, boxes from numpy's PCG64,
and hand-made degenerate cases.
"""
from __future__ import annotations

import numpy as np

from imbatracer_amd import scene as S


def mesh_cases() -> dict:
    cases = {}
    for name, n, seed in (("rand64", 64, 7), ("rand1000", 1000, 0x1234), ("rand4096", 4096, 0x1234),
                          ("rand20000", 20000, 99)):
        cases[name] = S.random_triangles(n, seed)
    cases["ico3"] = S.icosphere(3, 0.4, (1.5, 0.4, 2.5), 5)
    # 100 identical triangles: zero centroid range on every axis (the NaN bin path)
    v = np.tile(np.array([[0, 0, 0, 1], [1, 0, 0, 1], [0, 1, 0, 1]], np.float32), (100, 1))
    i = np.array([[3 * t, 3 * t + 1, 3 * t + 2, 0] for t in range(100)], np.uint32)
    cases["same100"] = (v, i)
    # flat grid of triangles in the plane z = 0.25 (zero extent in z)
    g = 12
    verts, idx = [], []
    for y in range(g):
        for x in range(g):
            b = len(verts)
            verts += [(x, y, 0.25, 1), (x + 1, y, 0.25, 1), (x, y + 1, 0.25, 1)]
            idx.append((b, b + 1, b + 2, 0))
    cases["flat144"] = (np.array(verts, np.float32) / np.float32(g), np.array(idx, np.uint32))
    # long thin triangles spanning the scene (many SBVH spatial splits)
    rng = np.random.default_rng(5)
    n = 300
    a = rng.random((n, 3), dtype=np.float32)
    bb = rng.random((n, 3), dtype=np.float32)
    c = a + np.float32(0.01) * rng.random((n, 3), dtype=np.float32)
    vv = np.ones((3 * n, 4), np.float32)
    vv[0::3, :3], vv[1::3, :3], vv[2::3, :3] = a, bb, c
    cases["slivers300"] = (vv, np.array([[3 * t, 3 * t + 1, 3 * t + 2, 0] for t in range(n)], np.uint32))
    return cases


def box_cases() -> dict:
    rng = np.random.default_rng(1)
    c = rng.random((500, 3), dtype=np.float32)
    e = (rng.random((500, 3), dtype=np.float32) * np.float32(0.05)).astype(np.float32)
    cases = {"rand500": (np.concatenate([c - e, c + e], 1).astype(np.float32), c)}
    cases["single"] = (cases["rand500"][0][:1].copy(), c[:1].copy())
    same = np.tile(np.array([[0, 0, 0, 1, 1, 1]], np.float32), (10, 1))
    cases["same10"] = (same, np.tile(np.array([[0.5, 0.5, 0.5]], np.float32), (10, 1)))
    return cases
