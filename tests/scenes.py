"""Seeded test scenes and ray sets shared by the tests (host side only)."""
from __future__ import annotations

import numpy as np

from imbatracer_amd import native as N
from imbatracer_amd import scene as S

from imbatracer_amd.workloads import C1_CAMERA, C1_LIGHT, primary_rays, random_tri_scene as _rts, shadow_rays  # noqa: F401


def random_tri_scene(n: int, builder=N.BUILDER_BINNED_SAH, seed: int = 0x1234):
    return _rts(n, builder, seed)


def random_rays(n: int, seed: int = 5, box=(-0.2, 1.2)) -> np.ndarray:
    """Incoherent rays: origins in a box around the unit cube, uniform directions."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(box[0], box[1], (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True).astype(np.float32)
    return S.make_rays(o, d.astype(np.float32), np.float32(0.0), np.float32(3.4028235e38))


def rot(axis: int, angle: float) -> np.ndarray:
    c, s = np.cos(angle), np.sin(angle)
    m = np.eye(4, dtype=np.float64)
    a, b = [(1, 2), (0, 2), (0, 1)][axis]
    m[a, a], m[a, b], m[b, a], m[b, b] = c, -s, s, c
    return m


def instanced_scene(builder=N.BUILDER_SBVH):
    """Three meshes, seven instances with rotations, scales and translations."""
    b = S.SceneBuild()
    v0, i0 = S.random_triangles(600, 11)
    v1, i1 = S.icosphere(2, 0.3, (0.0, 0.0, 0.0), 1)
    v2, i2 = S.random_triangles(300, 12)
    i2[:, 3] = 2
    m0, m1, m2 = b.add_mesh(v0, i0), b.add_mesh(v1, i1), b.add_mesh(v2, i2)
    rng = np.random.default_rng(3)
    for k, m in enumerate([m0, m1, m2, m1, m0, m1, m2]):
        t = np.eye(4)
        t[:3, 3] = rng.uniform(-1.5, 1.5, 3)
        sc = np.diag([rng.uniform(0.5, 1.5)] * 3 + [1.0])
        mat = t @ rot(k % 3, rng.uniform(0, 6.28)) @ sc
        b.add_instance(m, mat.astype(np.float32))
    b.run(builder, 2)
    return b


def masked_scene():
    """Random triangles with texcoords and a checkerboard opacity mask on material 0."""
    v, i = S.random_triangles(3000, 21)
    rng = np.random.default_rng(4)
    tc = rng.random((len(v), 2), dtype=np.float32) * 3.0
    b = S.SceneBuild()
    b.add_mesh(v, i, tc)
    yy, xx = np.mgrid[0:8, 0:16]
    b.set_mask(0, ((xx + yy) % 2).astype(np.uint8))
    b.run(N.BUILDER_BINNED_SAH, 1)
    return b
