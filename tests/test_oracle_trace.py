"""Pins of the traversal oracle (oracle/trace.c), independent of the GPU.

- brute force in double precision over the same triangles: the closest-hit triangle id must
  agree except where two candidate hits are within a relative 1e-6 of each other (near-ties
  the float and double arithmetic may order differently);
- closed-form single-triangle cases: centre hit at a known distance, edges and vertices are
  inside, a ray in the triangle's plane (det = 0) misses, tmin / tmax bounds are inclusive,
  axis-parallel rays (zero direction components) hit;
- instancing: an identity instance gives the same results as the mesh; a translated
  instance equals translated geometry; a scaled instance keeps t in world units;
- any hit agrees with closest hit on hit / no-hit; visit counts are consistent.
"""
import numpy as np
import pytest

import oracle
from imbatracer_amd import native as N
from imbatracer_amd import scene as S
from tests import scenes as T


def _one_tri_scene(tri, mat=None):
    v = np.array([[*p, 1.0] for p in tri], np.float32)
    b = S.SceneBuild()
    b.add_mesh(v, np.array([[0, 1, 2, 0]], np.uint32))
    if mat is not None:
        b.add_instance(0, mat)
    b.run(N.BUILDER_SBVH, 1)
    return b.arrays()


@pytest.mark.parametrize("builder", [N.BUILDER_BINNED_SAH, N.BUILDER_SBVH])
def test_closest_hit_equals_brute_force(builder):
    n = 3000
    rng = np.random.default_rng(77)  # triangles ~0.1 wide: most rays hit several
    c = rng.random((n, 1, 3), dtype=np.float32)
    v = np.ones((3 * n, 4), np.float32)
    v[:, :3] = (c + rng.uniform(-0.06, 0.06, (n, 3, 3)).astype(np.float32)).reshape(-1, 3)
    i = np.array([[3 * t, 3 * t + 1, 3 * t + 2, 0] for t in range(n)], np.uint32)
    b = S.SceneBuild()
    b.add_mesh(v, i)
    b.run(builder, 1)
    a = b.arrays()
    rays = np.concatenate([T.primary_rays(48, 48), T.random_rays(3000, seed=3)])
    hits, bad = oracle.trace(a, rays)
    assert bad == 0
    tv = v[:, :3].reshape(-1, 9)
    ids, ts = oracle.brute_closest(tv, rays)
    assert (ids >= 0).sum() > 500
    diff = np.nonzero(hits["tri_id"] != ids)[0]
    for k in diff:
        # allowed only for near-ties: the oracle's hit must also be a brute-force hit within 1e-6 t
        j = hits["tri_id"][k]
        assert j >= 0 and ids[k] >= 0, (k, j, ids[k])
        ij, tj = oracle.brute_closest(tv[j:j + 1], rays[k:k + 1])
        assert ij[0] == 0 and abs(tj[0] - ts[k]) <= 1e-6 * ts[k], (k, tj[0], ts[k])
    assert len(diff) <= 2
    ok = ids >= 0
    assert np.allclose(hits["tmax"][ok], ts[ok], rtol=1e-5)


def test_single_triangle_closed_form():
    # the test triangle (id 0) shares its leaf with a large triangle behind it (id 1), so the
    # leaf box is wider than the test triangle and the triangle test decides alone
    v = np.array([[0, 0, 1, 1], [1, 0, 1, 1], [0, 1, 1, 1],
                  [-10, -10, 5, 1], [10, -10, 5, 1], [0, 10, 5, 1]], np.float32)
    b = S.SceneBuild()
    b.add_mesh(v, np.array([[0, 1, 2, 0], [3, 4, 5, 0]], np.uint32))
    b.run(N.BUILDER_SBVH, 1)
    a = b.arrays()
    assert len(a["tris"]) == 6  # one leaf
    o = np.array([[0.25, 0.25, 0], [0, 0, 0], [1, 0, 0], [0, 1, 0], [0.5, 0.5, 0],  # centre, vertices, hypotenuse
                  [0.6, 0.6, 0], [-0.01, 0.2, 0],                                     # outside
                  [0.2, 0.2, 1], [0.25, 0.25, 0]], np.float32)                        # in plane; tmax short
    d = np.array([[0, 0, 1]] * 7 + [[1, 0, 0], [0, 0, 1]], np.float32)
    tmax = np.array([1e30] * 8 + [0.999], np.float32)
    rays = S.make_rays(o, d, np.float32(0.0), tmax)
    h, bad = oracle.trace(a, rays)
    assert bad == 0
    assert h["tri_id"].tolist() == [0, 0, 0, 0, 0, 1, 1, -1, -1]
    assert h["tmax"][0] == 1.0 and h["tmax"][5] == 5.0 and h["tmax"][7] == np.float32(1e30)
    # u is the weight of the second vertex: the ray through (1,0,1) has u = 1, through (0,1,1) u = 0
    assert h["u"][2] == 1.0 and h["u"][3] == 0.0 and h["u"][0] == 0.25
    # tmin/tmax inclusive: t exactly at tmin or tmax hits
    rays = S.make_rays(o[:1], d[:1], np.float32(1.0), np.float32(1.0))
    assert oracle.trace(a, rays)[0]["tri_id"][0] == 0
    rays = S.make_rays(o[:1], d[:1], np.float32(1.0000001), np.float32(2.0))
    assert oracle.trace(a, rays)[0]["tri_id"][0] == -1


def test_axis_parallel_and_zero_components():
    a = _one_tri_scene([(0, 0, 0), (0, 2, 0), (0, 0, 2)])  # triangle in the plane x = 0
    o = np.array([[-3, 0.5, 0.5], [3, 0.5, 0.5], [-3, 1.5, 1.5], [-3, 0.5, 0.5]], np.float32)
    d = np.array([[1, 0, 0], [-1, 0, 0], [1, 0, 0], [-1, 0, 0]], np.float32)
    h, _ = oracle.trace(a, S.make_rays(o, d, np.float32(0), np.float32(1e30)))
    assert h["tri_id"].tolist() == [0, 0, -1, -1]
    assert h["tmax"][0] == 3.0 and h["tmax"][1] == 3.0


def test_closest_of_stacked_triangles_and_equal_t_tie():
    # two coplanar overlapping triangles at z = 1 (ids 0, 1) and one nearer at z = 0.5 (id 2)
    v = np.array([[0, 0, 1, 1], [1, 0, 1, 1], [0, 1, 1, 1],
                  [0, 0, 1, 1], [1, 0, 1, 1], [0, 1, 1, 1],
                  [0, 0, .5, 1], [.3, 0, .5, 1], [0, .3, .5, 1]], np.float32)
    i = np.array([[0, 1, 2, 0], [3, 4, 5, 0], [6, 7, 8, 0]], np.uint32)
    b = S.SceneBuild()
    b.add_mesh(v, i)
    b.run(N.BUILDER_SBVH, 1)
    a = b.arrays()
    o = np.array([[0.1, 0.1, 0], [0.5, 0.2, 0]], np.float32)
    d = np.array([[0, 0, 1], [0, 0, 1]], np.float32)
    h, _ = oracle.trace(a, S.make_rays(o, d, np.float32(0), np.float32(1e30)))
    assert h["tri_id"].tolist() == [2, 0]  # nearer wins; equal t -> smaller id
    occ, _ = oracle.trace(a, S.make_rays(o, d, np.float32(0), np.float32(1e30)), any_hit=True)
    assert (occ["tri_id"] >= 0).tolist() == [True, True]


def test_instances_identity_translation_scale():
    tri = [(0, 0, 1), (1, 0, 1), (0, 1, 1)]
    o = np.array([[0.25, 0.25, 0], [5.25, 0.25, 0], [0.7, 0.7, 0]], np.float32)
    d = np.array([[0, 0, 1]] * 3, np.float32)
    rays = S.make_rays(o, d, np.float32(0), np.float32(1e30))
    plain = _one_tri_scene(tri)
    ident = _one_tri_scene(tri, np.eye(4, dtype=np.float32))
    hp, _ = oracle.trace(plain, rays)
    hi, _ = oracle.trace(ident, rays)
    assert hp.tobytes() == hi.tobytes()
    m = np.eye(4, dtype=np.float32)
    m[0, 3] = 5.0
    moved = _one_tri_scene(tri, m)
    hm, _ = oracle.trace(moved, rays)
    assert hm["tri_id"].tolist() == [-1, 0, -1] and hm["tmax"][1] == 1.0
    s = np.diag([2, 2, 2, 1]).astype(np.float32)
    scaled = _one_tri_scene(tri, s)  # triangle at z = 2, twice the size
    hs, _ = oracle.trace(scaled, rays)
    assert hs["tri_id"].tolist() == [0, -1, 0] and hs["tmax"][0] == 2.0


def test_instanced_scene_matches_flattened_geometry():
    """Seven transformed instances traced through the two-level BVH hit the same triangles as
    brute force over the geometry transformed to world space (up to near-ties)."""
    b = T.instanced_scene()
    a = b.arrays()
    rays = T.random_rays(3000, seed=21, box=(-2.0, 2.0))
    h, bad = oracle.trace(a, rays)
    assert bad == 0
    # world-space triangles per instance, ids = instance * 1e6 + tri
    inst = a["instances"]
    nodes = a["nodes"]
    tris = a["tris"]
    bits = tris[:, 3].view(np.uint32)
    world, ids = [], []
    for k in range(len(inst)):
        inv = np.vstack([inst["transf"][k], [0, 0, 0, 1]]).astype(np.float64)
        mat = np.linalg.inv(inv)
        # collect the triangles of the instanced mesh by walking its tree
        stack = [int(inst["next"][k])]
        seen = set()
        while stack:
            n = stack.pop()
            for c in (int(nodes[n]["left"]), int(nodes[n]["right"])):
                if c == N.DUMMY_CHILD:
                    continue
                if c >= 0:
                    stack.append(c)
                    continue
                t = ~c
                while True:
                    tid = int(bits[t + 1])
                    if tid not in seen:
                        seen.add(tid)
                        p = np.hstack([tris[t:t + 3, :3].astype(np.float64), np.ones((3, 1))]) @ mat.T
                        world.append(p[:, :3].reshape(9))
                        ids.append((int(inst["id"][k]), tid))
                    if bits[t + 2] == N.LEAF_END_BITS:
                        break
                    t += 3
    tv = np.array(world, np.float32)
    bi, bt = oracle.brute_closest(tv, rays)
    got = [(int(x["inst_id"]), int(x["tri_id"])) if x["tri_id"] >= 0 else None for x in h]
    want = [ids[j] if j >= 0 else None for j in bi]
    agree = sum(g == w for g, w in zip(got, want))
    assert sum(w is not None for w in want) > 80
    assert agree >= len(rays) - 3  # transformed-space float rounding may flip a grazing ray


def test_any_hit_agrees_with_closest_and_counts():
    b, v, i = T.random_tri_scene(5000, N.BUILDER_SBVH)
    a = b.arrays()
    rays = T.random_rays(2000, seed=4)
    hc, cc, _ = oracle.trace(a, rays, counts=True)
    ho, co, _ = oracle.trace(a, rays, any_hit=True, counts=True)
    assert np.array_equal(hc["tri_id"] >= 0, ho["tri_id"] >= 0)
    # any hit stops early: never more work than closest hit... per ray it may visit nodes
    # closest hit culled, so only the totals are compared
    assert co[:, 1].sum() <= cc[:, 1].sum() * 1.5
    # every ray tests the root; every hit needed an instance entry and a triangle test
    assert (cc[:, 0] >= 1).all()
    hit = hc["tri_id"] >= 0
    assert (cc[hit, 2] >= 1).all() and (cc[hit, 1] >= 1).all()
    mt, _ = oracle.trace(a, rays, threads=4)
    assert mt.tobytes() == hc.tobytes()
