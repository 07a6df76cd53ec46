"""The CPU layout (BVH4 nodes + 4-triangle packets, core/cpu_adapter.cpp) and the CPU-baseline
traversal over it (oracle/cpu_bvh4.c), host only.

Pins:
  - both layout writers (BVH2 and BVH4) decode back to the builders' emission streams, which
    equal the reference builders' own streams (directly when oracle/_ref is built);
  - packet contents: v0, e1 = v0 - v1, e2 = v2 - v0 exactly, ids, padding and sentinels;
  - the BVH4 traversal's closest hits equal the BVH2 oracle's bit for bit (and brute force),
    its occlusion flags equal the oracle's, on random, instanced and masked scenes.
"""
import numpy as np
import pytest

import oracle
from imbatracer_amd import native as N
from imbatracer_amd import scene as S
from tests import scenes as T

INT_MIN_BITS = 0x80000000


def _bits(x):
    return np.ascontiguousarray(np.asarray(x, np.float32)).view(np.int32)


def _decode_bvh2(a, root, tri_base=0):
    """Emission stream of a BVH2 mesh tree without the parent/leaf boxes (which the layout
    does not store): node -> [0, 2, child boxes], leaf -> [1, count, ids]."""
    nodes, tris = a["nodes"], a["tris"]
    bits = tris[:, 3].view(np.uint32)
    out = []

    def leaf(first):
        ids, t = [], first
        while True:
            ids.append(int(tris[t + 1:t + 2, 3].view(np.int32)[0]) - tri_base)
            if bits[t + 2] == N.LEAF_END_BITS:
                break
            t += 3
        out.extend([1, len(ids)] + ids)

    def box(bb):
        return _bits([bb["lo_x"], bb["lo_y"], bb["lo_z"], bb["hi_x"], bb["hi_y"], bb["hi_z"]]).tolist()

    r = nodes[root]
    if r["right"] == N.DUMMY_CHILD:
        leaf(~int(r["left"]))
        return out
    stack = [root]
    while stack:
        c = stack.pop()
        if c < 0:
            leaf(~c)
            continue
        n = nodes[c]
        out.extend([0, 2] + box(n["left_bb"]) + box(n["right_bb"]))
        stack.extend([int(n["right"]), int(n["left"])])
    return out


def _decode_bvh4(a, root, tri_base=0):
    nodes, tris = a["nodes"], a["tris"]
    ib = tris.view(np.int32)
    out = []

    def leaf(first):
        ids, p = [], first
        while not np.all(ib[p] == np.int32(-2 ** 31)):
            lanes = ib[p + 12]
            ids.extend(int(x) - tri_base for x in lanes if x != np.int32(-2 ** 31))
            p += 13
        out.extend([1, len(ids)] + ids)

    def used(n):
        return [j for j in range(4) if n["children"][j] != 0]

    r = nodes[root]
    if len(used(r)) == 1:
        assert r["children"][1] == 0
        leaf(~int(r["children"][0]))
        return out
    stack = [root]
    while stack:
        c = stack.pop()
        if c < 0:
            leaf(~c)
            continue
        n = nodes[c]
        u = used(n)
        assert u == list(range(len(u))), "used slots come first"
        rec = [0, len(u)]
        for j in u:
            rec += _bits([n["min_x"][j], n["min_y"][j], n["min_z"][j], n["max_x"][j], n["max_y"][j], n["max_z"][j]]).tolist()
        out.extend(rec)
        for j in range(len(u), 4):
            assert n["min_x"][j] == np.float32(3.4028235e38) and n["max_x"][j] == np.float32(-3.4028235e38)
        stack.extend(int(n["children"][j]) for j in reversed(u))
    return out


def _strip_boxes(stream):
    """Builder stream without the parent box of nodes and the box of leaves."""
    s, out, pos = stream.tolist(), [], 0
    while pos < len(s):
        kind, count = s[pos], s[pos + 1]
        if kind == 0:
            out += [0, count] + s[pos + 8:pos + 8 + 6 * count]
            pos += 8 + 6 * count
        else:
            out += [1, count] + s[pos + 8:pos + 8 + count]
            pos += 8 + count
    return out


def _mesh(n, seed):
    if n == "ico":
        return S.icosphere(3, 0.4, (1.5, 0.4, 2.5), 1)
    return S.random_triangles(n, seed)


@pytest.mark.parametrize("n,seed", [(1, 3), (2, 4), (5, 5), (700, 6), (4000, 7), ("ico", 0)])
@pytest.mark.parametrize("builder", [N.BUILDER_SBVH, N.BUILDER_BINNED_SAH])
def test_layouts_decode_to_builder_streams(n, seed, builder):
    v, i = _mesh(n, seed)
    b = S.SceneBuild()
    b.add_mesh(v, i)
    b.run(builder, 1)
    b.cpu_layout(builder, 1)
    a2, a4 = b.arrays(), b.cpu_arrays()
    s2 = S.emit_mesh(v, i, builder, 2, N.COST_GPU_MESH)
    s4 = S.emit_mesh(v, i, builder, 4, N.COST_CPU)
    assert _decode_bvh2(a2, int(a2["instances"]["next"][0])) == _strip_boxes(s2)
    assert _decode_bvh4(a4, int(a4["instances"]["next"][0])) == _strip_boxes(s4)
    if oracle.ref_available():
        assert np.array_equal(s4, oracle.ref_emit_mesh(v, i, builder, 4, N.COST_CPU))
        assert np.array_equal(s2, oracle.ref_emit_mesh(v, i, builder, 2, N.COST_GPU_MESH))


def test_packets_and_sentinels():
    v, i = S.random_triangles(3001, 21)
    b = S.SceneBuild()
    b.add_mesh(v, i)
    b.run(N.BUILDER_SBVH, 1)
    b.cpu_layout(N.BUILDER_SBVH, 1)
    a = b.cpu_arrays()
    tris, ib = a["tris"], a["tris"].view(np.int32)
    pos = v[:, :3][i[:, :3].astype(np.int64)]           # (n, 3 vertices, xyz)
    seen = []
    sentinel = np.int32(-2 ** 31)
    leaves = [~int(c) for n in a["nodes"][:a["root"]] for c in n["children"] if c < 0]
    packets = 0
    for first in leaves:
        p = first
        while not np.all(ib[p] == sentinel):
            packets += 1
            ids = ib[p + 12]
            for lane, tid in enumerate(ids):
                geo = tris[p:p + 12, lane]
                if tid == sentinel:
                    assert np.all(geo == 0)
                    continue
                seen.append(int(tid))
                v0, v1, v2 = pos[tid].astype(np.float32)
                assert _bits(geo[0:3]).tolist() == _bits(v0).tolist()
                assert _bits(geo[3:6]).tolist() == _bits(v0 - v1).tolist()
                assert _bits(geo[6:9]).tolist() == _bits(v2 - v0).tolist()
                # n = e1 x e2 up to the rounding of the two products (its bits are pinned by the
                # traversal tests below, which compute n on the fly in the BVH2 oracle)
                e1, e2 = (v0 - v1).astype(np.float64), (v2 - v0).astype(np.float64)
                mag = np.abs(np.roll(e1, -1) * np.roll(e2, -2)) + np.abs(np.roll(e1, -2) * np.roll(e2, -1))
                assert np.all(np.abs(geo[9:12] - np.cross(e1, e2)) <= 2.0 ** -23 * mag)
            # unused lanes only at the end of a leaf's last packet
            assert list(ids != sentinel) == sorted(ids != sentinel, reverse=True)
            p += 13
    assert sorted(set(seen)) == list(range(3001))
    # the packet array is exactly the leaves' packets, each leaf closed by one sentinel
    assert len(tris) == 13 * packets + len(leaves)


def test_single_triangle_dummy_parent_and_top_level():
    v = np.array([[-1, -1, 1, 1], [1, -1, 1, 1], [0, 1, 1, 1]], np.float32)
    i = np.array([[0, 1, 2, 0]], np.uint32)
    b = S.SceneBuild()
    b.add_mesh(v, i)
    b.run(N.BUILDER_SBVH, 1)
    b.cpu_layout(N.BUILDER_SBVH, 1)
    a = b.cpu_arrays()
    # mesh dummy parent (node 0) and top-level dummy parent (root), slots 1-3 zero
    assert a["root"] == 1 and len(a["nodes"]) == 2
    for n in a["nodes"]:
        assert n["children"][0] < 0 and list(n["children"][1:]) == [0, 0, 0]
        assert np.all(n["min_x"][1:] == 0) and np.all(n["max_z"][1:] == 0)
    assert a["instances"]["pad"][0].tolist() == [-1, -1]
    o = np.array([[0, 0, -1], [0, 0, -2], [0.2, -0.3, 0], [0, 0, 0]], np.float32)
    d = np.array([[0, 0, 1], [0, 0, 1], [0, 0, 1], [1, 0, 0]], np.float32)
    rays = S.make_rays(o, d, 0.0, 1e30)
    got, bad = oracle.trace_cpu4(a, rays)
    want, _ = oracle.trace(b.arrays(), rays)
    assert bad == 0 and list(got["tri_id"]) == [0, 0, 0, -1]
    assert np.array_equal(got.view(np.int32), want.view(np.int32))


def test_instanced_top_level_records():
    b = T.instanced_scene()
    b.cpu_layout(N.BUILDER_SBVH, 2)
    a4, a2 = b.cpu_arrays(), b.arrays()
    inst = a4["instances"]
    assert sorted(inst["id"].tolist()) == list(range(7))
    # each instance record carries the instance's world->object rows and its mesh's BVH4 root
    by_id = {int(r["id"]): r for r in a2["instances"]}
    mesh_of_root2 = {}
    for r in inst:
        r2 = by_id[int(r["id"])]
        assert np.array_equal(_bits(r["transf"]), _bits(r2["transf"]))
        mesh_of_root2.setdefault(int(r2["next"]), set()).add(int(r["next"]))
    assert all(len(s) == 1 for s in mesh_of_root2.values())
    # leaves end with pad {-1, -1}; top-level nodes follow the mesh nodes
    assert inst["pad"][-1].tolist() == [-1, -1]
    assert a4["root"] > inst["next"].max()
    top = a4["nodes"][a4["root"]:]
    for k, n in enumerate(top):
        for c in n["children"]:
            assert c == 0 or c < 0 or c > a4["root"] + k


def _same(a, b):
    return np.array_equal(a.view(np.int32), b.view(np.int32))


@pytest.mark.parametrize("builder", [N.BUILDER_BINNED_SAH, N.BUILDER_SBVH])
def test_closest_hits_equal_bvh2_oracle(builder):
    b, v, i = T.random_tri_scene(65536, builder)
    b.cpu_layout(builder, 2)
    a2, a4 = b.arrays(), b.cpu_arrays()
    for rays in (T.primary_rays(256, 256), T.random_rays(20000)):
        want, bad2 = oracle.trace(a2, rays, threads=8)
        got, bad4 = oracle.trace_cpu4(a4, rays, threads=8)
        assert bad2 == 0 and bad4 == 0
        assert _same(got, want)
    # occlusion flags of the shadow rays
    prim = T.primary_rays(256, 256)
    sh = T.shadow_rays(prim, oracle.trace(a2, prim, threads=8)[0])
    o2, _ = oracle.trace(a2, sh, any_hit=True, threads=8)
    o4, bad = oracle.trace_cpu4(a4, sh, any_hit=True, threads=8)
    assert bad == 0 and np.array_equal(o2["tri_id"] >= 0, o4["tri_id"] >= 0)
    assert 0.2 < (o4["tri_id"] >= 0).mean() < 0.95


def test_closest_hits_equal_brute_force():
    b, v, i = T.random_tri_scene(1500, N.BUILDER_SBVH, seed=77)
    b.cpu_layout(N.BUILDER_SBVH, 1)
    rays = T.random_rays(3000, seed=78)
    got, bad = oracle.trace_cpu4(b.cpu_arrays(), rays)
    ids, ts = oracle.brute_closest(v[:, :3][i[:, :3].astype(np.int64)].reshape(-1, 9), rays)
    hit = ids >= 0
    assert bad == 0 and np.array_equal(got["tri_id"] >= 0, hit)
    assert np.allclose(got["tmax"][hit], ts[hit], rtol=1e-5)
    # ids agree except where another triangle lies at (nearly) the same distance
    assert (got["tri_id"][hit] == ids[hit]).mean() > 0.999


def test_instanced_and_masked_scenes():
    b = T.instanced_scene()
    b.cpu_layout(N.BUILDER_SBVH, 2)
    rays = T.random_rays(40000, seed=9, box=(-2.0, 2.0))
    for any_hit in (False, True):
        want, _ = oracle.trace(b.arrays(), rays, any_hit=any_hit)
        got, bad = oracle.trace_cpu4(b.cpu_arrays(), rays, any_hit=any_hit, threads=4)
        assert bad == 0
        if any_hit:
            assert np.array_equal(got["tri_id"] >= 0, want["tri_id"] >= 0)
        else:
            assert _same(got, want) and (want["tri_id"] >= 0).sum() > 500
    m = T.masked_scene()
    m.cpu_layout(N.BUILDER_SBVH, 1)
    rays = T.primary_rays(128, 128)
    for any_hit in (False, True):
        want, _ = oracle.trace(m.arrays(), rays, any_hit=any_hit)
        got, bad = oracle.trace_cpu4(m.cpu_arrays(), rays, any_hit=any_hit)
        assert bad == 0
        if any_hit:
            assert np.array_equal(got["tri_id"] >= 0, want["tri_id"] >= 0)
        else:
            assert _same(got, want)


def test_cpu_layout_requires_a_build_and_is_thread_independent():
    b = S.SceneBuild()
    for n, seed in ((500, 5), (1, 6), (40, 7)):
        v, i = S.random_triangles(n, seed)
        b.add_mesh(v, i)
    with pytest.raises(N.ImbaError):
        b.cpu_layout(N.BUILDER_SBVH, 1)
    b.run(N.BUILDER_SBVH, 1)
    b.cpu_layout(N.BUILDER_SBVH, 1)
    one = b.cpu_arrays()
    b.cpu_layout(N.BUILDER_SBVH, 4)
    four = b.cpu_arrays()
    for k in ("nodes", "instances", "tris"):
        assert one[k].tobytes() == four[k].tobytes()
    # per-mesh roots: the one-triangle mesh is a dummy parent; node links are global
    roots = sorted(set(one["instances"]["next"].tolist()))
    assert len(roots) == 3 and roots[0] == 0
    for k, n in enumerate(one["nodes"]):
        assert all(c == 0 or c < 0 or c > k for c in n["children"])
