"""BVH2 traversal arrays written by the library (the layout of core/gpu_adapter.cpp) and the
structure checks applied before a scene is traversed (host only, no device)."""
import ctypes as C

import numpy as np
import pytest

from imbatracer_amd import native as N
from imbatracer_amd import scene as S
from tests import scenes as T


def _validate(arrays):
    d, keep = S.desc_from_arrays(arrays)
    ms, op = C.c_int(), C.c_int()
    code = N.lib.imba_scene_validate(C.byref(d), C.byref(ms), C.byref(op))
    return code, ms.value, op.value


def _boxes(bb):
    return np.stack([bb["lo_x"], bb["lo_y"], bb["lo_z"], bb["hi_x"], bb["hi_y"], bb["hi_z"]], -1)


def _walk(a, root, top):
    """Yields (node, parent box, depth) and leaf lists below root."""
    nodes = a["nodes"]
    stack = [(root, None)]
    leaves = []
    while stack:
        n, pbox = stack.pop()
        rec = nodes[n]
        for slot, key in ((0, "left"), (1, "right")):
            c = int(rec[key])
            box = _boxes(rec["left_bb" if slot == 0 else "right_bb"])
            if slot == 1 and c == N.DUMMY_CHILD:
                continue
            if c >= 0:
                assert c > n, "child %d of %d breaks the depth-first order" % (c, n)
                child = nodes[c]
                # the child's own child boxes lie inside the box its parent stores for it
                inner = [_boxes(child["left_bb"])]
                if child["right"] != N.DUMMY_CHILD:
                    inner.append(_boxes(child["right_bb"]))
                for cb in inner:
                    assert np.all(cb[:3] >= box[:3]) and np.all(cb[3:] <= box[3:])
                stack.append((c, box))
            else:
                leaves.append((~c, box))
    return leaves


@pytest.mark.parametrize("builder", [N.BUILDER_SBVH, N.BUILDER_BINNED_SAH])
def test_mesh_layout_invariants(builder):
    v, i = S.random_triangles(3000, 42)
    b = S.SceneBuild()
    b.add_mesh(v, i)
    b.run(builder, 1)
    a = b.arrays()
    tris = a["tris"]
    inst = a["instances"]
    assert len(inst) == 1 and inst["pad"][0].tolist() == [-1, -1] and inst["id"][0] == 0
    assert np.allclose(inst["transf"][0], np.eye(4, dtype=np.float32)[:3])
    leaves = _walk(a, int(inst["next"][0]), top=False)
    seen = []
    bits = tris[:, 3].view(np.uint32)
    for start, box in leaves:
        t = start
        while True:
            tri = tris[t:t + 3, :3]
            tid = int(bits[t + 1])
            seen.append(tid)
            # the leaf box overlaps the triangle's box (SBVH leaves hold clipped references)
            lo, hi = tri.min(0), tri.max(0)
            assert np.all(lo <= box[3:] + 1e-6) and np.all(hi >= box[:3] - 1e-6)
            if builder == N.BUILDER_BINNED_SAH:
                assert np.all(lo >= box[:3]) and np.all(hi <= box[3:])
            assert bits[t] == 0 and (bits[t + 2] in (0, N.LEAF_END_BITS))
            if bits[t + 2] == N.LEAF_END_BITS:
                break
            t += 3
    if builder == N.BUILDER_BINNED_SAH:
        assert sorted(seen) == list(range(3000))
    else:
        assert set(seen) == set(range(3000)) and len(seen) >= 3000
    # top level: root after the mesh nodes, a dummy parent over one instance leaf
    root = a["nodes"][a["root"]]
    assert a["root"] == len(a["nodes"]) - 1 and root["right"] == N.DUMMY_CHILD and root["left"] == ~0
    rb = root["right_bb"]
    assert [rb["lo_x"], rb["lo_y"], rb["lo_z"]] == [0, 0, 0]
    assert all(np.signbit(rb[k]) for k in ("hi_x", "hi_y", "hi_z"))
    code, ms, op = _validate(a)
    assert code == 0 and op == 1 and ms > 2


def test_instanced_top_level_invariants():
    b = T.instanced_scene()
    a = b.arrays()
    inst = a["instances"]
    assert sorted(inst["id"].tolist()) == list(range(7))
    leaves = _walk(a, a["root"], top=True)
    covered = []
    for start, box in leaves:
        k = start
        while True:
            covered.append(int(inst["id"][k]))
            if inst["pad"][k][0] == -1:
                break
            assert inst["pad"][k].tolist() == [0, 0]
            k += 1
    assert sorted(covered) == list(range(7))
    code, ms, op = _validate(a)
    assert code == 0


def _scene():
    v, i = S.random_triangles(500, 3)
    b = S.SceneBuild()
    b.add_mesh(v, i)
    b.run(N.BUILDER_BINNED_SAH, 1)
    return b.arrays()


def test_validation_rejects_malformed_scenes():
    a = _scene()
    assert _validate(a)[0] == 0
    bad = dict(a)
    # a cycle: a child pointing back to its parent
    nodes = a["nodes"].copy()
    k = np.nonzero(nodes["left"] > 0)[0][0]
    nodes["left"][k] = k
    bad["nodes"] = nodes
    assert _validate(bad)[0] == 4
    assert "depth-first order" in N.lib.imba_last_error().decode()
    # a leaf past the triangle array
    nodes = a["nodes"].copy()
    k = np.nonzero(nodes["left"] < 0)[0][0]
    nodes["left"][k] = ~(len(a["tris"]) + 3)
    bad["nodes"] = nodes
    assert _validate(bad)[0] == 4
    # a missing leaf sentinel at the end of the triangle array
    tris = a["tris"].copy()
    tris[-1, 3] = 0.0
    bad = dict(a)
    bad["tris"] = tris
    assert _validate(bad)[0] == 4
    # an instance whose root is out of range
    inst = a["instances"].copy()
    inst["next"][0] = len(a["nodes"]) + 10
    bad = dict(a)
    bad["instances"] = inst
    assert _validate(bad)[0] == 4
    # root out of range
    bad = dict(a)
    bad["root"] = len(a["nodes"])
    assert _validate(bad)[0] == 4


def test_validation_reports_masks():
    a = T.masked_scene().arrays()
    code, ms, op = _validate(a)
    assert code == 0 and op == 0
    bad = dict(a)
    m = a["masks"].copy()
    m["offset"][0] = len(a["mask_buffer"])
    bad["masks"] = m
    assert _validate(bad)[0] == 4


def test_empty_mesh_rejected():
    b = S.SceneBuild()
    with pytest.raises(N.ImbaError):
        b.add_mesh(np.zeros((0, 4), np.float32), np.zeros((0, 4), np.uint32))
    with pytest.raises(N.ImbaError):
        b.add_mesh(np.zeros((3, 4), np.float32), np.array([[0, 1, 5, 0]], np.uint32))
