"""The 4-wide tree derived from a scene's BVH2 at upload (host only).

Pins: every leaf of the BVH2 (triangle list or instance list) appears exactly once in the
4-wide tree, under the same box the BVH2 stores for it; every inner 4-wide node holds 2..4
children (1 only for the dummy parent of a single-leaf tree), unused slots are child 0 with an
empty box, children follow their parent; instance records keep transform and id and point at
their mesh's 4-wide root; the stack bound covers the pushes along every path.
"""
import numpy as np
import pytest

from imbatracer_amd import native as N
from imbatracer_amd import scene as S
from tests import scenes as T

FMAX = np.float32(3.4028235e38)


def _bvh2_leaves(a, root):
    """{leaf link: box bits} of the BVH2 tree at root."""
    out = {}
    stack = [root]
    nodes = a["nodes"]
    while stack:
        n = nodes[stack.pop()]
        for key, bb in (("left", "left_bb"), ("right", "right_bb")):
            c = int(n[key])
            if key == "right" and c == N.DUMMY_CHILD:
                continue
            if c >= 0:
                stack.append(c)
            else:
                b = n[bb]
                box = np.array([b["lo_x"], b["lo_y"], b["lo_z"], b["hi_x"], b["hi_y"], b["hi_z"]], np.float32)
                assert c not in out
                out[c] = box.view(np.int32).tolist()
    return out


def _wide_leaves(w, root):
    out, pushes = {}, 0
    stack = [(root, 0)]
    nodes = w["nodes"]
    while stack:
        k, depth = stack.pop()
        n = nodes[k]
        used = [j for j in range(4) if n["children"][j] != 0]
        assert used == list(range(len(used))) and 1 <= len(used) <= 4
        here = depth + len(used) - 1
        pushes = max(pushes, here)
        for j in range(len(used), 4):
            assert n["min_x"][j] == FMAX and n["max_x"][j] == -FMAX
        for j in used:
            c = int(n["children"][j])
            if c > 0:
                assert c > k
                stack.append((c, here))
            else:
                box = np.array([n["min_x"][j], n["min_y"][j], n["min_z"][j], n["max_x"][j], n["max_y"][j],
                                n["max_z"][j]], np.float32)
                assert c not in out
                out[c] = box.view(np.int32).tolist()
    return out, pushes


def _scenes():
    yield "sbvh", T.random_tri_scene(5000, N.BUILDER_SBVH)[0]
    yield "binned", T.random_tri_scene(5000, N.BUILDER_BINNED_SAH)[0]
    yield "instanced", T.instanced_scene()
    yield "masked", T.masked_scene()
    b = S.SceneBuild()
    b.add_mesh(np.array([[-1, -1, 1, 1], [1, -1, 1, 1], [0, 1, 1, 1]], np.float32), np.array([[0, 1, 2, 0]], np.uint32))
    b.run(N.BUILDER_SBVH, 1)
    yield "one-triangle", b


@pytest.mark.parametrize("name,build", list(_scenes()), ids=lambda x: x if isinstance(x, str) else "")
def test_same_leaves_under_same_boxes(name, build):
    a = build.arrays()
    w = S.wide_tree(a)
    # top level
    top2 = _bvh2_leaves(a, a["root"])
    top4, top_push = _wide_leaves(w, w["root"])
    assert top4 == top2
    # meshes
    roots2 = sorted(set(a["instances"]["next"].tolist()))
    mesh_push = 0
    for i, rec in enumerate(a["instances"]):
        r4 = int(w["instances"]["next"][i])
        leaves4, p = _wide_leaves(w, r4)
        mesh_push = max(mesh_push, p)
        assert leaves4 == _bvh2_leaves(a, int(rec["next"]))
        assert w["instances"][i]["transf"].tobytes() == rec["transf"].tobytes()
        assert w["instances"][i]["id"] == rec["id"] and list(w["instances"][i]["pad"]) == list(rec["pad"])
    assert w["max_stack"] >= top_push + mesh_push
    # fewer, wider nodes
    assert len(w["nodes"]) <= len(a["nodes"])
    if name == "sbvh":
        assert len(w["nodes"]) < 0.5 * len(a["nodes"])
    assert len(roots2) == len(set(w["instances"]["next"].tolist()))


def test_largest_child_is_opened_first():
    """A node's slots come from opening, largest surface area first, the inner children of the
    BVH2 node until four (or only leaves) remain."""
    b, _, _ = T.random_tri_scene(3000, N.BUILDER_SBVH)
    a = b.arrays()
    w = S.wide_tree(a)
    nodes2 = a["nodes"]
    mesh_root2 = int(a["instances"]["next"][0])
    root4 = int(w["instances"]["next"][0])

    def area(bb):
        dx, dy, dz = bb["hi_x"] - bb["lo_x"], bb["hi_y"] - bb["lo_y"], bb["hi_z"] - bb["lo_z"]
        return np.float32(dx * dy + dy * dz + dz * dx)

    n2 = nodes2[mesh_root2]
    slots = [(n2["left_bb"], int(n2["left"])), (n2["right_bb"], int(n2["right"]))]
    while len(slots) < 4:
        inner = [k for k, (_, c) in enumerate(slots) if c >= 0]
        if not inner:
            break
        k = max(inner, key=lambda j: (area(slots[j][0]), -j))
        c = nodes2[slots[k][1]]
        slots[k] = (c["left_bb"], int(c["left"]))
        slots.append((c["right_bb"], int(c["right"])))
    n4 = w["nodes"][root4]
    for j, (bb, c) in enumerate(slots):
        assert n4["min_x"][j] == bb["lo_x"] and n4["max_z"][j] == bb["hi_z"]
        if c < 0:
            assert n4["children"][j] == c


def _dequant(q, axis, j):
    """The kernel's child box on one axis, fma(q, 2^e, origin) rounded once to float32 (the
    float64 sum of these operands is exact)."""
    e = float(2.0 ** int(q["exp"][axis]))
    o = np.float64(q["origin"][axis])
    lo = np.float32(q["lo_" + "xyz"[axis]][j] * e + o)
    hi = np.float32(q["hi_" + "xyz"[axis]][j] * e + o)
    return lo, hi


@pytest.mark.parametrize("name,build", list(_scenes()), ids=lambda x: x if isinstance(x, str) else "")
def test_quantized_boxes_contain_the_exact_boxes(name, build):
    w = S.wide_tree(build.arrays())
    nodes, qn = w["nodes"], w["qnodes"]
    assert len(qn) == len(nodes)
    for n, q in zip(nodes, qn):
        assert list(q["children"]) == list(n["children"])
        k = int((n["children"] != 0).sum())
        assert q["count"] == k
        for j in range(k):
            for axis, (mn, mx) in enumerate((("min_x", "max_x"), ("min_y", "max_y"), ("min_z", "max_z"))):
                lo, hi = _dequant(q, axis, j)
                assert lo <= n[mn][j] and hi >= n[mx][j]
                # at most about one quantization step larger on each side
                parent = float(max(n[mx][:k]) - min(n[mn][:k]))
                step = 2.0 ** int(q["exp"][axis])
                assert n[mn][j] - lo <= step * 1.01 + 1e-6 and hi - n[mx][j] <= step * 1.01 + 1e-6
                assert step <= max(parent / 255.0 * 2.0, 2.0 ** -100)


@pytest.mark.parametrize("name,build", list(_scenes()), ids=lambda x: x if isinstance(x, str) else "")
def test_top_levels_breadth_first(name, build):
    """Each tree's first 256 nodes (the range a launch may stage in LDS) are its top levels in
    breadth-first order, children in slot order: node root + i is the i-th node a FIFO walk
    from the root meets."""
    w = S.wide_tree(build.arrays())
    nodes = w["nodes"]
    roots = {int(w["root"])} | {int(r) for r in w["instances"]["next"] if r >= 0}
    for root in roots:
        order, queue = [], [root]
        while queue and len(order) < 256:
            k = queue.pop(0)
            order.append(k)
            queue += [int(c) for c in nodes[k]["children"] if c > 0]
        assert order == list(range(root, root + len(order)))
