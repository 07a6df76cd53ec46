"""The library's BVH builders against the reference's own builders.

With the reference sources present, oracle/_ref/libref_builders.so (compiled from
core/fast_bvh_builder.h, core/sbvh_builder.h) is run on the same inputs and the node/leaf
emission streams must be identical int32 for int32.  Everywhere, the streams must match the
SHA-256 recorded from the reference in tests/golden/builders.json.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
from imbatracer_amd import native as N
from imbatracer_amd import scene as S
from tests import golden_inputs as G

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "builders.json")))
MESH_VARIANTS = [(0, 2, 0), (0, 4, 2), (0, 2, 2), (0, 4, 0), (1, 2, 0), (1, 4, 2), (1, 2, 2), (1, 4, 0)]
BOX_VARIANTS = [(2, 1), (4, 2), (2, 0)]
MESH_CASES = G.mesh_cases()
BOX_CASES = G.box_cases()


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("name", sorted(MESH_CASES))
def test_mesh_inputs_are_the_recorded_ones(name):
    v, i = MESH_CASES[name]
    assert sha(v) + ":" + sha(i) == GOLDEN["mesh"][name]["input"]


@pytest.mark.parametrize("name", sorted(MESH_CASES))
@pytest.mark.parametrize("variant", MESH_VARIANTS, ids=lambda x: "b%d_w%d_c%d" % x)
def test_mesh_builders_match_reference_golden(name, variant):
    v, i = MESH_CASES[name]
    s = S.emit_mesh(v, i, *variant)
    assert [sha(s), len(s)] == GOLDEN["mesh"][name]["%d/%d/%d" % variant]


@pytest.mark.parametrize("name", sorted(BOX_CASES))
@pytest.mark.parametrize("variant", BOX_VARIANTS, ids=lambda x: "w%d_c%d" % x)
def test_box_builder_matches_reference_golden(name, variant):
    b, c = BOX_CASES[name]
    assert sha(b) + ":" + sha(c) == GOLDEN["boxes"][name]["input"]
    s = S.emit_boxes(b, c, *variant)
    assert [sha(s), len(s)] == GOLDEN["boxes"][name]["%d/%d" % variant]


def test_full_stream_readable():
    """One stream kept in full: 64 random triangles, binned SAH, BVH2."""
    v, i = MESH_CASES["rand64"]
    s = S.emit_mesh(v, i, 1, 2, 0)
    assert s.tolist() == GOLDEN["full"]["rand64/1/2/0"]
    # the stream parses: nodes have 2 children, leaves list every triangle exactly once
    pos, leaves = 0, []
    while pos < len(s):
        kind, count = s[pos], s[pos + 1]
        if kind == 0:
            assert count == 2
            pos += 2 + 6 + 6 * count
        else:
            leaves += s[pos + 8:pos + 8 + count].tolist()
            pos += 2 + 6 + count
    assert sorted(leaves) == list(range(64))


needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference harness not built (no reference sources)")


@needs_ref
@pytest.mark.parametrize("seed,n", [(11, 10), (12, 333), (13, 5000)])
@pytest.mark.parametrize("variant", MESH_VARIANTS, ids=lambda x: "b%d_w%d_c%d" % x)
def test_mesh_builders_match_reference_directly(seed, n, variant):
    v, i = S.random_triangles(n, seed)
    assert np.array_equal(S.emit_mesh(v, i, *variant), oracle.ref_emit_mesh(v, i, *variant))


@needs_ref
@pytest.mark.parametrize("level", [0, 2, 4])
def test_sbvh_icosphere_matches_reference_directly(level):
    v, i = S.icosphere(level, 0.4, (3.5, 0.4, 7.5), 1)
    for variant in ((0, 2, 0), (0, 4, 2)):
        assert np.array_equal(S.emit_mesh(v, i, *variant), oracle.ref_emit_mesh(v, i, *variant))


@needs_ref
def test_random_triangle_generator_matches_reference():
    """The C1 scene generator draws the same numbers as the reference's MWC64X RNG."""
    for seed, n in ((0x1234, 65536), (7, 64)):
        v, _ = S.random_triangles(n, seed)
        assert np.array_equal(v, oracle.ref_random_triangles(seed, n))


def test_random_triangle_generator_matches_golden():
    for key, digest in GOLDEN["random_triangles"].items():
        seed, n = (int(x) for x in key.split("/"))
        v, _ = S.random_triangles(n, seed)
        assert sha(v) == digest


def test_mwc64x_matches_golden():
    """Python restatement used by the workloads (render/random.h:19-46) against the draws the
    reference's RNG produced."""
    from imbatracer_amd.workloads import mwc_uniform

    for seed, rec in GOLDEN["rng"].items():
        got = mwc_uniform(int(seed), 64)
        assert got.view(np.uint32).tolist() == rec["random_float_bits"]


def test_scene_build_identity_instances_and_layout_offsets():
    """Several meshes: global triangle ids, per-mesh roots (layout) and the top level after
    the mesh nodes (render/scene.cpp:49-117)."""
    b = S.SceneBuild()
    sizes = [(50, 1), (7, 2), (1, 3)]
    for n, seed in sizes:
        v, i = S.random_triangles(n, seed)
        b.add_mesh(v, i)
    b.run(N.BUILDER_SBVH, 2)
    a = b.arrays()
    inst = a["instances"]
    assert len(inst) == 3 and a["root"] > max(inst["next"])
    ids = a["tris"][1::3, 3].view(np.int32)
    assert sorted(set(ids.tolist())) == list(range(58))
    # the third mesh is one triangle: its root is a dummy parent
    last = a["nodes"][inst["next"].max()]
    assert last["right"] == N.DUMMY_CHILD
    # threads do not change the result
    b2 = S.SceneBuild()
    for n, seed in sizes:
        v, i = S.random_triangles(n, seed)
        b2.add_mesh(v, i)
    b2.run(N.BUILDER_SBVH, 1)
    a2 = b2.arrays()
    for k in ("nodes", "instances", "tris", "indices"):
        assert a[k].tobytes() == a2[k].tobytes()
