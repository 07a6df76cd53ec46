"""Acceleration-structure cache files (the container of loaders/load_bvh.cpp), host only.

Pins: the byte layout of the container (magic, block size/type/counts, records), the rebasing
rules of load_bvh.cpp:66-95 (BVH4 children and packet ids) and :121-139 (BVH2 links and v1.w
ids) and their inverses on store, one block per type per file, and Scene::build_mesh_accels'
load-or-build-and-store flow (render/scene.cpp:49-70) giving arrays identical to a fresh build.
"""
import struct

import numpy as np
import pytest

from imbatracer_amd import native as N
from imbatracer_amd import scene as S

SENT = np.int32(-2 ** 31)


def _meshes():
    out = []
    for n, seed in ((700, 1), (1, 2), (90, 3)):
        out.append(S.random_triangles(n, seed))
    out.append(S.icosphere(2, 0.3, (0.5, 0.5, 0.5), 4))
    return out


def _build(meshes, paths=None, cpu=True):
    b = S.SceneBuild()
    for k, (v, i) in enumerate(meshes):
        b.add_mesh(v, i)
        if paths:
            b.set_cache(k, paths[k])
    b.run(N.BUILDER_SBVH, 2)
    stats = [b.cache_stats()]
    if cpu:
        b.cpu_layout(N.BUILDER_SBVH, 2)
        stats.append(b.cache_stats())
    return b, stats


def _mesh_slices(a, four):
    """Per mesh (in mesh order): node range and Vec4 range of its tree in the global arrays."""
    roots = sorted(set(a["instances"]["next"].tolist()))
    ends = roots[1:] + [a["root"]]
    out = []
    for r, e in zip(roots, ends):
        nodes = a["nodes"][r:e]
        kids = nodes["children"].ravel() if four else np.concatenate([nodes["left"], nodes["right"]])
        leaves = [~int(c) for c in kids if c < 0]
        out.append((r, e, min(leaves)))
    return [(r, e, t0, out[k + 1][2] if k + 1 < len(out) else len(a["tris"])) for k, (r, e, t0) in enumerate(out)]


def test_container_bytes(tmp_path):
    v, i = S.random_triangles(300, 9)
    b = S.SceneBuild()
    b.add_mesh(v, i)
    b.run(N.BUILDER_SBVH, 1)
    a = b.arrays()
    nodes = a["nodes"][:a["root"]]
    f = tmp_path / "m.bvh"
    S.bvh_store(f, N.CACHE_BVH2, nodes, a["tris"], 0, 0, 5)
    raw = f.read_bytes()
    magic, size, btype, nc, pc = struct.unpack_from("<IQIII", raw, 0)
    assert magic == 0x313F1A57 and btype == 1 and nc == len(nodes) and pc == len(a["tris"])
    assert size == 4 + 8 + 64 * nc + 16 * pc and len(raw) == 4 + 8 + size
    got_nodes = np.frombuffer(raw, N.NODE_DTYPE, nc, 24)
    assert got_nodes.tobytes() == nodes.tobytes()
    tris = np.frombuffer(raw, np.float32, 4 * pc, 24 + 64 * nc).reshape(-1, 4)
    ids_file = tris[1::3, 3].view(np.int32)
    assert np.array_equal(ids_file, a["tris"][1::3, 3].view(np.int32) - 5)
    # everything but the ids is stored verbatim
    mask = np.ones_like(tris, bool)
    mask[1::3, 3] = False
    assert np.array_equal(tris[mask].view(np.int32), a["tris"][mask].view(np.int32))


def test_round_trip_and_one_block_per_type(tmp_path):
    b, _ = _build(_meshes()[:1])
    a2, a4 = b.arrays(), b.cpu_arrays()
    f = tmp_path / "mesh0.bvh"
    S.bvh_store(f, N.CACHE_BVH2, a2["nodes"][:a2["root"]], a2["tris"])
    S.bvh_store(f, N.CACHE_BVH4, a4["nodes"][:a4["root"]], a4["tris"])
    with pytest.raises(N.ImbaError):
        S.bvh_store(f, N.CACHE_BVH2, a2["nodes"][:a2["root"]], a2["tris"])
    assert N.lib.imba_bvh_has_block(str(f).encode(), 1) == 1 and N.lib.imba_bvh_has_block(str(f).encode(), 2) == 1
    n2, t2 = S.bvh_load(f, N.CACHE_BVH2)
    n4, t4 = S.bvh_load(f, N.CACHE_BVH4)
    assert n2.tobytes() == a2["nodes"][:a2["root"]].tobytes() and t2.tobytes() == a2["tris"].tobytes()
    assert n4.tobytes() == a4["nodes"][:a4["root"]].tobytes() and t4.tobytes() == a4["tris"].tobytes()


@pytest.mark.parametrize("four", [False, True])
def test_rebasing_of_a_mesh_inside_a_scene(tmp_path, four):
    """Storing mesh k's slice of the global arrays with its offsets and loading it back at the
    same offsets reproduces the slice; in the file its links and ids start from zero."""
    b, _ = _build(_meshes())
    a = b.cpu_arrays() if four else b.arrays()
    btype = N.CACHE_BVH4 if four else N.CACHE_BVH2
    tri_first = np.cumsum([0] + [len(i) for _, i in _meshes()])
    for k, (n0, n1, t0, t1) in enumerate(_mesh_slices(a, four)):
        f = tmp_path / ("m%d_%d.bvh" % (k, four))
        S.bvh_store(f, btype, a["nodes"][n0:n1], a["tris"][t0:t1], n0, t0, int(tri_first[k]))
        local_nodes, local_tris = S.bvh_load(f, btype)
        if four:
            kids = local_nodes["children"].ravel()
            assert kids.max() < len(local_nodes) and (kids[kids < 0] >= ~len(local_tris) + 0).all()
            ib, p, ids = local_tris.view(np.int32), 0, []
            while p < len(ib):
                if np.all(ib[p] == SENT):
                    p += 1
                    continue
                ids.extend(int(x) for x in ib[p + 12] if x != SENT)
                p += 13
            assert min(ids) == 0 and max(ids) == len(_meshes()[k][1]) - 1
        else:
            ids = local_tris[1::3, 3].view(np.int32)
            assert ids.min() == 0 and ids.max() == len(_meshes()[k][1]) - 1
        g_nodes, g_tris = S.bvh_load(f, btype, n0, t0, int(tri_first[k]))
        assert g_nodes.tobytes() == a["nodes"][n0:n1].tobytes()
        assert g_tris.tobytes() == a["tris"][t0:t1].tobytes()


def test_build_loads_what_it_stored(tmp_path):
    meshes = _meshes()
    paths = [tmp_path / ("mesh%d.bvh" % k) for k in range(len(meshes))]
    b1, stats1 = _build(meshes, paths)
    assert stats1 == [(0, 4), (0, 4)]
    assert all(p.exists() for p in paths)
    b2, stats2 = _build(meshes, paths)
    assert stats2 == [(4, 0), (4, 0)]
    for get in ("arrays", "cpu_arrays"):
        x, y = getattr(b1, get)(), getattr(b2, get)()
        for k in ("nodes", "instances", "tris", "indices"):
            assert x[k].tobytes() == y[k].tobytes(), (get, k)
        assert x["root"] == y["root"]
    # the single-triangle mesh keeps its dummy parent through the cache
    a = b2.arrays()
    roots = sorted(set(a["instances"]["next"].tolist()))
    assert a["nodes"][roots[1]]["right"] == N.DUMMY_CHILD


def test_unreadable_cache_files_are_rebuilt(tmp_path):
    meshes = _meshes()[:2]
    paths = [tmp_path / "a.bvh", tmp_path / "b.bvh"]
    paths[0].write_bytes(b"not a cache file")
    paths[1].write_bytes(struct.pack("<IQI", 0x313F1A57, 1 << 40, 1))   # block larger than the file
    b, stats = _build(meshes, paths, cpu=False)
    assert stats == [(0, 2)]
    fresh, _ = _build(meshes, None, cpu=False)
    assert b.arrays()["nodes"].tobytes() == fresh.arrays()["nodes"].tobytes()
    # the unreadable first file was replaced; the second kept its bad block and got a new one
    b3, stats3 = _build(meshes, paths, cpu=False)
    assert stats3[0][0] >= 1
    with pytest.raises(N.ImbaError):
        S.bvh_load(tmp_path / "missing.bvh", N.CACHE_BVH2)
