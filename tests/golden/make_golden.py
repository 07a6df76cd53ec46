"""Writes tests/golden/builders.json from the reference's own code.

Run in a container that has the reference sources (it builds and uses
oracle/_ref/libref_builders.so, compiled from the reference's core/fast_bvh_builder.h,
core/sbvh_builder.h and render/random.h).  The fixture holds, per case, the SHA-256 and
length of the reference builder's node/leaf emission stream (format of imba_bvh_emit_*),
the SHA-256 of the input, reference random-number draws, and one short stream in full.
tests/test_builders.py compares the library against it where the reference is absent.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from tests import golden_inputs as G  # noqa: E402

VARIANTS = [(0, 2, 0), (0, 4, 2), (0, 2, 2), (0, 4, 0), (1, 2, 0), (1, 4, 2), (1, 2, 2), (1, 4, 0)]
BOX_VARIANTS = [(2, 1), (4, 2), (2, 0)]


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    if not oracle.ref_available():
        oracle.build(with_ref=True)
    out = {"mesh": {}, "boxes": {}, "rng": {}, "random_triangles": {}, "full": {}}
    for name, (v, i) in G.mesh_cases().items():
        case = {"input": sha(v) + ":" + sha(i)}
        for builder, width, cost in VARIANTS:
            s = oracle.ref_emit_mesh(v, i, builder, width, cost)
            case["%d/%d/%d" % (builder, width, cost)] = [sha(s), int(len(s))]
        out["mesh"][name] = case
    for name, (b, c) in G.box_cases().items():
        case = {"input": sha(b) + ":" + sha(c)}
        for width, cost in BOX_VARIANTS:
            s = oracle.ref_emit_boxes(b, c, width, cost)
            case["%d/%d" % (width, cost)] = [sha(s), int(len(s))]
        out["boxes"][name] = case
    for seed in (0, 1, 0x1234, 0xDEADBEEFCAFE):
        raw, uni = oracle.ref_rng(seed, 64)
        out["rng"][str(seed)] = {"random_int_0_7fffffff": raw.tolist(),
                                 "random_float_bits": uni.view(np.uint32).tolist()}
    for seed, n in ((0x1234, 65536), (7, 64)):
        out["random_triangles"]["%d/%d" % (seed, n)] = sha(oracle.ref_random_triangles(seed, n))
    v, i = G.mesh_cases()["rand64"]
    out["full"]["rand64/1/2/0"] = oracle.ref_emit_mesh(v, i, 1, 2, 0).tolist()
    path = os.path.join(HERE, "builders.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", path)


if __name__ == "__main__":
    main()
