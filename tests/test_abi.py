"""The C ABI: the library loads without a device and exports every function the public
headers declare; the record layouts agree between the headers and the Python views."""
import ctypes as C
import os
import re

import numpy as np

from imbatracer_amd import native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M)))


def test_every_declared_function_is_exported():
    names = declared("imba_hip.h") + declared("traversal_gpu.h")
    assert len(names) > 50
    missing = [n for n in names if not hasattr(N.lib, n)]
    assert not missing, missing
    for n in ("intersect_gpu_masked_instanced", "occluded_gpu_masked_instanced", "imba_intersect",
              "imba_hashgrid_query", "imba_compact_hits"):
        assert n in names


def test_oracle_exports():
    import oracle

    text = open(os.path.join(ROOT, "oracle", "oracle.h")).read()
    names = re.findall(r"\b(oracle_[a-z_]+)\s*\(", text)
    L = oracle.lib()
    assert names and all(hasattr(L, n) for n in names)


def test_version_and_errors_without_device():
    assert b"gfx950" in N.lib.imba_version()
    # an argument error is reported through imba_last_error on this thread
    assert N.lib.imba_compact_hits(None, None, None, 3, 10, None, None, None, None, None) == 1
    assert b"multiple of 4" in N.lib.imba_last_error()
    assert N.lib.imba_hashgrid_query(None, None, 16, 0, 1, 10, None, None, None, None) == 1


def test_record_layouts():
    assert N.NODE_DTYPE.fields["left"][1] == 48 and N.NODE_DTYPE.fields["right"][1] == 52
    assert N.NODE_DTYPE.fields["right_bb"][1] == 24
    assert N.INSTANCE_DTYPE.fields["id"][1] == 48 and N.INSTANCE_DTYPE.fields["next"][1] == 52
    assert N.HIT_DTYPE.fields["tmax"][1] == 8 and N.RAY_DTYPE.fields["dir"][1] == 16
    # the leaf-end sentinel is the bit pattern of -0.0f
    assert np.array([-0.0], np.float32).view(np.uint32)[0] == N.LEAF_END_BITS
    assert C.sizeof(N.SceneDesc) == 8 + 7 * 16
