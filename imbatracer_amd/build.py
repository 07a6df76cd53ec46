"""Builds libimba_hip.so (host C++ + HIP kernels for gfx950) in-tree.

Host sources are compiled with g++, HIP sources with hipcc --offload-arch=gfx950, both with
-ffp-contract=off (the traversal arithmetic relies on explicit fma only) and the link is done
by hipcc.  Every call rebuilds from scratch.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import tempfile

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB_NAME = "libimba_hip.so"
LIB_PATH = os.path.join(PKG, LIB_NAME)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

HOST_SOURCES = ["host/vec.cpp", "host/status.cpp", "host/scene_build.cpp", "host/gen.cpp", "host/bvh_emit.cpp",
                "host/abi_check.cpp", "host/bvh_cache.cpp", "host/wide_bvh.cpp"]
HIP_SOURCES = ["hip/runtime.hip", "hip/traverse.hip", "hip/compact.hip", "hip/hashgrid.hip", "hip/pt.hip",
               "hip/sort.hip", "hip/photon.hip", "hip/shade_batch.hip", "hip/probe.hip",
               "hip/vcm.hip"]

HOST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-Wall", "-I" + INCLUDE, "-pthread"]
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-std=c++17", "-O3", "-fPIC", "-ffp-contract=off",
    # No packed fp32 math (v_pk_*_f32) anywhere in the device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I" + INCLUDE,
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("build step failed:\n  %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose: bool = False) -> str:
    """Compiles every source and links LIB_PATH. Returns the library path."""
    tmp = tempfile.mkdtemp(prefix="imba_build_")
    try:
        jobs = []
        objs = []
        for src in HOST_SOURCES:
            obj = os.path.join(tmp, os.path.basename(src) + ".o")
            objs.append(obj)
            jobs.append([CXX] + HOST_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        for src in HIP_SOURCES:
            obj = os.path.join(tmp, os.path.basename(src) + ".o")
            objs.append(obj)
            jobs.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        workers = min(len(jobs), max(1, (os.cpu_count() or 4)))
        with cf.ThreadPoolExecutor(max_workers=workers) as ex:
            for out in ex.map(_run, jobs):
                if verbose and out.strip():
                    print(out)
        tmp_lib = LIB_PATH + ".tmp"
        _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp_lib] + objs + ["-pthread"])
        os.replace(tmp_lib, LIB_PATH)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
