"""The reference-shaped C++ interfaces (include/imba/ray_queue.hpp: TraversalData, RayQueue,
HashGrid) driven by a C++ program (tests/cpp/ray_queue_test.cpp) against the oracle."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "ray_queue_test.bin")


def test_ray_queue_and_hash_grid_cpp(gpu):
    assert os.path.exists(BIN), "tests/cpp/ray_queue_test.bin is not built (run __graft_entry__.build())"
    p = subprocess.run([BIN], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout
    assert "ray_queue_test OK" in p.stdout
