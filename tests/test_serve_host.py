"""Host-side parts of the request loop: the device arena's first-fit allocation and merging,
request footprints and the Poisson request generator (no device)."""
import ctypes as C

import numpy as np

from imbatracer_amd import serve as V


class _FakeBuf:
    ptr = C.c_void_p(0)


def test_arena_first_fit_and_merge():
    a = V.DeviceArena(10 * 256, buf=_FakeBuf())
    x = a.alloc(300)            # rounds up to 512
    y = a.alloc(256)
    z = a.alloc(1000)           # 1024
    assert x == (0, 512) and y == (512, 256) and z == (768, 1024)
    assert a.alloc(2000) is None
    a.release(*y)
    assert a.alloc(200) == (512, 256)    # first fit reuses the hole
    a.release(512, 256)
    a.release(*x)
    a.release(*z)
    assert a.free == [(0, 2560)]           # everything merged back


def test_footprint_covers_the_buffers():
    r = V.Request(0, 0, 0, 640, 480, 3, (0, 0, 0), (0, 0, 1), 1, 0.0)
    n = 640 * 480 * 3
    assert V.footprint(r) >= n * (2 * 32 + 2 * 32 + 16) + 640 * 480 * 12


def test_vcm_footprint_covers_the_buffers():
    r = V.Request(0, 0, 0, 640, 480, 2, (0, 0, 0), (0, 0, 1), 1, 0.0, integrator="vcm")
    p = 640 * 480
    cap, slots = max(p, 2 * p), V.VCM_DEPTH * p
    # rays, states, hits, shadow slots (ray, state, hit), vertex slots, cache, merge records,
    # merge queries, image
    need = (2 * 32 * cap + 2 * 48 * cap + 16 * cap + (1 + V.VCM_C) * cap * 64 + 5 * 32 * slots + 64 * cap +
            12 * p)
    assert V.footprint(r) >= need
    assert V.first_rays(r) == p + 2 * p


def test_poisson_requests():
    reqs = V.poisson_requests(4000, 200.0, 16, 8, [(0.5, 0.4, 0.5, 3.0)] * 8, seed=3)
    t = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(t) > 0)
    assert abs(len(reqs) / t[-1] - 200.0) < 15.0          # mean rate
    assert {r.session for r in reqs} == set(range(16)) and {r.scene for r in reqs} == set(range(8))
    assert {(r.width, r.height) for r in reqs} == {(256, 256), (640, 480), (1280, 720), (1920, 1080)}
    assert {r.spp for r in reqs} == {1, 2, 3, 4}
