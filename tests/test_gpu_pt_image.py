"""A path-traced frame on the GPU against a numpy restatement of the reference's path tracer
with the same per-pixel seeds.

The camera rays are restated bit for bit (PixelRayGen seeding, render/ray_gen/ray_gen.h:51-72;
the pinhole of render/ray_gen/camera.h:57-79 through the raster-to-world matrix; division and
square root are correctly rounded on both sides).  Every bounce of
PathTracer::process_primary_rays (render/integrators/pt.cpp:82-160) is restated over the GPU's
own queue and the oracle's closest-hit traversal, for Lambertian surfaces and a triangle light:
  - emitter hits: AreaEmitter::radiance (render/light.h:19-39) with the MIS weight of
    pt.cpp:120-136;
  - next-event estimation (compute_direct_illum, pt.cpp:17-43; TriangleLight::sample_direct,
    light.h:143-170; the Lambertian BSDF of render/materials/bsdf.h:125-219): the shadow ray of
    every slot (origin, direction, tmin, tmax) and its contribution bit for bit, and the
    GPU's any-hit results on those rays bit for bit against the oracle's;
  - the bounce (pt.cpp:46-80): path-length cut, Russian roulette drawn on every bounce and
    applied from bounce rr_start on (0 = the reference), BSDF::sample (component draw,
    cosine-weighted direction, bsdf.h:147-199, random.h:80-83): which paths continue, their
    random-number states, bounce counts and last pdfs and the next queue's order and origins
    bit for bit; directions and throughputs within 1e-5 (sin / cos are the device's).
The image summed from the restated emitter and unoccluded shadow contributions is within 1e-4
mean relative error of the GPU's (the GPU adds a pixel's samples with float atomics in no fixed
order).  The two draws of sample_cos_hemisphere and sample_uniform_triangle are taken left to
right (C++ leaves the order of a call's arguments open).
"""
import numpy as np
import pytest

import oracle
from imbatracer_amd import native as N
from imbatracer_amd import workloads as W
from tests.test_gpu_photon_bounces import FLT_MAX, Rng

pytestmark = pytest.mark.gpu

F = np.float32
PI = F(3.14159265359)


def _seeded(seed_base, i, pixel):
    seed = seed_base & 0xFFFFFFFF
    for _ in range(4):
        seed = ((33 * seed) & 0xFFFFFFFF) ^ i
    sseed = seed - (1 << 32) if seed >= 1 << 31 else seed
    rng = Rng(np.array([sseed & 0xFFFFFFFFFFFFFFFF], np.uint64))
    for _ in range(int(np.fmod(sseed, 5)) + 16 + pixel % 5):
        rng.draw(np.array([True]))
    return rng


def _camera(pt):
    n = pt.n
    rays = np.zeros(n, N.RAY_DTYPE)
    rng_s = np.zeros(n, np.uint64)
    m = pt.r2w
    for i in range(n):
        pixel, y, x = i // pt.spp, (i // pt.spp) // pt.w, (i // pt.spp) % pt.w
        rng = _seeded(pt.seed_base, i, pixel)
        one = np.array([True])
        sx = F(x) + rng.draw(one)[0]
        sy = F(y) + rng.draw(one)[0]
        t = [((m[4 * r] * sy + m[4 * r + 1] * sx) + m[4 * r + 2] * F(0)) + m[4 * r + 3] for r in range(4)]
        d = np.array([t[0] / t[3] - pt.cam_pos[0], t[1] / t[3] - pt.cam_pos[1], t[2] / t[3] - pt.cam_pos[2]], F)
        inv = F(1) / np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
        rays["org"][i, :3] = pt.cam_pos
        rays["dir"][i, :3] = d * inv
        rays["dir"][i, 3] = FLT_MAX
        rng_s[i] = rng.s[0]
    return rays, rng_s


def _dot(a, b):
    return (a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1]) + a[:, 2] * b[:, 2]


def _normalize(a):
    return a * (F(1) / np.sqrt(_dot(a, a)))[:, None]


def _cross(a, b):
    return np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1], a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                     a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], 1)


def _local_frame(n):
    """local_coordinates (render/random.h:57-72): (u_tangent, v_tangent), both normalized."""
    xbig = np.abs(n[:, 0]) > np.abs(n[:, 1])
    n0 = np.where(xbig, n[:, 0], n[:, 1])
    sig = np.where(xbig, F(-1), F(1))
    inv_len = F(1) / np.sqrt(n0 * n0 + n[:, 2] * n[:, 2])
    t = np.zeros_like(n)
    t[:, 0] = np.where(xbig, n[:, 2] * sig * inv_len, F(0))
    t[:, 1] = np.where(xbig, F(0), n[:, 2] * sig * inv_len)
    t[:, 2] = n0 * F(-1) * sig * inv_len
    b = _cross(n, t)
    return _normalize(t), _normalize(b)


def _raw(rng, mask):
    """One MWC64X draw as the raw 32-bit value (random_int), for the entries of mask."""
    s = rng.s[mask]
    c = s >> np.uint64(32)
    x = s & np.uint64(0xFFFFFFFF)
    with np.errstate(over="ignore"):
        rng.s[mask] = x * np.uint64(4294883355) + c
    out = np.zeros(len(rng.s), np.uint64)
    out[mask] = (x ^ c) & np.uint64(0xFFFFFFFF)
    return out


def _shade(pt, rays, states, hits):
    """One shading step of every entry; returns the restated outcome as a dict."""
    s = pt.s
    sh_lights = np.asarray(s.lights, F).reshape(-1, 12)
    lc = len(sh_lights)
    tri = hits["tri_id"]
    hit = (tri >= 0) & (tri < s.tri_count)
    t_ = np.maximum(tri, 0)
    mat = np.where(hit, s.tri_materials[t_], -1)
    ok = hit & (mat >= 0) & (mat < len(s.albedo))
    mat = np.maximum(mat, 0)
    o, d = rays["org"][:, :3], rays["dir"][:, :3]
    t = hits["tmax"]
    pos = o + t[:, None] * d
    out = -_normalize(d)
    n = _normalize(s.tri_normals[t_])
    tp = np.ascontiguousarray(states["tp"])
    em = s.emission[mat]
    emit = ok & (em != 0).any(1)
    # emitter hits
    cos_o = _dot(n, out)
    front = cos_o > F(0)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        pdf_direct_a = np.where(front, F(1) / s.emitter_area[mat], F(1))
        li = np.where(front[:, None], em, F(0))
        pdf_di = pdf_direct_a / F(max(lc, 1))
        pdf_di = pdf_di * ((t * t) / _dot(n, out))
        first = (states["bounces"] == 0) | (states["last_specular"] != 0)
        mis_e = np.where(first, F(1), states["last_pdf"] / (states["last_pdf"] + pdf_di))
        emit_contrib = (tp * li) * mis_e[:, None]
    surf = ok & ~emit
    rng = Rng(states["rng"])
    ut, vt = _local_frame(n)
    alb = s.albedo[mat] * (F(1) / PI)
    # next-event estimation
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        pick = (_raw(rng, surf) % np.uint64(max(lc, 1))).astype(np.int64)
        r1, r2 = rng.draw(surf), rng.draw(surf)
        L = sh_lights[pick]
        p0, p1, p2, inten = L[:, 0:3], L[:, 3:6], L[:, 6:9], L[:, 9:12]
        ln = _cross(p1 - p0, p2 - p0)
        area = np.sqrt(_dot(ln, ln)) * F(0.5)
        lnn = ln * (F(1) / np.sqrt(_dot(ln, ln)))[:, None]
        sr1 = np.sqrt(r1)
        lu = F(1) - sr1
        lv = r2 * sr1
        lw = (F(1) - lu) - lv
        lp = (lu[:, None] * p0 + lv[:, None] * p1) + lw[:, None] * p2
        dv = lp - pos
        distsq = _dot(dv, dv)
        dist = np.sqrt(distsq)
        dv = dv * (F(1) / dist)[:, None]
        cos_out = _dot(lnn, F(-1) * dv)
        lit = (cos_out > F(0)) & (cos_out < F(1))
        k = area / distsq
        rad = np.where(lit[:, None], (inten * cos_out[:, None]) * k[:, None], F(0))
        pdf_dw = np.where(lit, ((F(1) / area) * distsq) / cos_out, F(1))
        side = (_dot(dv, n) * _dot(out, n)) > F(0)
        lo, li_z = _dot(n, out), _dot(n, dv)
        refl = side & (lo * li_z > F(0))
        bsdf = np.where(refl[:, None], alb, F(0))
        pdf_hit = np.where(refl, np.abs(li_z) / PI, F(0))
        pick_pdf = F(1) / F(max(lc, 1))
        pdf_di2 = pick_pdf * pdf_dw
        mis = pdf_di2 / (pdf_di2 + pdf_hit)
        shadow = surf & (pdf_hit != F(0)) & (pdf_di2 != F(0))
        c = np.abs(_dot(n, dv))
        stp = ((((tp * bsdf) * c[:, None]) * rad) * mis[:, None]) / pick_pdf
        off = F(1e-3) * dist
    # bounce
    alive = surf & (states["bounces"] + 1 < pt.max_path_len)
    u = rng.draw(alive)
    rr = alive & (states["bounces"] >= pt.rr_start)
    lum = (tp[:, 0] * F(0.2126) + tp[:, 1] * F(0.7152)) + tp[:, 2] * F(0.0722)
    rr_pdf = np.where(rr, np.minimum(F(1), lum * F(2)), F(1))
    alive &= ~(rr & ~(u < rr_pdf))
    lo_z = _dot(n, out)
    rng.draw(alive)  # BSDF component choice
    u1, u2 = rng.draw(alive), rng.draw(alive)
    sin_t, cos_t = np.sqrt(F(1) - u2), np.sqrt(u2)
    phi = F(2) * PI * u1
    lx, ly, lz = sin_t * np.cos(phi), sin_t * np.sin(phi), cos_t
    pdf = lz / PI
    lz = np.where(lo_z < F(0), -lz, lz)
    same = lo_z * lz > F(0)
    w = (vt * lx[:, None] + ut * ly[:, None]) + n * lz[:, None]
    black = (alb <= F(0)).all(1)
    alive &= ~((pdf == F(0)) | (_dot(w, n) * _dot(out, n) <= F(0)) | ~same | black)
    with np.errstate(divide="ignore", invalid="ignore"):
        cos_term = np.abs(_dot(n, w))
        q = pdf * rr_pdf
        tp_new = tp * ((alb * cos_term[:, None]) / q[:, None])
    return {"emit": emit, "emit_contrib": emit_contrib, "shadow": shadow, "pos": pos, "dir": dv, "off": off,
            "dist": dist, "stp": stp, "alive": alive, "w": w, "tp": tp_new, "pdf": pdf, "rng": rng.s, "t": t}


@pytest.fixture(scope="module", params=[0, 3], ids=["rr_every_bounce", "rr_from_bounce_3"])
def frame(gpu, request):
    scene = W.SphereGridScene(side=4, level=2, builder=N.BUILDER_SBVH)
    pt = W.PathTracer(scene, 48, 32, spp=16, max_path_len=8, rr_start=request.param)
    N.check(N.lib.imba_memset(pt.image.ptr, 0, pt.w * pt.h * 12, None), "memset")
    counts, _, queues = pt.frame(keep_queues=8)
    pt.stream.sync()
    image = pt.image.download(np.float32, pt.w * pt.h * 3).reshape(-1, 3)
    return pt, counts, queues, list(pt.kept_states), list(pt.kept_shadow), image


def test_camera_rays(frame):
    pt, counts, queues, states, _, _ = frame
    rays, rng_s = _camera(pt)
    assert np.array_equal(queues[0].view(np.uint8), rays.view(np.uint8))
    assert np.array_equal(states[0]["rng"], rng_s)
    assert np.array_equal(states[0]["pixel"], np.arange(pt.n) // pt.spp)
    assert np.array_equal(states[0]["sample"], np.arange(pt.n) % pt.spp)
    assert np.all(states[0]["bounces"] == 0) and np.all(states[0]["tp"] == 1)


def test_bounces_shadow_rays_and_image(frame):
    pt, counts, queues, states, shadows, image = frame
    arrays = pt.s.build.arrays()
    want = np.zeros((pt.w * pt.h, 3), np.float64)
    assert len(queues) == len(counts) >= 3
    n_shadow = 0
    for b in range(len(queues)):
        rays, st = queues[b], states[b]
        hits = oracle.trace(arrays, rays, threads=8)[0]
        r = _shade(pt, rays, st, hits)
        # emitter contributions
        for i in np.nonzero(r["emit"])[0]:
            want[st["pixel"][i]] += r["emit_contrib"][i].astype(np.float64)
        # shadow slots, bit for bit
        srays, sst, shits = shadows[b]
        sh = r["shadow"]
        assert np.array_equal(sst["pixel"] >= 0, sh)
        assert np.array_equal(sst["pixel"][sh], st["pixel"][sh])
        assert np.array_equal(np.ascontiguousarray(sst["tp"][sh]).view(np.int32),
                              np.ascontiguousarray(r["stp"][sh]).view(np.int32))
        assert np.array_equal(srays["org"][sh, :3], r["pos"][sh])
        assert np.array_equal(srays["org"][sh, 3], r["off"][sh])
        assert np.array_equal(srays["dir"][sh, :3], r["dir"][sh])
        assert np.array_equal(srays["dir"][sh, 3], r["dist"][sh] - r["off"][sh])
        # the GPU's any-hit results on its shadow rays against the oracle's
        want_sh = oracle.trace(arrays, srays, any_hit=True, threads=8)[0]
        assert np.array_equal(shits[sh].view(np.int32), want_sh[sh].view(np.int32))
        assert np.all(shits["tri_id"][~sh] < 0)  # empty slots end at the root without a hit
        free = sh & (want_sh["tri_id"] < 0)
        for i in np.nonzero(free)[0]:
            want[st["pixel"][i]] += r["stp"][i].astype(np.float64)
        n_shadow += int(sh.sum())
        # the next queue
        idx = np.nonzero(r["alive"])[0]
        if b + 1 < len(queues):
            nr, ns = queues[b + 1], states[b + 1]
            assert len(nr) == len(idx)
            assert np.array_equal(ns["pixel"], st["pixel"][idx])
            assert np.array_equal(ns["sample"], st["sample"][idx])
            assert np.array_equal(ns["bounces"], st["bounces"][idx] + 1)
            assert np.array_equal(ns["rng"], r["rng"][idx])
            assert np.array_equal(ns["last_pdf"], r["pdf"][idx])
            assert np.all(ns["last_specular"] == 0)
            assert np.allclose(ns["tp"], r["tp"][idx], rtol=1e-5, atol=0)
            assert np.array_equal(nr["org"][:, :3], r["pos"][idx])
            assert np.array_equal(nr["org"][:, 3], F(1e-3) * r["t"][idx])
            assert np.all(nr["dir"][:, 3] == FLT_MAX)
            assert np.allclose(nr["dir"][:, :3], r["w"][idx], atol=1e-5, rtol=0)
        else:
            assert b + 1 == pt.max_path_len or len(idx) == 0
    assert n_shadow > pt.n  # shadow rays over the bounces
    assert pt.shadow_counts[:len(shadows)] == [int((x[1]["pixel"] >= 0).sum()) for x in shadows]
    lit = want.sum(1) > 0
    assert lit.sum() >= 0.5 * pt.w * pt.h  # next-event estimation lights most pixels
    assert np.array_equal(image.sum(1) > 0, lit)
    rel = np.abs(image[lit] - want[lit]) / np.maximum(np.abs(want[lit]), 1e-30)
    assert float(rel.mean()) <= 1e-4
