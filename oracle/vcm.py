"""Numpy restatement of one VCM iteration's surface steps — TEST INFRASTRUCTURE ONLY.

VCMIntegrator<algo> (render/integrators/vcm.cpp) for the materials of the measured scenes
(Lambertian, mirror, glass of index 1.5) and triangle lights, vectorized over a queue in
float32 with the reference's operation order, so that everything but sin / cos rounds as the
device does.  Each function takes one queue (rays, hits, states as the GPU read them back) and
returns what the step writes:

  light_start    trace_light_paths' sampler (vcm.cpp:64-96; TriangleLight::sample_emit,
                 render/light.h:109-141; LightRayGen seeding, render/ray_gen/ray_gen.h:138-145)
  light_bounce   process_light_rays (vcm.cpp:196-256): partial weights at the hit, the light
                 vertex, connect_to_camera (:258-318), bounce (:129-193, adjoint)
  camera_start   trace_camera_paths' sampler (vcm.cpp:98-127; PixelRayGen seeding, ray_gen.h:51-72)
  camera_bounce  process_camera_rays (vcm.cpp:320-443): partial weights, emitter hits with their
                 MIS weight, direct_illum (:445-488), connect (:490-570), the merge query
                 (vertex_merging :572-611 is restated in merge()), bounce (non-adjoint)
  merge          vertex_merging over given k-nearest lists

Params.algo picks the sub-algorithm as the reference's template parameter does (vcm.h:37-43):
"vcm"; "bpt" (mis_eta_vm 0, vcm.cpp:45; no merge queries, :427); "lt" (no stored vertices,
:238; camera connections with weight 1, :301); "pt" (next event with weight 1, :473; no
connections, :424, or merge queries).

Only tests import this module.
"""
from __future__ import annotations

import numpy as np

F = np.float32
PI = F(3.14159265359)  # core/common.h
FLT_MAX = F(3.4028234663852886e38)
RAY_DTYPE = np.dtype([("org", "<f4", (4,)), ("dir", "<f4", (4,))])
VCM_STATE_DTYPE = np.dtype([("pixel", "<i4"), ("sample", "<i4"), ("rng", "<u8"), ("tp", "<f4", (3,)),
                            ("path_length", "<i4"), ("dVC", "<f4"), ("dVCM", "<f4"), ("dVM", "<f4"), ("pad", "<i4")])
VTX_A_DTYPE = np.dtype([("pos", "<f4", (3,)), ("path_length", "<i4"), ("out", "<f4", (3,)), ("tri", "<i4")])
VTX_B_DTYPE = np.dtype([("key", "<i4"), ("tp", "<f4", (3,)), ("dVCM", "<f4"), ("dVM", "<f4"), ("dVC", "<f4"),
                        ("mat", "<i4")])
SHADOW_STATE_DTYPE = np.dtype([("pixel", "<i4"), ("tp", "<f4", (3,))])
QUERY_DTYPE = np.dtype([("pos", "<f4", (3,)), ("pixel", "<i4"), ("n", "<f4", (3,)), ("dVCM", "<f4"),
                        ("out", "<f4", (3,)), ("dVM", "<f4"), ("tp", "<f4", (3,)), ("mat", "<i4")])


class Params:
    """Scene, camera and integrator settings of one iteration (the fields of ImbaVcmFrame)."""

    def __init__(self, scene, r2w, w2r, cam_pos, cam_dir, ipd, width, height, spp, light_paths, radius,
                 max_path_len, light_max_depth, rr_start, num_connections, algo: str = "vcm"):
        self.tri_normals = np.asarray(scene.tri_normals, F)
        self.tri_materials = np.asarray(scene.tri_materials, np.int32)
        self.albedo = np.asarray(scene.albedo, F)
        self.emission = np.asarray(scene.emission, F)
        self.emitter_area = np.asarray(scene.emitter_area, F)
        self.kinds = np.asarray(scene.kinds, np.int32)
        self.lights = np.asarray(scene.lights, F).reshape(-1, 12)
        self.r2w, self.w2r = np.asarray(r2w, F), np.asarray(w2r, F)
        self.cam_pos, self.cam_dir, self.ipd = np.asarray(cam_pos, F), np.asarray(cam_dir, F), F(ipd)
        self.width, self.height, self.spp = width, height, spp
        self.P = light_paths
        r = F(radius)
        eta_vcm = PI * (r * r) * F(light_paths)  # render(): pi * sqr(pm_radius) * light_path_count
        self.algo = algo
        self.eta_vc = F(1) / eta_vcm
        self.eta_vm = F(0) if algo == "bpt" else eta_vcm
        self.radius, self.radius_sqr = r, r * r
        self.max_path_len, self.light_max_depth, self.rr_start, self.C = (max_path_len, light_max_depth, rr_start,
                                                                           num_connections)


class Rng:
    """MWC64X (render/random.h:15-47), one generator per entry."""

    def __init__(self, s):
        self.s = np.asarray(s, np.uint64).copy()

    def raw(self, mask):
        s = self.s[mask]
        c = s >> np.uint64(32)
        x = s & np.uint64(0xFFFFFFFF)
        with np.errstate(over="ignore"):
            self.s[mask] = x * np.uint64(4294883355) + c
        out = np.zeros(len(self.s), np.uint64)
        out[mask] = (x ^ c) & np.uint64(0xFFFFFFFF)
        return out

    def uniform(self, mask):
        """random_float(): next() / 2^32."""
        return (self.raw(mask).astype(np.uint32).astype(F) / F(4294967296.0)).astype(F)

    def integer(self, mask, lo, hi):
        """random_int(lo, hi) (random.h:31-33)."""
        if hi == lo:
            return np.full(len(self.s), lo, np.int64)
        return (self.raw(mask) % np.uint64(hi - lo)).astype(np.int64) + lo


def seeded(seed_base: int, index: np.ndarray, pixel: np.ndarray) -> Rng:
    """PixelRayGen / LightRayGen seeding (ray_gen.h:63-70, 138-145): four rounds of
    seed = 33 * seed ^ index in int arithmetic, RNG(seed), then (seed % 5) + 16 + pixel % 5
    discarded draws."""
    index = np.asarray(index, np.int64)
    seed = np.full(len(index), seed_base & 0xFFFFFFFF, np.uint64)
    for _ in range(4):
        seed = ((np.uint64(33) * seed) & np.uint64(0xFFFFFFFF)) ^ (index.astype(np.uint64) & np.uint64(0xFFFFFFFF))
    sseed = seed.astype(np.int64)
    sseed = np.where(sseed >= 1 << 31, sseed - (1 << 32), sseed)
    rng = Rng(sseed.astype(np.uint64))  # (uint64_t)(int64_t) seed: sign-extended
    n = np.fmod(sseed, 5) + 16 + np.asarray(pixel, np.int64) % 5
    for k in range(int(n.max()) if len(n) else 0):
        rng.raw(n > k)
    return rng


def dot(a, b):
    return (a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1]) + a[:, 2] * b[:, 2]


def scale(a, s):
    return a * np.asarray(s, F)[:, None]


def normalize(a):
    return scale(a, F(1) / np.sqrt(dot(a, a)))


def cross(a, b):
    return np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1], a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                     a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], 1)


def local_frame(n):
    """local_coordinates (render/random.h:57-72): (tangent, binormal), both normalized."""
    xbig = np.abs(n[:, 0]) > np.abs(n[:, 1])
    n0 = np.where(xbig, n[:, 0], n[:, 1])
    sig = np.where(xbig, F(-1), F(1))
    inv_len = F(1) / np.sqrt(n0 * n0 + n[:, 2] * n[:, 2])
    t = np.zeros_like(n)
    t[:, 0] = np.where(xbig, n[:, 2] * sig * inv_len, F(0))
    t[:, 1] = np.where(xbig, F(0), n[:, 2] * sig * inv_len)
    t[:, 2] = n0 * F(-1) * sig * inv_len
    b = cross(n, t)
    return normalize(t), normalize(b)


def lambert_ok(n, a, b):
    """The Lambertian BSDF's value / pdf for the pair (a, b) is non-zero: both directions on one
    side of the surface (the BRDF is chosen, bsdf.h:125-145) and in one shading hemisphere."""
    return (dot(b, n) * dot(a, n) > F(0)) & (dot(n, a) * dot(n, b) > F(0))


def _tri_light(L):
    p0, p1, p2 = L[:, 0:3], L[:, 3:6], L[:, 6:9]
    ln = cross(p1 - p0, p2 - p0)
    area = np.sqrt(dot(ln, ln)) * F(0.5)
    return p0, p1, p2, ln, area


def _uniform_triangle_point(p0, p1, p2, r1, r2):
    sr1 = np.sqrt(r1)
    u = F(1) - sr1
    v = r2 * sr1
    w = (F(1) - u) - v
    return (scale(p0, u) + scale(p1, v)) + scale(p2, w)


def _cos_hemisphere(u1, u2):
    sin_t, cos_t = np.sqrt(F(1) - u2), np.sqrt(u2)
    phi = F(2) * PI * u1
    return sin_t * np.cos(phi).astype(F), sin_t * np.sin(phi).astype(F), cos_t


def light_start(p: Params, first: int, count: int, seed_base: int):
    i = np.arange(first, first + count)
    lc = len(p.lights)
    per = p.P // lc
    firstn = per + p.P % lc
    light = np.where(i < firstn, 0, np.minimum(lc - 1, 1 + (i - firstn) // max(per, 1)))
    L = p.lights[light]
    rng = seeded(seed_base, i, i)
    p0, p1, p2, ln, area = _tri_light(L)
    n = normalize(ln)
    tg, bn = local_frame(n)
    all_ = np.ones(count, bool)
    r1, r2 = rng.uniform(all_), rng.uniform(all_)
    pos = _uniform_triangle_point(p0, p1, p2, r1, r2)
    u1, u2 = rng.uniform(all_), rng.uniform(all_)
    lx, ly, lz = _cos_hemisphere(u1, u2)
    dpdf = lz / PI
    d = (scale(bn, lx) + scale(tg, ly)) + scale(n, lz)
    zero = dpdf <= F(0)
    with np.errstate(divide="ignore", invalid="ignore"):
        rad = (L[:, 9:12] * area[:, None]) * PI
        rad[zero] = F(0)
        cos_out = np.where(zero, F(0), lz)
        pdf_emit_w = np.where(zero, F(1), dpdf / area)
        pdf_direct_a = np.where(zero, F(1), F(1) / area)
    pick = F(1) / F(lc)
    rays = np.zeros(count, RAY_DTYPE)
    rays["org"][:, :3] = pos
    rays["org"][:, 3] = F(1e-3)
    rays["dir"][:, :3] = d
    rays["dir"][:, 3] = FLT_MAX
    st = np.zeros(count, VCM_STATE_DTYPE)
    st["pixel"] = i
    st["rng"] = rng.s
    st["tp"] = rad / pick
    st["path_length"] = 1
    st["dVCM"] = pdf_direct_a / pdf_emit_w
    st["dVC"] = cos_out / (pdf_emit_w * pick)
    st["dVM"] = st["dVC"] * p.eta_vc
    return rays, st


def camera_start(p: Params, first: int, count: int, seed_base: int):
    i = np.arange(first, first + count)
    pixel = i // p.spp
    y, x = pixel // p.width, pixel % p.width
    rng = seeded(seed_base, i, pixel)
    all_ = np.ones(count, bool)
    sx = x.astype(F) + rng.uniform(all_)
    sy = y.astype(F) + rng.uniform(all_)
    m = p.r2w
    t = [((m[4 * r] * sy + m[4 * r + 1] * sx) + m[4 * r + 2] * F(0)) + m[4 * r + 3] for r in range(4)]
    d = np.stack([t[0] / t[3] - p.cam_pos[0], t[1] / t[3] - p.cam_pos[1], t[2] / t[3] - p.cam_pos[2]], 1)
    d = normalize(d)
    rays = np.zeros(count, RAY_DTYPE)
    rays["org"][:, :3] = p.cam_pos
    rays["dir"][:, :3] = d
    rays["dir"][:, 3] = FLT_MAX
    cos_o = dot(d, np.broadcast_to(p.cam_dir, d.shape))
    r = p.ipd / cos_o
    pdf_cam_w = r * r / cos_o
    st = np.zeros(count, VCM_STATE_DTYPE)
    st["pixel"] = pixel
    st["sample"] = i % p.spp
    st["rng"] = rng.s
    st["tp"] = F(1)
    st["path_length"] = 1
    st["dVCM"] = F(p.P) / pdf_cam_w
    return rays, st


def _surface(p: Params, rays, hits):
    tri = hits["tri_id"]
    ok = (tri >= 0) & (tri < len(p.tri_materials))
    t_ = np.maximum(tri, 0)
    mat = np.where(ok, p.tri_materials[t_], -1)
    ok &= (mat >= 0) & (mat < len(p.albedo))
    mat = np.maximum(mat, 0)
    o, d = rays["org"][:, :3], rays["dir"][:, :3]
    t = hits["tmax"]
    pos = o + scale(d, t)
    out = -normalize(d)
    n = normalize(p.tri_normals[t_])
    return ok, mat, pos, out, n, d, t, t_


def _specular(kind, d, nf, entering, rng, mask):
    out = np.zeros_like(d)
    cos_i = -dot(d, nf)
    refl = d + scale(nf, F(2) * cos_i)
    out[mask & (kind == 1)] = refl[mask & (kind == 1)]
    gl = mask & (kind == 2)
    eta = np.where(entering, F(1) / F(1.5), F(1.5)).astype(F)
    with np.errstate(invalid="ignore", divide="ignore"):
        sin2 = (eta * eta) * (F(1) - cos_i * cos_i)
        tir = gl & (sin2 >= F(1))
        fr = gl & ~(sin2 >= F(1))
        cos_t = np.sqrt(F(1) - sin2)
        rs = (eta * cos_i - cos_t) / (eta * cos_i + cos_t)
        rp = (cos_i - eta * cos_t) / (cos_i + eta * cos_t)
        f = F(0.5) * (rs * rs + rp * rp)
    u = rng.uniform(fr)
    take = tir | (fr & (u < f))
    out[take] = refl[take]
    refr = fr & ~(u < f)
    out[refr] = (scale(d, eta) + scale(nf, eta * cos_i - cos_t))[refr]
    return out


def _bounce(p: Params, st, rng, mask, mat, kind, d, out, n, adjoint):
    """VCMIntegrator::bounce for the entries of mask; returns (alive, direction)."""
    tp = st["tp"]
    u = rng.uniform(mask)
    lum = (tp[:, 0] * F(0.2126) + tp[:, 1] * F(0.7152)) + tp[:, 2] * F(0.0722)
    pr = np.minimum(F(1), lum * F(2))
    rr_apply = mask & (st["path_length"] - 1 >= p.rr_start)
    rr_pdf = np.where(rr_apply, pr, F(1)).astype(F)
    alive = mask & ~(rr_apply & ~(u < rr_pdf))
    new_in = np.zeros_like(d)
    alb = p.albedo[mat]
    # Lambertian
    dif = alive & (kind == 0)
    ut, vt = local_frame(n)
    lo_z = dot(n, out)
    rng.uniform(dif)
    u1, u2 = rng.uniform(dif), rng.uniform(dif)
    lx, ly, lz = _cos_hemisphere(u1, u2)
    pdf = lz / PI
    lz = np.where(lo_z < F(0), -lz, lz)
    din = (scale(vt, lx) + scale(ut, ly)) + scale(n, lz)
    v = alb * (F(1) / PI)
    black = (v <= F(0)).all(1)
    fail = dif & ((pdf == F(0)) | ~lambert_ok(n, out, din) | black)
    dif_ok = dif & ~fail
    alive = alive & ~fail
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        pdf_rev = np.abs(dot(n, out)) / PI
        cos_i = np.abs(dot(out, n) * dot(din, n) / dot(out, n)) if adjoint else np.abs(dot(din, n))
        c = cos_i / pdf
        dvc = c * ((st["dVC"] * pdf_rev + st["dVCM"]) + p.eta_vm)
        dvm = c * ((st["dVM"] * pdf_rev + st["dVCM"] * p.eta_vc) + F(1))
        dvcm = F(1) / pdf
        q = rr_pdf * pdf
        tp_d = tp * ((v * cos_i[:, None]) / q[:, None])
    st["dVC"] = np.where(dif_ok, dvc, st["dVC"])
    st["dVM"] = np.where(dif_ok, dvm, st["dVM"])
    st["dVCM"] = np.where(dif_ok, dvcm, st["dVCM"])
    st["tp"] = np.where(dif_ok[:, None], tp_d, st["tp"])
    new_in[dif_ok] = din[dif_ok]
    # mirror / glass
    spec = alive & (kind != 0)
    cos_dn = dot(d, n)
    entering = cos_dn < F(0)
    nf = np.where(entering[:, None], n, -n).astype(F)
    sd = _specular(kind, d, nf, entering, rng, spec)
    with np.errstate(divide="ignore", invalid="ignore"):
        cos_s = np.abs(dot(out, n) * dot(sd, n) / dot(out, n)) if adjoint else np.abs(dot(sd, n))
        tp_m = (st["tp"] * alb) / rr_pdf[:, None]
        tp_g = st["tp"] / rr_pdf[:, None]
    st["dVCM"] = np.where(spec, F(0), st["dVCM"])
    st["dVC"] = np.where(spec, st["dVC"] * cos_s, st["dVC"])
    st["dVM"] = np.where(spec, st["dVM"] * cos_s, st["dVM"])
    st["tp"] = np.where((spec & (kind == 1))[:, None], tp_m, st["tp"])
    st["tp"] = np.where((spec & (kind == 2))[:, None], tp_g, st["tp"])
    new_in[spec] = sd[spec]
    st["path_length"] = np.where(alive, st["path_length"] + 1, st["path_length"])
    return alive, new_in


def _empty_shadow(n):
    rays = np.zeros(n, RAY_DTYPE)
    rays["org"][:, 3] = F(1)
    rays["dir"][:, 2] = F(1)
    st = np.zeros(n, SHADOW_STATE_DTYPE)
    st["pixel"] = -1
    return rays, st


def light_bounce(p: Params, rays, hits, states):
    """process_light_rays for one queue: dict of states, rays_out (rows of continuing paths),
    vertex slots vtx_a / vtx_b, and the camera-connection shadow slots."""
    n = len(rays)
    st = states.copy()
    ok, mat, pos, out, nrm, d, t, tri = _surface(p, rays, hits)
    cos_o = np.abs(dot(out, nrm))
    live = ok & (cos_o != F(0))
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        dvcm = st["dVCM"] * (t * t)
        ic = F(1) / cos_o
        st["dVCM"] = np.where(live, dvcm * ic, st["dVCM"])
        st["dVC"] = np.where(live, st["dVC"] * ic, st["dVC"])
        st["dVM"] = np.where(live, st["dVM"] * ic, st["dVM"])
    kind = p.kinds[mat]
    dif = live & (kind == 0)
    va = np.zeros(n, VTX_A_DTYPE)
    vb = np.zeros(n, VTX_B_DTYPE)
    vb["key"] = -1
    sv = dif & (p.algo != "lt")  # light tracing stores no vertices
    va["pos"][sv] = pos[sv]
    va["path_length"][sv] = st["path_length"][sv] + 1
    va["out"][sv] = out[sv]
    va["tri"][sv] = tri[sv]
    vb["key"][sv] = st["pixel"][sv]
    vb["tp"][sv] = st["tp"][sv]
    vb["dVCM"][sv], vb["dVM"][sv], vb["dVC"][sv] = st["dVCM"][sv], st["dVM"][sv], st["dVC"][sv]
    vb["mat"][sv] = mat[sv]
    # connect_to_camera
    sh_rays, sh_st = _empty_shadow(n)
    cp = np.broadcast_to(p.cam_pos, pos.shape)
    cd = np.broadcast_to(p.cam_dir, pos.shape)
    dtc = cp - pos
    front = dot(-dtc, cd) >= F(0)
    m = p.w2r
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    r = [((m[4 * k] * x + m[4 * k + 1] * y) + m[4 * k + 2] * z) + m[4 * k + 3] * F(1) for k in range(4)]
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        rx, ry = r[1] / r[3], r[0] / r[3]
        px, py = np.floor(rx), np.floor(ry)
        inside = (px >= 0) & (px < p.width) & (py >= 0) & (py < p.height)
        pxi = np.where(inside, px, 0).astype(np.int64)
        pyi = np.where(inside, py, 0).astype(np.int64)
        d2 = dot(dtc, dtc)
        dist = np.sqrt(d2)
        dtc = dtc / dist[:, None]
        cos_cam = np.abs(dot(cd, -dtc))
        cos_surf = np.abs(dot(out, nrm) * dot(dtc, nrm) / dot(out, nrm))
        conn = dif & front & inside & lambert_ok(nrm, out, dtc)
        pdf_rev = np.abs(dot(nrm, out)) / PI
        img_to_surf = ((p.ipd * p.ipd) * cos_surf) / ((d2 * cos_cam) * (cos_cam * cos_cam))
        wl = (img_to_surf / F(p.P)) * ((p.eta_vm + st["dVCM"]) + st["dVC"] * pdf_rev)
        mis = F(1) / (wl + F(1)) if p.algo != "lt" else np.ones(n, F)
        alb = p.albedo[mat] * (F(1) / PI)
        tps = st["tp"] * (((mis[:, None] * alb) * img_to_surf[:, None]) / F(p.P))
        off = dist * F(1e-3)
    sh_st["pixel"][conn] = (pyi * p.width + pxi)[conn]
    sh_st["tp"][conn] = tps[conn]
    sh_rays["org"][conn, :3] = pos[conn]
    sh_rays["org"][conn, 3] = off[conn]
    sh_rays["dir"][conn, :3] = dtc[conn]
    sh_rays["dir"][conn, 3] = (dist - off)[conn]
    # bounce
    rng = Rng(st["rng"])
    go = live & (st["path_length"] < p.light_max_depth)
    alive, din = _bounce(p, st, rng, go, mat, kind, d, out, nrm, adjoint=True)
    st["rng"] = rng.s
    st["pixel"] = np.where(alive, st["pixel"], -1)
    out_rays = rays.copy()
    out_rays["org"][alive, :3] = pos[alive]
    out_rays["org"][alive, 3] = (t * F(1e-4))[alive]
    out_rays["dir"][alive, :3] = din[alive]
    out_rays["dir"][alive, 3] = FLT_MAX
    return {"states": st, "rays": out_rays, "alive": alive, "vtx_a": va, "vtx_b": vb, "shadow_rays": sh_rays,
            "shadow_states": sh_st}


def camera_bounce(p: Params, rays, hits, states, cache_a, cache_b):
    """process_camera_rays for one queue: dict of states, rays_out, shadow slots ((1 + C) per
    entry), merge queries and the emitter contributions (pixel, rgb) added this step."""
    n = len(rays)
    C = p.C
    st = states.copy()
    ok, mat, pos, out, nrm, d, t, tri = _surface(p, rays, hits)
    cos_o = np.abs(dot(out, nrm))
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        st["dVCM"] = np.where(ok, st["dVCM"] * ((t * t) / cos_o), st["dVCM"])
        st["dVC"] = np.where(ok, st["dVC"] * (F(1) / cos_o), st["dVC"])
        st["dVM"] = np.where(ok, st["dVM"] * (F(1) / cos_o), st["dVM"])
    live = ok & (cos_o != F(0))
    em = p.emission[mat]
    emit = live & (em != F(0)).any(1)
    lc = len(p.lights)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        c = dot(nrm, out)
        front = c > F(0)
        ia = F(1) / p.emitter_area[mat]
        pdf_direct_a = np.where(front, ia, F(1))
        pdf_emit_w = np.where(front, ia * (np.abs(c) / PI), F(1))
        li = np.where(front[:, None], em, F(0))
        pick = F(1) / F(lc)
        wc = (pdf_direct_a * pick) * st["dVCM"] + (pdf_emit_w * pick) * st["dVC"]
        mis = np.where(st["path_length"] == 1, F(1), F(1) / (wc + F(1)))
        emit_rgb = (st["tp"] * li) * mis[:, None]
    surf = live & ~emit & (st["path_length"] < p.max_path_len)
    kind = p.kinds[mat]
    alb = p.albedo[mat] * (F(1) / PI)
    rng = Rng(st["rng"])
    sh_rays, sh_st = _empty_shadow((1 + C) * n)
    rows = np.arange(n) * (1 + C)
    # direct_illum
    l = rng.integer(surf, 0, lc)
    L = p.lights[np.minimum(l, lc - 1)]
    p0, p1, p2, ln, area = _tri_light(L)
    lnn = normalize(ln)
    r1, r2 = rng.uniform(surf), rng.uniform(surf)
    lp = _uniform_triangle_point(p0, p1, p2, r1, r2)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        ldir = lp - pos
        distsq = dot(ldir, ldir)
        dist = np.sqrt(distsq)
        ldir = scale(ldir, F(1) / dist)
        lcos = dot(lnn, F(-1) * ldir)
        good = (lcos > F(0)) & (lcos < F(1))
        k_ = area / distsq
        rad = np.where(good[:, None], (L[:, 9:12] * lcos[:, None]) * k_[:, None], F(0))
        cos_out = np.where(good, lcos, F(1))
        pdf_emit_l = np.where(good, (np.abs(lcos) / PI) / area, F(1))
        pdf_direct_w = np.where(good, ((F(1) / area) * distsq) / lcos, F(1))
        cos_i = np.abs(dot(nrm, ldir))
        pdf_dir = np.abs(dot(nrm, ldir)) / PI
        pdf_rev = np.abs(dot(nrm, out)) / PI
        nee = surf & (kind == 0) & lambert_ok(nrm, out, ldir) & (pdf_dir != F(0)) & (pdf_rev != F(0))
        pick_inv = F(lc)
        wl = (pdf_dir * pick_inv) / pdf_direct_w
        wcn = ((pdf_emit_l * cos_i) / (pdf_direct_w * cos_out)) * ((p.eta_vm + st["dVCM"]) + st["dVC"] * pdf_rev)
        misn = F(1) / ((wcn + F(1)) + wl) if p.algo != "pt" else np.ones(n, F)
        tpn = (((st["tp"] * misn[:, None]) * alb) * cos_i[:, None]) * rad * pick_inv
        offn = F(1e-3) * np.where(dist == FLT_MAX, F(1), dist)
    sl = rows[nee]
    sh_st["pixel"][sl] = st["pixel"][nee]
    sh_st["tp"][sl] = tpn[nee]
    sh_rays["org"][sl, :3] = pos[nee]
    sh_rays["org"][sl, 3] = offn[nee]
    sh_rays["dir"][sl, :3] = ldir[nee]
    sh_rays["dir"][sl, 3] = (dist - offn)[nee]
    # connect
    dif = surf & (kind == 0)
    nv = len(cache_a)
    vc_weight = F(nv) / (F(p.P) * F(C))
    for cc in range(C):
        if nv <= 0 or p.algo not in ("vcm", "bpt"):
            break
        vi = np.minimum(rng.integer(dif, 0, nv), nv - 1)
        A, B = cache_a[vi], cache_b[vi]
        lpos, lout = A["pos"], A["out"]
        lnrm = normalize(p.tri_normals[A["tri"]])
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            cdir = lpos - pos
            d2 = dot(cdir, cdir)
            cdist = np.sqrt(d2)
            cdir = scale(cdir, F(1) / cdist)
            ncdir = -cdir
            keep = dif & (A["path_length"] + st["path_length"] <= p.max_path_len) & ~(cdist < p.radius)
            keep &= lambert_ok(nrm, out, cdir) & lambert_ok(lnrm, lout, ncdir)
            pdc, prc = np.abs(dot(nrm, cdir)) / PI, np.abs(dot(nrm, out)) / PI
            pdl, prl = np.abs(dot(lnrm, ncdir)) / PI, np.abs(dot(lnrm, lout)) / PI
            keep &= (pdc != F(0)) & (pdl != F(0)) & (prc != F(0)) & (prl != F(0))
            cos_cam = np.abs(dot(nrm, cdir))
            cos_light = np.abs(dot(lout, lnrm) * dot(ncdir, lnrm) / dot(lout, lnrm))
            geom = cos_cam * cos_light / d2
            keep &= geom > F(0)
            pdf_cam_a = pdc * cos_light / d2
            pdf_light_a = pdl * cos_cam / d2
            wlc = pdf_cam_a * ((p.eta_vm + B["dVCM"]) + B["dVC"] * prl)
            wcc = pdf_light_a * ((p.eta_vm + st["dVCM"]) + st["dVC"] * prc)
            misc = F(1) / ((wcc + F(1)) + wlc)
            lalb = p.albedo[B["mat"]] * (F(1) / PI)
            tpc = (((((st["tp"] * vc_weight) * misc[:, None]) * geom[:, None]) * alb) * lalb) * B["tp"]
            offc = F(1e-3) * cdist
        sl = rows[keep] + 1 + cc
        sh_st["pixel"][sl] = st["pixel"][keep]
        sh_st["tp"][sl] = tpc[keep]
        sh_rays["org"][sl, :3] = pos[keep]
        sh_rays["org"][sl, 3] = offc[keep]
        sh_rays["dir"][sl, :3] = cdir[keep]
        sh_rays["dir"][sl, 3] = (cdist - offc)[keep]
    # merge query
    q = np.zeros(n, QUERY_DTYPE)
    q["pos"] = np.float32(np.inf)
    q["pixel"] = -1
    if p.algo != "vcm":
        dif = dif & False
    q["pos"][dif] = pos[dif]
    q["pixel"][dif] = st["pixel"][dif]
    q["n"][dif] = nrm[dif]
    q["dVCM"][dif] = st["dVCM"][dif]
    q["out"][dif] = out[dif]
    q["dVM"][dif] = st["dVM"][dif]
    q["tp"][dif] = st["tp"][dif]
    q["mat"][dif] = mat[dif]
    # bounce
    alive, din = _bounce(p, st, rng, surf, mat, kind, d, out, nrm, adjoint=False)
    st["rng"] = rng.s
    st["pixel"] = np.where(alive, st["pixel"], -1)
    out_rays = rays.copy()
    out_rays["org"][alive, :3] = pos[alive]
    out_rays["org"][alive, 3] = (t * F(1e-4))[alive]
    out_rays["dir"][alive, :3] = din[alive]
    out_rays["dir"][alive, 3] = FLT_MAX
    emitters = np.nonzero(emit)[0]
    return {"states": st, "rays": out_rays, "alive": alive, "shadow_rays": sh_rays, "shadow_states": sh_st,
            "queries": q, "emit_pixel": states["pixel"][emitters], "emit_rgb": emit_rgb[emitters]}


def merge(p: Params, q, idx, dist, cnt, cache_a, cache_b, k: int, image):
    """vertex_merging (vcm.cpp:572-611) for the queries q with their k-nearest lists (indices
    into the cache, squared distances, counts), added to image (pixels x 3, float64)."""
    for j in np.nonzero((q["pixel"] >= 0) & (cnt > 0))[0]:
        c = cnt[j]
        n, out = q["n"][j][None], q["out"][j][None]
        rs = dist[j, k - 1] if c == k else p.radius_sqr
        alb = p.albedo[q["mat"][j]] * (F(1) / PI)
        acc = np.zeros(3, F)
        for l in range(c):
            v = idx[j, l]
            pin = cache_a["out"][v][None]
            if not lambert_ok(n, out, pin)[0]:
                continue
            pdf_dir = np.abs(dot(n, pin))[0] / PI
            pdf_rev = np.abs(dot(n, out))[0] / PI
            wl = cache_b["dVCM"][v] * p.eta_vc + cache_b["dVM"][v] * pdf_dir
            wc = q["dVCM"][j] * p.eta_vc + q["dVM"][j] * pdf_rev
            mis = F(1) / ((wl + F(1)) + wc)
            kern = F(1) - dist[j, l] / rs
            acc = acc + ((mis * alb) * kern) * cache_b["tp"][v]
        scale_ = F(2) / ((PI * rs) * F(p.P))
        image[q["pixel"][j]] += q["tp"][j] * (acc * scale_)
