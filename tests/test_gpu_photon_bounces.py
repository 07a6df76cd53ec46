"""Every bounce of the photon-mapping frame against a numpy restatement of the device's surface
step, on the GPU's own queues (rays and states read back before each traversal) and the
oracle's traversal.

Light paths (render/integrators/vcm.cpp:196-256 with bounce :129-193, adjoint): the photons a
bounce deposits (position, direction back along the ray, path, path length, throughput), the
Russian-roulette decisions (render/random.h:183-188 with the MWC64X draws of random.h:15-47),
the throughput update, the random-number state and the order of the next queue are compared
bit for bit; the next ray's origin and tmin bit for bit; its direction bit for bit after a
mirror, within 1e-5 after a Lambertian bounce (sin / cos and the frame are the device's) and
after glass (its Fresnel choice must match).  Camera paths (process_camera_rays :372-441, PPM):
emitters end the path, non-specular hits write their merge query and end it, mirror and glass
continue it; compared the same way, with every written query bit for bit.
"""
import numpy as np
import pytest

import oracle
from imbatracer_amd import native as N
from imbatracer_amd import workloads as W

pytestmark = pytest.mark.gpu

F = np.float32
FLT_MAX = F(3.4028234663852886e38)


class Rng:
    """MWC64X (render/random.h:15-47), one generator per entry."""

    def __init__(self, s):
        self.s = np.asarray(s, np.uint64).copy()

    def draw(self, mask):
        s = self.s[mask]
        c = s >> np.uint64(32)
        x = s & np.uint64(0xFFFFFFFF)
        with np.errstate(over="ignore"):
            self.s[mask] = x * np.uint64(4294883355) + c
        out = (x ^ c).astype(np.uint32)
        u = np.zeros(len(self.s), np.float32)
        u[mask] = out.astype(np.float32) / F(4294967296.0)
        return u


def _dot(a, b):
    return (a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1]) + a[:, 2] * b[:, 2]


def _frame(n):
    """Tangent and binormal of the unit normals n (render/random.h:57-72 as the device has it)."""
    xbig = np.abs(n[:, 0]) > np.abs(n[:, 1])
    n0 = np.where(xbig, n[:, 0], n[:, 1])
    sig = np.where(xbig, F(-1), F(1))
    il = F(1) / np.sqrt(n0 * n0 + n[:, 2] * n[:, 2])
    t = np.zeros_like(n)
    t[:, 0] = np.where(xbig, n[:, 2] * sig * il, F(0))
    t[:, 1] = np.where(xbig, F(0), n[:, 2] * sig * il)
    t[:, 2] = n0 * F(-1) * sig * il
    b = np.stack([n[:, 1] * t[:, 2] - n[:, 2] * t[:, 1], n[:, 2] * t[:, 0] - n[:, 0] * t[:, 2],
                  n[:, 0] * t[:, 1] - n[:, 1] * t[:, 0]], 1)
    b = b * (F(1) / np.sqrt(_dot(b, b)))[:, None]
    return t, b


def _specular(kind, d, nf, entering, rng, mask):
    """New directions for the mirror (1) / glass (2) entries of mask; (dir, exact) arrays."""
    out = np.zeros_like(d)
    cos_i = -_dot(d, nf)
    refl = d + nf * (F(2) * cos_i)[:, None]
    mir = mask & (kind == 1)
    out[mir] = refl[mir]
    gl = mask & (kind == 2)
    eta = np.where(entering, F(1) / F(1.5), F(1.5)).astype(np.float32)
    with np.errstate(invalid="ignore"):
        sin2 = (eta * eta) * (F(1) - cos_i * cos_i)
        tir = gl & (sin2 >= F(1))
        fr = gl & ~(sin2 >= F(1))
        cos_t = np.sqrt(F(1) - sin2)
        rs = (eta * cos_i - cos_t) / (eta * cos_i + cos_t)
        rp = (cos_i - eta * cos_t) / (cos_i + eta * cos_t)
        f = F(0.5) * (rs * rs + rp * rp)
    u = rng.draw(fr)
    take_refl = tir | (fr & (u < f))
    out[take_refl] = refl[take_refl]
    refr = fr & ~(u < f)
    out[refr] = (d * eta[:, None] + nf * (eta * cos_i - cos_t)[:, None])[refr]
    return out


def _roulette(states, tp, alive, rng, rr_start):
    rr = alive & (states["bounces"] >= rr_start)
    u = rng.draw(rr)
    lum = (F(0.2126) * tp[:, 0] + F(0.7152) * tp[:, 1]) + F(0.0722) * tp[:, 2]
    p = np.minimum(F(1), lum * F(2))
    dead = rr & ~(u < p)
    alive = alive & ~dead
    keep = rr & ~dead
    with np.errstate(divide="ignore"):
        ip = F(1) / p
    tp[keep] = tp[keep] * ip[keep, None]
    return alive


def _surface(m, rays, states, hits):
    s = m.s
    tri = hits["tri_id"]
    ok = (tri >= 0) & (tri < s.tri_count)
    t_ = np.maximum(tri, 0)
    mat = np.where(ok, s.tri_materials[t_], -1)
    ok &= (mat >= 0) & (mat < len(s.albedo))
    mat = np.maximum(mat, 0)
    o, d = rays["org"][:, :3], rays["dir"][:, :3]
    t = hits["tmax"]
    p = o + t[:, None] * d
    nrm = s.tri_normals[t_]
    cos_o = -_dot(d, nrm)
    return ok, mat, p, nrm, cos_o, d, t


def light_step(m, rays, states, hits):
    s = m.s
    ok, mat, p, nrm, cos_o, d, t = _surface(m, rays, states, hits)
    ok &= cos_o != 0
    kind = s.kinds[mat]
    deposit = ok & (kind == 0)
    photons = (np.nonzero(deposit)[0], p, -d)
    alive = ok & (states["bounces"] + 1 < m.L)
    rng = Rng(states["rng"])
    tp = states["tp"].copy()
    alive = _roulette(states, tp, alive, rng, m.rr_start)
    entering = cos_o > 0
    nf = np.where(entering[:, None], nrm, -nrm).astype(np.float32)
    w = np.zeros_like(d)
    dif = alive & (kind == 0)
    u1 = rng.draw(dif)
    u2 = rng.draw(dif)
    tg, bn = _frame(nf)
    sin_t, cos_t = np.sqrt(F(1) - u2), np.sqrt(u2)
    ang = F(2) * F(3.14159265359) * u1
    lx, ly, lz = sin_t * np.cos(ang), sin_t * np.sin(ang), cos_t
    wd = tg * lx[:, None] + bn * ly[:, None] + nf * lz[:, None]
    w[dif] = wd[dif]
    spec = alive & (kind != 0)
    ws = _specular(kind, d, nf, entering, rng, spec)
    w[spec] = ws[spec]
    notglass = alive & (kind != 2)
    tp[notglass] = tp[notglass] * s.albedo[mat[notglass]]
    return photons, alive, w, p, t, tp, rng.s, kind


def camera_step(m, rays, states, hits):
    s = m.s
    ok, mat, p, nrm, cos_o, d, t = _surface(m, rays, states, hits)
    ok &= cos_o != 0
    emit = ok & (s.emission[mat] > 0).any(1)
    ok &= ~emit
    ok &= states["bounces"] + 1 < m.L
    kind = s.kinds[mat]
    query = ok & (kind == 0)
    alive = ok & (kind != 0)
    rng = Rng(states["rng"])
    tp = states["tp"].copy()
    alive = _roulette(states, tp, alive, rng, m.rr_start)
    entering = cos_o > 0
    nf = np.where(entering[:, None], nrm, -nrm).astype(np.float32)
    w = _specular(kind, d, nf, entering, rng, alive)
    mir = alive & (kind == 1)
    tp[mir] = tp[mir] * s.albedo[mat[mir]]
    c = F(1) / F(3.14159265359)
    weight = states["tp"] * s.albedo[mat] * c
    return query, (p, nrm, -d, weight), alive, w, p, t, tp, rng.s, kind


def _check_next(nxt_rays, nxt_states, states, alive, w, p, t, tp, rng_s, kind):
    idx = np.nonzero(alive)[0]
    assert len(nxt_rays) == len(idx)
    assert np.array_equal(nxt_states["pixel"], states["pixel"][idx])
    assert np.array_equal(nxt_states["bounces"], states["bounces"][idx] + 1)
    assert np.array_equal(nxt_states["rng"], rng_s[idx])
    assert np.array_equal(np.ascontiguousarray(nxt_states["tp"]).view(np.int32), tp[idx].view(np.int32))
    assert np.array_equal(nxt_rays["org"][:, :3], p[idx])
    assert np.array_equal(nxt_rays["org"][:, 3], F(1e-4) * t[idx])
    assert np.all(nxt_rays["dir"][:, 3] == FLT_MAX)
    got = nxt_rays["dir"][:, :3]
    mir = kind[idx] == 1
    assert np.array_equal(got[mir], w[idx][mir])
    assert np.allclose(got[~mir], w[idx][~mir], atol=1e-5, rtol=0)


@pytest.fixture(scope="module")
def pm(gpu):
    scene = W.SphereGridScene(side=4, level=2, builder=N.BUILDER_SBVH, specular_fraction=0.6)
    m = W.PhotonMapper(scene, 64, 48, light_paths=20000, max_depth=6, rr_start=1, radius_factor=0.02, k=10)
    m.clear_image()
    m.keep_queues = True
    m.frame()
    return m


def test_light_bounces(pm):
    arrays = pm.s.build.arrays()
    ph_rays, ph_states = pm.photons_host()
    queues = pm.kept["light"]
    assert len(queues) == pm.L and len(queues[0][0]) == pm.P
    for b, (rays, states) in enumerate(queues):
        if len(rays) == 0:
            continue
        hits = oracle.trace(arrays, rays, threads=8)[0]
        (dep, p, back), alive, w, pos, t, tp, rng_s, kind = light_step(pm, rays, states, hits)
        mine = ph_rays["length"] == b + 2
        assert np.array_equal(ph_states["path"][mine], states["pixel"][dep])
        assert np.array_equal(ph_rays["pos"][mine], p[dep])
        assert np.array_equal(ph_rays["out_dir"][mine], back[dep])
        assert np.array_equal(ph_states["tp"][mine].view(np.int32), states["tp"][dep].view(np.int32))
        if b + 1 < pm.L:
            nr, ns = queues[b + 1]
            _check_next(nr, ns, states, alive, w, pos, t, tp, rng_s, kind)
    assert len(ph_rays) == sum(int((ph_rays["length"] == b + 2).sum()) for b in range(pm.L))


def test_camera_bounces(pm):
    arrays = pm.s.build.arrays()
    q = pm.queries_host()
    queues = pm.kept["camera"]
    written = np.zeros(len(q), bool)
    for b, (rays, states) in enumerate(queues):
        if len(rays) == 0:
            continue
        hits = oracle.trace(arrays, rays, threads=8)[0]
        query, (p, nrm, back, weight), alive, w, pos, t, tp, rng_s, kind = camera_step(pm, rays, states, hits)
        i = np.nonzero(query)[0]
        slot = states["pixel"][i] * pm.spp + states["sample"][i]
        assert np.array_equal(q["pixel"][slot], states["pixel"][i])
        assert np.array_equal(q["pos"][slot], p[i])
        assert np.array_equal(q["normal"][slot], nrm[i])
        assert np.array_equal(q["out_dir"][slot], back[i])
        assert np.array_equal(q["weight"][slot].view(np.int32), weight[i].view(np.int32))
        written[slot] = True
        if b + 1 < pm.L:
            nr, ns = queues[b + 1]
            _check_next(nr, ns, states, alive, w, pos, t, tp, rng_s, kind)
    assert np.all(q["pixel"][~written] == -1)
    assert written.sum() > 0.5 * len(q)
