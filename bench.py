"""Benchmark: incoherent closest-hit traversal throughput on one MI355X.

Default workload (the configuration the headline metric is quoted on): path tracing at
1280x720, 1 spp, maximum path length 8 with Russian roulette from bounce 3, over the
sphere-grid scene (32 x 32 level-5 icospheres + open box + triangle light, 20,971,533
triangles, SBVH per mesh).  One step = one frame: camera rays, then per bounce closest-hit
traversal, Lambertian bounce and queue compaction, all on the device, captured once into a
graph whose queue sizes stay in device memory (--mode eager: one launch per call, the host
reading every bounce's queue size).  `value` is the number of closest-hit rays traversed per
second of frame time (inputs resident in HBM; the scene build and upload are outside the
timed region).

--workload c1 runs the CPU-runnable configuration instead: 65,536 random triangles
(binned SAH), 1,048,576 primary closest-hit rays plus 1,048,576 occlusion rays per step.
--workload c3: one VCM iteration at 1280x720 with 921,600 light paths: light and camera
traversal with next-event, camera-connection and vertex-connection shadow rays, the light vertex
cache, hash-grid build and a k-nearest merge at every Lambertian camera vertex per frame
(--workload c3-ppm: the integrator's PPM mode, one merge query per pixel).
--workload c4: 256 queues over 8 scenes, about 35M closest-hit rays per step in one batched
launch, against the same queues launched one at a time.
--workload c5: serving; Poisson request arrivals over 8 resident scenes, the request loop with
cross-session batching against the plain per-request path (p50/p99 latency and Mrays/s).

Also reported: the roofline of the traversal kernel (algorithmic bytes per ray
32 + 16 + 64 N_node + 48 N_tri + 64 N_inst with N_node the inner nodes of the reference's binary
tree (BVH2) whose child boxes a ray tests, N_tri its triangle tests and N_inst its instance
visits, counted on the same rays by the BVH2 stats variant, whose counts equal the oracle's,
over the kernel time measured with HIP events on its stream; the same figure with the visit
counts of the tree the kernel walks is reported beside it)
and a CPU baseline: the reference's CPU path restated (oracle/cpu_bvh4.c over the reference's
CPU layout, BVH4 nodes + 4-triangle packets built from the same scene) on host threads over a
bounded sample of the same rays.
"""
from __future__ import annotations

import argparse
import ctypes as C
import glob
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s incoherent closest-hit; served p50/p99 frame latency vs plain GPU"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def ray_bytes(visits, rays, node_bytes=64):
    """Algorithmic bytes of `rays` rays with summed visit counts (nodes, tris, instances); a
    BVH2 node is 64 bytes, a 4-wide node 128."""
    nodes, tris, inst = (float(x) for x in visits)
    return 32.0 * rays + 16.0 * rays + node_bytes * nodes + 48.0 * tris + 64.0 * inst


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` on `workload` from the rocprofv3 PMC summaries under
    profiles/ (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by a device copy of known size),
    as (bytes, file), or (None, None) when no summary covers this workload."""
    pat = re.compile(r"traffic: workload=(\S+) kernel=(\S+) hbm_bytes_per_launch=(\d+)")
    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(f) as fh:
            for line in fh:
                m = pat.search(line)
                if m and m.group(1) == workload and m.group(2) == kernel:
                    return int(m.group(3)), os.path.relpath(f, ROOT)
    return None, None


def fixed_config(args):
    """The launch configuration given on the command line, or None."""
    from imbatracer_amd import native as N

    if args.block or args.persistent or args.blocks_per_cu or args.bvh_width or args.quantized or args.treelet:
        return N.TraceConfig(args.block, args.stack, args.persistent, args.blocks_per_cu, args.bvh_width or
                             (4 if args.quantized or args.treelet else 0), int(args.quantized or args.treelet > 0),
                             args.treelet)
    return None


def cfg_dict(cfg):
    if cfg is None:
        return "default"
    return {"block": cfg.block_size or 256, "stack_lds": cfg.stack_lds or 16, "persistent": cfg.persistent,
            "bvh_width": cfg.bvh_width or 2, "quantized": cfg.quantized, "treelet": cfg.treelet}


BVH2_CFG_ARGS = (256, 16, 0, 0, 2, 0, 0)  # TraceConfig of the reference's binary tree


def bvh2_visits(scene, rays, any_hit=False):
    """Summed visit counts (inner nodes, triangles, instances) of rays on the reference's binary
    tree (BVH2, nearer child first, ties to the left child: the oracle's traversal rule and
    counts).  rays: a host ray array."""
    from imbatracer_amd import native as N

    n = len(rays)
    if n == 0:
        return np.zeros(3, np.int64)
    rb = N.DeviceBuffer.from_array(rays)
    hb, cb = N.DeviceBuffer(n * 16), N.DeviceBuffer(n * 12)
    scene.trace_stats(rb, hb, cb, n, any_hit=any_hit, cfg=N.TraceConfig(*BVH2_CFG_ARGS))
    return cb.download(np.uint32, 3 * n).reshape(-1, 3).astype(np.int64).sum(0)


BYTES_RULE = ("per ray 32 (ray) + 16 (hit) + 64 x inner nodes of the reference's binary tree (BVH2) whose "
              "child boxes the ray tests + 48 x triangles tested + 64 x instances visited, counted on the same "
              "rays by the BVH2 stats variant (the oracle's counts); the kernel itself walks the quantized "
              "4-wide tree derived from the BVH2 (own_tree: the same formula with that tree's counts, 64 B per "
              "quantized node)")


def node_bytes_of(cfg):
    """Bytes of one node of the tree a configuration walks: BVH2 64, 4-wide 128, quantized 64."""
    return 128 if cfg is not None and cfg.bvh_width == 4 and not cfg.quantized else 64


def kernel_label(cfg, any_hit=False):
    suffix = ""
    if cfg is not None and cfg.bvh_width == 4:
        suffix = "_w4q" if cfg.quantized else "_w4"
        if cfg.quantized and cfg.treelet:
            suffix += "_tl%d" % cfg.treelet
    return ("k_trace_occluded" if any_hit else "k_trace_closest") + suffix


def autotune(scene, rays, any_hit=False):
    """The library's autotuner on one device queue (every shape, persistent or not, both tree
    widths): the fastest configuration and its time per launch."""
    from imbatracer_amd import native as N

    rb = N.DeviceBuffer.from_array(rays)
    hb = N.DeviceBuffer(len(rays) * 16)
    cfg, ms = scene.autotune(rb, hb, len(rays), any_hit=any_hit, reps=3)
    log("autotune (%s, %d rays): %s, %.3f ms" % ("occluded" if any_hit else "closest", len(rays), cfg_dict(cfg), ms))
    return cfg


def read_bandwidths(stream):
    """Measured read bandwidth (GB/s) of the level that holds a buffer of each size: 2 MB (L2),
    128 MB (Infinity Cache, above the 32 MB of all L2s), 4 GiB (HBM); best of three runs."""
    from imbatracer_amd import native as N

    out = {}
    sink = N.DeviceBuffer(32 * 1024 * 256)
    for name, size, passes in (("l2_2MB", 2 << 20, 2000), ("mall_128MB", 128 << 20, 32), ("hbm_4GiB", 4 << 30, 2)):
        buf = N.DeviceBuffer(size)
        N.check(N.lib.imba_memset(buf.ptr, 0, size, stream.handle), "memset")
        ms = C.c_float()
        best = None
        for _ in range(3):
            N.check(N.lib.imba_read_bandwidth(buf.ptr, size, passes, sink.ptr, sink.nbytes, stream.handle,
                                              C.byref(ms)), "read_bandwidth")
            gbs = size * passes / (ms.value * 1e-3) / 1e9
            best = gbs if best is None else max(best, gbs)
        out[name] = best
        del buf
    log("read bandwidth (GB/s): %s" % {k: round(v) for k, v in out.items()})
    return out


def shadow_tune(pt, frames=3):
    """Launch configuration of the shadow-ray (any-hit) traversal, per bounce: the library's
    autotuner on the first bounces' shadow slots plus a few plain launches, each timed over
    whole frames; the tree is the fastest candidate's, and every bounce takes the fastest
    candidate on that tree."""
    from imbatracer_amd import native as N

    pt.frame(keep_queues=8)
    shadows = [sr for sr, _, _ in pt.kept_shadow]
    cands = [autotune(pt.gpu_scene, shadows[b], any_hit=True) for b in sorted({0, len(shadows) // 2})
             if b < len(shadows) and len(shadows[b])]
    cands += [N.TraceConfig(256, 16, 0, 0, 4, 1, 0), N.TraceConfig(256, 16, 0, 0, 2, 0, 0),
              N.TraceConfig(128, 16, 0, 0, 4, 1, 0), N.TraceConfig(64, 32, 0, 0, 4, 1, 0)]
    seen, timed = set(), []
    for c in cands:
        key = tuple(getattr(c, f) for f, _ in c._fields_)
        if key in seen:
            continue
        seen.add(key)
        pt.shadow_cfg, pt.shadow_bounce_cfg = c, None
        pt.frame()
        per_bounce = None
        for _ in range(frames):
            pt.frame(timed=True)
            pt.stream.sync()
            t = [ms for ms, _ in pt.shadow_times_ms()]
            per_bounce = t if per_bounce is None else [x + y for x, y in zip(per_bounce, t)]
        log("shadow tune: %s %.3f ms of shadow traversal per frame" % (cfg_dict(c), sum(per_bounce) / frames))
        timed.append((sum(per_bounce), c, per_bounce))
    best = min(timed, key=lambda x: x[0])
    tree = (best[1].bvh_width, best[1].quantized)
    same = [x for x in timed if (x[1].bvh_width, x[1].quantized) == tree]
    bounce_cfg = [min(same, key=lambda x: x[2][b])[1] for b in range(len(best[2]))]
    pt.shadow_cfg, pt.shadow_bounce_cfg = best[1], bounce_cfg
    log("shadow tune: per bounce %s" % [cfg_dict(c) for c in bounce_cfg])
    return best[1], bounce_cfg


def frame_tune(pt, queues, frames=3):
    """Per-scene launch configuration for whole frames: the library's autotuner on queues of
    three sizes (a frame's queues shrink from bounce to bounce), each distinct winner and a few
    plain quantized launches (with and without a 64-node treelet, persistent or not) timed over
    whole frames (the traversal kernels' times per bounce).  The frame's tree is the one of the
    fastest candidate; each bounce then takes the fastest candidate on that tree for its own
    queue, so every bounce walks the same tree (same visits, same algorithmic bytes)."""
    from imbatracer_amd import native as N

    cands = [autotune(pt.gpu_scene, queues[b]) for b in sorted({1, len(queues) // 2, len(queues) - 2})
             if 0 < b < len(queues) and len(queues[b])]
    cands += [N.TraceConfig(256, 16, 0, 0, 4, 1, 0), N.TraceConfig(256, 16, 0, 0, 4, 1, 64),
              N.TraceConfig(128, 16, 0, 0, 4, 1, 0), N.TraceConfig(64, 32, 0, 0, 4, 1, 0),
              N.TraceConfig(256, 16, 32, 0, 4, 1, 0), N.TraceConfig(256, 16, 8, 0, 4, 1, 0)]
    seen, timed = set(), []
    for c in cands:
        key = tuple(getattr(c, f) for f, _ in c._fields_)
        if key in seen:
            continue
        seen.add(key)
        pt.cfg, pt.bounce_cfg = c, None
        pt.frame()
        per_bounce = None
        for _ in range(frames):
            pt.frame(timed=True)
            pt.stream.sync()
            t = [ms for ms, _ in pt.trace_times_ms()]
            per_bounce = t if per_bounce is None else [x + y for x, y in zip(per_bounce, t)]
        log("frame tune: %s %.3f ms of traversal per frame" % (cfg_dict(c), sum(per_bounce) / frames))
        timed.append((sum(per_bounce), c, per_bounce))
    best = min(timed, key=lambda x: x[0])
    tree = (best[1].bvh_width, best[1].quantized)
    same = [x for x in timed if (x[1].bvh_width, x[1].quantized) == tree]
    bounce_cfg = [min(same, key=lambda x: x[2][b])[1] for b in range(len(best[2]))]
    total = sum(min(x[2][b] for x in same) for b in range(len(best[2])))
    log("frame tune: per bounce %s, %.3f ms of traversal per frame" % (
        [cfg_dict(c) for c in bounce_cfg], total / frames))
    pt.cfg, pt.bounce_cfg = best[1], bounce_cfg
    return best[1], bounce_cfg


def cgroup_cpus():
    """The CPU quota of this process's control group in cores (None if unlimited or unknown)."""
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()[:2]
        if quota != "max":
            return float(quota) / float(period)
    except (OSError, ValueError):
        pass
    try:
        with open("/sys/fs/cgroup/cpu/cpu.cfs_quota_us") as f:
            quota = float(f.read())
        with open("/sys/fs/cgroup/cpu/cpu.cfs_period_us") as f:
            period = float(f.read())
        if quota > 0:
            return quota / period
    except (OSError, ValueError):
        pass
    return None


def affinity_cores() -> int:
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:
        return max(1, os.cpu_count() or 1)


def host_cores() -> int:
    """Host cores this process may run on: the affinity mask, capped by the control group's CPU
    quota when there is one.  TileScheduler starts one render thread per hardware thread
    (render/scheduling/tile_scheduler.h:69-78); more threads than the quota allows would only
    queue for it."""
    n = affinity_cores()
    q = cgroup_cpus()
    if q is not None:
        n = min(n, max(1, int(q + 0.5)))
    return n


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def sample_queues(queues, per_set: int):
    """An every-k-th-ray sample of each queue with the same k for all, so the sample keeps the
    queues' proportions (about per_set rays from the largest one)."""
    k = max(1, max(len(q) for q in queues) // max(per_set, 1))
    return [np.ascontiguousarray(q[::k]) for q in queues if len(q[::k])]


def cpu_baseline(build, builder, threads_build, ray_sets, any_flags, budget_s=12.0, note=""):
    """Times the reference's CPU path restated (oracle/cpu_bvh4.c: BVH4 nodes + 4-triangle
    packets of core/cpu_adapter.cpp, 4-wide SSE/FMA box and triangle tests, one ray per host
    thread at a time) on every host core the process may use over a bounded sample of the
    benchmarked rays: one untimed pass, then whole passes until the time budget is used."""
    import oracle

    t0 = time.perf_counter()
    build.cpu_layout(builder, threads_build)
    t_layout = time.perf_counter() - t0
    desc = build.cpu_desc()
    threads = host_cores()
    od = oracle.SceneDesc(*[getattr(desc, f) for f, _ in oracle.SceneDesc._fields_])
    L = oracle.lib()
    sets = [(np.ascontiguousarray(r), a) for r, a in zip(ray_sets, any_flags)]
    hits = np.empty(max(len(r) for r, _ in sets), oracle.HIT_DTYPE)

    def one_pass():
        bad = 0
        t = time.perf_counter()
        for rays, any_hit in sets:
            bad += L.cpu4_trace_mt(C.byref(od), C.c_void_p(rays.ctypes.data), len(rays), int(any_hit),
                                   C.c_void_p(hits.ctypes.data), threads)
        if bad:
            raise RuntimeError("CPU traversal met a malformed structure")
        return time.perf_counter() - t

    one_pass()
    n_pass = sum(len(r) for r, _ in sets)
    times = []
    while sum(times) < budget_s and len(times) < 1000:
        times.append(one_pass())
    elapsed = sum(times)
    return {"value": n_pass * len(times) / elapsed / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "affinity_cores": affinity_cores(), "cgroup_cpu_quota": cgroup_cpus(),
            "cpu_model": cpu_model(), "threads": threads,
            "sample": "%d timed passes (after 1 warm-up) over %s rays (%s)%s; oracle/cpu_bvh4.c (the reference CPU "
                      "layout: BVH4 + 4-triangle packets, 4-wide SSE/FMA) on %d host threads (one per core of the "
                      "affinity mask, capped by the CPU quota), %.1f s; median pass %.1f Mrays/s; CPU layout built in %.1f s" % (
                          len(times), "+".join(str(len(r)) for r, _ in sets),
                          "/".join("occlusion" if a else "closest" for _, a in sets), note, threads, elapsed,
                          n_pass / float(np.median(times)) / 1e6, t_layout)}


def run_pt(args):
    from imbatracer_amd import native as N
    from imbatracer_amd import workloads as W

    t0 = time.time()
    threads = host_cores()
    scene = W.SphereGridScene(side=args.side, level=5, builder=N.BUILDER_SBVH, threads=threads)
    log("scene: %d triangles, %d nodes, built in %.1f s" % (scene.tri_count, scene.desc.node_count, time.time() - t0))
    cfg = fixed_config(args)
    pt = W.PathTracer(scene, args.width, args.height, spp=1, max_path_len=8, rr_start=3, trace_cfg=cfg)
    log("uploaded %.2f GB" % (pt.gpu_scene.footprint() / 1e9))
    # the queues of a frame (closest hits, hence the queues, do not depend on the configuration)
    _, _, queues = pt.frame(keep_queues=8)
    shadow_queues = [sr[ss["pixel"] >= 0] for sr, ss, _ in pt.kept_shadow]
    bounce_cfg = shadow_bounce_cfg = None
    if cfg is None and not args.no_autotune:
        cfg, bounce_cfg = frame_tune(pt, queues)
        _, shadow_bounce_cfg = shadow_tune(pt)
    node_bytes = node_bytes_of(cfg)

    # untimed: visit counts of every bounce's closest-hit and shadow launches (same seeds -> same
    # rays as the timed frames)
    counts, visits, _ = pt.frame(stats=True)
    shadow_counts, shadow_visits = list(pt.shadow_counts), list(pt.shadow_visit)
    stat_bytes = [ray_bytes(v, n, node_bytes) for v, n in zip(visits, counts)]
    shadow_bytes = [ray_bytes(v, n, node_bytes_of(pt.shadow_cfg_of(b)))
                    for b, (v, n) in enumerate(zip(shadow_visits, counts))]
    # the algorithmic bytes of the roofline: the reference's BVH2 visit counts on the same rays
    bvh2_vis = [bvh2_visits(pt.gpu_scene, q) for q in queues]
    bvh2_shadow_vis = [bvh2_visits(pt.gpu_scene, q, any_hit=True) for q in shadow_queues]
    contract_bytes = [ray_bytes(v, len(q), 64) for v, q in zip(bvh2_vis, queues)]
    contract_shadow_bytes = [ray_bytes(v, len(q), 64) for v, q in zip(bvh2_shadow_vis, shadow_queues)]
    log("rays per bounce %s, shadow rays %s" % (counts, shadow_counts))

    # eager frames (host reads each bounce's queue size): traversal kernel times for the roofline
    for _ in range(args.warmup):
        pt.frame()
    pt.stream.sync()
    eager_ms, trace_ms, rays_eager, bytes_total, contract_total = [], 0.0, 0, 0.0, 0.0
    shadow_ms, shadow_bytes_total, shadow_rays_eager, contract_shadow_total = 0.0, 0.0, 0, 0.0
    per_bounce_ms = np.zeros(len(counts))
    per_bounce_shadow_ms = np.zeros(len(counts))
    t_eager = time.perf_counter()
    for _ in range(args.steps):
        f0 = time.perf_counter()
        c, _, _ = pt.frame(timed=True)
        pt.stream.sync()
        eager_ms.append((time.perf_counter() - f0) * 1e3)
        if c != counts or pt.shadow_counts != shadow_counts:
            raise RuntimeError("frame ray counts changed between frames: %s vs %s" % (c, counts))
        for b, ((ms, n), sb) in enumerate(zip(pt.trace_times_ms(), stat_bytes)):
            trace_ms += ms
            bytes_total += sb
            contract_total += contract_bytes[b]
            per_bounce_ms[b] += ms
        for b, ((ms, n), sb) in enumerate(zip(pt.shadow_times_ms(), shadow_bytes)):
            shadow_ms += ms
            shadow_bytes_total += sb
            contract_shadow_total += contract_shadow_bytes[b] if b < len(contract_shadow_bytes) else 0.0
            per_bounce_shadow_ms[b] += ms
        rays_eager += sum(c)
        shadow_rays_eager += sum(shadow_counts)
    wall_eager = time.perf_counter() - t_eager
    per_bounce_ms /= args.steps
    per_bounce_shadow_ms /= args.steps

    # the measured frames: graph launches, `frames_in_flight` frame loops (each with its own
    # queues, stream and graph over the one uploaded scene) taking turns, so that one frame's
    # short last bounces overlap the next frame's first bounces on the device (progressive
    # rendering renders frame after frame, frontend/render_window.cpp:50-91)
    frame_ms, rays_total = [], 0
    loops = []
    if args.mode == "graph":
        loops = [pt] + [W.PathTracer(scene, args.width, args.height, spp=1, max_path_len=8, rr_start=3, trace_cfg=cfg,
                                     shadow_cfg=pt.shadow_cfg, gpu_scene=pt.gpu_scene)
                        for _ in range(args.frames_in_flight - 1)]
        for lp in loops:
            lp.bounce_cfg, lp.shadow_bounce_cfg = pt.bounce_cfg, pt.shadow_bounce_cfg
            lp.build_graph(use_batch=args.trace_kernel == "batch")
            for _ in range(args.warmup):
                lp.frame_graph()
        started = {}
        t_start = time.perf_counter()
        for step in range(args.steps + len(loops)):
            lp = loops[step % len(loops)]
            if id(lp) in started:
                c = lp.finish_graph()
                frame_ms.append((time.perf_counter() - started.pop(id(lp))) * 1e3)
                if c != counts:
                    raise RuntimeError("frame ray counts changed between frames: %s vs %s" % (c, counts))
                rays_total += sum(c)
            if step < args.steps:
                lp.launch_graph()
                started[id(lp)] = time.perf_counter()
        wall = time.perf_counter() - t_start
    else:
        t_start = time.perf_counter()
        for _ in range(args.steps):
            f0 = time.perf_counter()
            c, _, _ = pt.frame()
            pt.stream.sync()
            frame_ms.append((time.perf_counter() - f0) * 1e3)
            if c != counts:
                raise RuntimeError("frame ray counts changed between frames: %s vs %s" % (c, counts))
            rays_total += sum(c)
        wall = time.perf_counter() - t_start
    if pt.gpu_scene.error_count():
        raise RuntimeError("traversal reported a malformed scene")
    achieved = contract_total / (trace_ms * 1e-3) / 1e9
    shadow_achieved = contract_shadow_total / (shadow_ms * 1e-3) / 1e9
    own_achieved = bytes_total / (trace_ms * 1e-3) / 1e9
    own_shadow_achieved = shadow_bytes_total / (shadow_ms * 1e-3) / 1e9
    workload = "pt-%dx%d-1spp-depth8-rr3-nee-sphere-grid-%dtri-sbvh" % (args.width, args.height, scene.tri_count)
    traffic, traffic_src = measured_traffic(workload, kernel_label(cfg))
    shadow_frame = sum(shadow_counts) * args.steps
    # stack use per bounce (the stats variant on the same queues): rays whose deepest stack entry
    # lies beyond the 16 entries kept in LDS spill to the global stack buffer
    depth = []
    for b, rays in enumerate(queues):
        h = pt.gpu_scene.stack_depths(rays, cfg=pt.cfg_of(b))
        hs = pt.gpu_scene.stack_depths(shadow_queues[b], any_hit=True, cfg=pt.shadow_cfg_of(b)) \
            if b < len(shadow_queues) else np.zeros(64, np.uint32)
        depth.append((h, hs))

    def depth_summary(h):
        n = max(int(h.sum()), 1)
        nz = np.nonzero(h)[0]
        return {"max": int(nz.max()) if len(nz) else 0, "mean": float((np.arange(64) * h).sum() / n),
                "frac_beyond_16": float(h[17:].sum() / n), "frac_beyond_8": float(h[9:].sum() / n)}

    table = [{"bounce": b, "rays": int(n), "closest_ms": float(per_bounce_ms[b]),
              "closest_stack_depth": depth_summary(depth[b][0]) if b < len(depth) else None,
              "occluded_stack_depth": depth_summary(depth[b][1]) if b < len(depth) else None,
              "closest_mrays_s": n / (per_bounce_ms[b] * 1e-3) / 1e6,
              "closest_cfg": cfg_dict(pt.cfg_of(b)),
              "shadow_rays": int(shadow_counts[b]), "occluded_ms": float(per_bounce_shadow_ms[b]),
              "occluded_mrays_s": shadow_counts[b] / (per_bounce_shadow_ms[b] * 1e-3) / 1e6,
              "occluded_cfg": cfg_dict(pt.shadow_cfg_of(b))} for b, n in enumerate(counts)]
    res = {
        "metric": METRIC, "value": rays_total / wall / 1e6, "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded procedural scene and camera)",
        "config": {"workload": workload,
                   "width": args.width, "height": args.height, "spp": 1, "max_path_len": 8,
                   "integrator": "PathTracer::process_primary_rays: emitter hits with MIS, one next-event "
                                 "shadow ray per surface hit (traversed any-hit), Lambertian bounce; Russian "
                                 "roulette drawn every bounce, applied from bounce 3",
                   "rays_per_frame": int(sum(counts)), "rays_per_bounce": counts,
                   "shadow_rays_per_frame": int(sum(shadow_counts)), "shadow_rays_per_bounce": shadow_counts,
                   "triangles": scene.tri_count, "nodes": scene.desc.node_count,
                   "frame": ("one graph launch per frame (queue sizes on the device), %d frames in flight on "
                             "their own streams" % args.frames_in_flight) if args.mode == "graph"
                            else "eager launches, host reads each bounce's queue size",
                   "frames_in_flight": args.frames_in_flight if args.mode == "graph" else 1,
                   "trace_kernel": args.trace_kernel,
                   "trace_config": cfg_dict(cfg), "autotuned": cfg is not None and not args.block,
                   "trace_config_per_bounce": [cfg_dict(c) for c in bounce_cfg] if bounce_cfg else None,
                   "shadow_config_per_bounce": [cfg_dict(c) for c in shadow_bounce_cfg] if shadow_bounce_cfg
                   else None},
        "value_note": "closest-hit rays per second of frame wall time; a frame also traverses its shadow rays",
        "rays_incl_shadow_mrays_s": (rays_total + shadow_frame) / wall / 1e6,
        "eager_frames": {"value": rays_eager / wall_eager / 1e6, "ms_per_step": wall_eager * 1e3 / args.steps,
                         "frame_latency_ms": {"p50": float(np.percentile(eager_ms, 50)),
                                              "p99": float(np.percentile(eager_ms, 99))}},
        "trace_mrays_s": rays_eager / (trace_ms * 1e-3) / 1e6,
        "occluded": {"mrays_s": shadow_rays_eager / (shadow_ms * 1e-3) / 1e6,
                     "ms_per_frame": shadow_ms / args.steps, "rays_per_frame": int(sum(shadow_counts)),
                     "roofline": {"bound": "hbm", "achieved": shadow_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": shadow_achieved / HBM_PEAK_GBS,
                                  "kernel": "k_trace (any hit) over the shadow slots",
                                  "traffic": measured_traffic(workload, kernel_label(pt.shadow_cfg_of(0), any_hit=True))[0],
                                  "algorithmic_bytes_per_launch": sum(contract_shadow_bytes) / len(shadow_bytes),
                                  "bytes_rule": BYTES_RULE + "; the shadow rays only, not the empty slots",
                                  "bvh2_visits_per_ray": [float(x) / max(sum(shadow_counts), 1)
                                                          for x in np.sum(bvh2_shadow_vis, 0)],
                                  "own_tree": {"achieved": own_shadow_achieved,
                                               "frac": own_shadow_achieved / HBM_PEAK_GBS,
                                               "visits_per_slot": [float(x) / sum(counts)
                                                                   for x in np.sum(shadow_visits, 0)]}}},
        "frame_traversal_bandwidth": {
            "achieved": (sum(contract_bytes) + sum(contract_shadow_bytes)) / (wall / args.steps) / 1e9, "unit": "GB/s",
            "frac": (sum(contract_bytes) + sum(contract_shadow_bytes)) / (wall / args.steps) / 1e9 / HBM_PEAK_GBS,
            "own_tree_frac": (sum(stat_bytes) + sum(shadow_bytes)) / (wall / args.steps) / 1e9 / HBM_PEAK_GBS,
            "rule": "algorithmic bytes of a frame's closest-hit and shadow traversals over the measured frame "
                    "time (frames overlapping as configured): the traversal kernels' combined rate"},
        "per_bounce": table,
        "frame_latency_ms": {"p50": float(np.percentile(frame_ms, 50)), "p99": float(np.percentile(frame_ms, 99))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "k_trace (closest hit)",
                     "algorithmic_bytes_per_launch": sum(contract_bytes) / len(contract_bytes),
                     "bytes_per_ray": contract_total / max(rays_eager, 1),
                     "bytes_rule": BYTES_RULE,
                     "bvh2_visits_per_ray": [float(x) / sum(counts) for x in np.sum(bvh2_vis, 0)],
                     "own_tree": {"achieved": own_achieved, "frac": own_achieved / HBM_PEAK_GBS,
                                  "algorithmic_bytes_per_launch": sum(stat_bytes) / len(stat_bytes),
                                  "bytes_per_ray": bytes_total / max(rays_eager, 1),
                                  "visits_per_ray": [float(x) / sum(counts) for x in np.sum(visits, 0)]}},
    }
    if not args.no_serving:
        # the metric's second half: served vs plain per-request throughput and p99 latency on the
        # serving mix (the 21M-triangle scene reused), and on small requests
        loops.clear()  # the extra frame loops' queues are not needed any more
        c5 = run_c5(args, big_scene=scene, saturated=False)
        res["serving"] = {k: c5[k] for k in ("config", "served", "plain_per_request", "served_vs_plain",
                                              "small_requests")}
    if not args.no_cpu:
        sample = sample_queues(queues, 200_000)
        res["cpu_baseline"] = cpu_baseline(scene.build, N.BUILDER_SBVH, threads, sample, [False] * len(sample),
                                           budget_s=args.cpu_seconds,
                                           note=": every k-th ray of each of the frame's %d closest-hit queues" %
                                                len(sample))
    return res


def run_c1(args):
    from imbatracer_amd import native as N
    from imbatracer_amd import scene as S
    from imbatracer_amd import workloads as T

    b, _, _ = T.random_tri_scene(65536, N.BUILDER_BINNED_SAH)
    desc = b.desc()
    sc = S.Scene(desc)
    prim = T.primary_rays(1024, 1024)
    n = len(prim)
    rb = N.DeviceBuffer.from_array(prim)
    hb = N.DeviceBuffer(n * 16)
    cb = N.DeviceBuffer(n * 12)
    cfg = cfg_any = fixed_config(args)
    hits = sc.trace_host(prim)
    shadow = T.shadow_rays(prim, hits)
    if cfg is None and not args.no_autotune:
        cfg = autotune(sc, prim)
        cfg_any = autotune(sc, shadow, any_hit=True)
    node_bytes = node_bytes_of(cfg)
    sc.trace_stats(rb, hb, cb, n, cfg=cfg)
    vis_p = cb.download(np.uint32, 3 * n).reshape(-1, 3).astype(np.int64).sum(0)
    sb = N.DeviceBuffer.from_array(shadow)
    ob = N.DeviceBuffer(n * 16)
    sc.trace_stats(sb, ob, cb, n, any_hit=True, cfg=cfg_any)
    vis_s = cb.download(np.uint32, 3 * n).reshape(-1, 3).astype(np.int64).sum(0)
    st = N.Stream()
    for _ in range(args.warmup):
        sc.intersect(rb, hb, n, cfg, st.handle)
        sc.occluded(sb, ob, n, cfg_any, st.handle)
    st.sync()
    ev = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e0, e1, e2 = N.Event(), N.Event(), N.Event()
        e0.record(st)
        sc.intersect(rb, hb, n, cfg, st.handle)
        e1.record(st)
        sc.occluded(sb, ob, n, cfg_any, st.handle)
        e2.record(st)
        ev.append((e0, e1, e2))
    st.sync()
    wall = time.perf_counter() - t0
    t_closest = sum(a.elapsed_ms(b) for a, b, c in ev)
    vis_p2 = bvh2_visits(sc, prim)
    achieved = ray_bytes(vis_p2, n, 64) * args.steps / (t_closest * 1e-3) / 1e9
    own_achieved = ray_bytes(vis_p, n, node_bytes) * args.steps / (t_closest * 1e-3) / 1e9
    workload = "c1-65536-random-tris-binned-sah-1M-primary-1M-occlusion"
    traffic, traffic_src = measured_traffic(workload, kernel_label(cfg))
    res = {
        "metric": METRIC, "value": 2 * n * args.steps / wall / 1e6, "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded random triangles, pinhole camera, point light)",
        "config": {"workload": workload, "trace_config": cfg_dict(cfg), "occluded_config": cfg_dict(cfg_any)},
        "closest_mrays_s": n * args.steps / (t_closest * 1e-3) / 1e6,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "k_trace (closest hit)", "algorithmic_bytes_per_launch": ray_bytes(vis_p2, n, 64),
                     "bytes_per_ray": ray_bytes(vis_p2, n, 64) / n, "bytes_rule": BYTES_RULE,
                     "bvh2_visits_per_ray": [float(x) / n for x in vis_p2],
                     "own_tree": {"achieved": own_achieved, "frac": own_achieved / HBM_PEAK_GBS,
                                  "bytes_per_ray": ray_bytes(vis_p, n, node_bytes) / n,
                                  "visits_per_ray": [float(x) / n for x in vis_p]}},
    }
    # the PCIe-inclusive path (RayQueue::traverse_gpu, ray_queue.h:269-293: host rays copied in,
    # hits copied out, per call, from pageable host memory): reported beside `value`, never as it
    host_hits = np.empty(n, N.HIT_DTYPE)
    st.sync()
    t_pcie = time.perf_counter()
    for _ in range(args.steps):
        N.check(N.lib.imba_memcpy_h2d(rb.ptr, N.ptr(prim), prim.nbytes, st.handle), "h2d")
        sc.intersect(rb, hb, n, cfg, st.handle)
        N.check(N.lib.imba_memcpy_d2h(N.ptr(host_hits), hb.ptr, host_hits.nbytes, st.handle), "d2h")
        st.sync()
    t_pcie = time.perf_counter() - t_pcie
    res["pcie_inclusive"] = {"closest_mrays_s": n * args.steps / t_pcie / 1e6,
                             "rule": "primary rays copied from pageable host memory, traversed, hits copied back, "
                                     "one call at a time (the reference's RayQueue::traverse_gpu)"}
    # the scene (a few MB) is served from L2 / Infinity Cache: the same figure against their
    # measured read bandwidth
    bw = read_bandwidths(st)
    res["cache_rooflines"] = {k: {"read_gbs": v, "frac": achieved / v} for k, v in bw.items()}
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(b, N.BUILDER_BINNED_SAH, 1, [prim, shadow], [False, True],
                                           budget_s=args.cpu_seconds)
    return res


def run_c4(args):
    """256 ray queues over 8 scenes (~40M incoherent closest-hit rays per step): one batched
    launch per step against one launch per queue."""
    from imbatracer_amd import native as N
    from imbatracer_amd import scene as S
    from imbatracer_amd import workloads as W

    sizes = W.C4_SCENE_TRIS
    builds, scenes, geometry = [], [], []
    t0 = time.time()
    for k, n in enumerate(sizes):
        b, v, i = W.random_tri_scene(n, N.BUILDER_SBVH if n <= 250_000 else N.BUILDER_BINNED_SAH, seed=0x1234 + k)
        builds.append(b)
        geometry.append((v, i))
        scenes.append(S.Scene(b.desc()))
    log("8 scenes built in %.1f s" % (time.time() - t0))
    q_sizes = W.c4_queue_sizes(256)
    if args.c4_rays:
        q_sizes = np.maximum(1, (q_sizes * (args.c4_rays / q_sizes.sum())).astype(np.int64))
    n_queues = len(q_sizes)
    queues = []
    t0 = time.time()
    for q in range(n_queues):
        n = int(q_sizes[q])
        v, i = geometry[q % 8]
        rays = W.c4_rays(v, i, n, coherent=q % 2 == 0, seed=5000 + q)
        queues.append((scenes[q % 8], N.DeviceBuffer.from_array(rays), N.DeviceBuffer(n * 16), n, None, False))
        if q < 2:
            queues[-1] += (rays,)
    per_queue = int(q_sizes.max())
    log("%d queues, %d rays (%d..%d per queue) generated and uploaded in %.1f s" % (
        n_queues, int(q_sizes.sum()), int(q_sizes.min()), int(q_sizes.max()), time.time() - t0))
    cfg = fixed_config(args)
    batch = S.TraceBatch([q[:6] for q in queues])
    st = N.Stream()
    if cfg is None and not args.no_autotune:
        # the batched launch over every queue, timed for each tree (quantized also with a 64-node
        # LDS treelet per wave) and a few shapes
        best, tried = None, []
        for width, quant, tl in ((2, 0, 0), (4, 0, 0), (4, 1, 0), (4, 1, 64)):
            for block, stack in ((256, 16), (256, 8), (128, 32)):
                c = N.TraceConfig(block, stack, 0, 0, width, quant, tl)
                batch.launch(c, st)
                e0, e1 = N.Event(), N.Event()
                e0.record(st)
                batch.launch(c, st)
                e1.record(st)
                st.sync()
                ms = e0.elapsed_ms(e1)
                tried.append("%d/%d/w%d%s%s %.2f" % (block, stack, width, "q" if quant else "",
                                                     "/tl%d" % tl if tl else "", ms))
                if best is None or ms < best[0]:
                    best = (ms, c)
        cfg = best[1]
        log("batch autotune: %s, %.2f ms (tried, ms: %s)" % (cfg_dict(cfg), best[0], ", ".join(tried)))
    # visit counts per queue (untimed)
    cb = N.DeviceBuffer(per_queue * 12)
    bytes_step, own_bytes_step = 0.0, 0.0
    node_bytes = node_bytes_of(cfg)
    bvh2 = N.TraceConfig(*BVH2_CFG_ARGS)
    for sc, rb, hb, n, _, _ in [q[:6] for q in queues]:
        sc.trace_stats(rb, hb, cb, n, cfg=cfg)
        own_bytes_step += ray_bytes(cb.download(np.uint32, 3 * n).reshape(-1, 3).astype(np.int64).sum(0), n,
                                    node_bytes)
        sc.trace_stats(rb, hb, cb, n, cfg=bvh2)
        bytes_step += ray_bytes(cb.download(np.uint32, 3 * n).reshape(-1, 3).astype(np.int64).sum(0), n, 64)

    def per_queue_launches():
        for sc, rb, hb, n, _, _ in [q[:6] for q in queues]:
            sc.intersect(rb, hb, n, cfg, st.handle)

    # parity: the batch's hits equal the per-queue launches' hits
    per_queue_launches()
    st.sync()
    ref = [q[2].download(N.HIT_DTYPE, q[3]) for q in queues]
    batch.launch(cfg, st)
    st.sync()
    for q, r in zip(queues, ref):
        if not np.array_equal(q[2].download(N.HIT_DTYPE, q[3]).view(np.int32), r.view(np.int32)):
            raise RuntimeError("batched hits differ from per-queue hits")

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        st.sync()
        ev = []
        t = time.perf_counter()
        for _ in range(args.steps):
            e0, e1 = N.Event(), N.Event()
            e0.record(st)
            fn()
            e1.record(st)
            ev.append((e0, e1))
        st.sync()
        return time.perf_counter() - t, sum(a.elapsed_ms(b) for a, b in ev)

    wall_b, gpu_b = timed(lambda: batch.launch(cfg, st))
    wall_q, gpu_q = timed(per_queue_launches)
    rays_step = int(q_sizes.sum())
    achieved = bytes_step * args.steps / (gpu_b * 1e-3) / 1e9
    own_achieved = own_bytes_step * args.steps / (gpu_b * 1e-3) / 1e9
    c4_traffic, c4_traffic_src = measured_traffic("c4-256-queues-%d-rays-8-scenes" % rays_step, "k_trace_batch")
    res = {
        "metric": METRIC, "value": rays_step * args.steps / wall_b / 1e6, "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall_b * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded random triangles and incoherent rays)",
        "config": {"workload": "c4-256-queues-%d-rays-8-scenes" % rays_step, "queues": n_queues,
                   "rays_per_queue": "log-uniform in [1K, 1M]; half pinhole, half cosine-hemisphere from surfaces",
                   "scene_triangles": sizes,
                   "trace": "one batched launch per step (imba_trace_batch_launch)", "trace_config": cfg_dict(cfg)},
        "per_queue_launches": {"value": rays_step * args.steps / wall_q / 1e6, "ms_per_step": wall_q * 1e3 / args.steps,
                               "gpu_ms_per_step": gpu_q / args.steps},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": c4_traffic, "traffic_source": c4_traffic_src,
                     "kernel": "k_trace_batch (closest hit)",
                     "algorithmic_bytes_per_launch": bytes_step, "bytes_per_ray": bytes_step / rays_step,
                     "bytes_rule": BYTES_RULE,
                     "own_tree": {"achieved": own_achieved, "frac": own_achieved / HBM_PEAK_GBS,
                                  "bytes_per_ray": own_bytes_step / rays_step}},
    }
    if args.c4_bounces > 0:
        res.update(run_c4_loop(args, builds, scenes, queues, cfg, rays_step))
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(builds[0], N.BUILDER_BINNED_SAH, 1, [queues[0][6]], [False],
                                           budget_s=args.cpu_seconds)
    return res


def run_c4_loop(args, builds, scenes, queues, cfg, rays_first):
    """The same 256 queues through a Lambertian bounce loop of args.c4_bounces bounces: per bounce
    one batched traversal and one batched bounce + compaction of every queue, the whole loop one
    graph launch per step; against the same loop launched eagerly per queue."""
    from imbatracer_amd import native as N
    from imbatracer_amd import workloads as W

    shading = []
    for b in builds:
        normals, mats = b.tri_attrs()
        n_mat = int(mats.max()) + 1 if len(mats) else 1
        shading.append(W.PtShadingData(normals, mats, np.full((n_mat, 3), 0.5, np.float32),
                                       np.zeros((n_mat, 3), np.float32), np.zeros(n_mat, np.float32),
                                       np.zeros((0, 12), np.float32)))
    loop = W.QueueLoop(scenes, shading, [(k % 8, q[1], q[3]) for k, q in enumerate(queues)],
                       max_path_len=args.c4_bounces, rr_start=3, cfg=cfg)
    loop.build_graph()
    st = loop.stream
    loop.step(graph=False)
    eager_counts = loop.rays_per_bounce()
    loop.step(graph=True)
    graph_counts = loop.rays_per_bounce()
    if not np.array_equal(eager_counts, graph_counts):
        raise RuntimeError("graph step and eager step traverse different queues: %s vs %s" % (graph_counts,
                                                                                               eager_counts))
    if int(graph_counts[0]) != rays_first:
        raise RuntimeError("the loop's first bounce does not hold the queues' rays")
    rays_step = int(graph_counts.sum())
    log("c4 loop: rays per bounce %s" % graph_counts.tolist())

    def timed(graph):
        for _ in range(args.warmup):
            loop.step(graph)
        st.sync()
        t = time.perf_counter()
        for _ in range(args.steps):
            loop.step(graph)
        st.sync()
        return time.perf_counter() - t

    wall_g = timed(True)
    wall_e = timed(False)

    # the plain path: every queue on its own, launch by launch (eager, one step)
    L = args.c4_bounces
    ws = N.DeviceBuffer(int(N.lib.imba_compact_workspace_size(max(q["n"] for q in loop.q))))
    N.check(N.lib.imba_memcpy_d2d(loop.counts.ptr, loop.counts0.ptr, 4 * loop.n_counts, st.handle), "counts")
    st.sync()
    t = time.perf_counter()
    for k, e in enumerate(loop.q):
        n = e["n"]
        if n == 0:
            continue
        sh = loop.sh[e["scene"]]
        N.check(N.lib.imba_memcpy_d2d(e["rays"][0].ptr, e["rays0"].ptr, 32 * n, st.handle), "rays")
        N.check(N.lib.imba_memcpy_d2d(e["states"][0].ptr, e["states0"].ptr, W.PT_STATE_BYTES * n, st.handle),
                "states")
        for b in range(L):
            cur, nxt = b % 2, (b + 1) % 2
            cnt = loop._cnt(k, b)
            N.check(N.lib.imba_intersect_dev(scenes[e["scene"]].h, e["rays"][cur].ptr, e["hits"].ptr, n, cnt,
                                             C.byref(cfg) if cfg else None, st.handle), "intersect_dev")
            N.check(N.lib.imba_pt_shade(C.byref(sh), e["rays"][cur].ptr, e["hits"].ptr, e["states"][cur].ptr, n, cnt,
                                        None, e["rays"][cur].ptr, None, None, st.handle), "pt_shade")
            N.check(N.lib.imba_compact_rays_dev(e["rays"][cur].ptr, e["states"][cur].ptr, W.PT_STATE_BYTES, n, cnt,
                                                e["rays"][nxt].ptr, e["states"][nxt].ptr, loop._cnt(k, b + 1),
                                                ws.ptr, st.handle), "compact_dev")
    st.sync()
    wall_p = time.perf_counter() - t
    if not np.array_equal(loop.rays_per_bounce(), graph_counts):
        raise RuntimeError("per-queue launches traverse different queues than the batched loop")
    return {
        "bounce_loop": {"bounces": L, "rays_per_step": rays_step, "rays_per_bounce": graph_counts.tolist(),
                        "graph": {"value": rays_step * args.steps / wall_g / 1e6,
                                  "ms_per_step": wall_g * 1e3 / args.steps},
                        "eager_batched": {"value": rays_step * args.steps / wall_e / 1e6,
                                          "ms_per_step": wall_e * 1e3 / args.steps},
                        "per_queue_launches": {"value": rays_step / wall_p / 1e6, "ms_per_step": wall_p * 1e3,
                                               "steps": 1},
                        "step": "one graph launch: per bounce imba_trace_batch_launch + imba_shade_batch_launch "
                                "over all 256 queues (Lambertian, albedo 0.5, no lights, Russian roulette from "
                                "bounce 3)"},
    }


def run_c5(args, big_scene=None, saturated: bool = True):
    """Serving: Poisson arrivals from 16 sessions over 8 resident sphere-grid scenes (20K to 21M
    triangles), random cameras, resolutions {256^2, 640x480, 1280x720, 1920x1080}, 1-4 spp, PT or
    VCM (photon mapping, the PPM mode of VCM, on request);
    the request loop with cross-session batching against the plain per-request path on the
    same arrivals.  big_scene: the 32 x 32 sphere grid when already built."""
    from imbatracer_amd import native as N
    from imbatracer_amd import serve as V
    from imbatracer_amd import workloads as W

    threads = host_cores()
    grids = [(1, 5), (2, 5), (4, 5), (6, 5), (8, 5), (12, 5), (20, 5), (32, 5)]
    if args.c5_small:
        grids = [(1, 3), (2, 3), (2, 4), (3, 4), (4, 4), (4, 5), (6, 5), (8, 5)]
    t0 = time.time()
    scenes, centres, tris = [], [], []
    for side, level in grids:
        reuse = big_scene is not None and side == big_scene.side and level == 5
        sg = big_scene if reuse else W.SphereGridScene(side=side, level=level, threads=threads)
        scenes.append(V.ServedScene(sg))
        centres.append(W.served_view(side))
        tris.append(sg.tri_count)
    log("8 scenes (%s triangles) built in %.1f s" % (tris, time.time() - t0))
    integrators = tuple(args.c5_integrators.split(","))
    reqs = V.poisson_requests(args.c5_requests, args.c5_rate, 16, len(scenes), centres, seed=11,
                              integrators=integrators, ceiling=W.SERVED_CEILING)
    cfg = fixed_config(args)
    arena = args.c5_arena_gb << 30

    def compare(request_list, label):
        out = {}
        for name, batched in (("batched", True), ("plain", False)):
            server = V.Server(scenes, arena, max_active=args.c5_max_active, batched=batched, cfg=cfg,
                              max_rays_in_flight=args.c5_max_rays)
            # warm-up on a few requests (kernels loaded, batch tables sized)
            server.run(V.poisson_requests(6, 1000.0, 1, len(scenes), centres, seed=99, integrators=integrators,
                                          ceiling=W.SERVED_CEILING))
            server = V.Server(scenes, arena, max_active=args.c5_max_active, batched=batched, cfg=cfg,
                              max_rays_in_flight=args.c5_max_rays)
            t = time.perf_counter()
            done = server.run([V.Request(**{**r.__dict__}) for r in request_list], progress=log)
            wall = time.perf_counter() - t
            st = V.latency_stats(done)
            st["steps"] = server.steps
            st["wall_s"] = wall
            st["gpu_busy_fraction"] = server.gpu_ms * 1e-3 / wall
            out[name] = st
            log("%s %s: %s" % (label, name, st))
        return out

    results = compare(reqs, "poisson")
    b, p = results["batched"], results["plain"]
    # capacity: the same requests all waiting at the start (the loop is never idle), and a mix
    # of small requests (256x256, 1 spp) that leave the GPU underfilled when run one at a time
    sat = compare([V.Request(**{**r.__dict__, "arrival": 0.0}) for r in reqs], "saturated") if saturated else None
    small = [V.Request(**{**r.__dict__, "width": 256, "height": 256, "spp": 1, "arrival": 0.0}) for r in reqs]
    sm = compare(small, "small")

    def ratio(d):
        return {"throughput": d["batched"]["mrays_s"] / d["plain"]["mrays_s"],
                "requests_per_s": d["batched"]["requests_per_s"] / d["plain"]["requests_per_s"],
                "p99": d["batched"]["p99_ms"] / d["plain"]["p99_ms"]}
    res = {
        "metric": METRIC, "value": b["mrays_s"], "unit": "Mrays/s", "n_gpus": 1,
        "steps": len(reqs), "warmup": 4, "ms_per_step": b["wall_s"] * 1e3 / len(reqs),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded procedural scenes, Poisson arrivals, random cameras)",
        "config": {"workload": "c5-serving-%d-requests-%.0f-per-s-16-sessions-8-scenes" % (len(reqs), args.c5_rate),
                   "scene_triangles": tris, "integrators": list(integrators),
                   "requests_per_integrator": {i: sum(r.integrator == i for r in reqs) for i in integrators},
                   "arena_gb": args.c5_arena_gb,
                   "max_active": args.c5_max_active, "max_rays_in_flight": args.c5_max_rays, "own_launch_rays": V.Server.OWN_LAUNCH_RAYS,
                   "resolutions": ["256x256", "640x480", "1280x720", "1920x1080"], "spp": "1-4"},
        "served": b, "plain_per_request": p,
        "served_vs_plain": {"throughput": b["mrays_s"] / p["mrays_s"], "p99": b["p99_ms"] / p["p99_ms"]},
        "saturated": {"served": sat["batched"], "plain_per_request": sat["plain"], "served_vs_plain": ratio(sat),
                      "arrivals": "all %d requests waiting at the start" % len(reqs)} if sat else None,
        "small_requests": {"served": sm["batched"], "plain_per_request": sm["plain"], "served_vs_plain": ratio(sm),
                           "arrivals": "the same %d requests at 256x256, 1 spp, all waiting at the start" % len(reqs)},
        "frame_latency_ms": {"p50": b["p50_ms"], "p99": b["p99_ms"]},
    }
    return res


def grid_candidates(info, qpos, limit=None):
    """Photons visited per query by HashGrid::query (rangesearch.h:77-129): the sizes of the 8
    probed buckets (a bucket probed twice counts twice); 0 for queries outside the bounds.
    With `limit` (per query: the squared distance of its k-th result when its list is full,
    else the squared radius) also the photons of the probed buckets whose photon box lies within
    that final limit: buckets the scan must read whatever the order (it skips a bucket only when
    the box lies beyond the limit at the bucket's turn, and the limit only shrinks), a lower
    bound of what the kernel reads."""
    from imbatracer_amd import native as N

    ends = np.empty(info.table_size, np.int32)
    N.check(N.lib.imba_memcpy_d2h(N.ptr(ends), C.c_void_p(info.cell_ends), ends.nbytes, None), "d2h")
    size = np.diff(np.concatenate([[0], ends])).astype(np.int64)
    bmin = np.array(info.bbox_min, np.float32)
    bmax = np.array(info.bbox_max, np.float32)
    q = np.ascontiguousarray(qpos, np.float32)
    inside = np.all((q >= bmin) & (q <= bmax), axis=1)
    qi = q[inside]
    c = np.float32(info.inv_cell_size) * (qi - bmin)
    f = np.floor(c)
    p = f.astype(np.int64)
    o = p + np.where((c - f) < np.float32(0.5), -1, 1)
    total = np.zeros(len(qi), np.int64)
    needed = np.zeros(len(qi), np.int64)
    if limit is not None:
        pos = np.empty((info.photon_count, 4), np.float32)
        N.check(N.lib.imba_memcpy_d2h(N.ptr(pos), C.c_void_p(info.positions), pos.nbytes, None), "d2h")
        nonempty = np.nonzero(size)[0]
        starts = (ends[nonempty] - size[nonempty]).astype(np.int64)
        lo = np.full((info.table_size, 3), np.inf, np.float32)
        hi = np.full((info.table_size, 3), -np.inf, np.float32)
        if len(nonempty):
            lo[nonempty] = np.minimum.reduceat(pos[:, :3], starts, axis=0)
            hi[nonempty] = np.maximum.reduceat(pos[:, :3], starts, axis=0)
        lim = np.asarray(limit, np.float32)[inside]
    for j in range(8):
        x = o[:, 0] if j & 4 else p[:, 0]
        y = o[:, 1] if j & 2 else p[:, 1]
        z = o[:, 2] if j & 1 else p[:, 2]
        h = ((x.astype(np.uint32) * np.uint32(73856093)) ^ (y.astype(np.uint32) * np.uint32(19349663))
             ^ (z.astype(np.uint32) * np.uint32(83492791))) % np.uint32(info.table_size)
        h = h.astype(np.int64)
        total += size[h]
        if limit is not None:
            with np.errstate(invalid="ignore"):
                g = np.maximum(np.maximum(lo[h] - qi, qi - hi[h]), np.float32(0))
                lb = (g[:, 0] * g[:, 0] + g[:, 1] * g[:, 1]) + g[:, 2] * g[:, 2]
            needed += np.where(lb <= lim, size[h], 0)
    out = np.zeros(len(qpos), np.int64)
    out[inside] = total
    if limit is None:
        return out
    out_needed = np.zeros(len(qpos), np.int64)
    out_needed[inside] = needed
    return out, out_needed


def run_c3_ppm(args):
    """Photon mapping (the PPM mode of the VCM integrator) at 1280x720 over the sphere-grid scene
    with 30% mirror and glass spheres: 921,600 light paths of up to 6 hits deposit photons, a
    hash grid is built over them (r = 0.005 x the scene diagonal), one merge query per pixel
    finds its 10 nearest photons.  One step = one frame (light pass, photon compaction, grid
    build, camera pass, queries, merging); value = closest-hit rays traversed per second of
    frame time."""
    import oracle
    from imbatracer_amd import native as N
    from imbatracer_amd import workloads as W

    t0 = time.time()
    threads = host_cores()
    scene = W.SphereGridScene(side=args.side, level=5, builder=N.BUILDER_SBVH, threads=threads,
                              specular_fraction=0.3)
    log("scene: %d triangles, built in %.1f s" % (scene.tri_count, time.time() - t0))
    cfg = fixed_config(args)
    pm = W.PhotonMapper(scene, args.width, args.height, light_paths=args.c3_paths, max_depth=6, rr_start=3,
                        radius_factor=0.005, k=10, trace_cfg=cfg, merge="fused")
    if cfg is None and not args.no_autotune:
        lr, ls = N.DeviceBuffer(pm.P * 32), N.DeviceBuffer(pm.P * 32)
        N.check(N.lib.imba_ppm_light(N.ptr(pm.light_tri), N.ptr(pm.light_int), pm.light_seed, 0, pm.P, lr.ptr,
                                     ls.ptr, None), "ppm_light")
        hb = N.DeviceBuffer(pm.P * 16)
        cfg, ms = pm.gpu_scene.autotune(lr, hb, pm.P, reps=3)
        log("autotune (light rays, %d): %s, %.3f ms" % (pm.P, cfg_dict(cfg), ms))
        del lr, ls, hb
        # the winner on the first light queue and the plain quantized launches (with and
        # without a 64-node treelet), each timed over whole frames (light and camera passes,
        # where the traversal runs); the fastest wins
        best, seen = None, set()
        for c in (cfg, N.TraceConfig(256, 16, 0, 0, 4, 1, 0), N.TraceConfig(256, 16, 0, 0, 4, 1, 64),
                  N.TraceConfig(256, 8, 0, 0, 4, 1, 0)):
            key = tuple(getattr(c, f) for f, _ in c._fields_)
            if key in seen:
                continue
            seen.add(key)
            pm.cfg = c
            pm.frame()
            t = 0.0
            for _ in range(2):
                pm.frame(timed=True)
                ph = pm.phase_ms()
                t += ph["light"] + ph["camera"]
            log("frame tune: %s %.3f ms of light + camera passes per frame" % (cfg_dict(c), t / 2))
            if best is None or t < best[0]:
                best = (t, c)
        cfg = best[1]
        pm.cfg = cfg
    light, cam, n_ph = pm.frame()
    log("light queues %s, camera queues %s, photons %d" % (light, cam, n_ph))
    for _ in range(args.warmup):
        pm.frame()
    phases, frame_ms, rays_total = [], [], 0
    t_start = time.perf_counter()
    for _ in range(args.steps):
        f0 = time.perf_counter()
        l, c, n = pm.frame(timed=True)
        frame_ms.append((time.perf_counter() - f0) * 1e3)
        if (l, c, n) != (light, cam, n_ph):
            raise RuntimeError("frame queue sizes changed between frames")
        rays_total += sum(l) + sum(c)
        phases.append(pm.phase_ms())
    wall = time.perf_counter() - t_start
    if pm.gpu_scene.error_count():
        raise RuntimeError("traversal reported a malformed scene")
    phase = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}
    log("phases (ms): %s" % {k: round(v, 3) for k, v in phase.items()})

    # parity sample: the grid over the GPU's photons and a sample of the queries, against the oracle
    # (the frame's fused merge keeps its lists in the waves: the index query writes them, untimed)
    ph_rays, _ = pm.photons_host()
    pos = np.ascontiguousarray(ph_rays["pos"])
    q = pm.queries_host()
    qpos = np.ascontiguousarray(q["pos"])
    pm.query()
    idx, dist, cnt = pm.results_host()
    og = oracle.Grid()
    tb = time.perf_counter()
    og.build(pos, pm.radius)
    cpu_build_s = time.perf_counter() - tb
    valid = np.nonzero(q["pixel"] >= 0)[0]
    rng = np.random.default_rng(3)
    sample = np.sort(rng.choice(valid, size=min(len(valid), args.c3_check), replace=False))
    qs = np.ascontiguousarray(qpos[sample])
    tq = time.perf_counter()
    o_idx, o_dist, o_cnt = og.query_many(qs, pm.k)
    cpu_query_s = time.perf_counter() - tq
    bad = 0
    for r, j in enumerate(sample):
        c = cnt[j]
        if c != o_cnt[r] or not np.array_equal(idx[j, :c], o_idx[r, :c]) or \
                not np.array_equal(dist[j, :c].view(np.int32), o_dist[r, :c].view(np.int32)):
            bad += 1
    if bad:
        raise RuntimeError("%d of %d sampled photon queries differ from the oracle grid" % (bad, len(sample)))
    info = pm.grid.info()
    found = cnt.astype(np.int64)
    r2 = np.float32(pm.radius) * np.float32(pm.radius)
    lim = np.where(found == pm.k, dist[:, pm.k - 1], r2)
    cand, needed = grid_candidates(info, qpos, lim)
    # fused query + merge: 16 B of every query record, 8 x 8 B bucket ends and 8 x 24 B bucket
    # photon boxes per query inside the grid bounds, 16 B per photon of a bucket the scan must
    # read (slot-ordered position), and for a query with results the rest of its record (48 B),
    # 32 B per result (slot-ordered direction and state) and 12 B of image
    inside = np.isfinite(qpos[:, 0]) & (q["pixel"] >= 0)
    q_bytes = float(np.sum(16 + np.where(inside, 8 * 8 + 8 * 24, 0) + needed * 16 +
                           np.where(found > 0, 48 + 12, 0) + found * 32))
    q_ms = phase["merge"]
    achieved = q_bytes / (q_ms * 1e-3) / 1e9
    n_q = pm.n_cam
    # build: bounds 12 B, count 12 + 4 B, scan 2 x 4 B per bucket, scatter 16 + 4 + 4 B per photon
    b_bytes = float(n_ph * 12 + n_ph * 16 + 2 * info.table_size * 4 + n_ph * 24)
    workload = "c3-ppm-%dx%d-%d-light-paths-depth6-sphere-grid-%dtri-30pct-specular" % (
        args.width, args.height, pm.P, scene.tri_count)
    q_traffic, q_traffic_src = measured_traffic(workload, "k_merge_grid")
    res = {
        "metric": METRIC, "value": rays_total / wall / 1e6, "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded procedural scene, light and camera)",
        "config": {"workload": workload, "width": args.width, "height": args.height, "light_paths": pm.P,
                   "max_depth": 6, "rr_start": 3, "k": pm.k, "radius": pm.radius, "photons": n_ph,
                   "light_rays_per_bounce": light, "camera_rays_per_bounce": cam,
                   "trace_config": cfg_dict(cfg),
                   "frame": "eager launches with device-side queue sizes; one host read (photon count) per frame; "
                            "grid query and merge fused in one launch"},
        "phase_ms": phase,
        "roofline": None,
        "frame_latency_ms": {"p50": float(np.percentile(frame_ms, 50)), "p99": float(np.percentile(frame_ms, 99))},
        "grid": {"photons": n_ph, "table_size": info.table_size, "build_ms": phase["grid_build"],
                 "queries": n_q, "query_merge_ms": q_ms, "mqueries_s": n_q / (q_ms * 1e-3) / 1e6,
                 "photons_visited_per_query": float(cand.mean()),
                 "photons_read_per_query_lower_bound": float(needed.mean()),
                 "results_per_query": float(found.mean()),
                 "parity": {"queries_checked": int(len(sample)), "mismatches": bad,
                            "rule": "photon index lists and squared distances bit-exact against the oracle grid "
                                    "built over the same photons"},
                 "query_roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                    "frac": achieved / HBM_PEAK_GBS,
                                    "kernel": "k_merge_grid (fused grid query + merge)",
                                    "traffic": q_traffic, "traffic_source": q_traffic_src,
                                    "algorithmic_bytes_per_launch": q_bytes,
                                    "bytes_rule": "16 per query record; 8x8 (bucket ends) + 8x24 (bucket photon "
                                                  "boxes) per query inside the bounds; 16 per photon of a bucket "
                                                  "the scan must read (its box within the query's final limit; "
                                                  "slot-ordered position); per query with results 48 (rest of "
                                                  "its record) + 12 (image) + 32 per result (slot-ordered "
                                                  "direction and state)"},
                 "build_algorithmic_bytes": b_bytes},
        "cpu_grid": {"build_ms": cpu_build_s * 1e3, "mqueries_s": len(sample) / cpu_query_s / 1e6, "cores": 1,
                     "kind": "port", "sample": "oracle/hashgrid.c (serial restatement of HashGrid) building over "
                                               "the frame's %d photons and answering %d sampled queries "
                                               "(k = 10, one C loop)" % (n_ph, len(sample))},
    }
    res["roofline"] = res["grid"]["query_roofline"]
    if not args.no_cpu:
        lr, ls = N.DeviceBuffer(pm.P * 32), N.DeviceBuffer(pm.P * 32)
        N.check(N.lib.imba_ppm_light(N.ptr(pm.light_tri), N.ptr(pm.light_int), pm.light_seed, 0, pm.P, lr.ptr,
                                     ls.ptr, None), "ppm_light")
        light_rays = lr.download(N.RAY_DTYPE, pm.P)
        res["cpu_baseline"] = cpu_baseline(scene.build, N.BUILDER_SBVH, threads, [light_rays], [False],
                                           budget_s=args.cpu_seconds)
    return res


def run_c3(args):
    """One VCM iteration (VCMIntegrator<ALGO_VCM>::render, render/integrators/vcm.cpp:24-56) at
    1280x720 over the sphere-grid scene with 30% mirror and glass spheres: 921,600 light paths of
    up to 6 hits store light vertices and connect them to the camera, the hash grid is built over
    the cache (r = 0.005 x the scene diagonal), camera paths of up to 6 vertices add emitter hits,
    next-event estimation, one connection to a cached light vertex and a k = 10 merge at every
    Lambertian vertex.  One step = one iteration; value = closest-hit rays (light + camera) per
    second of frame time."""
    import oracle
    from imbatracer_amd import native as N
    from imbatracer_amd import vcm as VC
    from imbatracer_amd import workloads as W

    t0 = time.time()
    threads = host_cores()
    scene = W.SphereGridScene(side=args.side, level=5, builder=N.BUILDER_SBVH, threads=threads,
                              specular_fraction=0.3)
    log("scene: %d triangles, built in %.1f s" % (scene.tri_count, time.time() - t0))
    cfg = fixed_config(args)
    vr = VC.VcmRenderer(scene, args.width, args.height, light_paths=args.c3_paths, max_path_len=6, light_max_depth=6,
                        rr_start=0, radius_factor=0.005, k=10, trace_cfg=cfg,
                        shadow_cfg=N.TraceConfig(256, 16, 0, 0, 4, 1, 0))
    if cfg is None and not args.no_autotune:
        best = None
        for c in (N.TraceConfig(256, 16, 0, 0, 4, 1, 0), N.TraceConfig(256, 16, 0, 0, 4, 1, 64),
                  N.TraceConfig(256, 8, 0, 0, 4, 1, 0), N.TraceConfig(128, 16, 0, 0, 4, 1, 0)):
            vr.cfg = c
            vr.frame()
            t = 0.0
            for _ in range(2):
                vr.frame(timed=True)
                ph = vr.phase_ms()
                t += ph["light"] + ph["camera"]
            log("frame tune: %s %.3f ms of light + camera passes per frame" % (cfg_dict(c), t / 2))
            if best is None or t < best[0]:
                best = (t, c)
        cfg = best[1]
        vr.cfg = cfg
    light, cam, nv = vr.frame()
    log("light queues %s, camera queues %s, light vertices %d" % (light, cam, nv))
    for _ in range(args.warmup):
        vr.frame()
    # phases of one iteration at a time
    phases = []
    for _ in range(3):
        vr.frame(timed=True)
        phases.append(vr.phase_ms())
    phase = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}
    log("phases (ms): %s" % {k: round(v, 3) for k, v in phase.items()})
    # the measured iterations: `frames_in_flight` renderers (own buffers and streams, one scene)
    # take turns, one's light pass overlapping the other's camera pass
    loops = [vr] + [VC.VcmRenderer(scene, args.width, args.height, light_paths=args.c3_paths, max_path_len=6,
                                   light_max_depth=6, rr_start=0, radius_factor=0.005, k=10, trace_cfg=cfg,
                                   shadow_cfg=vr.shadow_cfg, gpu_scene=vr.gpu_scene)
                    for _ in range(args.frames_in_flight - 1)]
    for lp in loops[1:]:
        for _ in range(max(1, args.warmup)):
            lp.frame()
    frame_ms, rays_total = [], 0
    pending = {}

    def done(lp):
        l, c, n = lp.finish()
        frame_ms.append((time.perf_counter() - pending.pop(id(lp))) * 1e3)
        if (l, c, n) != (light, cam, nv):
            raise RuntimeError("frame queue sizes changed between frames")
        return sum(l) + sum(c)

    t_start = time.perf_counter()
    loops[0].launch_light()
    pending[id(loops[0])] = time.perf_counter()
    for step in range(args.steps):
        lp, other = loops[step % len(loops)], loops[(step + 1) % len(loops)]
        if step + 1 < args.steps and other is not lp:
            if id(other) in pending:
                rays_total += done(other)
            other.launch_light()
            pending[id(other)] = time.perf_counter()
        lp.launch_camera()
        if len(loops) == 1:
            rays_total += done(lp)
            if step + 1 < args.steps:
                lp.launch_light()
                pending[id(lp)] = time.perf_counter()
    for lp in loops:
        if id(lp) in pending:
            rays_total += done(lp)
    wall = time.perf_counter() - t_start
    loops = loops[:1]
    if vr.gpu_scene.error_count():
        raise RuntimeError("traversal reported a malformed scene")
    # kernel time per kind of launch (events around every launch, separate frames)
    vr.time_kernels = True
    kms = []
    for _ in range(3):
        vr.frame()
        kms.append(vr.kernel_ms())
    vr.time_kernels = False
    kernel_ms = {k: float(np.mean([d[k] for d in kms])) for k in kms[0]}
    log("kernel ms per frame: %s" % {k: round(v, 3) for k, v in kernel_ms.items()})

    # one more iteration with every camera bounce's merge queries read back: a query sample
    # against the oracle grid over the frame's light vertex cache, and the merge launches'
    # algorithmic bytes
    vr.keep_queues = True
    vr.frame()
    vr.keep_queues = False
    a, _ = vr.cache_host()
    og = oracle.Grid()
    tb = time.perf_counter()
    og.build(np.ascontiguousarray(a["pos"]), vr.radius)
    cpu_build_s = time.perf_counter() - tb
    info = vr.grid.info()
    qall = np.concatenate([k["queries"] for k in vr.kept["camera"]])
    vr.kept = {"light": [], "camera": []}
    valid = np.nonzero(qall["pixel"] >= 0)[0]
    rng = np.random.default_rng(3)
    sample = np.sort(rng.choice(valid, size=min(len(valid), args.c3_check), replace=False))
    qs = np.ascontiguousarray(qall["pos"][sample])
    qb = N.DeviceBuffer.from_array(qs)
    oi, od, oc = N.DeviceBuffer(len(qs) * 40), N.DeviceBuffer(len(qs) * 40), N.DeviceBuffer(len(qs) * 4)
    vr.grid.query(qb, 12, 0, len(qs), vr.k, oi, oc, od)
    g_idx = oi.download(np.int32, len(qs) * vr.k).reshape(-1, vr.k)
    g_dist = od.download(np.float32, len(qs) * vr.k).reshape(-1, vr.k)
    g_cnt = oc.download(np.int32, len(qs))
    tq = time.perf_counter()
    o_idx, o_dist, o_cnt = og.query_many(qs, vr.k)
    cpu_query_s = time.perf_counter() - tq
    bad = 0
    for r in range(len(qs)):
        c = g_cnt[r]
        if c != o_cnt[r] or not np.array_equal(g_idx[r, :c], o_idx[r, :c]) or \
                not np.array_equal(g_dist[r, :c].view(np.int32), o_dist[r, :c].view(np.int32)):
            bad += 1
    if bad:
        raise RuntimeError("%d of %d sampled merge queries differ from the oracle grid" % (bad, len(qs)))
    # every valid query's list on the device (its length and k-th distance bound the scan)
    qv = np.ascontiguousarray(qall["pos"][valid])
    qb = N.DeviceBuffer.from_array(qv)
    oi, od, oc = N.DeviceBuffer(len(qv) * 4 * vr.k), N.DeviceBuffer(len(qv) * 4 * vr.k), N.DeviceBuffer(len(qv) * 4)
    vr.grid.query(qb, 12, 0, len(qv), vr.k, oi, oc, od)
    found = oc.download(np.int32, len(qv)).astype(np.int64)
    kd = od.download(np.float32, len(qv) * vr.k).reshape(-1, vr.k)[:, vr.k - 1]
    r2 = np.float32(vr.radius) * np.float32(vr.radius)
    cand, needed = grid_candidates(info, qv, np.where(found == vr.k, kd, r2))
    # fused query + merge, per valid query: 64 B query record, 8 x 8 B bucket ends, 8 x 24 B
    # bucket photon boxes, 16 B per photon of a bucket the scan must read (slot-ordered
    # position), 32 B merge record per result, 12 B image add; 16 B per empty query slot
    m_bytes = float(np.sum(64 + 8 * 8 + 8 * 24 + needed * 16 + found * 32 + 12) + (len(qall) - len(valid)) * 16)
    m_ms = kernel_ms["merge"]
    m_achieved = m_bytes / (m_ms * 1e-3) / 1e9
    workload = "c3-vcm-%dx%d-%d-light-paths-depth6-sphere-grid-%dtri-30pct-specular" % (
        args.width, args.height, vr.P, scene.tri_count)
    res = {
        "metric": METRIC, "value": rays_total / wall / 1e6, "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded procedural scene, light and camera)",
        "config": {"workload": workload, "width": args.width, "height": args.height, "light_paths": vr.P,
                   "light_max_depth": 6, "max_path_len": 6, "rr_start": 0, "k": vr.k, "num_connections": vr.C,
                   "radius": vr.radius, "light_vertices": nv,
                   "light_rays_per_bounce": light, "camera_rays_per_bounce": cam,
                   "trace_config": cfg_dict(cfg),
                   "frame": "eager launches with device-side queue sizes; one host read (cache size) per "
                            "iteration; %d iterations in flight on their own streams" % args.frames_in_flight,
                   "frames_in_flight": args.frames_in_flight},
        "phase_ms": phase,
        "kernel_ms_per_frame": kernel_ms,
        "frame_latency_ms": {"p50": float(np.percentile(frame_ms, 50)), "p99": float(np.percentile(frame_ms, 99))},
        "merge": {"queries": int(len(valid)), "ms_per_frame": m_ms, "photons_visited_per_query": float(cand.mean()) if len(cand) else 0.0,
                  "photons_read_per_query_lower_bound": float(needed.mean()) if len(needed) else 0.0,
                  "results_per_query": float(found.mean()) if len(found) else 0.0,
                  "algorithmic_bytes_per_frame": m_bytes,
                  "parity": {"queries_checked": int(len(qs)), "mismatches": bad,
                             "rule": "photon index lists and squared distances of the grid query bit-exact against "
                                     "the oracle grid built over the frame's light vertex cache"}},
        "cpu_grid": {"build_ms": cpu_build_s * 1e3, "mqueries_s": len(qs) / cpu_query_s / 1e6, "cores": 1,
                     "kind": "port", "sample": "oracle/hashgrid.c building over the frame's %d light vertices and "
                                               "answering %d sampled merge queries (k = 10, one C loop)"
                                               % (nv, len(qs))},
        "roofline": {"bound": "hbm", "achieved": m_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": m_achieved / HBM_PEAK_GBS, "kernel": "k_vcm_merge (fused grid query + merge)",
                     "traffic": measured_traffic(workload, "k_vcm_merge")[0],
                     "traffic_per": "iteration: the six camera bounces' merge launches together, as achieved",
                     "algorithmic_bytes_per_frame": m_bytes,
                     "bytes_rule": "per merge query 64 (record) + 8 x 8 (bucket ends) + 8 x 24 (bucket photon "
                                   "boxes) + 16 per light vertex of a bucket the scan must read (its box within "
                                   "the query's final limit; slot-ordered position) + 32 per result (slot-ordered "
                                   "merge record) + 12 (image); 16 per empty query slot"},
    }
    if not args.no_c3_modes:
        # the integrator's sub-algorithms on the same scene and launch shapes: one iteration at
        # a time (eager, host waits for each), closest-hit rays per iteration time
        del loops
        modes = {}
        for algo in ("bpt", "lt", "pt"):
            mr = VC.VcmRenderer(scene, args.width, args.height, light_paths=args.c3_paths, max_path_len=6,
                                light_max_depth=6, rr_start=0, radius_factor=0.005, k=10, trace_cfg=cfg,
                                shadow_cfg=vr.shadow_cfg, gpu_scene=vr.gpu_scene, algo=algo)
            mr.frame()
            t = []
            for _ in range(3):
                t0m = time.perf_counter()
                ml, mc, mv = mr.frame()
                t.append(time.perf_counter() - t0m)
            ms = float(np.median(t)) * 1e3
            modes[algo] = {"ms_per_iteration": ms, "mrays_s": (sum(ml) + sum(mc)) / ms / 1e3,
                           "light_rays": int(sum(ml)), "camera_rays": int(sum(mc)), "light_vertices": int(mv)}
            log("%s iteration: %.2f ms, %.1f Mrays/s" % (algo.upper(), ms, modes[algo]["mrays_s"]))
            del mr
        res["sub_algorithms"] = modes
    if not args.no_cpu:
        lr, ls = N.DeviceBuffer(vr.P * 32), N.DeviceBuffer(vr.P * 48)
        N.check(N.lib.imba_vcm_light(C.byref(vr.frame_args), vr.light_seed, 0, vr.P, lr.ptr, ls.ptr, None),
                "vcm_light")
        light_rays = lr.download(N.RAY_DTYPE, vr.P)
        res["cpu_baseline"] = cpu_baseline(scene.build, N.BUILDER_SBVH, threads, [light_rays], [False],
                                           budget_s=args.cpu_seconds)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="pt", choices=["pt", "c1", "c3", "c3-ppm", "c4", "c5"])
    ap.add_argument("--c5-requests", type=int, default=300, help="c5: requests in the run")
    ap.add_argument("--c5-rate", type=float, default=200.0, help="c5: mean arrivals per second")
    ap.add_argument("--c5-arena-gb", type=int, default=48, help="c5: device memory for request buffers")
    ap.add_argument("--c5-small", action="store_true", help="c5: eight small scenes (quick runs)")
    ap.add_argument("--c5-max-active", type=int, default=64, help="c5: requests in flight at once (batched loop)")
    ap.add_argument("--c5-max-rays", type=int, default=32 << 20,
                    help="c5: first-traversal rays of the requests in flight at once (batched loop; 0 = no limit)")
    ap.add_argument("--c5-integrators", default="pt,vcm",
                    help="c5: integrators drawn per request (pt: path tracing, vcm: vertex connection and merging, "
                         "ppm: photon mapping)")
    ap.add_argument("--c3-paths", type=int, default=921_600, help="c3: light paths per frame")
    ap.add_argument("--c3-check", type=int, default=20_000, help="c3: queries checked against the oracle grid")
    ap.add_argument("--no-c3-modes", action="store_true", help="c3: skip the BPT / LT / PT iterations")
    ap.add_argument("--c4-bounces", type=int, default=8, help="c4: bounces of the graph-captured loop (0: off)")
    ap.add_argument("--c4-rays", type=int, default=0,
                    help="c4: rescale the 256 log-uniform queue sizes to this many rays (0: as drawn, about 37M)")
    ap.add_argument("--width", type=int, default=1280)
    ap.add_argument("--height", type=int, default=720)
    ap.add_argument("--side", type=int, default=32, help="spheres per side of the grid")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--stack", type=int, default=0)
    ap.add_argument("--persistent", type=int, default=0)
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--frames-in-flight", type=int, default=2,
                    help="pt (graph mode) and c3: frame loops taking turns on their own streams")
    ap.add_argument("--mode", default="graph", choices=["graph", "eager"],
                    help="pt: frames as one captured graph (default) or as eager launches")
    ap.add_argument("--trace-kernel", default="plain", choices=["plain", "batch"],
                    help="pt graph frames: k_trace with a device count, or the batched queue kernel")
    ap.add_argument("--bvh-width", type=int, default=0, choices=[0, 2, 4],
                    help="tree the traversal walks: 2 = the scene's BVH2, 4 = the derived 4-wide tree (0 = default)")
    ap.add_argument("--quantized", action="store_true", help="walk the quantized 4-wide tree")
    ap.add_argument("--treelet", type=int, default=0,
                    help="quantized tree: nodes of its top levels each block stages in LDS (0..256)")
    ap.add_argument("--no-autotune", action="store_true", help="use the library's default launch configuration")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serving", action="store_true", help="pt: skip the served-vs-plain comparison")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (reference point for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    from imbatracer_amd import native as N

    if N.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    N.check(N.lib.imba_set_device(0), "set device")
    if args.calibrate_copy:
        nbytes = args.calibrate_copy << 20
        a, b = N.DeviceBuffer(nbytes), N.DeviceBuffer(nbytes)
        N.check(N.lib.imba_memcpy_d2d(b.ptr, a.ptr, nbytes, None), "copy")
        print(json.dumps({"calibrate_copy_bytes": nbytes}))
        return
    res = {"pt": run_pt, "c1": run_c1, "c3": run_c3, "c3-ppm": run_c3_ppm, "c4": run_c4,
           "c5": run_c5}[args.workload](args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
