"""Benchmark: incoherent closest-hit traversal throughput on one MI355X.

Default workload (the configuration the headline metric is quoted on): path tracing at
1280x720, 1 spp, maximum path length 8 with Russian roulette from bounce 3, over the
sphere-grid scene (32 x 32 level-5 icospheres + open box + triangle light, 20,971,544
triangles, SBVH per mesh).  One step = one frame: camera rays, then per bounce closest-hit
traversal, Lambertian bounce and queue compaction, all on the device.  `value` is the number of
closest-hit rays traversed per second of frame time (inputs resident in HBM; the scene build
and upload are outside the timed region).

--workload c1 runs the CPU-runnable configuration instead: 65,536 random triangles
(binned SAH), 1,048,576 primary closest-hit rays plus 1,048,576 occlusion rays per step.

Also reported: the roofline of the traversal kernel (algorithmic bytes per ray
32 + 16 + 64 N_node + 48 N_tri + 64 N_inst with visit counts measured by the stats variant of
the same kernel on the same rays, over the kernel time measured with HIP events on its stream)
and a CPU baseline (the C restatement of the traversal, oracle/trace.c, on host threads over a
bounded sample of the same rays).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s incoherent closest-hit; served p50/p99 frame latency vs plain GPU"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def ray_bytes(visits, rays):
    """Algorithmic bytes of `rays` rays with summed visit counts (nodes, tris, instances)."""
    nodes, tris, inst = (float(x) for x in visits)
    return 32.0 * rays + 16.0 * rays + 64.0 * nodes + 48.0 * tris + 64.0 * inst


def cpu_baseline(desc, ray_sets, any_flags, budget_s=12.0):
    """Times the oracle traversal on host threads over a bounded sample of the given rays."""
    import oracle

    threads = max(1, min(16, os.cpu_count() or 1))
    od = oracle.SceneDesc(*[getattr(desc, f) for f, _ in oracle.SceneDesc._fields_])
    L = oracle.lib()
    done, elapsed, passes = 0, 0.0, 0
    sets = [(np.ascontiguousarray(r), a) for r, a in zip(ray_sets, any_flags)]
    hits = np.empty(max(len(r) for r, _ in sets), oracle.HIT_DTYPE)
    # whole passes over the sample until the time budget is used
    while elapsed < budget_s and passes < 100:
        for rays, any_hit in sets:
            t0 = time.perf_counter()
            L.oracle_trace_mt(C.byref(od), C.c_void_p(rays.ctypes.data), len(rays), int(any_hit),
                              C.c_void_p(hits.ctypes.data), threads)
            elapsed += time.perf_counter() - t0
            done += len(rays)
        passes += 1
    return {"value": done / elapsed / 1e6 if elapsed > 0 else None, "unit": "Mrays/s", "cores": threads,
            "kind": "port",
            "sample": "%d passes over %s rays of the benchmarked queues (%d rays), traced by oracle/trace.c "
                      "(scalar, BVH2) on %d host threads in %.1f s" % (
                          passes, "+".join(str(len(r)) for r, _ in sets), done, threads, elapsed)}


def run_pt(args):
    from imbatracer_amd import native as N
    from imbatracer_amd import workloads as W

    t0 = time.time()
    threads = max(1, min(16, os.cpu_count() or 1))
    scene = W.SphereGridScene(side=args.side, level=5, builder=N.BUILDER_SBVH, threads=threads)
    log("scene: %d triangles, %d nodes, built in %.1f s" % (scene.tri_count, scene.desc.node_count, time.time() - t0))
    cfg = N.TraceConfig(args.block, args.stack, args.persistent, args.blocks_per_cu) if (args.block or args.persistent or args.blocks_per_cu) else None
    pt = W.PathTracer(scene, args.width, args.height, spp=1, max_path_len=8, rr_start=3, trace_cfg=cfg)
    log("uploaded %.2f GB" % (pt.gpu_scene.footprint() / 1e9))

    # untimed: visit counts of every bounce's queue (same seeds -> same rays as the timed frames)
    counts, visits, queues = pt.frame(stats=True, keep_queues=2)
    stat_bytes = [ray_bytes(v, n) for v, n in zip(visits, counts)]
    log("rays per bounce %s" % counts)

    for _ in range(args.warmup):
        pt.frame()
    pt.stream.sync()
    frame_ms, trace_ms, rays_total, bytes_total = [], 0.0, 0, 0.0
    t_start = time.perf_counter()
    for _ in range(args.steps):
        f0 = time.perf_counter()
        c, _, _ = pt.frame(timed=True)
        pt.stream.sync()
        frame_ms.append((time.perf_counter() - f0) * 1e3)
        if c != counts:
            raise RuntimeError("frame ray counts changed between frames: %s vs %s" % (c, counts))
        for (ms, n), b in zip(pt.trace_times_ms(), stat_bytes):
            trace_ms += ms
            bytes_total += b
        rays_total += sum(c)
    wall = time.perf_counter() - t_start
    if pt.gpu_scene.error_count():
        raise RuntimeError("traversal reported a malformed scene")
    achieved = bytes_total / (trace_ms * 1e-3) / 1e9
    res = {
        "metric": METRIC, "value": rays_total / wall / 1e6, "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded procedural scene and camera)",
        "config": {"workload": "pt-1280x720-1spp-depth8-rr3-sphere-grid-%dtri-sbvh" % scene.tri_count,
                   "width": args.width, "height": args.height, "spp": 1, "max_path_len": 8,
                   "rays_per_frame": int(sum(counts)), "rays_per_bounce": counts,
                   "triangles": scene.tri_count, "nodes": scene.desc.node_count,
                   "trace_shape": [args.block, args.stack] if args.block else "default"},
        "trace_mrays_s": rays_total / (trace_ms * 1e-3) / 1e6,
        "frame_latency_ms": {"p50": float(np.percentile(frame_ms, 50)), "p99": float(np.percentile(frame_ms, 99))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_trace (closest hit)",
                     "bytes_per_ray": bytes_total / max(rays_total, 1),
                     "visits_per_ray": [float(x) / sum(counts) for x in np.sum(visits, 0)]},
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(scene.desc, queues[:2], [False, False], budget_s=args.cpu_seconds)
    return res


def run_c1(args):
    from imbatracer_amd import native as N
    from imbatracer_amd import scene as S
    from tests import scenes as T

    v, i = S.random_triangles(65536, 0x1234)
    b = S.SceneBuild()
    b.add_mesh(v, i)
    b.run(N.BUILDER_BINNED_SAH, 1)
    desc = b.desc()
    sc = S.Scene(desc)
    prim = T.primary_rays(1024, 1024)
    n = len(prim)
    rb = N.DeviceBuffer.from_array(prim)
    hb = N.DeviceBuffer(n * 16)
    cb = N.DeviceBuffer(n * 12)
    cfg = N.TraceConfig(args.block, args.stack, args.persistent, args.blocks_per_cu) if (args.block or args.persistent or args.blocks_per_cu) else None
    sc.intersect_stats(rb, hb, cb, n)
    hits = hb.download(N.HIT_DTYPE, n)
    vis_p = cb.download(np.uint32, 3 * n).reshape(-1, 3).astype(np.int64).sum(0)
    shadow = T.shadow_rays(prim, hits)
    sb = N.DeviceBuffer.from_array(shadow)
    ob = N.DeviceBuffer(n * 16)
    sc.intersect_stats(sb, ob, cb, n, any_hit=True)
    vis_s = cb.download(np.uint32, 3 * n).reshape(-1, 3).astype(np.int64).sum(0)
    st = N.Stream()
    for _ in range(args.warmup):
        sc.intersect(rb, hb, n, cfg, st.handle)
        sc.occluded(sb, ob, n, cfg, st.handle)
    st.sync()
    ev = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e0, e1, e2 = N.Event(), N.Event(), N.Event()
        e0.record(st)
        sc.intersect(rb, hb, n, cfg, st.handle)
        e1.record(st)
        sc.occluded(sb, ob, n, cfg, st.handle)
        e2.record(st)
        ev.append((e0, e1, e2))
    st.sync()
    wall = time.perf_counter() - t0
    t_closest = sum(a.elapsed_ms(b) for a, b, c in ev)
    achieved = ray_bytes(vis_p, n) * args.steps / (t_closest * 1e-3) / 1e9
    res = {
        "metric": METRIC, "value": 2 * n * args.steps / wall / 1e6, "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded random triangles, pinhole camera, point light)",
        "config": {"workload": "c1-65536-random-tris-binned-sah-1M-primary-1M-occlusion"},
        "closest_mrays_s": n * args.steps / (t_closest * 1e-3) / 1e6,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_trace (closest hit)",
                     "bytes_per_ray": ray_bytes(vis_p, n) / n},
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(desc, [prim, shadow], [False, True], budget_s=args.cpu_seconds)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="pt", choices=["pt", "c1"])
    ap.add_argument("--width", type=int, default=1280)
    ap.add_argument("--height", type=int, default=720)
    ap.add_argument("--side", type=int, default=32, help="spheres per side of the grid")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--stack", type=int, default=0)
    ap.add_argument("--persistent", type=int, default=0)
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
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
    res = run_pt(args) if args.workload == "pt" else run_c1(args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
