"""Benchmark of the MI355X ray-cast + lightmap-gather path (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload bake|collider]

Prints ONE JSON line.  Workloads (synthetic, seeded):
  bake      plain lightmap bake: closed 10x10x4 m box room + 20 boxes (~50k
            triangles), 21 lightmaps of 256x256, 1024 hemisphere rays per texel
            plus 103 shadow-ray rounds to the point light; one step = the gather
            of every covered texel.  value = baked texels/s.
  collider  closest hit on a 100,820-triangle displaced icosphere, 1M random
            rays resident in HBM; one step = one batch.  value = Mrays/s.
Timing: W untimed steps, then K steps between device synchronisations; the
dominant kernel is also timed with HIP events on its own stream.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "closest-hit Mrays/s; baked texels/s at ≤2% RMSE, plain vs learned gather"


def _events():
    import ctypes
    from lightsprintsdk_amd import native as N
    L = N.lib()
    s = ctypes.c_void_p(); a = ctypes.c_void_p(); b = ctypes.c_void_p()
    N.check(L.lsg_stream_create(ctypes.byref(s)), "stream")
    N.check(L.lsg_event_create(ctypes.byref(a)), "event")
    N.check(L.lsg_event_create(ctypes.byref(b)), "event")
    return L, s, a, b


def collider_workload(steps, warmup, cpu_threads):
    import ctypes
    from lightsprintsdk_amd import native as N, scenes
    from lightsprintsdk_amd.collider import RRCollider, DeviceRays
    V, I = scenes.icosphere(scenes.ICOSPHERE_100K_FREQUENCY, seed=1)
    R = scenes.random_rays(V, 1_000_000, seed=2)
    col = RRCollider.create(V, I)
    info = col.info()
    rays = DeviceRays(R)
    L, s, e0, e1 = _events()
    for _ in range(warmup):
        col.intersect_device(rays, s)
    N.check(L.lsg_synchronize(), "sync")
    t0 = time.perf_counter()
    N.check(L.lsg_event_record(e0, s), "ev")
    for _ in range(steps):
        col.intersect_device(rays, s)
    N.check(L.lsg_event_record(e1, s), "ev")
    N.check(L.lsg_synchronize(), "sync")
    wall = time.perf_counter() - t0
    ms = ctypes.c_float()
    N.check(L.lsg_event_elapsed_ms(e0, e1, ctypes.byref(ms)), "elapsed")
    kernel_ms = ms.value / steps
    n_node, n_tri = col.trace_stats(rays)
    b_ray = 48 + 64 * n_node + 48 * n_tri
    mrays = len(R) / (wall / steps) / 1e6
    achieved = b_ray * len(R) / (kernel_ms * 1e-3) / 1e9
    hits, _ = rays.download_hits()
    res = dict(
        metric=METRIC, value=round(mrays, 3), unit="Mrays/s", ms_per_step=round(wall / steps * 1e3, 4),
        dtype="f32", config=dict(workload="collider: displaced icosphere 100,820 tris, 1M random rays, closest hit",
                                 triangles=int(len(I)), rays=int(len(R)), leaf_size=int(info.leaf_size),
                                 bvh_depth=int(info.depth), build_ms=round(info.build_ms, 3),
                                 hit_ratio=round(float((hits["tri"] >= 0).mean()), 4)),
        roofline=dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                      frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                      kernel="k_trace_closest", kernel_ms=round(kernel_ms, 4),
                      bytes_per_ray=round(b_ray, 2), n_node=round(n_node, 3), n_tri=round(n_tri, 3)),
    )
    res["cpu_baseline"] = collider_cpu_baseline(V, I, R, cpu_threads, col)
    return res


def collider_cpu_baseline(V, I, R, threads, col):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as O
    if O.have_reference():
        _, tm = O.reference_intersect(V, I, R, technique=3, threads=threads, repeats=3, flags=(1 | 4 | 16))
        return dict(value=round(tm["mrays_per_s"], 4), unit="Mrays/s", cores=int(tm["threads"]), kind="reference",
                    sample=f"IT_BSP_FASTER (compiled reference collider) on the same mesh and all {len(R)} rays, "
                           f"FILL_DISTANCE|FILL_POINT2D|FILL_TRIANGLE, best of 3; build {tm['build_s']:.2f} s not included")
    tree = col.download_tree()
    n = min(len(R), 500_000)
    t0 = time.perf_counter()
    O.bvh_traverse(tree, V, I, R[:n], threads=threads)
    dt = time.perf_counter() - t0
    return dict(value=round(n / dt / 1e6, 4), unit="Mrays/s", cores=threads, kind="port",
                sample=f"oracle C traversal of the same LBVH, {n} of the rays")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default=os.environ.get("LSG_BENCH_WORKLOAD", "auto"))
    ap.add_argument("--cpu-threads", type=int, default=0)
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 (rays and texels shard "
                                   "by batch/object across GPUs, see DESIGN.md)", "n_gpus": a.gpus}))
        sys.exit(2)
    threads = a.cpu_threads or min(os.cpu_count() or 1, 16)
    workload = a.workload
    if workload == "auto":
        try:
            from lightsprintsdk_amd import bake  # noqa: F401
            workload = "bake"
        except ImportError:
            workload = "collider"
    if workload == "bake":
        from lightsprintsdk_amd import bake
        res = bake.bench_workload(a.steps, a.warmup, threads)
    elif workload == "batched":
        from lightsprintsdk_amd import bake
        res = bake.bench_batched(a.steps, a.warmup, threads)
    else:
        res = collider_workload(a.steps, a.warmup, threads)
    out = dict(metric=res.pop("metric"), value=res.pop("value"), unit=res.pop("unit"), n_gpus=1, steps=a.steps,
               warmup=a.warmup, ms_per_step=res.pop("ms_per_step"), higher_is_better=True, scaling="weak",
               vs_baseline=None, dtype=res.pop("dtype"), data="synthetic (seeded procedural scene and rays)")
    out.update(res)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
