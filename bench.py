"""Benchmark of the MI355X ray-cast + lightmap-gather path (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload bake|collider|batched|learned|scale]

Prints ONE JSON line.  Workloads (synthetic, seeded):
  bake      (default) plain lightmap bake: closed 10x10x4 m box room + 20 boxes
            (~50k triangles), 21 lightmaps of 256x256, 1024 hemisphere rays per texel
            plus 103 shadow-ray rounds to the point light; one step = the final gather
            of every covered texel + the filter chain + RGBF writeback.
            value = baked texels/s.  After the timed region the line also carries the
            learned-gather comparison on the 5 held-out scenes of the same family
            ("learned") and the closest-hit rate on the 100k-triangle icosphere
            ("collider"); --no-extras leaves them out.
  collider  closest hit on a 100,820-triangle displaced icosphere, 1M random rays
            resident in HBM; one step = one batch.  value = Mrays/s.
  batched   200 objects (~2M triangles), lightmaps 32..1024, sun + 4 point lights:
            first gather + persistent final gather + filters replayed from a hipGraph.
  learned   the 5 held-out box-room scenes: one step = the learned+fallback bake of
            one scene (cycled); value = its texels/s, the plain bake timed beside it.
  scale     city of 5000 boxes + 2048^2-vertex terrain (~10M triangles), 64 lightmaps
            of 2048^2; one step = the re-bake for one sun direction of a 60-step sweep
            (learned+fallback); p50/p99 step latency, plain re-bake timed once.
Timing: W untimed steps, then K steps between device synchronisations; the dominant
kernel is also timed with HIP events on the stream it is launched on.  The CPU
baseline is the oracle (or the compiled reference collider where it exists) on a
bounded sample, host threads stated.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NODE_BYTES = 128          # one 4-wide node per traversal step (n_node counts these)
# L2-to-fabric bytes per launch from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, scaled by
# the same counters on a 1 GiB copy (x2.0 and x1.0):
# profiles/pmc_traffic_k_gather_persistent_plan_bake_box_room_50k.md,
# profiles/pmc_traffic_k_trace_closest_collider_ico100k.md
TRAFFIC_BYTES = {"k_gather_persistent": 3.372e9, "k_trace_closest": 1.068e9}
# VALU issue: a SIMD issues one wave64 VALU instruction per 4 cycles (16 lanes wide; the
# SQ counters give one quad-cycle of SQ_ACTIVE_INST_VALU per SQ_INSTS_VALU), so MI355X
# issues 256 CUs x 4 SIMDs x 2.4 GHz / 4 = 614.4 G wave64 VALU instructions per second.
VALU_PEAK_G = 256 * 4 * 2.4 / 4
# wave64 VALU instructions per launch and active lanes per instruction, from rocprofv3
# --pmc SQ_INSTS_VALU SQ_ACTIVE_INST_VALU SQ_THREAD_CYCLES_VALU:
# profiles/pmc_sq_valu_k_gather_persistent_plan_bake_box_room_50k.md,
# profiles/pmc_sq_valu_k_trace_closest_collider_ico100k.md
VALU_INSTS = {"k_gather_persistent": 47.8033e9, "k_trace_closest": 207.612e6}
LANE_UTIL = {"k_gather_persistent": 0.7067, "k_trace_closest": 0.1960}
METRIC = "closest-hit Mrays/s; baked texels/s at ≤2% RMSE, plain vs learned gather"


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as O
    return O


def _events():
    from lightsprintsdk_amd import native as N
    L = N.lib()
    s = ctypes.c_void_p(); a = ctypes.c_void_p(); b = ctypes.c_void_p()
    N.check(L.lsg_stream_create(ctypes.byref(s)), "stream")
    N.check(L.lsg_event_create(ctypes.byref(a)), "event")
    N.check(L.lsg_event_create(ctypes.byref(b)), "event")
    return L, s, a, b


# ---- closest hit: 100,820-triangle displaced icosphere, 1M rays ---------------------------------------------------------------

def collider_workload(steps, warmup, cpu_threads, with_cpu=True):
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
    b_ray = 48 + NODE_BYTES * n_node + 48 * n_tri
    mrays = len(R) / (wall / steps) / 1e6
    achieved = b_ray * len(R) / (kernel_ms * 1e-3) / 1e9
    hits, _ = rays.download_hits()
    res = dict(
        metric=METRIC, value=round(mrays, 3), unit="Mrays/s", ms_per_step=round(wall / steps * 1e3, 4),
        dtype="f32", config=dict(workload="collider: displaced icosphere 100,820 tris, 1M random rays, closest hit",
                                 triangles=int(len(I)), rays=int(len(R)), leaf_size=int(info.leaf_size),
                                 bvh_depth=int(info.depth), bvh4_nodes=int(info.num_nodes4),
                                 stack_need=int(info.stack_need), build_ms=round(info.build_ms, 3),
                                 hit_ratio=round(float((hits["tri"] >= 0).mean()), 4)),
        roofline=dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                      frac=round(achieved / HBM_PEAK_GBS, 4), traffic=TRAFFIC_BYTES["k_trace_closest"],
                      kernel="k_trace_closest", kernel_ms=round(kernel_ms, 4),
                      bytes_per_ray=round(b_ray, 2), n_node=round(n_node, 3), n_tri=round(n_tri, 3),
                      hbm_counters_gbs=round(TRAFFIC_BYTES["k_trace_closest"] / (kernel_ms * 1e-3) / 1e9, 1),
                      valu=_valu("k_trace_closest", kernel_ms)),
    )
    if with_cpu:
        res["cpu_baseline"] = collider_cpu_baseline(V, I, R, cpu_threads, col)
    return res


def _valu(kernel, kernel_ms):
    """VALU issue of a kernel: its profiled instruction count per launch over the live
    launch time, against the chip's issue peak."""
    if kernel not in VALU_INSTS:
        return dict(achieved=None, peak=VALU_PEAK_G, unit="G VALU instr/s", frac=None,
                    note="no profiled instruction count for this kernel")
    g = VALU_INSTS[kernel] / (kernel_ms * 1e-3) / 1e9
    return dict(achieved=round(g, 1), peak=VALU_PEAK_G, unit="G VALU instr/s", frac=round(g / VALU_PEAK_G, 4),
                insts_per_launch=VALU_INSTS[kernel], lanes_active=LANE_UTIL[kernel])


def cpu_calibration(sample: int = 200_000):
    """rho = the oracle's CPU traversal over the compiled reference collider (IT_BSP_FASTER),
    one core each, on the same 100,820-triangle icosphere and rays: the ratio that turns a
    time of this project's CPU statement into a reference-equivalent one where the
    reference cannot run (its RRSolver needs boost)."""
    O = _oracle()
    if not O.have_reference():
        return None
    from lightsprintsdk_amd import scenes
    from lightsprintsdk_amd.collider import RRCollider
    V, I = scenes.icosphere(scenes.ICOSPHERE_100K_FREQUENCY, seed=1)
    R = scenes.random_rays(V, sample, seed=2)
    _, tm = O.reference_intersect(V, I, R, technique=3, threads=1, repeats=3, flags=(1 | 4 | 16))
    tree = RRCollider.create(V, I).download_tree()
    O.bvh_traverse(tree, V, I, R[:10_000], threads=1)
    best = 0.0
    for _ in range(3):
        t0 = time.perf_counter()
        O.bvh_traverse(tree, V, I, R, threads=1)
        best = max(best, len(R) / (time.perf_counter() - t0) / 1e6)
    return dict(rho=round(best / tm["mrays_per_s"], 4), reference_mrays_per_core=round(tm["mrays_per_s"], 4),
                port_mrays_per_core=round(best, 4), rays=sample,
                note="one host core each, closest hit on the same icosphere and rays")


def collider_cpu_baseline(V, I, R, threads, col):
    O = _oracle()
    if O.have_reference():
        _, tm = O.reference_intersect(V, I, R, technique=3, threads=threads, repeats=3, flags=(1 | 4 | 16))
        return dict(value=round(tm["mrays_per_s"], 4), unit="Mrays/s", cores=int(tm["threads"]), kind="reference",
                    sample=f"IT_BSP_FASTER (the reference collider compiled from its sources) on the same mesh and "
                           f"all {len(R)} rays, all {threads} host cores, FILL_DISTANCE|FILL_POINT2D|FILL_TRIANGLE, "
                           f"best of 3; its BSP build ({tm['build_s']:.2f} s) not included")
    tree = col.download_tree()
    n = min(len(R), 500_000)
    t0 = time.perf_counter()
    O.bvh_traverse(tree, V, I, R[:n], threads=threads)
    dt = time.perf_counter() - t0
    return dict(value=round(n / dt / 1e6, 4), unit="Mrays/s", cores=threads, kind="port",
                sample=f"oracle C traversal of the same tree, {n} of the rays")


# ---- plain bake: box room, ~50k triangles, 21 lightmaps of 256x256 ----------------------------------------------------------------

def bake_workload(steps: int, warmup: int, cpu_threads: int):
    """Lightmap bake of the box room through the bake plan (lsg_bake_*, the path the solver
    adapter's updateLightmaps takes): one step = final gather (persistent, XCD-local tile
    queues, launch shape from the autotuner) + filter chain + RGBF writeback, replayed from a
    hipGraph.  The exitance table comes from the plan's first gather, run once before."""
    from lightsprintsdk_amd import native as N, scenes as S
    from lightsprintsdk_amd.bake import BakePlan, GatherParams, box_room_scene, autotune, TEXEL_DTYPE, SUBTEXEL_DTYPE
    sc = box_room_scene()
    t0 = time.perf_counter()
    plan = BakePlan(sc, GatherParams(quality=1024, ray_length_min=S.minimal_safe_distance(sc.verts), seed=1))
    setup_s = time.perf_counter() - t0
    L = N.lib()
    t0 = time.perf_counter()
    plan.run(first=True, graph=False)
    N.check(L.lsg_synchronize(), "sync")
    first_s = time.perf_counter() - t0
    cache = os.environ.get("LSG_TUNING_CACHE", os.path.expanduser("~/.cache/lsgpu_tuning.json"))
    os.makedirs(os.path.dirname(cache), exist_ok=True)
    t0 = time.perf_counter()
    tune = autotune(plan, sample_texels=150_000, cache_path=cache)
    tune_s = time.perf_counter() - t0
    for _ in range(warmup):
        plan.run(first=False, graph=True)
    N.check(L.lsg_synchronize(), "sync")
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.run(first=False, graph=True)      # final gather + filter chain + RGBF writeback
    N.check(L.lsg_synchronize(), "sync")
    wall = time.perf_counter() - t0
    # the dominant kernel alone (every texel, the tuned shape), HIP events on its stream
    kernel_ms = float(np.mean([plan.sample_ms(1) for _ in range(3)]))
    inf = plan.info()
    ntex, nsub = int(inf.num_texels), int(inf.num_subtexels)
    texels = np.empty(ntex, TEXEL_DTYPE)
    subs = np.empty(nsub, SUBTEXEL_DTYPE)
    N.check(L.lsg_memcpy_d2h(texels.ctypes.data_as(ctypes.c_void_p), inf.texels, texels.nbytes), "d2h")
    N.check(L.lsg_memcpy_d2h(subs.ctypes.data_as(ctypes.c_void_p), inf.subtexels, subs.nbytes), "d2h")
    texels_per_s = ntex / (wall / steps)
    # traversal work of the same gather (separate counting launch)
    st = (ctypes.c_double * 6)()
    p = plan.params.c()
    N.check(L.lsg_gather_stats(plan.gs._h, inf.texels, ntex, inf.subtexels, nsub, ctypes.byref(p), inf.rgba, st),
            "gather_stats")
    plan.run(first=False, graph=False)            # the counting launch wrote the work buffer
    hr, hn, ht, sr, sn, stt = list(st)
    b_ray = 48 + NODE_BYTES * hn / hr + 48 * ht / hr
    b_shadow = 48 + NODE_BYTES * sn / max(sr, 1) + 48 * stt / max(sr, 1)
    n_sub = nsub / ntex
    R = hr / ntex
    Lr = sr / ntex
    b_texel = R * (b_ray + 12) + Lr * b_shadow + 32 * n_sub + 12
    achieved = b_texel * ntex / (kernel_ms * 1e-3) / 1e9
    rays_per_s = (hr + sr) / (kernel_ms * 1e-3)
    # the filtered lightmaps in slot order (objects' slots are consecutive), at the covered texels
    lit = np.concatenate([plan.lightmap(k).reshape(-1, 3) for k in range(len(sc.objects))])[texels["out_index"]]
    kern = "k_gather_persistent"
    res = dict(
        metric=METRIC, value=round(texels_per_s, 1), unit="texels/s", ms_per_step=round(wall / steps * 1e3, 3),
        dtype="f32",
        config=dict(workload="lightmap bake (bake plan): box room 10x10x4 m + 20 boxes, 21 lightmaps 256x256 RGBF, "
                             "1024 hemisphere rays + 103 shadow rounds per texel (point light + emissive quad); "
                             "step = final gather + filter chain + RGBF writeback, hipGraph replay",
                    triangles=int(sc.ntri), lightmaps=len(sc.objects), texels=ntex, subtexels=nsub,
                    quality=1024, hemisphere_rays=int(hr), shadow_rays=int(sr), rays_per_s=round(rays_per_s, 1),
                    mrays_per_s=round(rays_per_s / 1e6, 2), first_run_s=round(first_s, 4), setup_s=round(setup_s, 3),
                    tuning=dict(tile=inf.tuning.tile, waves_per_block=inf.tuning.waves_per_block,
                                blocks_per_cu=inf.tuning.blocks_per_cu, xcd_queues=inf.tuning.xcd_queues,
                                lane_refill=inf.tuning.lane_refill),
                    autotune=tune, autotune_s=round(tune_s, 2), mean_irradiance=float(lit.mean())),
        roofline=dict(bound="valu", **{k: v for k, v in _valu(kern, kernel_ms).items()},
                      traffic=TRAFFIC_BYTES.get(kern), kernel=kern, kernel_ms=round(kernel_ms, 4),
                      hbm=dict(achieved_gbs=round(TRAFFIC_BYTES.get(kern, 0) / (kernel_ms * 1e-3) / 1e9, 1), peak_gbs=HBM_PEAK_GBS,
                               frac=round(TRAFFIC_BYTES.get(kern, 0) / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
                               note="L2-to-fabric bytes per launch (FETCH_SIZE/WRITE_SIZE); the scene stays in L2/MALL"),
                      model=dict(bytes_per_texel=round(b_texel, 1), bytes_per_hemisphere_ray=round(b_ray, 2),
                                 bytes_per_shadow_ray=round(b_shadow, 2), n_node_hemi=round(hn / hr, 3),
                                 n_tri_hemi=round(ht / hr, 3), n_node_shadow=round(sn / max(sr, 1), 3),
                                 n_tri_shadow=round(stt / max(sr, 1), 3), subtexels_per_texel=round(n_sub, 3),
                                 algorithmic_bytes=round(b_texel * ntex), model_gbs=round(achieved, 1),
                                 note="node and triangle bytes the traversal reads, most of them cache hits: "
                                      "a work measure, not HBM traffic")),
    )
    res["cpu_baseline"] = bake_cpu_baseline(plan.gs, texels, subs, plan.params, plan.exitance(), cpu_threads)
    return res


def bake_cpu_baseline(gs, texels, subs, params, exitance, threads: int, target_s: float = 12.0):
    """The CPU oracle's FP64 statement of processTexel (a port: the reference's RRSolver
    needs boost and is not built here) on a bounded, evenly spread sample of the same
    texels, all host cores; with the per-core rate and, where the compiled reference
    collider exists, the reference-equivalent rate port / rho (rho measured on the collider,
    which is most of a gather's work)."""
    O = _oracle()
    tree = gs.col.download_tree()
    arrays = gs.oracle_arrays(exitance)
    n0 = 64 * threads
    sel = np.linspace(0, len(texels) - 1, n0).astype(np.int64)
    t0 = time.perf_counter()
    O.gather(arrays, params.as_dict(), texels[sel].view(np.uint32).reshape(-1, 4), subs, tree=tree, threads=threads)
    dt = time.perf_counter() - t0
    n1 = int(min(len(texels), max(n0, n0 * target_s / max(dt, 1e-3))))
    sel = np.linspace(0, len(texels) - 1, n1).astype(np.int64)
    t0 = time.perf_counter()
    O.gather(arrays, params.as_dict(), texels[sel].view(np.uint32).reshape(-1, 4), subs, tree=tree, threads=threads)
    dt = time.perf_counter() - t0
    out = dict(value=round(n1 / dt, 1), unit="texels/s", cores=threads, kind="port",
               sample=f"{n1} of {len(texels)} texels evenly spread, oracle C gather (FP64 sums) over the same tree, "
                      f"{dt:.1f} s on all {threads} host cores", per_core=round(n1 / dt / threads, 1))
    cal = cpu_calibration()
    if cal:
        out["calibration"] = cal
        out["reference_equivalent"] = round(n1 / dt / cal["rho"], 1)
    return out


# ---- learned + fallback vs plain on the held-out box-room scenes -------------------------

def learned_summary(reps: int = 3, log=None):
    """The committed surrogate on the 5 held-out box-room scenes (the plain bake of each
    is the reference and is timed beside it; median of `reps` runs of each)."""
    from lightsprintsdk_amd import learned as Lr
    got = Lr.load_weights()
    if got is None:
        return dict(error="no trained surrogate (python -m lightsprintsdk_amd.learned train)")
    net, bound, meta = got
    rows = Lr.heldout_report(net, bound, reps=reps, log=log or (lambda m: None))
    tex = sum(r["texels"] for r in rows)
    tp = sum(r["plain_ms"] for r in rows) / 1e3
    tl = sum(r["learned_ms"] for r in rows) / 1e3
    return dict(workload="learned gather: 5 held-out box-room scenes (seeds 2000-2004), 16-ray cheap gather + "
                         "MLP 32-64x4-4 (BF16 MFMA, FP32 accumulate) + per-scene error bound from a second "
                         "16-ray gather + fallback to the 1024-ray plain gather; "
                         "RMSE of the filtered lightmaps against the plain bake of the same scene",
                texels_per_s=round(tex / tl, 1), plain_texels_per_s=round(tex / tp, 1), speedup=round(tp / tl, 3),
                rmse_max=round(max(r["rmse"] for r in rows), 5), rmse_bound=0.02,
                fallback_mean=round(float(np.mean([r["fallback"] for r in rows])), 4),
                bound=dict(kind="per-scene: largest tau whose estimated kept-texel RMSE (second 16-ray gather) "
                                "is within target", target=bound["target"]),
                timing=f"median of {reps} runs per scene, plain and learned alternating",
                training=dict(source="metadata of the shipped surrogate (trained offline by python -m "
                                     "lightsprintsdk_amd.learned train; not run or timed by this command)",
                              scenes=20, samples=meta.get("samples"), epochs=meta.get("epochs"),
                              train_s=meta.get("train_s")),
                scenes=[{k: (round(v, 5) if isinstance(v, float) else v) for k, v in r.items()} for r in rows])


def learned_workload(steps: int, warmup: int):
    from lightsprintsdk_amd import learned as Lr, native as N
    got = Lr.load_weights()
    if got is None:
        raise SystemExit("no trained surrogate: run python -m lightsprintsdk_amd.learned train")
    net, bound, meta = got
    scenes = [Lr.prepare_eval(sd) for sd in Lr.HELDOUT_SEEDS]
    L = N.lib()
    for k in range(warmup):
        scenes[k % len(scenes)].learned(net, target=bound["target"])
    N.check(L.lsg_synchronize(), "sync")
    t0 = time.perf_counter()
    tex = 0
    fb = 0
    for k in range(steps):
        ls = scenes[k % len(scenes)]
        fb += ls.learned(net, target=bound["target"])
        tex += ls.n
    N.check(L.lsg_synchronize(), "sync")
    wall = time.perf_counter() - t0
    summ = learned_summary()
    return dict(metric=METRIC, value=round(tex / wall, 1), unit="texels/s", ms_per_step=round(wall / steps * 1e3, 3),
                dtype="bf16 MLP + f32 gather",
                config=dict(workload=summ.pop("workload"), steps_cycle_scenes=len(scenes),
                            fallback_fraction=round(fb / max(tex, 1), 4)),
                learned=summ)


# ---- large re-bake with a sun sweep (10M-triangle city and terrain) ---------------------------------------------

def scale_workload(steps: int, warmup: int, sweep: int = 60, log=None):
    """Large re-bake: city of 5000 extruded boxes + 2048^2-vertex terrain (~10M triangles),
    64 lightmaps of 2048^2.  One step = a full re-bake for the next sun direction of a
    60-step sweep: first gather (exitance) + learned+fallback final gather + filters.
    `steps` directions spread over the sweep are timed (p50/p99); the plain re-bake of
    the first direction is timed once and is the RMSE reference."""
    from lightsprintsdk_amd import learned as Lr, native as N, scenes as S
    log = log or (lambda m: print(m, file=sys.stderr, flush=True))
    L = N.lib()
    got = Lr.load_weights(Lr.weights_path("city"))
    if got is None:
        raise SystemExit("no trained city surrogate: python -m lightsprintsdk_amd.learned train --family city")
    net, bound, meta = got
    t0 = time.perf_counter()
    idx = [int(round(k * (sweep - 1) / max(steps - 1, 1))) for k in range(steps)]
    sc = S.city_terrain(seed=5, sun_direction=S.sun_sweep(idx[0], sweep))
    gen_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    ls = Lr.LearnedScene(sc, seed=1)
    setup_s = time.perf_counter() - t0
    ntex = ls.n
    log(f"city re-bake: {sc.ntri} triangles, {len(sc.objects)} lightmaps, {ntex} texels; generate {gen_s:.1f}s, setup {setup_s:.1f}s")

    def set_sun(i):
        lt = dict(sc.lights[0]); lt["direction"] = S.sun_sweep(i, sweep)
        ls.sb.gs.set_lights([lt])

    def sync():
        N.check(L.lsg_synchronize(), "sync")
    # plain re-bake of the first direction (reference and plain timing)
    set_sun(idx[0]); sync()
    t0 = time.perf_counter()
    ls.sb.first_gather()
    ls.plain()
    sync()
    plain_s = time.perf_counter() - t0
    ref = ls.final_rgb()
    log(f"plain re-bake: {plain_s:.2f}s")
    for _ in range(warmup):
        ls.learned(net, target=bound["target"])
    sync()
    lat, fbs, rmse0 = [], [], None
    for k, i in enumerate(idx):
        set_sun(i); sync()
        t0 = time.perf_counter()
        ls.sb.first_gather()
        nfb = ls.learned(net, target=bound["target"])
        sync()
        lat.append(time.perf_counter() - t0)
        fbs.append(nfb / max(ntex, 1))
        if k == 0:
            rmse0 = Lr.rmse_rel(ls.final_rgb(), ref)
        log(f"step {k} (sweep {i}): {lat[-1]:.2f}s, fallback {fbs[-1]:.3f}" + (f", rmse {rmse0:.4f}" if k == 0 else ""))
    lat = np.array(lat)
    p50 = float(np.percentile(lat, 50))
    return dict(metric=METRIC, value=round(ntex / float(lat.mean()), 1), unit="texels/s",
                ms_per_step=round(float(lat.mean()) * 1e3, 1), dtype="bf16 MLP + f32 gather",
                config=dict(workload="large re-bake: city of 5000 extruded boxes + 2048^2-vertex terrain, "
                                     "64 lightmaps 2048^2 RGBF, directional sun swept over 60 steps; step = first "
                                     "gather + learned+fallback final gather (per-scene error bound) + filters",
                            triangles=int(sc.ntri), lightmaps=len(sc.objects), texels=int(ntex),
                            sweep_steps_timed=idx, generate_s=round(gen_s, 1), setup_s=round(setup_s, 1)),
                scale=dict(step_latency_p50_s=round(p50, 3), step_latency_p99_s=round(float(np.percentile(lat, 99)), 3),
                           plain_step_s=round(plain_s, 3), plain_texels_per_s=round(ntex / plain_s, 1),
                           speedup=round(plain_s / float(lat.mean()), 3), rmse_first_step=round(rmse0, 5),
                           fallback=[round(f, 4) for f in fbs], target=bound["target"],
                           surrogate=dict(family="city", samples=meta.get("samples"), train_s=meta.get("train_s"))))


# ---- batched persistent bake: 200 objects, ~2M triangles ----------------------------------------------------

def batched_workload(steps: int, warmup: int, n_objects: int = 200, target_tris: int = 2_000_000):
    """Batched bake: 200 objects (~2M triangles), lightmaps 32..1024, 1024 rays per texel,
    sun + 4 point lights; one step = first gather + persistent final gather + filters
    replayed from a hipGraph, launch shape from the autotuner."""
    from lightsprintsdk_amd import native as N, scenes as S
    from lightsprintsdk_amd.bake import BakePlan, GatherParams, autotune
    t0 = time.perf_counter()
    sc = S.object_field(seed=4, n_objects=n_objects, target_tris=target_tris)
    gen_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    plan = BakePlan(sc, GatherParams(quality=1024, ray_length_min=S.minimal_safe_distance(sc.verts), seed=1))
    setup_s = time.perf_counter() - t0
    plan.run(first=True, graph=False)
    t0 = time.perf_counter()
    cache = os.environ.get("LSG_TUNING_CACHE", os.path.expanduser("~/.cache/lsgpu_tuning.json"))
    os.makedirs(os.path.dirname(cache), exist_ok=True)
    tune = autotune(plan, sample_texels=150_000, cache_path=cache)
    tune_s = time.perf_counter() - t0
    for _ in range(warmup):
        plan.run(first=True, graph=True)
    N.check(N.lib().lsg_synchronize(), "sync")
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.run(first=True, graph=True)
    N.check(N.lib().lsg_synchronize(), "sync")
    wall = (time.perf_counter() - t0) / steps
    inf = plan.info()
    return dict(metric=METRIC, value=round(inf.num_texels / wall, 1), unit="texels/s", ms_per_step=round(wall * 1e3, 2),
                dtype="f32",
                config=dict(workload="batched bake: 200 objects (displaced boxes, cylinders, terrain patches), "
                                     "lightmaps 32..1024 RGBF, 1024 rays/texel, sun + 4 point lights, hipGraph replay",
                            triangles=int(sc.ntri), objects=len(sc.objects), texels=int(inf.num_texels),
                            texel_slots=int(inf.num_slots), subtexels=int(inf.num_subtexels),
                            tuning=dict(tile=inf.tuning.tile, waves_per_block=inf.tuning.waves_per_block,
                                        blocks_per_cu=inf.tuning.blocks_per_cu, xcd_queues=inf.tuning.xcd_queues),
                            autotune=tune, autotune_s=round(tune_s, 2), generate_s=round(gen_s, 2),
                            setup_s=round(setup_s, 2), grow_steps=int(inf.grow_steps)))


def copy_workload(steps: int, warmup: int):
    """A plain 1 GiB device-to-device copy: the known-size kernel whose FETCH_SIZE and
    WRITE_SIZE readings scale the same counters of the bake and collider kernels."""
    import torch
    n = 1 << 28
    a = torch.empty(n, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(warmup):
        b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / max(steps, 1)
    return dict(metric="device copy GB/s (counter calibration)", value=round(2 * 4 * n / dt / 1e9, 1), unit="GB/s",
                ms_per_step=round(dt * 1e3, 4), dtype="f32",
                config=dict(workload="1 GiB device-to-device copy (reads 1 GiB, writes 1 GiB)", bytes=4 * n))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default=os.environ.get("LSG_BENCH_WORKLOAD", "bake"),
                    choices=["bake", "collider", "batched", "learned", "scale", "copy"])
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-extras", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "liblsgpu runs on one GPU per process; --gpus must be 1 (rays and texels "
                                   "would shard by batch/object across GPUs; no multi-GPU path is implemented)",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    threads = a.cpu_threads or len(os.sched_getaffinity(0))      # every host core this process may use
    if a.workload == "bake":
        res = bake_workload(a.steps, a.warmup, threads)
        if not a.no_extras:
            res["learned"] = learned_summary()
            bt = batched_workload(2, 1)
            res["batched"] = dict(workload=bt["config"]["workload"], texels_per_s=bt["value"], ms_per_step=bt["ms_per_step"],
                                  steps=2, warmup=1, **{k: bt["config"][k] for k in ("triangles", "objects", "texels", "tuning",
                                                                                      "setup_s", "autotune_s")})
            c = collider_workload(5, 1, threads, with_cpu=True)
            cb = c["cpu_baseline"]
            res["collider"] = dict(workload=c["config"]["workload"], mrays_per_s=c["value"],
                                   kernel_ms=c["roofline"]["kernel_ms"], roofline=c["roofline"], cpu_baseline=cb,
                                   vs_cpu=round(c["value"] / cb["value"], 2))
    elif a.workload == "batched":
        res = batched_workload(a.steps, a.warmup)
    elif a.workload == "learned":
        res = learned_workload(a.steps, a.warmup)
    elif a.workload == "scale":
        res = scale_workload(a.steps, a.warmup)
    elif a.workload == "copy":
        res = copy_workload(a.steps, a.warmup)
    else:
        res = collider_workload(a.steps, a.warmup, threads)
    out = dict(metric=res.pop("metric"), value=res.pop("value"), unit=res.pop("unit"), n_gpus=1, steps=a.steps,
               warmup=a.warmup, ms_per_step=res.pop("ms_per_step"), higher_is_better=True, scaling="weak",
               vs_baseline=None, dtype=res.pop("dtype"), data="synthetic (seeded procedural scene and rays)")
    out.update(res)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
