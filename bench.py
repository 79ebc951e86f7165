"""Benchmark of the PatchMatch depth path on one MI355X.

Workload (one "step"): plain single-view PatchMatch on a 1600x1200 float32
reference with 8 neighbours (seeded 64-facet piecewise-planar heightfield,
3x3 camera grid, depth range [4, 40] m, patch 7x7, 3 iterations):
random init + 3 x (red pass + black pass) + decode, all on the device with
the views already resident in HBM.  Metric: Mpix*views*iters/s =
(W*H/1e6) * N_src * iterations / seconds.

Also reported: the roofline of the candidate-scoring kernel
(candidates_kernel, fixed algorithmic work per launch, timed with HIP events
on the estimator's stream inside the timed steps) and of the whole colour
pass against the FP32 vector peak, per-stage device time, accuracy against
the synthetic ground truth, and the CPU oracle timed on the host cores on a
bounded sample of the same scene.  Each est.run() returns only after its HIP stream is idle
(hipStreamSynchronize), so the wall clock brackets complete GPU work.
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

METRIC = "Mpix·views·iters/s on 1 GPU vs host CPU; % pixels |Δd|/d<1% vs reference"
UNIT = "Mpix·views·iters/s"
PEAK_FP32_TFLOPS = 157.3          # MI355X published FP32 vector peak
PATCH = 7
FLOP_PER_UNIT = 15 * (21 * PATCH * PATCH + 75)   # 15 hypotheses x (21 FLOP/tap x P taps + 75)
# HBM bytes per candidates_kernel<3, float, 4> launch: rocprofv3 --pmc
# FETCH_SIZE (x 1.9999) + WRITE_SIZE (x 1.0) on this workload, each factor
# read off the same counter on a 1 GiB device-to-device copy
# (profiles/*_hbm_traffic.txt).
TRAFFIC_PER_LAUNCH = 2.135e9
TRAFFIC_TILE = 4


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


WORKLOADS = {
    # name: (scene factory, description)
    "c2": ("facet_heightfield", "plain PatchMatch, one 1600x1200 {storage} reference + 8 neighbours, 64-facet "
                                "heightfield (seed {seed}), patch 7x7, {iters} iterations, init+sweeps+decode per step"),
    "c4": ("sinusoidal_bumps", "PatchMatch + 10 Gauss-Newton steps, one 1024x768 {storage} reference + 6 neighbours, "
                               "sinusoidal bumps 0.5 m over a 10 m plane (seed {seed}), patch 7x7, {iters} iterations"),
    "c3": ("camera_walk", "batched PatchMatch, {views} references of 640x480..2000x1500 with 4-12 neighbours each "
                          "(camera walk around a 20x20 m box scene, seed {seed}), {storage} images, patch 7x7, "
                          "{iters} iterations, one HIP-graph launch per step"),
    "cluster": ("camera_walk", "cluster surface, {views} references of 640x480..2000x1500 with 4-12 neighbours each "
                               "(camera walk, seed {seed}), {levels} pyramid levels with geometric links "
                               "(weight 0.05), f32 images, patch 7x7, {iters} iterations per level, every level "
                               "one batch"),
    "c5": ("aerial_survey", "aerial batch, {views} nadir {width}x{height} {storage} views over a fractal terrain "
                            "(2x2 km, 0-80 m relief, 1 m grid, seed {seed}), 80%/70% overlap at 100 m, "
                            "{neighbours} neighbours each, depth [60,140] m, patch 7x7, {iters} iterations, "
                            "groups admitted by device memory, {refine} Gauss-Newton steps per reference"),
}


def make_scene(seed: int, workload: str = "c2"):
    from opensfm_amd import synthetic
    if workload == "c4":
        return synthetic.sinusoidal_bumps(seed=seed, width=1024, height=768, num_src=6)
    return synthetic.facet_heightfield(seed=seed, width=1600, height=1200, num_src=8)


def configure(est, scene, iters: int, storage: str = "f32"):
    """The workload's estimator: the dense pipeline's parameters at one
    pyramid level (the workloads are single-level), patch 7x7, the scene's
    depth range, no post-processing."""
    from opensfm_amd import native
    for im, K, R, t in zip(scene.images, scene.Ks, scene.Rs, scene.ts):
        est.add_view(K, R, t, im)
    est.set_params({**native().pipeline_params(), "hierarchy_levels": 1, "max_iterations": iters,
                    "patch_size": PATCH, "image_storage": storage, "speckle_min_size": 0, "gap_max_size": 0})
    est.set_depth_range(scene.depth_min, scene.depth_max, 0)


def storage_effect(pydense, scene, iters: int, depth_f16) -> dict:
    """Depth of the f16-storage run against the same scene run with f32
    storage (full-precision texels, the oracle's arithmetic otherwise)."""
    e = pydense.DepthmapEstimator()
    configure(e, scene, iters, "f32")
    d32 = e.compute()["depth"]
    m = d32 > 0
    return {"reference": "the same run with f32 image storage",
            "within_1pct": round(float(np.mean(np.abs(depth_f16[m] - d32[m]) < 0.01 * d32[m])), 5),
            "bit_equal": round(float(np.mean(depth_f16 == d32)), 5)}


def cpu_baseline(scene, iters: int, rows: int, out: dict = None, half: bool = False):
    """CPU oracle on a band of reference rows (all sources), init + iterations.
    ``half``: the images as the f16 storage holds them.  With ``out``, the
    band scene and the oracle's depth go into it."""
    from oracle import pm_oracle as po
    y0 = scene.gt_depth.shape[0] // 2 - rows // 2
    band = scene.crop_rows(y0, y0 + rows)
    p = po.Params(patch_size=PATCH, depth_min=scene.depth_min, depth_max=scene.depth_max, max_iterations=iters)
    imgs = [np.asarray(im, np.float32) for im in band.images]
    if half:
        imgs = [im.astype(np.float16).astype(np.float32) for im in imgs]
    sc = po.OracleScene(imgs, band.Ks, band.Rs, band.ts)
    st = po.OracleState(sc.width, sc.height)
    t0 = time.perf_counter()
    po.run(sc, p, st)
    depth, _, _, _ = po.decode(sc, p, st)
    dt = time.perf_counter() - t0
    if out is not None:
        out["band"], out["depth"] = band, depth
    units = sc.width * sc.height / 1e6 * band.num_src * iters
    return {"value": units / dt, "unit": UNIT, "cores": po.max_threads(), "kind": "port",
            "sample": f"reference rows {y0}..{y0 + rows} of the same {sc.width}x{scene.gt_depth.shape[0]} scene "
                      f"({sc.width}x{rows} px, "
                      f"{band.num_src} sources, {iters} iterations, init+iterations+decode), {dt:.1f} s"}


def run_c3(args, pydense) -> int:
    """Many references of mixed resolution in one batched, graph-captured run."""
    import torch
    from opensfm_amd import synthetic
    from oracle import pm_oracle as po

    seed = 3 if args.seed == 2 else args.seed
    storage = args.storage if args.storage_set else "f16"
    dev = "cuda" if torch.cuda.is_available() else "cpu"
    log(f"rendering camera walk ({args.views} views) on {dev} ...")
    t0 = time.perf_counter()
    ws = synthetic.camera_walk(seed=seed, num_views=args.views, device=dev)
    torch.cuda.empty_cache()
    log(f"rendered in {time.perf_counter() - t0:.1f} s")
    units = ws.units(args.iters)
    ests = []
    for i in range(len(ws.refs)):
        est = pydense.DepthmapEstimator()
        configure(est, ws.reference(i), args.iters, storage)
        ests.append(est)
    batch = pydense.DepthmapBatch()
    for e in ests:
        batch.add(e)
    t0 = time.perf_counter()
    batch.prepare()
    prep_s = time.perf_counter() - t0
    cfg = batch.config()
    log(f"staged + tuned + captured in {prep_s:.1f} s; config {cfg}")
    for i in range(args.warmup):
        log(f"warmup {i + 1}/{args.warmup}: {batch.run():.1f} ms")
    t0 = time.perf_counter()
    dev_ms = []
    for i in range(args.steps):
        dev_ms.append(batch.run())
        log(f"step {i + 1}/{args.steps}: {dev_ms[-1]:.1f} ms")
    elapsed = time.perf_counter() - t0
    value = units * args.steps / elapsed

    # accuracy against the ground truth, pooled over all references
    hit = tot = hit5 = 0
    outs = []
    for i, e in enumerate(ests):
        d = e.download()["depth"]
        outs.append(d)
        gt = ws.gt_depth[ws.refs[i][0]]
        m = gt > 0
        hit += int(np.sum(np.abs(d[m] - gt[m]) < 0.01 * gt[m]))
        hit5 += int(np.sum(np.abs(d[m] - gt[m]) < 0.05 * gt[m]))
        tot += int(m.sum())
    # the same references one after the other (single-reference path)
    seq_ms = 0.0
    for e in ests:
        e.run()
        seq_ms += e.last_timing()["run_ms"]
    log(f"sequential single-reference runs: {seq_ms:.1f} ms")

    # CPU oracle on the smallest reference, and parity of the batched result there
    small = min(range(len(ws.refs)), key=lambda i: ws.images[ws.refs[i][0]].size * len(ws.refs[i][1]))
    sc = ws.reference(small)
    cpu, parity = None, None
    if args.cpu_rows > 0:
        log(f"CPU oracle on reference {small} ...")
        p = po.Params(patch_size=PATCH, depth_min=sc.depth_min, depth_max=sc.depth_max, max_iterations=args.iters)
        imgs = [im.astype(np.float16).astype(np.float32) for im in sc.images] if storage == "f16" else sc.images
        t1 = time.perf_counter()
        d_o, _, _, _, _ = po.estimate(imgs, sc.Ks, sc.Rs, sc.ts, p)
        dt = time.perf_counter() - t1
        hh, ww = sc.gt_depth.shape
        u = ww * hh / 1e6 * sc.num_src * args.iters
        cpu = {"value": round(u / dt, 5), "unit": UNIT, "cores": po.max_threads(), "kind": "port",
               "sample": f"reference {small} of the walk ({ww}x{hh}, {sc.num_src} sources, {args.iters} iterations, "
                         f"init+iterations+decode), {dt:.1f} s"}
        m = d_o > 0
        parity = {"reference": small, "within_1pct_of_oracle": round(float(np.mean(
            np.abs(outs[small][m] - d_o[m]) < 0.01 * d_o[m])), 5),
            "bit_equal": round(float(np.mean(outs[small] == d_o)), 5)}

    f16_effect = storage_effect(pydense, sc, args.iters, outs[small]) if storage == "f16" else None
    if f16_effect:
        f16_effect["reference"] += f" (reference {small})"

    sweep_tflops = units * 1e6 * FLOP_PER_UNIT / (float(np.mean(dev_ms)) / 1e3) / 1e12
    line = {
        "metric": METRIC, "value": round(value, 3), "unit": UNIT, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32" if storage == "f32" else "f16-storage/f32",
        "data": "synthetic",
        "config": {"workload": WORKLOADS["c3"][1].format(views=len(ws.refs), seed=seed, storage=storage,
                                                         iters=args.iters),
                   "references": len(ws.refs), "megapixels": round(sum(ws.images[r[0]].size for r in ws.refs) / 1e6, 3),
                   "sources_per_reference": [len(r[1]) for r in ws.refs], "iterations": args.iters, "patch": PATCH,
                   "units_per_step": round(units, 3)},
        "roofline_path": {"scope": "whole batched run (init, colour passes, decode) of one step",
                          "achieved": round(sweep_tflops, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                          "frac": round(sweep_tflops / PEAK_FP32_TFLOPS, 4), "device_ms": round(float(np.mean(dev_ms)), 3)},
        "batch": {"config": cfg, "prepare_s": round(prep_s, 2), "sequential_ms": round(seq_ms, 3),
                  "sequential_value": round(units / (seq_ms / 1e3), 3),
                  "tune_log": batch.tune_log().round(3).tolist()},
        "cpu_baseline": cpu,
        "parity": parity,
        "accuracy": {"gt_within_1pct": round(hit / max(tot, 1), 4), "gt_within_5pct": round(hit5 / max(tot, 1), 4)},
        "f16_storage_vs_f32": f16_effect,
    }
    if cpu and cpu["value"] > 0:
        line["gpu_over_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(line), flush=True)
    return 0


def run_cluster(args, pydense) -> int:
    """The pipeline's caller surface: DepthmapClusterEstimator over the
    camera walk with hierarchy levels and geometric links between the
    references (depthmaps.py:154-286 drives it so)."""
    import torch
    from opensfm_amd import synthetic

    seed = 3 if args.seed == 2 else args.seed
    views = args.views or 64
    dev = "cuda" if torch.cuda.is_available() else "cpu"
    log(f"rendering camera walk ({views} views) on {dev} ...")
    ws = synthetic.camera_walk(seed=seed, num_views=views, device=dev)
    torch.cuda.empty_cache()
    refs = ws.refs
    index_of = {v: i for i, (v, _, _, _) in enumerate(refs)}

    def cluster():
        c = pydense.DepthmapClusterEstimator()
        c.set_params(pydense.pipeline_params())   # what depthmaps.py:154-175 sets
        c.set_hierarchy_levels(args.levels)
        c.set_geom_consistency_weight(0.05 if args.levels > 1 else 0.0)
        c.set_max_iterations(args.iters)
        c.set_patch_size(PATCH)
        for v, srcs, dmin, dmax in refs:
            c.begin_ref_view(ws.Ks[v], ws.Rs[v], ws.ts[v], ws.images[v], dmin, dmax)
            for j in srcs:
                c.add_source_view(ws.Ks[j], ws.Rs[j], ws.ts[j], ws.images[j])
        for i, (v, srcs, _, _) in enumerate(refs):
            for pos, j in enumerate(srcs):
                if j in index_of:
                    c.add_geom_link(i, pos, index_of[j])
        return c

    c = cluster()
    for i in range(args.warmup):
        c.run()
        log(f"warmup {i + 1}/{args.warmup} done")
    t0 = time.perf_counter()
    for k in range(args.steps):
        res = c.run()
        log(f"step {k + 1}/{args.steps} done")
    elapsed = time.perf_counter() - t0
    # work: every level's references, Mpix x sources x iterations
    units = 0.0
    for v, srcs, _, _ in refs:
        h, w = ws.images[v].shape
        for lv in range(args.levels):
            f = 2 ** (args.levels - 1 - lv)
            units += (w // f) * (h // f) / 1e6 * len(srcs) * args.iters
    value = units * args.steps / elapsed
    hit = tot = 0
    for i, (v, _, _, _) in enumerate(refs):
        d = np.asarray(res[i][0])
        gt = ws.gt_depth[v]
        m = gt > 0
        hit += int(np.sum(np.abs(d[m] - gt[m]) < 0.01 * gt[m]))
        tot += int(m.sum())
    # the same references' pyramids one after the other, each its own estimator (no links)
    ests = []
    for v, srcs, dmin, dmax in refs:
        e = pydense.DepthmapEstimator()
        configure(e, ws.reference(len(ests)), args.iters, "f32")
        e.set_hierarchy_levels(args.levels)
        ests.append(e)
    t1 = time.perf_counter()
    for e in ests:
        e.compute()
        e.release_device()
    seq_s = time.perf_counter() - t1
    line = {
        "metric": METRIC, "value": round(value, 3), "unit": UNIT, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": WORKLOADS["cluster"][1].format(views=len(refs), seed=seed, levels=args.levels,
                                                              iters=args.iters),
                   "references": len(refs), "levels": args.levels, "units_per_step": round(units, 3),
                   "value_scope": "DepthmapClusterEstimator.run(): staging, every level, downloads"},
        "sequential": {"scope": "each reference's pyramid through its own DepthmapEstimator, one after the "
                                "other, without links", "s": round(seq_s, 3),
                       "value": round(units / seq_s, 3)},
        "accuracy": {"gt_within_1pct": round(hit / max(tot, 1), 4)},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


def run_c5(args, pydense) -> int:
    """Nadir survey: every view a reference, in groups admitted by device
    memory, each group one batched run followed by Gauss-Newton refinement."""
    import torch
    from opensfm_amd import synthetic
    from opensfm_amd.admission import default_budget, run_admitted

    seed = 5 if args.seed == 2 else args.seed
    storage = args.storage if args.storage_set else "f16"
    views = args.views or 200
    dev = "cuda" if torch.cuda.is_available() else "cpu"
    log(f"rendering aerial survey ({views} views of {args.width}x{args.height}) on {dev} ...")
    t0 = time.perf_counter()
    ws = synthetic.aerial_survey(seed=seed, num_views=views, width=args.width, height=args.height,
                                 neighbours=args.neighbours, device=dev)
    torch.cuda.empty_cache()
    log(f"rendered in {time.perf_counter() - t0:.1f} s")
    units = ws.units(args.iters)
    budget = int(args.hbm_budget_gb * 1e9) if args.hbm_budget_gb > 0 else default_budget(pydense)

    def build(i):
        est = pydense.DepthmapEstimator()
        configure(est, ws.reference(i), args.iters, storage)
        return est

    def keep_depth(i, est):
        return est.download(["depth"])["depth"]

    def group_log(g):
        log(f"group of {len(g.refs)} (refs {g.refs[0]}..{g.refs[-1]}, {g.device_bytes / 1e9:.1f} GB): build "
            f"{g.build_s:.1f} s (waited {g.wait_s:.1f} s), prepare {g.prepare_s:.1f} s, PatchMatch {g.run_ms:.0f} ms, Gauss-Newton "
            f"{g.refine_ms:.0f} ms, results {g.results_s:.1f} s, release {g.release_s:.1f} s, wall {g.wall_s:.1f} s")

    log(f"device budget {budget / 1e9:.1f} GB; one reference needs {build(0).device_bytes() / 1e9:.2f} GB")
    for i in range(args.warmup):        # first use: autotuner, code objects
        run_admitted(pydense, min(2, views), build, budget=budget, refine_steps=args.refine_steps,
                     on_result=lambda i, e: None, overlap_staging=not args.no_overlap)
        log(f"warmup {i + 1}/{args.warmup} done")
    depths = None
    groups = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        depths, groups = run_admitted(pydense, views, build, budget=budget, refine_steps=args.refine_steps,
                                      on_result=keep_depth, on_group=group_log,
                                      overlap_staging=not args.no_overlap)
        log(f"step {k + 1}/{args.steps}: {len(groups)} groups")
    elapsed = time.perf_counter() - t0
    value = units * args.steps / elapsed
    pm_ms = sum(g.run_ms for g in groups)
    gn_ms = sum(g.refine_ms for g in groups)

    hit = hit5 = tot = 0
    for i, d in depths.items():
        gt = ws.gt_depth[ws.refs[i][0]]
        m = gt > 0
        hit += int(np.sum(np.abs(d[m] - gt[m]) < 0.01 * gt[m]))
        hit5 += int(np.sum(np.abs(d[m] - gt[m]) < 0.05 * gt[m]))
        tot += int(m.sum())
    del depths

    cpu, parity = None, None
    if args.cpu_rows > 0:
        log(f"CPU oracle on {args.cpu_rows} rows of reference 0 ...")
        keep = {}
        sc0 = ws.reference(0)
        cpu = cpu_baseline(sc0, args.iters, args.cpu_rows, keep)
        cpu["value"] = round(cpu["value"], 5)
        est = pydense.DepthmapEstimator()
        configure(est, keep["band"], args.iters, storage)
        d = est.compute()["depth"]
        m = keep["depth"] > 0
        parity = {"sample": "the same rows of reference 0 on the GPU (no refinement)",
                  "bit_equal": round(float(np.mean(d == keep["depth"])), 5),
                  "within_1pct_of_oracle": round(float(np.mean(np.abs(d[m] - keep["depth"][m])
                                                               < 0.01 * keep["depth"][m])), 5)}

    pm_tflops = units * 1e6 * FLOP_PER_UNIT / (pm_ms / 1e3) / 1e12
    w, h = args.width, args.height
    line = {
        "metric": METRIC, "value": round(value, 3), "unit": UNIT, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32" if storage == "f32" else "f16-storage/f32",
        "data": "synthetic",
        "config": {"workload": WORKLOADS["c5"][1].format(views=views, width=w, height=h, storage=storage, seed=seed,
                                                         neighbours=args.neighbours, iters=args.iters,
                                                         refine=args.refine_steps),
                   "references": views, "megapixels": round(views * w * h / 1e6, 3),
                   "iterations": args.iters, "patch": PATCH, "units_per_step": round(units, 3),
                   "staging": "serial" if args.no_overlap else "overlapped with the previous group",
                   "value_scope": "whole job: building, staging, batched PatchMatch, Gauss-Newton and depth "
                                  "download of every group"},
        "roofline_path": {"scope": "batched PatchMatch device time of all groups (init, colour passes, decode)",
                          "achieved": round(pm_tflops, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                          "frac": round(pm_tflops / PEAK_FP32_TFLOPS, 4), "device_ms": round(pm_ms, 3)},
        "patchmatch_device_value": round(units / (pm_ms / 1e3), 3),
        "refine": {"steps": args.refine_steps, "ms": round(gn_ms, 3),
                   "value": round(views * w * h / 1e6 * args.neighbours * args.refine_steps / (gn_ms / 1e3), 3)
                   if gn_ms > 0 else None, "unit": "Mpx·views·steps/s"},
        "admission": {"budget_bytes": budget, "groups": len(groups),
                      "per_group": [{"refs": len(g.refs), "device_gb": round(g.device_bytes / 1e9, 2),
                                     "build_s": round(g.build_s, 2), "wait_s": round(g.wait_s, 2),
                                     "prepare_s": round(g.prepare_s, 2),
                                     "run_ms": round(g.run_ms, 1), "refine_ms": round(g.refine_ms, 1),
                                     "results_s": round(g.results_s, 2), "release_s": round(g.release_s, 2),
                                     "wall_s": round(g.wall_s, 2)}
                                    for g in groups]},
        "cpu_baseline": cpu,
        "parity": parity,
        "accuracy": {"scope": "after refinement, all references", "gt_within_1pct": round(hit / max(tot, 1), 4),
                     "gt_within_5pct": round(hit5 / max(tot, 1), 4)},
    }
    if cpu and cpu["value"] > 0:
        line["gpu_over_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(line), flush=True)
    return 0


def copy_probe(mib: int) -> None:
    """One device-to-device copy of ``mib`` MiB through the HIP runtime (a
    plain copy kernel of known size, to scale the HBM traffic counters)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_size_t(mib << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), n) or hip.hipMalloc(ctypes.byref(b), n):
        raise RuntimeError("copy probe: hipMalloc failed")
    hip.hipMemset(a, 1, n)
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(b, a, n, 3):  # hipMemcpyDeviceToDevice
        raise RuntimeError("copy probe: hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy probe: {mib} MiB read + {mib} MiB written")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 5; 1 for c5)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--iters", type=int, default=3)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--storage", choices=["f32", "f16"], default=None,
                    help="image storage (default f32; f16 for the c3 workload)")
    ap.add_argument("--cpu-rows", type=int, default=384,
                    help="reference rows in the CPU-oracle sample, about 10 s on 16 host threads (0 = skip)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="c2",
                    help="c2: the headline single-view workload; c3: many references in one batched launch; "
                         "c4: PatchMatch + Gauss-Newton refinement; c5: aerial survey in memory-admitted groups; "
                         "cluster: the cluster surface with pyramid levels and geometric links")
    ap.add_argument("--refine-steps", type=int, default=10, help="Gauss-Newton steps of the c4 and c5 workloads")
    ap.add_argument("--views", type=int, default=None, help="references of the c3 (64) and c5 (200) workloads")
    ap.add_argument("--levels", type=int, default=3, help="pyramid levels of the cluster workload")
    ap.add_argument("--width", type=int, default=4000, help="c5 image width")
    ap.add_argument("--height", type=int, default=3000, help="c5 image height")
    ap.add_argument("--neighbours", type=int, default=10, help="c5 sources per reference")
    ap.add_argument("--no-overlap", action="store_true",
                    help="c5: stage each group after the previous one instead of while it runs")
    ap.add_argument("--hbm-budget-gb", type=float, default=0.0,
                    help="c5 device budget per group in GB (default: 90%% of the free device memory)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="after the timed steps, run one plain device copy of this size (counter calibration)")
    args = ap.parse_args()
    args.storage_set = args.storage is not None
    if args.steps is None:
        args.steps = 1 if args.workload in ("c5", "cluster") else 5
    if args.storage is None:
        args.storage = "f32"
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path shards by reference "
              f"view with no cross-GPU traffic at one hierarchy level, so N GPUs would run N independent "
              f"copies of this workload.", file=sys.stderr)
        return 2

    from opensfm_amd import native
    pydense = native()
    if not pydense.DepthmapEstimator.is_available():
        raise RuntimeError("no HIP device visible")

    if args.workload == "c3":
        if args.views is None:
            args.views = 64
        return run_c3(args, pydense)
    if args.workload == "c5":
        return run_c5(args, pydense)
    if args.workload == "cluster":
        return run_cluster(args, pydense)
    if args.workload == "c4" and args.seed == 2:
        args.seed = 4
    log("rendering scene ...")
    scene = make_scene(args.seed, args.workload)
    h, w = scene.gt_depth.shape
    n_src = scene.num_src
    units_per_step = w * h / 1e6 * n_src * args.iters

    est = pydense.DepthmapEstimator()
    configure(est, scene, args.iters, args.storage)
    est.upload()
    for i in range(args.warmup):
        est.run()
        log(f"warmup {i + 1}/{args.warmup}: {est.last_timing()['run_ms']:.1f} ms")

    stages = {}
    t0 = time.perf_counter()
    for i in range(args.steps):
        est.run()
        tm = est.last_timing()
        for k, (ms, n) in est.stage_timing().items():
            acc = stages.setdefault(k, [0.0, 0])
            acc[0] += ms
            acc[1] += n
        log(f"step {i + 1}/{args.steps}: {tm['run_ms']:.1f} ms (iterations {tm['sweep_ms']:.1f} ms)")
    elapsed = time.perf_counter() - t0
    value = units_per_step * args.steps / elapsed

    if args.copy_probe_mb > 0:
        copy_probe(args.copy_probe_mb)

    out = est.download()
    # Host-buffer rate: padding/conversion + upload + run + download.
    t1 = time.perf_counter()
    est.compute_patch_match()
    pcie_value = units_per_step / (time.perf_counter() - t1)
    d = out["depth"]
    gt = scene.gt_depth
    acc1 = float(np.mean(np.abs(d - gt) < 0.01 * gt))
    acc5 = float(np.mean(np.abs(d - gt) < 0.05 * gt))
    f16_effect = storage_effect(pydense, scene, args.iters, d) if args.storage == "f16" else None

    # Roofline of the candidates kernel (fixed work per launch): one launch scores the
    # 8 propagated candidates + the current plane of every pixel of one
    # colour, i.e. (W*H/2) x 9 hypotheses x N_src x (21 P + 75) FLOP.
    cand_ms, cand_n = stages["candidates"]
    avg_launch_s = cand_ms / max(cand_n, 1) / 1e3
    flop_per_launch = (w * h / 2.0) * 9 * n_src * (21 * PATCH * PATCH + 75)
    achieved = flop_per_launch / avg_launch_s / 1e12 if avg_launch_s > 0 else 0.0
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                "traffic": (TRAFFIC_PER_LAUNCH if args.workload == "c2" and args.storage == "f32"
                            and est.tiles()[0] == TRAFFIC_TILE else None),
                "kernel": f"candidates_kernel<3,{'float' if args.storage == 'f32' else '_Float16'},{est.tiles()[0]}>",
                "avg_launch_ms": round(avg_launch_s * 1e3, 3), "flop_per_launch": flop_per_launch}
    stage_ms = {k: round(v[0] / args.steps, 3) for k, v in stages.items()}
    # The whole iteration path (all colour-pass kernels, fallbacks included)
    # against the same peak: FLOP_unit x units / device time of the sweeps.
    sweep_ms = sum(v[0] for k, v in stages.items() if k != "init") / args.steps
    path_tflops = units_per_step * 1e6 * FLOP_PER_UNIT / (sweep_ms / 1e3) / 1e12
    roofline_path = {"scope": "all colour-pass kernels of one step", "achieved": round(path_tflops, 3),
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(path_tflops / PEAK_FP32_TFLOPS, 4),
                     "sweep_ms": round(sweep_ms, 3)}

    refine = None
    if args.workload == "c4":
        # Gauss-Newton after PatchMatch (device time of the steps, HIP events)
        gn_ms, acc_b, acc_a = [], [], []
        for i in range(args.steps):
            est.run()
            before = est.download()["depth"]
            est.gauss_newton(args.refine_steps, 1e-3)
            gn_ms.append(est.last_refine_ms())
            after = est.download()["depth"]
            acc_b.append(float(np.mean(np.abs(before - gt) < 0.01 * gt)))
            acc_a.append(float(np.mean(np.abs(after - gt) < 0.01 * gt)))
        ms = float(np.mean(gn_ms))
        refine = {"steps": args.refine_steps, "ms": round(ms, 3),
                  "value": round(w * h / 1e6 * n_src * args.refine_steps / (ms / 1e3), 3),
                  "unit": "Mpx·views·steps/s", "gt_within_1pct_before": round(acc_b[-1], 4),
                  "gt_within_1pct_after": round(acc_a[-1], 4)}
        log(f"Gauss-Newton {args.refine_steps} steps: {ms:.2f} ms")

    cpu, parity = None, None
    if args.cpu_rows > 0:
        log(f"CPU oracle on {args.cpu_rows} rows ...")
        keep = {}
        cpu = cpu_baseline(scene, args.iters, args.cpu_rows, keep, half=args.storage == "f16")
        cpu["value"] = round(cpu["value"], 5)
        # the same rows on the GPU: parity of the measured path with the oracle
        band_est = pydense.DepthmapEstimator()
        configure(band_est, keep["band"], args.iters, args.storage)
        d_band = band_est.compute()["depth"]
        d_o = keep["depth"]
        m = d_o > 0
        parity = {"sample": "the CPU-baseline rows on the GPU, same inputs",
                  "bit_equal": round(float(np.mean(d_band == d_o)), 6),
                  "within_1pct_of_oracle": round(float(np.mean(np.abs(d_band[m] - d_o[m]) < 0.01 * d_o[m])), 6)}

    line = {
        "metric": METRIC, "value": round(value, 3), "unit": UNIT, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32" if args.storage == "f32" else "f16-storage/f32",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[args.workload][1].format(storage=args.storage, seed=args.seed,
                                                                   iters=args.iters),
                   "width": w, "height": h, "sources": n_src, "iterations": args.iters, "patch": PATCH,
                   "units_per_step": units_per_step},
        "roofline": roofline,
        "roofline_path": roofline_path,
        "cpu_baseline": cpu,
        "parity": parity,
        "accuracy": {"gt_within_1pct": round(acc1, 4), "gt_within_5pct": round(acc5, 4)},
        "stage_ms_per_step": stage_ms,
        "pcie_inclusive_value": round(pcie_value, 3),
        "f16_storage_vs_f32": f16_effect,
    }
    if refine is not None:
        line["refine"] = refine
    if cpu and cpu["value"] > 0:
        line["gpu_over_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
