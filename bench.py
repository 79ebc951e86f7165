"""Benchmark of the PatchMatch depth path on one MI355X.

Default workload: the aerial survey -- 200 nadir views of 4000x3000 over a
seeded fractal terrain, every view a reference with its 10 nearest views as
sources, float16 image storage, patch 7x7, 3 iterations, then 10
Gauss-Newton steps of the differentiable cost per reference.  The views are
rendered on the GPU and stay in HBM.  References run in groups admitted by
device memory; one "step" is 10 references, taken in order (--steps 20 is
the whole survey): building their estimators, staging their views on the
device (the reference's UploadData: padding, conversion, quads, reference
taps), the batched PatchMatch (init, iterations, decode), Gauss-Newton and
the depth download, in groups of as many references as device memory holds.
Metric: Mpix*views*iters/s = sum over references of (W*H/1e6) * N_src *
iterations / seconds.

Also reported in the same line: the roofline of the candidate-scoring kernel
(fixed algorithmic work per launch; HIP events around every launch on the
batch stream, over the timed steps), the whole PatchMatch against the FP32
vector peak, per-stage device time, the rate from host buffers (PCIe
included, never the headline), accuracy against the synthetic ground truth,
the CPU oracle on a bounded band of reference 0 (host cores) and the GPU's
parity with it, and the plain single-view float32 path (1600x1200, 8
neighbours; upload timed per step) with its own CPU baseline.  --workload
c2/c3/c4/cluster run the other workloads alone.  Every run call returns
after its HIP stream is idle, so the wall clock brackets complete GPU work.
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
# The reference scores 15 hypotheses per pixel and iteration (8 candidates,
# the current plane, 6 refinements); the current plane's per-view costs are
# carried over from its last scoring here, so 14 are scored.
CAND_HYPS = 8                     # hypotheses one candidates launch scores
FLOP_PER_UNIT = (CAND_HYPS + 6) * (21 * PATCH * PATCH + 75)   # x (21 FLOP/tap x P taps + 75)
# HBM bytes per candidates launch: rocprofv3 --pmc FETCH_SIZE (x 1.9999) +
# WRITE_SIZE (x 1.0), each factor read off the same counter on a 1 GiB
# device-to-device copy (profiles/*_fetch.txt, profiles/*_write.txt).
# candidates_kernel<3, float, 4>, 1600x1200 with 8 sources:
TRAFFIC_PER_LAUNCH = 1.245e9
TRAFFIC_TILE = 4
# candidates_batch<3, _Float16, 16> per 4000x3000 reference with 10 sources
# in the launch (a launch covers every reference of an admitted group):
TRAFFIC_PER_REF_C5 = 1.012e10


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


def configure(est, scene, iters: int, storage: str = "f32", images=None):
    """The workload's estimator: the dense pipeline's parameters at one
    pyramid level (the workloads are single-level), patch 7x7, the scene's
    depth range, no post-processing; views from host memory (``images``
    overrides the scene's; [] adds none)."""
    from opensfm_amd import native
    imgs = scene.images if images is None else images
    for im, K, R, t in zip(imgs, scene.Ks, scene.Rs, scene.ts):
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


def host_cpu() -> str:
    """The host CPU's model name (/proc/cpuinfo) and its logical CPU count."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model}, {os.cpu_count()} logical CPUs visible"


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
    return {"value": units / dt, "unit": UNIT, "cores": po.max_threads(), "kind": "port", "cpu": host_cpu(),
            "sample": f"reference rows {y0}..{y0 + rows} of the same {sc.width}x{scene.gt_depth.shape[0]} scene "
                      f"({sc.width}x{rows} px, "
                      f"{band.num_src} sources, {iters} iterations, init+iterations+decode), {dt:.1f} s"}


def fp64_error(pydense, band, iters: int) -> dict:
    """Depth of ``band`` run on the GPU with float32 and with float16 image
    storage, against the FP64 oracle on the same band (the float32 texels
    widened exactly, float64 cameras, double arithmetic throughout)."""
    from oracle import pm_oracle as po
    p = po.Params(patch_size=PATCH, depth_min=band.depth_min, depth_max=band.depth_max, max_iterations=iters)
    imgs = [np.asarray(im, np.float32) for im in band.images]
    t0 = time.perf_counter()
    d64, _, _, _ = po.estimate_f64(imgs, band.Ks, band.Rs, band.ts, p)
    dt = time.perf_counter() - t0
    h, w = d64.shape
    m = d64 > 0
    out = {"reference": f"FP64 oracle on the CPU-baseline rows ({w}x{h}, {band.num_src} sources, {iters} "
                        f"iterations), {dt:.1f} s", "valid_pixels": int(m.sum())}
    for storage in ("f32", "f16"):
        e = pydense.DepthmapEstimator()
        configure(e, band, iters, storage)
        d = e.compute()["depth"]
        rel = np.abs(d[m] - d64[m]) / d64[m]
        out[storage] = {"within_1pct": round(float(np.mean(rel < 0.01)), 5),
                        "median_rel": float(f"{np.median(rel):.3g}"), "mean_rel": float(f"{np.mean(rel):.3g}")}
    log(f"FP64 depth error: f32 {out['f32']}, f16 {out['f16']}")
    return out


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
    cpu, parity, fp64 = None, None, None
    if args.cpu_rows > 0:
        log(f"CPU oracle on reference {small} ...")
        p = po.Params(patch_size=PATCH, depth_min=sc.depth_min, depth_max=sc.depth_max, max_iterations=args.iters)
        imgs = [im.astype(np.float16).astype(np.float32) for im in sc.images] if storage == "f16" else sc.images
        t1 = time.perf_counter()
        d_o, _, _, _, _ = po.estimate(imgs, sc.Ks, sc.Rs, sc.ts, p)
        dt = time.perf_counter() - t1
        hh, ww = sc.gt_depth.shape
        u = ww * hh / 1e6 * sc.num_src * args.iters
        cpu = {"value": round(u / dt, 5), "unit": UNIT, "cores": po.max_threads(), "kind": "port", "cpu": host_cpu(),
               "sample": f"reference {small} of the walk ({ww}x{hh}, {sc.num_src} sources, {args.iters} iterations, "
                         f"init+iterations+decode), {dt:.1f} s"}
        m = d_o > 0
        parity = {"reference": small, "within_1pct_of_oracle": round(float(np.mean(
            np.abs(outs[small][m] - d_o[m]) < 0.01 * d_o[m])), 5),
            "bit_equal": round(float(np.mean(outs[small] == d_o)), 5)}
        fp64 = fp64_error(pydense, sc, args.iters)
        fp64["reference"] += f" (reference {small})"

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
        "fp64_depth_error": fp64,
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


def _device_views(est, images, Ks, Rs, ts, storage: str):
    """Adds device-resident views (torch tensors, (H, W) contiguous) to ``est``."""
    for im, K, R, t in zip(images, Ks, Rs, ts):
        dt = {2: "f16", 4: "f32", 1: "u8"}[im.element_size()]
        est.add_view_device(K, R, t, im.data_ptr(), dt, im.shape[1], im.shape[0], 0)


def plain_leg(pydense, steps: int, warmup: int, cpu_rows: int, seed: int = 2, iters: int = 3) -> dict:
    """The plain single-view float32 path on the 1600x1200 facet heightfield
    with 8 neighbours: views resident in HBM, each step = upload (padding,
    quads, reference taps) + init + iterations + decode; CPU oracle on a band."""
    import torch
    log("plain single-view leg: rendering the 1600x1200 facet heightfield ...")
    scene = make_scene(seed, "c2")
    dev = [torch.from_numpy(np.asarray(im, np.float32)).cuda() for im in scene.images]
    torch.cuda.synchronize()
    h, w = scene.gt_depth.shape
    units = w * h / 1e6 * scene.num_src * iters
    est = pydense.DepthmapEstimator()
    configure(est, scene, iters, "f32", images=[])
    _device_views(est, dev, scene.Ks, scene.Rs, scene.ts, "f32")
    for _ in range(max(warmup, 1)):
        est.upload()
        est.run()
    t0 = time.perf_counter()
    cand = [0.0, 0]
    for i in range(steps):
        est.upload()
        est.run()
        ms, n = est.stage_timing()["candidates"]
        cand[0] += ms
        cand[1] += n
    elapsed = time.perf_counter() - t0
    value = units * steps / elapsed
    d = est.download(["depth"])["depth"]
    gt = scene.gt_depth
    out = {"workload": "plain PatchMatch, one 1600x1200 f32 reference + 8 neighbours, 64-facet heightfield "
                       f"(seed {seed}), patch 7x7, {iters} iterations; per step upload (views resident in HBM) "
                       "+ init + iterations + decode",
           "value": round(value, 3), "ms_per_step": round(elapsed / steps * 1e3, 3), "steps": steps,
           "units_per_step": units,
           "candidates_avg_launch_ms": round(cand[0] / max(cand[1], 1), 4),
           "accuracy": {"gt_within_1pct": round(float(np.mean(np.abs(d - gt) < 0.01 * gt)), 4)}}
    if cpu_rows > 0:
        log(f"plain leg: CPU oracle on {cpu_rows} rows ...")
        keep = {}
        cpu = cpu_baseline(scene, iters, cpu_rows, keep)
        cpu["value"] = round(cpu["value"], 5)
        band = pydense.DepthmapEstimator()
        configure(band, keep["band"], iters, "f32")
        d_band = band.compute()["depth"]
        m = keep["depth"] > 0
        out["cpu_baseline"] = cpu
        out["gpu_over_cpu"] = round(value / cpu["value"], 1)
        out["parity"] = {"sample": "the CPU-baseline rows on the GPU", "bit_equal": round(float(np.mean(
            d_band == keep["depth"])), 6), "within_1pct_of_oracle": round(float(np.mean(
                np.abs(d_band[m] - keep["depth"][m]) < 0.01 * keep["depth"][m])), 6)}
        out["fp64_depth_error"] = fp64_error(pydense, keep["band"], iters)
    del dev
    torch.cuda.empty_cache()
    return out


def run_c5(args, pydense) -> int:
    """Nadir survey: every view a reference, in groups admitted by device
    memory, each group one batched run followed by Gauss-Newton refinement.

    The 200 views are rendered on the GPU and stay there (float16 tensors);
    a step is a fixed number of references (10), in order, so 20 steps run
    the whole 200-view survey; they run in groups admitted by device memory
    (a group's references are built, their views staged on the device --
    padding, quads, reference taps: the reference's UploadData --, run as
    one batch, refined by Gauss-Newton, and their depths downloaded)."""
    import torch
    from opensfm_amd import synthetic
    from opensfm_amd.admission import default_budget, run_admitted

    seed = 5 if args.seed == 2 else args.seed
    storage = args.storage if args.storage_set else "f16"
    views = args.views or 200
    w, h = args.width, args.height
    log(f"rendering aerial survey ({views} views of {w}x{h}) on the GPU, kept in HBM ...")
    t0 = time.perf_counter()
    ws = synthetic.aerial_survey(seed=seed, num_views=views, width=w, height=h, neighbours=args.neighbours,
                                 device="cuda", images_on_device=True)
    if storage == "f32":
        ws.images = [im.float().contiguous() for im in ws.images]
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    log(f"rendered in {time.perf_counter() - t0:.1f} s")

    def build_at(i):
        v, srcs, dmin, dmax = ws.refs[i % views]
        order = [v] + list(srcs)
        est = pydense.DepthmapEstimator()
        est.set_params({**pydense.pipeline_params(), "hierarchy_levels": 1, "max_iterations": args.iters,
                        "patch_size": PATCH, "image_storage": storage, "speckle_min_size": 0, "gap_max_size": 0})
        est.set_depth_range(dmin, dmax, 0)
        _device_views(est, [ws.images[j] for j in order], [ws.Ks[j] for j in order], [ws.Rs[j] for j in order],
                      [ws.ts[j] for j in order], storage)
        return est

    budget = int(args.hbm_budget_gb * 1e9) if args.hbm_budget_gb > 0 else default_budget(pydense)
    ref_bytes = build_at(0).device_bytes()
    group = args.refs_per_step
    per_ref_units = w * h / 1e6 * args.neighbours * args.iters
    log(f"device budget {budget / 1e9:.1f} GB; one reference needs {ref_bytes / 1e9:.2f} GB; a step is "
        f"{group} references")

    def keep_depth(i, est):
        return est.download(["depth"])["depth"]

    def group_log(g):
        log(f"group of {len(g.refs)} (refs {g.refs[0]}..{g.refs[-1]}): build {g.build_s:.1f} s (waited "
            f"{g.wait_s:.1f} s), staging {g.prepare_s:.1f} s, PatchMatch {g.run_ms:.0f} ms, Gauss-Newton "
            f"{g.refine_ms:.0f} ms, results {g.results_s:.1f} s, release {g.release_s:.1f} s, wall {g.wall_s:.1f} s")

    cfg = {"use_graph": False}     # per-launch HIP events on the batch stream (the roofline below)
    if args.tiles:
        cfg["cost_tile_x"], cfg["pixel_tile_x"] = args.tiles
    for i in range(args.warmup):  # first use: autotuner, code objects
        run_admitted(pydense, group, build_at, budget=budget, refine_steps=args.refine_steps,
                     on_result=lambda i, e: None, overlap_staging=not args.no_overlap, batch_config=cfg)
        log(f"warmup {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    depths, groups = run_admitted(pydense, group * args.steps, build_at, budget=budget,
                                  refine_steps=args.refine_steps, on_result=keep_depth, on_group=group_log,
                                  overlap_staging=not args.no_overlap, batch_config=cfg)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    if args.copy_probe_mb > 0:
        copy_probe(args.copy_probe_mb)
    nrefs = group * args.steps
    units = per_ref_units * nrefs
    value = units / elapsed
    if len(groups) != args.steps:
        log(f"{len(groups)} admitted groups ran the {args.steps} steps")
    pm_ms = sum(g.run_ms for g in groups)
    gn_ms = sum(g.refine_ms for g in groups)
    stages = {}
    for g in groups:
        for k, (ms, n) in g.stage_ms.items():
            acc = stages.setdefault(k, [0.0, 0])
            acc[0] += ms
            acc[1] += n

    hit = hit5 = tot = 0
    for i, d in depths.items():
        gt = ws.gt_depth[ws.refs[i % views][0]]
        m = gt > 0
        hit += int(np.sum(np.abs(d[m] - gt[m]) < 0.01 * gt[m]))
        hit5 += int(np.sum(np.abs(d[m] - gt[m]) < 0.05 * gt[m]))
        tot += int(m.sum())
    del depths

    # Roofline of the candidate-scoring kernel (fixed algorithmic work per
    # launch): one batched launch scores the 8 candidates of every pixel of
    # one colour of every reference of the group (the current plane's costs
    # are carried over), 21 FLOP per tap and 75 per (hypothesis, view).  Launch times: HIP events on the batch stream
    # around every launch, over the timed steps.
    # A launch covers one colour of every reference of an admitted group.
    # Groups differ in size (the last one is smaller), so the roofline takes
    # the launches of the groups of the most common size: one launch scores
    # refs x (W*H/2) pixels x 8 hypotheses x N sources.
    sizes = [len(g.refs) for g in groups]
    modal = max(sorted(set(sizes)), key=sizes.count)
    cand_ms = sum(g.stage_ms.get("candidates", (0.0, 0))[0] for g in groups if len(g.refs) == modal)
    cand_n = sum(g.stage_ms.get("candidates", (0.0, 0))[1] for g in groups if len(g.refs) == modal)
    flop_per_launch = modal * (w * h / 2.0) * CAND_HYPS * args.neighbours * (21 * PATCH * PATCH + 75)
    avg_launch_s = cand_ms / max(cand_n, 1) / 1e3
    achieved = flop_per_launch / avg_launch_s / 1e12 if avg_launch_s > 0 else 0.0
    sweep_ms = sum(v[0] for k, v in stages.items() if k != "init")
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                "traffic": (TRAFFIC_PER_REF_C5 * modal if storage == "f16" and (w, h) == (4000, 3000)
                            and args.neighbours == 10 else None),
                "kernel": f"candidates_batch<3,{'_Float16' if storage == 'f16' else 'float'},*>",
                "references_per_launch": modal, "launches": cand_n,
                "avg_launch_ms": round(avg_launch_s * 1e3, 3), "flop_per_launch": flop_per_launch}
    path_tflops = units * 1e6 * FLOP_PER_UNIT / (pm_ms / 1e3) / 1e12 if pm_ms > 0 else 0.0

    # the same job from host buffers (PCIe and host staging included): one group
    log("one group from host buffers ...")
    hviews = {}

    def build_host(i):
        v, srcs, dmin, dmax = ws.refs[i % views]
        est = pydense.DepthmapEstimator()
        est.set_params({**pydense.pipeline_params(), "hierarchy_levels": 1, "max_iterations": args.iters,
                        "patch_size": PATCH, "image_storage": storage, "speckle_min_size": 0, "gap_max_size": 0})
        est.set_depth_range(dmin, dmax, 0)
        for j in [v] + list(srcs):
            if j not in hviews:
                hviews[j] = ws.images[j].cpu().numpy()
            est.add_view(ws.Ks[j], ws.Rs[j], ws.ts[j], hviews[j])
        return est

    t1 = time.perf_counter()
    run_admitted(pydense, group, build_host, budget=budget, refine_steps=args.refine_steps, on_result=keep_depth,
                 overlap_staging=False)
    pcie_value = per_ref_units * group / (time.perf_counter() - t1)
    hviews.clear()

    cpu, parity, fp64 = None, None, None
    if args.cpu_rows > 0:
        log(f"CPU oracle on {args.cpu_rows} rows of reference 0 ...")
        v, srcs, dmin, dmax = ws.refs[0]
        order = [v] + list(srcs)
        sc0 = synthetic.Scene([ws.images[j].float().cpu().numpy() for j in order], [ws.Ks[j] for j in order],
                              [ws.Rs[j] for j in order], [ws.ts[j] for j in order], ws.gt_depth[v], dmin, dmax,
                              "aerial reference 0")
        keep = {}
        cpu = cpu_baseline(sc0, args.iters, args.cpu_rows, keep)
        cpu["value"] = round(cpu["value"], 5)
        est = pydense.DepthmapEstimator()
        configure(est, keep["band"], args.iters, storage)
        d = est.compute()["depth"]
        m = keep["depth"] > 0
        parity = {"sample": "the CPU-baseline rows of reference 0 on the GPU (no refinement)",
                  "bit_equal": round(float(np.mean(d == keep["depth"])), 5),
                  "within_1pct_of_oracle": round(float(np.mean(np.abs(d[m] - keep["depth"][m])
                                                               < 0.01 * keep["depth"][m])), 5)}
        fp64 = fp64_error(pydense, keep["band"], args.iters)
    del ws
    torch.cuda.empty_cache()

    plain = None
    if not args.no_plain_leg:
        plain = plain_leg(pydense, steps=5, warmup=1, cpu_rows=192 if args.cpu_rows > 0 else 0)

    line = {
        "metric": METRIC, "value": round(value, 3), "unit": UNIT, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32" if storage == "f32" else "f16-storage/f32",
        "data": "synthetic (rendered on the GPU, resident in HBM before the timed region)",
        "config": {"workload": WORKLOADS["c5"][1].format(views=views, width=w, height=h, storage=storage, seed=seed,
                                                         neighbours=args.neighbours, iters=args.iters,
                                                         refine=args.refine_steps),
                   "references": nrefs, "references_per_step": group, "admitted_groups": len(groups),
                   "megapixels": round(nrefs * w * h / 1e6, 3),
                   "iterations": args.iters, "patch": PATCH, "units_per_step": round(per_ref_units * group, 3),
                   "staging": "serial" if args.no_overlap else "overlapped with the previous group",
                   "value_scope": "per step: building the group's estimators, staging its device-resident views "
                                  "(padding, f16 conversion, quads, reference taps), batched PatchMatch (init, "
                                  "iterations, decode), Gauss-Newton, depth download"},
        "roofline": roofline,
        "roofline_path": {"scope": "batched PatchMatch device time of all groups (init, colour passes, decode)",
                          "achieved": round(path_tflops, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                          "frac": round(path_tflops / PEAK_FP32_TFLOPS, 4), "device_ms": round(pm_ms, 3),
                          "colour_pass_ms": round(sweep_ms, 3)},
        "cpu_baseline": cpu,
        "parity": parity,
        "fp64_depth_error": fp64,
        "patchmatch_device_value": round(units / (pm_ms / 1e3), 3) if pm_ms > 0 else None,
        "stage_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in stages.items()},
        "refine": {"steps": args.refine_steps, "ms": round(gn_ms, 3),
                   "value": round(nrefs * w * h / 1e6 * args.neighbours * args.refine_steps / (gn_ms / 1e3), 3)
                   if gn_ms > 0 else None, "unit": "Mpx·views·steps/s"},
        "pcie_inclusive_value": round(pcie_value, 3),
        "admission": {"budget_bytes": budget, "groups": len(groups), "reference_bytes": ref_bytes,
                      "per_group_wall_s": [round(g.wall_s, 2) for g in groups]},
        "accuracy": {"scope": "after refinement, all references", "gt_within_1pct": round(hit / max(tot, 1), 4),
                     "gt_within_5pct": round(hit5 / max(tot, 1), 4)},
    }
    if cpu and cpu["value"] > 0:
        line["gpu_over_cpu"] = round(value / cpu["value"], 1)
    if plain is not None:
        line["plain_single_view"] = plain
        if cpu and plain.get("gpu_over_cpu"):
            line["speedup_ratio_vs_plain"] = {
                "scope": "(this workload's GPU/CPU) / (plain single-view f32 GPU/CPU)",
                "value": round(line["gpu_over_cpu"] / plain["gpu_over_cpu"], 3)}
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
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: c5 6 groups, 20 = the whole survey; cluster 1; others 5)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--iters", type=int, default=3)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--storage", choices=["f32", "f16"], default=None,
                    help="image storage (default f32; f16 for the c3 workload)")
    ap.add_argument("--cpu-rows", type=int, default=None,
                    help="reference rows in the CPU-oracle sample (default 128 for c5, the whole frame for c2 "
                         "and c4, else 384; 10-30 s on 16 host threads; 0 = skip)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="c5",
                    help="c5 (default): the aerial survey, batched f16 PatchMatch + Gauss-Newton in memory-admitted "
                         "groups, with the plain single-view leg; c2: the plain single-view workload alone; c3: "
                         "many references in one batched launch; c4: PatchMatch + Gauss-Newton refinement; "
                         "cluster: the cluster surface with pyramid levels and geometric links")
    ap.add_argument("--no-plain-leg", action="store_true", help="c5: skip the plain single-view leg")
    ap.add_argument("--refs-per-step", type=int, default=10, help="c5: references per timed step")
    ap.add_argument("--tiles", type=lambda v: tuple(int(x) for x in v.split(",")), default=None,
                    help="c5: cost and pixel tile widths of the batches, e.g. 4,32 (default: the autotuner's)")
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
        args.steps = {"c5": 6, "cluster": 1}.get(args.workload, 5)
    if args.cpu_rows is None:  # c2 and c4: the whole frame (1600x1200, 1024x768)
        args.cpu_rows = {"c5": 128, "c4": 768, "c2": 1200}.get(args.workload, 384)
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

    # views resident in HBM; each step stages them (padding, conversion,
    # quads, reference taps) and runs init + iterations + decode
    import torch
    dev_views = [torch.from_numpy(np.asarray(im, np.float32)).cuda() for im in scene.images]
    torch.cuda.synchronize()
    est = pydense.DepthmapEstimator()
    configure(est, scene, args.iters, args.storage, images=[])
    _device_views(est, dev_views, scene.Ks, scene.Rs, scene.ts, args.storage)
    for i in range(args.warmup):
        est.upload()
        est.run()
        log(f"warmup {i + 1}/{args.warmup}: {est.last_timing()['run_ms']:.1f} ms")

    stages = {}
    t0 = time.perf_counter()
    for i in range(args.steps):
        est.upload()
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
    # Host-buffer rate: host padding/conversion + PCIe upload + run + download.
    host_est = pydense.DepthmapEstimator()
    configure(host_est, scene, args.iters, args.storage)
    host_est.compute_patch_match()
    t1 = time.perf_counter()
    host_est.compute_patch_match()
    pcie_value = units_per_step / (time.perf_counter() - t1)
    del host_est
    d = out["depth"]
    gt = scene.gt_depth
    acc1 = float(np.mean(np.abs(d - gt) < 0.01 * gt))
    acc5 = float(np.mean(np.abs(d - gt) < 0.05 * gt))
    f16_effect = storage_effect(pydense, scene, args.iters, d) if args.storage == "f16" else None

    # Roofline of the candidates kernel (fixed work per launch): one launch scores the
    # 8 propagated candidates of every pixel of one colour (the current
    # plane's costs are carried over), i.e. (W*H/2) x 8 hypotheses x N_src x
    # (21 P + 75) FLOP.
    cand_ms, cand_n = stages["candidates"]
    avg_launch_s = cand_ms / max(cand_n, 1) / 1e3
    flop_per_launch = (w * h / 2.0) * CAND_HYPS * n_src * (21 * PATCH * PATCH + 75)
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

    cpu, parity, fp64 = None, None, None
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
        fp64 = fp64_error(pydense, keep["band"], args.iters)

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
        "fp64_depth_error": fp64,
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
