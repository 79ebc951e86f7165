"""Shared helpers of the test-suite (estimator set-up on both sides)."""
import numpy as np

from oracle import pm_oracle as po


def params_for(scene, **kw):
    p = po.Params(depth_min=scene.depth_min, depth_max=scene.depth_max)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def gpu_estimator(pydense, scene, params, storage="f32", images=None):
    est = pydense.DepthmapEstimator()
    imgs = scene.images if images is None else images
    for im, K, R, t in zip(imgs, scene.Ks, scene.Rs, scene.ts):
        est.add_view(K, R, t, im)
    est.set_depth_range(params.depth_min, params.depth_max, 0)
    est.set_patchmatch_iterations(params.max_iterations)
    est.set_patch_size(params.patch_size)
    est.set_sigma_spatial(params.sigma_spatial)
    est.set_sigma_color(params.sigma_color)
    est.set_top_k(params.top_k)
    est.set_use_census(params.use_census)
    est.set_smooth_weight(params.smooth_weight)
    est.set_anchor_views(params.anchor_views)
    est.set_checkerboard_filter(params.checkerboard_filter)
    est.set_var_threshold(params.var_threshold)
    est.set_speckle_min_size(params.speckle_min_size)
    est.set_gap_max_size(params.gap_max_size)
    est.set_image_storage(storage)
    return est


def half_round(images):
    return [np.asarray(im, np.float32).astype(np.float16).astype(np.float32) for im in images]


def random_planes(scene, n, seed, spread=0.05):
    """Planes near the ground truth (and some random ones) at random pixels."""
    rng = np.random.default_rng(seed)
    h, w = scene.gt_depth.shape
    xs = rng.integers(0, w, n)
    ys = rng.integers(0, h, n)
    K = scene.Ks[0]
    d = scene.gt_depth[ys, xs].astype(np.float64)
    d = np.where(d > 0, d, 0.5 * (scene.depth_min + scene.depth_max))
    d = d * (1.0 + rng.uniform(-spread, spread, n))
    rnd = rng.random(n) < 0.3
    d[rnd] = np.exp(rng.uniform(np.log(scene.depth_min), np.log(scene.depth_max), rnd.sum()))
    nrm = np.stack([rng.uniform(-0.5, 0.5, n), rng.uniform(-0.5, 0.5, n), -np.ones(n)], 1)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    pt = np.stack([d * (xs - K[0, 2]) / K[0, 0], d * (ys - K[1, 2]) / K[1, 1], d], 1)
    planes = np.concatenate([nrm, -(nrm * pt).sum(1, keepdims=True)], 1).astype(np.float32)
    return np.stack([xs, ys], 1).astype(np.int32), planes
