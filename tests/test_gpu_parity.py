"""HIP path vs the CPU oracle, through the C ABI (pydense module).

Both sides follow the same numerics contract, so the expected agreement is
bit-exact; the asserted bars are the stated tolerances: per-view scores
within 1e-5 absolute, full depth maps within 1 % relative depth on >= 99 %
of the oracle's valid pixels.
"""
import numpy as np
import pytest

from oracle import pm_oracle as po
from tests.helpers import gpu_estimator, half_round, params_for, random_planes

pytestmark = pytest.mark.gpu


def _report(tag, a, b):
    eq = np.mean(a == b)
    print(f"[{tag}] bit-equal fraction {eq:.6f} max|diff| {np.max(np.abs(a.astype(np.float64) - b)):.3g}", flush=True)


@pytest.mark.parametrize("patch", [5, 7])
@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_score_planes_match_oracle(pydense, scene_small, patch, storage):
    print(f"score parity patch={patch} storage={storage}", flush=True)
    p = params_for(scene_small, patch_size=patch)
    imgs = scene_small.images if storage == "f32" else half_round(scene_small.images)
    est = gpu_estimator(pydense, scene_small, p, storage, images=imgs)
    est.upload()
    xy, planes = random_planes(scene_small, 4096, seed=patch)
    g_cost, g_init, g_sel = est.score_planes(xy, planes)
    osc = po.OracleScene(imgs, scene_small.Ks, scene_small.Rs, scene_small.ts)
    o_cost, o_init, o_sel = po.score_planes(osc, p, xy, planes)
    _report("score", g_cost, o_cost)
    assert np.max(np.abs(g_cost - o_cost)) <= 1e-5
    assert np.max(np.abs(g_init - o_init)) <= 1e-5
    assert np.mean(g_sel == o_sel) >= 0.999


def test_init_state_matches_oracle(pydense, scene_small):
    p = params_for(scene_small)
    est = gpu_estimator(pydense, scene_small, p)
    est.upload()
    est.init_state()
    planes, costs, rng, sel = est.read_state()
    osc = po.OracleScene(scene_small.images, scene_small.Ks, scene_small.Rs, scene_small.ts)
    st = po.OracleState(osc.width, osc.height)
    po.random_init(osc, p, st)
    assert np.array_equal(rng, st.rng)
    assert np.array_equal(planes, st.planes)
    _report("init cost", costs, st.costs)
    assert np.max(np.abs(costs - st.costs)) <= 1e-5
    assert np.mean(sel == st.sel) >= 0.999


@pytest.mark.parametrize("color", [0, 1])
def test_single_pass_matches_oracle(pydense, scene_small, color):
    """Same state in, one colour pass, same state out."""
    p = params_for(scene_small)
    osc = po.OracleScene(scene_small.images, scene_small.Ks, scene_small.Rs, scene_small.ts)
    st = po.OracleState(osc.width, osc.height)
    po.random_init(osc, p, st)
    po.sweep(osc, p, st, 0, 0)  # start from a partly converged state
    po.sweep(osc, p, st, 1, 0)
    est = gpu_estimator(pydense, scene_small, p)
    est.upload()
    est.write_state(st.planes, st.costs, st.rng, st.sel)
    est.iterate(1, color)
    planes, costs, rng, sel = est.read_state()
    po.sweep(osc, p, st, color, 1)
    assert np.array_equal(rng, st.rng)
    _report("pass planes", planes, st.planes)
    d_g = planes[..., 3]
    d_o = st.planes[..., 3]
    assert np.mean(np.abs(d_g - d_o) <= 0.01 * np.abs(d_o)) >= 0.99
    assert np.max(np.abs(costs - st.costs)) <= 1e-5 or np.mean(costs == st.costs) >= 0.99


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_full_run_depth_parity(pydense, scene_small, storage):
    p = params_for(scene_small, var_threshold=(1.0 / 255.0) ** 2)
    imgs = scene_small.images if storage == "f32" else half_round(scene_small.images)
    est = gpu_estimator(pydense, scene_small, p, storage, images=imgs)
    out = est.compute()
    d_o, n_o, c_o, conf_o, _ = po.estimate(imgs, scene_small.Ks, scene_small.Rs, scene_small.ts, p)
    d_g = out["depth"]
    valid = d_o > 0
    within = np.abs(d_g[valid] - d_o[valid]) < 0.01 * d_o[valid]
    gt = scene_small.gt_depth
    gt_ok = np.abs(d_g - gt) < 0.01 * gt
    print(f"[full {storage}] bit-equal {np.mean(d_g == d_o):.6f} within1% oracle {within.mean():.6f} "
          f"within1% GT {gt_ok.mean():.4f}", flush=True)
    assert within.mean() >= 0.99
    assert gt_ok.mean() >= 0.8


def test_compute_patch_match_surface(pydense, scene_small):
    """compute_patch_match returns [depth, normal, score, views] with the documented shapes."""
    p = params_for(scene_small)
    est = gpu_estimator(pydense, scene_small, p)
    depth, normal, score, views = est.compute_patch_match()
    h, w = scene_small.gt_depth.shape
    assert depth.shape == (h, w) and depth.dtype == np.float32
    assert normal.shape == (h, w, 3)
    assert score.shape == (h, w)
    assert views.shape == (h, w) and views.dtype == np.uint32
    assert np.all(views < (1 << scene_small.num_src))
    nz = depth > 0
    assert np.allclose(np.linalg.norm(normal[nz], axis=1), 1.0, atol=1e-5)
    assert np.all((depth[nz] >= p.depth_min) & (depth[nz] <= p.depth_max))


def _run_both(pydense, scene, p, prior=None, prev=None):
    """Full single-level run on the GPU and in the oracle; returns both depth maps."""
    est = gpu_estimator(pydense, scene, p)
    h, w = scene.gt_depth.shape
    kw = {}
    if prior is not None:
        est.set_prior(prior[0], prior[1])
        kw.update(prior_planes=prior[0], plane_masks=prior[1])
    if prev is not None:
        est.set_geom_consistency(p.geom_weight, prev)
        stack = np.zeros((scene.num_src, h, w), np.float32)
        mask = 0
        for i, d in enumerate(prev):
            if d is not None:
                stack[i] = d
                mask |= 1 << i
        kw.update(prev_depths=stack, prev_depth_mask=mask)
    out = est.compute()
    sc = po.OracleScene(scene.images, scene.Ks, scene.Rs, scene.ts)
    st = po.OracleState(w, h, **kw)
    po.run(sc, p, st)
    d_o, _, _, _ = po.decode(sc, p, st)
    return out["depth"], d_o


def _assert_parity(tag, d_g, d_o):
    valid = d_o > 0
    within = np.mean(np.abs(d_g[valid] - d_o[valid]) < 0.01 * d_o[valid])
    print(f"[{tag}] bit-equal {np.mean(d_g == d_o):.6f} within1% {within:.6f}", flush=True)
    assert within >= 0.99


def test_prior_run_matches_oracle(pydense, scene_small):
    # planar prior from the ground truth on every third row (estimator.cc:574-614)
    s = scene_small
    h, w = s.gt_depth.shape
    K = np.asarray(s.Ks[0])
    ys, xs = np.mgrid[0:h, 0:w]
    d = np.where(s.gt_depth > 0, s.gt_depth, 10.0).astype(np.float64)
    n = np.array([0.05, -0.05, -1.0]) / np.linalg.norm([0.05, -0.05, -1.0])
    pt = np.stack([d * (xs - K[0, 2]) / K[0, 0], d * (ys - K[1, 2]) / K[1, 1], d], -1)
    planes = np.concatenate([np.broadcast_to(n, (h, w, 3)), -(pt @ n)[..., None]], -1).astype(np.float32)
    masks = np.zeros((h, w), np.uint32)
    masks[::3] = 1
    p = params_for(s, patch_size=5)
    d_g, d_o = _run_both(pydense, s, p, prior=(planes, masks))
    _assert_parity("prior", d_g, d_o)


def test_geom_consistency_run_matches_oracle(pydense, scene_small):
    # previous depth maps in source slots 0 and 2 (SetPreviousDepths, estimator.cc:59-116)
    s = scene_small
    rng = np.random.default_rng(7)
    base = np.where(s.gt_depth > 0, s.gt_depth, 0.0).astype(np.float32)
    prev = [base * np.float32(1.0 + 0.02 * rng.standard_normal()), None,
            (base * (1.0 + 0.01 * rng.standard_normal(base.shape))).astype(np.float32), None]
    p = params_for(s, patch_size=5, geom_weight=0.1)
    d_g, d_o = _run_both(pydense, s, p, prev=prev)
    _assert_parity("geom", d_g, d_o)


@pytest.mark.parametrize("variant", ["checkerboard", "no_census", "patch3_topk2", "no_smooth_no_anchor",
                                     "sigmas"])
def test_parameter_variants_match_oracle(pydense, scene_tiny, variant):
    kw = {"checkerboard": dict(checkerboard_filter=True),
          "no_census": dict(use_census=False),
          "patch3_topk2": dict(patch_size=3, top_k=2),
          "no_smooth_no_anchor": dict(smooth_weight=0.0, anchor_views=0),
          "sigmas": dict(sigma_spatial=3.0, sigma_color=3.0 / 255.0, max_iterations=4)}[variant]
    p = params_for(scene_tiny, **kw)
    d_g, d_o = _run_both(pydense, scene_tiny, p)
    _assert_parity(variant, d_g, d_o)
