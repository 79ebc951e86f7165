"""The textureless fallback on the GPU against the oracle, on scenes built so
that the strided NCC decides the cost.

The scenes of tests/test_oracle_strided.py (reference pixels with a flat
stride-1 window and a textured stride-2 or stride-3 lattice, sources seen
through rolled and zoomed affine warps, and an integer-disparity stereo pair)
send nearly every (pixel, hypothesis, view) through the strided NCC or census:
  * pm_score_planes runs the fallback inline (the init's path);
  * a full PatchMatch run sends every colour pass's pending views through
    the fallback kernel (per (view, tile) waves, reference texels in LDS).
The bar is the parity bar of the other GPU tests: scores within 1e-5 and
depth within 1 % on >= 99 % of the oracle's valid pixels; both sides follow
one numerics contract, so bit-equality is expected and reported.
"""
import numpy as np
import pytest

from oracle import pm_oracle as po
from tests.helpers import gpu_estimator
from tests.test_oracle_strided import SIGMA_C, SIGMA_S, _K, _decimation_case

pytestmark = pytest.mark.gpu


def _stereo_case(k, patch):
    fine, _, _ = _decimation_case(11 * k, k, patch, 0.05)
    ref = fine["images"][0]
    H, W = ref.shape
    f, b, disp = 80.0, 0.25, 4
    K = _K(f, (W - 1) / 2.0, (H - 1) / 2.0)
    src = np.full_like(ref, ref[0, 0])
    src[:, :W - disp] = ref[:, disp:]
    return dict(images=[ref, src], Ks=[K, K], Rs=[np.eye(3)] * 2, ts=[np.zeros(3), np.array([-b, 0.0, 0.0])],
                xy=fine["xy"], depth=f * b / disp)


class _Scene:
    """The attributes tests.helpers.gpu_estimator reads."""

    def __init__(self, d, dmin, dmax):
        self.images, self.Ks, self.Rs, self.ts = d["images"], d["Ks"], d["Rs"], d["ts"]
        self.depth_min, self.depth_max = dmin, dmax


def _params(patch, **kw):
    p = po.Params(patch_size=patch, depth_min=2.0, depth_max=12.0, sigma_spatial=SIGMA_S, sigma_color=SIGMA_C)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _cases():
    out = []
    for k, patch in ((2, 7), (2, 5), (3, 5)):
        for seed in (0, 1):
            fine, _, _ = _decimation_case(500 + 10 * k + seed, k, patch, 0.04)
            out.append((f"decimation k={k} p={patch} seed={seed}", fine, patch))
        out.append((f"stereo k={k} p={patch}", _stereo_case(k, patch), patch))
    return out


@pytest.mark.parametrize("census", [True, False])
def test_strided_scores_match_oracle(pydense, census):
    rng = np.random.default_rng(3)
    for name, case, patch in _cases():
        print(f"strided scores: {name} census={census}", flush=True)
        sc = _Scene(case, 2.0, 12.0)
        p = _params(patch, use_census=census)
        est = gpu_estimator(pydense, sc, p)
        est.upload()
        # every pixel of the designed lattice, at depths around the scene's
        # plane, fronto-parallel and tilted
        xy = np.repeat(case["xy"], 12, axis=0).astype(np.int32)
        z = rng.uniform(3.0, 8.0, len(xy))
        nrm = np.stack([rng.uniform(-0.3, 0.3, len(xy)), rng.uniform(-0.3, 0.3, len(xy)), -np.ones(len(xy))], 1)
        nrm[::3] = (0.0, 0.0, -1.0)
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
        K = sc.Ks[0]
        pts = np.stack([z * (xy[:, 0] - K[0, 2]) / K[0, 0], z * (xy[:, 1] - K[1, 2]) / K[1, 1], z], 1)
        planes = np.concatenate([nrm, -(nrm * pts).sum(1, keepdims=True)], 1).astype(np.float32)
        g_cost, g_init, g_sel = est.score_planes(xy, planes)
        osc = po.OracleScene(sc.images, sc.Ks, sc.Rs, sc.ts)
        o_cost, o_init, o_sel = po.score_planes(osc, p, xy, planes)
        # every scored pixel has a flat stride-1 window (its variance is 0), so
        # every cost comes from the strided NCC or census
        ref = sc.images[0]
        hp = patch // 2
        assert all(np.ptp(ref[y - hp:y + hp + 1, x - hp:x + hp + 1]) == 0 for x, y in case["xy"])
        print(f"  bit-equal {np.mean(g_cost == o_cost):.6f}, max|diff| {np.max(np.abs(g_cost - o_cost)):.2e}, "
              f"census (1.0) {np.mean(o_cost == 1.0):.2f}, 2.0 {np.mean(o_cost == 2.0):.2f}", flush=True)
        assert np.max(np.abs(g_cost - o_cost)) <= 1e-5
        assert np.max(np.abs(g_init - o_init)) <= 1e-5
        assert np.array_equal(g_sel, o_sel)


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_strided_scenes_full_run_matches_oracle(pydense, storage):
    for name, case, patch in _cases():
        print(f"strided full run: {name} storage={storage}", flush=True)
        imgs = case["images"]
        if storage == "f16":
            imgs = [im.astype(np.float16).astype(np.float32) for im in imgs]
        sc = _Scene(dict(case, images=imgs), 2.0, 12.0)
        p = _params(patch)
        est = gpu_estimator(pydense, sc, p, storage=storage)
        r = est.compute()
        d_o, _, c_o, _, _ = po.estimate(imgs, sc.Ks, sc.Rs, sc.ts, p)
        valid = d_o > 0
        within = np.mean(np.abs(r["depth"][valid] - d_o[valid]) < 0.01 * d_o[valid])
        print(f"  bit-equal {np.mean(r['depth'] == d_o):.6f}, within 1% {within:.6f}", flush=True)
        assert within >= 0.99
