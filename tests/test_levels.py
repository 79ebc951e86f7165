"""Hierarchy levels: image pyramid, coarse-to-fine transfer, multi-level runs.

CPU tests pin the oracle's pyramid (BuildImagePyramid, estimator.cc:727-756),
resize of previous depths (estimator.cc:92-101) and the coarse-to-fine
transfer (acmmp_upsample, opencl_kernels.h:1792-1856 with
UpsampleDepthNormal, estimator.cc:775-791); the GPU tests compare a
multi-level run with the oracle's.
"""
import numpy as np
import pytest

from oracle import pm_oracle as po
from tests.helpers import gpu_estimator, params_for


def _overlap_matrix(n_in, n_out):
    """A[o, i] = length of input cell i inside output cell o's footprint."""
    s = n_in / n_out
    A = np.zeros((n_out, n_in))
    for o in range(n_out):
        lo, hi = o * s, (o + 1) * s
        for i in range(n_in):
            A[o, i] = max(0.0, min(hi, i + 1) - max(lo, i))
    return A


@pytest.mark.parametrize("shape,new", [((8, 12), (4, 6)), ((9, 7), (4, 3)), ((16, 16), (4, 4)), ((5, 11), (2, 5))])
def test_resize_area_is_the_footprint_mean(shape, new):
    rng = np.random.default_rng(1)
    img = rng.random(shape).astype(np.float32)
    h, w = shape
    nh, nw = new
    got = po.resize_area(img, nw, nh)
    Ay, Ax = _overlap_matrix(h, nh), _overlap_matrix(w, nw)
    want = Ay @ img.astype(np.float64) @ Ax.T / (Ay.sum(1)[:, None] * Ax.sum(1)[None, :])
    assert np.max(np.abs(got - want)) < 1e-6
    if h % nh == 0 and w % nw == 0:  # integer factor: plain block mean
        fy, fx = h // nh, w // nw
        blocks = img.astype(np.float64).reshape(nh, fy, nw, fx).mean(axis=(1, 3))
        assert np.max(np.abs(got - blocks)) < 1e-6


def test_resize_area_constant_and_u8_rounding():
    c = np.full((10, 14), 0.37, np.float32)
    assert np.all(po.resize_area(c, 7, 5) == np.float32(0.37))
    u8 = np.random.default_rng(2).integers(0, 256, (10, 14)).astype(np.float32)
    r = po.resize_area(u8, 5, 3, round_int=True)
    assert np.all(r == np.round(r)) and r.min() >= 0 and r.max() <= 255
    exact = po.resize_area(u8, 5, 3)
    assert np.max(np.abs(r - exact)) <= 0.5


def test_resize_linear_reproduces_linear_fields():
    h, w, nh, nw = 12, 16, 24, 20
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    f = (0.3 * xx + 0.7 * yy + 2.0).astype(np.float32)
    out = po.resize_linear(f, nw, nh)
    # output pixel centre (dx + 0.5) * w / nw - 0.5 samples the linear field exactly inside the image
    dy, dx = np.mgrid[0:nh, 0:nw]
    sx = (dx + 0.5) * w / nw - 0.5
    sy = (dy + 0.5) * h / nh - 0.5
    inside = (sx >= 0) & (sx <= w - 1) & (sy >= 0) & (sy <= h - 1)
    want = 0.3 * sx + 0.7 * sy + 2.0
    assert np.max(np.abs(out - want)[inside]) < 1e-5
    assert np.array_equal(po.resize_linear(f, w, h), f)


def test_level_views_scale_sizes_and_intrinsics():
    rng = np.random.default_rng(3)
    imgs = [rng.random((241, 321)).astype(np.float32), rng.integers(0, 256, (241, 321)).astype(np.uint8)]
    K = np.array([[270.0, 0, 160.0], [0, 265.0, 120.0], [0, 0, 1]])
    li, lk = po.level_views(imgs, [K, K], 0.25)
    assert li[0].shape == (60, 80) and li[1].shape == (60, 80)       # int(321 * 0.25), int(241 * 0.25)
    sx, sy = float(np.float32(80) / np.float32(321)), float(np.float32(60) / np.float32(241))
    assert np.allclose(lk[0], [[270 * sx, 0, 160 * sx], [0, 265 * sy, 120 * sy], [0, 0, 1]], rtol=0, atol=1e-12)
    # u8 inputs are resampled as u8 then scaled to [0, 1]
    q = li[1] / np.float32(1.0 / 255.0)
    assert np.max(np.abs(q - np.round(q))) < 1e-3


def test_upsample_transfers_planes_and_reseeds():
    # a coarse level decoded from a known tilted plane: the fine planes must
    # reproduce that plane's depth at every fine pixel; invalid coarse pixels
    # get a random plane in the depth range; costs 2.0, sel 0, rng re-seeded.
    rng = np.random.default_rng(4)
    W, H, cw, ch = 64, 48, 32, 24
    K = np.array([[70.0, 0, 31.5], [0, 72.0, 23.5], [0, 0, 1]])
    imgs = [rng.random((H, W)).astype(np.float32) for _ in range(2)]
    sc = po.OracleScene(imgs, [K, K], [np.eye(3)] * 2, [np.zeros(3), np.array([-0.3, 0, 0])])
    p = po.Params(depth_min=2.0, depth_max=20.0)
    n = np.array([0.2, -0.15, -1.0]); n /= np.linalg.norm(n)
    dplane = 6.0  # n . X + d = 0 with d = 6
    f32 = np.float32
    fx, fy = f32(K[0, 0]) * f32(cw) / f32(W), f32(K[1, 1]) * f32(ch) / f32(H)
    cx, cy = f32(K[0, 2]) * f32(cw) / f32(W), f32(K[1, 2]) * f32(ch) / f32(H)
    ys, xs = np.mgrid[0:ch, 0:cw].astype(np.float64)
    ray = np.stack([(xs - cx) / fx, (ys - cy) / fy, np.ones_like(xs)], -1)
    cdepth = (-dplane / (ray @ n)).astype(np.float32)
    cnorm = np.broadcast_to(n.astype(np.float32), (ch, cw, 3)).copy()
    cdepth[0, 0] = 0.0; cnorm[0, 0] = 0.0   # one invalid coarse pixel
    st = po.OracleState(W, H)
    st.sel[:] = 7
    po.upsample(sc, p, st, cdepth, cnorm)
    assert np.all(st.costs == 2.0) and np.all(st.sel == 0)
    fys, fxs = np.mgrid[0:H, 0:W].astype(np.float64)
    fray = np.stack([(fxs - K[0, 2]) / K[0, 0], (fys - K[1, 2]) / K[1, 1], np.ones_like(fxs)], -1)
    want = -dplane / (fray @ n)
    pl = st.planes.astype(np.float64)
    got = -pl[..., 3] / np.einsum("hwk,hwk->hw", pl[..., :3], fray)
    valid = np.ones((H, W), bool)
    valid[:2, :2] = False                     # fine pixels fed by the invalid coarse pixel
    assert np.max(np.abs(got - want)[valid] / want[valid]) < 1e-5
    assert np.all((got[:2, :2] >= 2.0 * (1 - 1e-5)) & (got[:2, :2] <= 20.0 * (1 + 1e-5)))
    idx = np.arange(W * H, dtype=np.uint64)
    s0 = ((idx * 1099087573 + 2654435769) % 2 ** 32).astype(np.uint32).reshape(H, W)
    assert np.array_equal(st.rng[..., 0][valid], s0[valid])


def test_two_levels_reach_ground_truth(scene_small):
    s = scene_small
    p = params_for(s, patch_size=5)
    d, _, _, _, _ = po.estimate_levels(s.images, s.Ks, s.Rs, s.ts, p, 2)
    gt = s.gt_depth
    assert np.mean(np.abs(d - gt) < 0.01 * gt) > 0.6


@pytest.mark.gpu
@pytest.mark.parametrize("levels", [2, 3])
def test_gpu_levels_match_oracle(pydense, scene_small, levels):
    s = scene_small
    p = params_for(s, patch_size=5)
    est = gpu_estimator(pydense, s, p)
    est.set_hierarchy_levels(levels)
    depth, normal, score, nghbr = est.compute_patch_match()
    d_o, n_o, c_o, _, st = po.estimate_levels(s.images, s.Ks, s.Rs, s.ts, p, levels)
    valid = d_o > 0
    within = np.mean(np.abs(depth[valid] - d_o[valid]) < 0.01 * d_o[valid])
    print(f"levels {levels}: within 1% {within:.6f}, bit-equal {np.mean(depth == d_o):.6f}", flush=True)
    assert within >= 0.99
    assert np.mean(depth == d_o) > 0.99


@pytest.mark.gpu
def test_gpu_levels_u8_input_match_oracle(pydense, scene_small):
    s = scene_small
    u8 = [np.clip(np.round(im * 255), 0, 255).astype(np.uint8) for im in s.images]
    p = params_for(s, patch_size=5)
    est = gpu_estimator(pydense, s, p, images=u8)
    est.set_hierarchy_levels(2)
    depth, _, _, _ = est.compute_patch_match()
    d_o, _, _, _, _ = po.estimate_levels(u8, s.Ks, s.Rs, s.ts, p, 2)
    valid = d_o > 0
    within = np.mean(np.abs(depth[valid] - d_o[valid]) < 0.01 * d_o[valid])
    print(f"u8 levels 2: within 1% {within:.6f}, bit-equal {np.mean(depth == d_o):.6f}", flush=True)
    assert within >= 0.99


@pytest.mark.gpu
@pytest.mark.parametrize("levels", [1, 2])
def test_gpu_postprocessing_matches_oracle(pydense, scene_small, levels):
    # speckle removal + gap filling on every level (ReadBackResults, estimator.cc:1194-1222)
    s = scene_small
    p = params_for(s, patch_size=5, speckle_min_size=40, gap_max_size=5)
    est = gpu_estimator(pydense, s, p)
    est.set_hierarchy_levels(levels)
    out = est.compute()
    d_o, n_o, c_o, conf_o, _ = po.estimate_levels(s.images, s.Ks, s.Rs, s.ts, p, levels)
    d_g = out["depth"]
    print(f"postprocess levels {levels}: bit-equal depth {np.mean(d_g == d_o):.6f} "
          f"conf {np.mean(out['confidence'] == conf_o):.6f} zero-depth {np.mean(d_g == 0):.4f}", flush=True)
    valid = d_o > 0
    assert np.mean(np.abs(d_g[valid] - d_o[valid]) < 0.01 * d_o[valid]) >= 0.99
    assert np.mean((d_g > 0) == valid) >= 0.99
