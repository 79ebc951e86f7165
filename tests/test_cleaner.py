"""Depth-map cleaner (GPUDepthmapCleaner, cleaner.cc:95-242; kernel
acmmp_clean_depthmap, opencl_kernels.h:2116-2279).

The oracle's restatement is pinned on a fronto-parallel plane seen by
sideways-translated cameras, where the projection of every reference pixel
into a neighbour is a known shift, so the number of consistent views per
pixel has a closed form; floaters check the carving vote and the budget
formula; normals check the grazing rule; a neighbour given at another size
checks the resize to the reference size.  The GPU kernel must then match
the oracle bit for bit on noisy depth maps of a rendered scene.
"""
import numpy as np
import pytest

from oracle import pm_oracle as po

W, H, F = 48, 32, 40.0
K = np.array([[F, 0, (W - 1) / 2], [0, F, (H - 1) / 2], [0, 0, 1.0]])


def _cam(bx):
    """Camera at (bx, 0, 0) looking down +z: R = I, t = -C."""
    return K, np.eye(3), np.array([-bx, 0.0, 0.0])


def _shift_baseline(k):
    # a neighbour whose projection of the z = 10 plane is x - (k + 0.5)
    return (k + 0.5) * 10.0 / F


def _expected_counts(ks):
    cnt = np.ones((H, W), int)
    x = np.arange(W)[None, :]
    y = np.arange(H)[:, None]
    for k in ks:
        cnt += ((x >= k + 1) & (x <= W + k - 1) & (y <= H - 2)).astype(int)
    return cnt


def _plane_views(ks, ref_depth=None):
    ref = np.full((H, W), 10.0, np.float32) if ref_depth is None else ref_depth
    views = [_cam(0.0)] + [(*_cam(_shift_baseline(k)), np.full((H, W), 10.0, np.float32)) for k in ks]
    return ref, views


@pytest.mark.parametrize("min_views", [2, 3, 4])
def test_consistent_view_count_closed_form(min_views):
    ks = [3, -6, 10]
    ref, views = _plane_views(ks)
    out = po.clean_depthmap(ref, None, views, same_depth_threshold=0.01, min_consistent_views=min_views)
    keep = _expected_counts(ks) >= min_views
    assert np.array_equal(out > 0, keep)
    assert np.all(out[keep] == 10.0)


def test_empty_reference_pixels_stay_empty():
    ks = [3, -6, 10]
    ref, views = _plane_views(ks)
    ref = ref.copy()
    ref[5:9, 20:30] = 0.0
    ref[12, 12] = -1.0
    out = po.clean_depthmap(ref, None, views, min_consistent_views=1)
    assert np.all(out[5:9, 20:30] == 0) and out[12, 12] == 0
    assert np.all(out[20, 20:30] == 10.0)


def _floater_case(n_with, n_without, max_carved):
    # a 4 x 4 block of the reference at depth 8 (in front of the z = 10
    # plane); n_with neighbours also see it (their depth is 8 over its
    # projection), n_without see the plane behind it and vote to carve
    ref = np.full((H, W), 10.0, np.float32)
    ref[12:16, 20:24] = 8.0
    ks = [1, -2, 3, -3, 2, -1][: n_with + n_without]
    views = [_cam(0.0)]
    for j, k in enumerate(ks):
        b = _shift_baseline(k)
        d = np.full((H, W), 10.0, np.float32)
        if j < n_with:
            shift = F * b / 8.0           # projection of depth-8 points: x - shift
            x0 = int(np.floor(20 - shift)) - 1
            x1 = int(np.ceil(23 - shift)) + 2
            d[10:18, x0:x1] = 8.0
        views.append((*_cam(b), d))
    out = po.clean_depthmap(ref, None, views, same_depth_threshold=0.01, min_consistent_views=3,
                            carving_threshold=0.05, max_carved_views=max_carved)
    return out[12:16, 20:24]


def test_carving_budget():
    # consistent = 1 + 3, carved = 2: budget = max_carved + (4 - 3)
    assert np.all(_floater_case(3, 2, max_carved=1) == 8.0)    # 2 <= 1 + 1
    assert np.all(_floater_case(3, 2, max_carved=0) == 0.0)    # 2 >  0 + 1
    # no neighbour agrees: too few consistent views whatever the budget
    assert np.all(_floater_case(0, 3, max_carved=99) == 0.0)
    # every neighbour agrees: kept
    assert np.all(_floater_case(5, 0, max_carved=0) == 8.0)


def test_grazing_normals_need_one_more_view():
    ks = [3, -6]                  # interior pixels: 3 consistent views
    ref, views = _plane_views(ks)
    cnt = _expected_counts(ks)
    base = po.clean_depthmap(ref, None, views, min_consistent_views=3)
    frontal = np.zeros((H, W, 3), np.float32)
    frontal[..., 2] = -1.0
    assert np.array_equal(po.clean_depthmap(ref, frontal, views, min_consistent_views=3), base)
    side = np.zeros((H, W, 3), np.float32)
    side[..., 0] = 1.0            # perpendicular to the optical axis
    out = po.clean_depthmap(ref, side, views, min_consistent_views=3)
    xs = np.arange(W)[None, :].repeat(H, 0)
    ys = np.arange(H)[:, None].repeat(W, 1)
    rx, ry = (xs - K[0, 2]) / F, (ys - K[1, 2]) / F
    cosang = np.abs(rx) / np.sqrt(rx * rx + ry * ry + 1.0)
    grazing = cosang < 0.2
    want = np.where(grazing, cnt >= 4, cnt >= 3)
    clear = np.abs(cosang - 0.2) > 1e-4
    assert np.array_equal((out > 0)[clear], want[clear])
    assert np.any(grazing & (cnt == 3))


def test_neighbour_of_another_size_is_resized():
    ks = [3, -6, 10]
    ref, views = _plane_views(ks)
    big = [views[0]] + [(Kv, Rv, tv, np.full((2 * H, 2 * W), 10.0, np.float32)) for Kv, Rv, tv, _ in views[1:]]
    a = po.clean_depthmap(ref, None, views)
    b = po.clean_depthmap(ref, None, big)
    assert np.array_equal(a, b)


def test_effective_minimum_is_capped_by_the_view_count():
    # cleaner.cc:222: min(min_consistent_views, num_views)
    ref, views = _plane_views([3])
    out = po.clean_depthmap(ref, None, views, min_consistent_views=5)
    assert np.array_equal(out > 0, _expected_counts([3]) >= 2)


# ---------------------------------------------------------------------------
# GPU against the oracle
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def noisy_views():
    from opensfm_amd import synthetic
    ws = synthetic.camera_walk(7, num_views=6, sizes=((96, 72),))
    rng = np.random.default_rng(3)
    depths, normals = [], []
    for v, d in enumerate(ws.gt_depth):
        d = d * (1.0 + 0.004 * rng.standard_normal(d.shape)).astype(np.float32)
        d[rng.random(d.shape) < 0.08] = 0.0                            # holes
        m = rng.random(d.shape) < 0.05
        d[m] *= 0.8                                                    # floaters
        depths.append(d.astype(np.float32))
        n = rng.standard_normal(d.shape + (3,)).astype(np.float32)
        n[..., 2] = -np.abs(n[..., 2]) - 0.3
        normals.append(n / np.linalg.norm(n, axis=-1, keepdims=True))
    return ws, depths, normals


@pytest.mark.gpu
@pytest.mark.parametrize("with_normal", [False, True])
def test_gpu_cleaner_matches_oracle(pydense, noisy_views, with_normal):
    ws, depths, normals = noisy_views
    c = pydense.GPUDepthmapCleaner()
    c.set_same_depth_threshold(0.02)
    c.set_min_consistent_views(3)
    c.set_carving_threshold(0.03)
    c.set_max_carved_views(1)
    c.set_grazing_cos_threshold(0.3)
    for v in range(len(depths)):
        if with_normal:
            n3 = np.asfortranarray(normals[v].reshape(-1, 3).T)      # (3, H*W) as the pipeline passes it
            c.add_view_with_normal(ws.Ks[v], ws.Rs[v], ws.ts[v], depths[v], n3)
        else:
            c.add_view(ws.Ks[v], ws.Rs[v], ws.ts[v], depths[v])
    kept = []
    for ref in range(len(depths)):
        nbrs = [j for j in range(len(depths)) if j != ref]
        got = c.clean(ref, np.array(nbrs, np.int32))
        views = [(ws.Ks[ref], ws.Rs[ref], ws.ts[ref])] + [(ws.Ks[j], ws.Rs[j], ws.ts[j], depths[j]) for j in nbrs]
        want = po.clean_depthmap(depths[ref], normals[ref] if with_normal else None, views, 0.02, 3, 0.03, 1, 0.3)
        assert got.shape == depths[ref].shape
        assert np.array_equal(got, want), (ref, np.mean(got == want))
        kept.append(np.mean(got > 0) / max(np.mean(depths[ref] > 0), 1e-9))
    print(f"cleaner (normals={with_normal}): kept fraction per view {np.round(kept, 3).tolist()}, "
          f"{c.last_ms():.3f} ms last kernel", flush=True)
    assert 0.3 < np.mean(kept) < 1.0


@pytest.mark.gpu
def test_gpu_cleaner_resizes_neighbours(pydense):
    ks = [3, -6, 10]
    ref, views = _plane_views(ks)
    c = pydense.GPUDepthmapCleaner()
    c.add_view(*views[0], ref)
    for Kv, Rv, tv, d in views[1:]:
        c.add_view(Kv, Rv, tv, np.full((2 * H, 2 * W), 10.0, np.float32))
    got = c.clean(0, np.array([1, 2, 3], np.int32))
    assert np.array_equal(got, po.clean_depthmap(ref, None, views))


@pytest.mark.gpu
def test_gpu_cleaner_errors(pydense):
    c = pydense.GPUDepthmapCleaner()
    with pytest.raises((IndexError, RuntimeError), match="out of range"):
        c.clean(0, np.array([], np.int32))
    c.add_view(K, np.eye(3), np.zeros(3), np.ones((H, W), np.float32))
    with pytest.raises(RuntimeError, match=r"neighbor_ids\[0\] = 4 out of range"):
        c.clean(0, np.array([4], np.int32))
    c.clear()
    with pytest.raises((IndexError, RuntimeError), match="out of range"):
        c.clean(0, np.array([], np.int32))
    assert pydense.GPUDepthmapCleaner.is_available()
