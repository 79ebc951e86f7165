"""SfM planar prior (ComputeSfMPlanarPrior, estimator.cc:838-1107) and
neighbour selection (SelectNeighbors, dense_helpers.cc:331-568).

The triangulation is pinned against Qhull (scipy.spatial.Delaunay); the
prior against the numpy restatement in the oracle (planes to float rounding,
masks exactly) and against closed forms: points on one world plane give
that plane at every masked pixel; the depth-ratio and area gates drop the
triangles they name.  Neighbour selection is checked against the oracle's
plain-Python restatement and a configuration with known viewing angles.
The GPU test runs PatchMatch with the points and with the equivalent
explicit prior.
"""
import numpy as np
import pytest

from oracle import pm_oracle as po


@pytest.fixture(scope="module")
def pd():
    from opensfm_amd import native
    return native()


def test_delaunay_matches_qhull(pd):
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(5)
    for trial in range(60):
        n = int(rng.integers(3, 400))
        xy = rng.random((n, 2)) * rng.choice([1.0, 640.0, 4000.0])
        if trial % 3 == 0:                                   # clustered
            xy = np.concatenate([xy[: n // 2] * 0.01, xy[n // 2:]])
        got = pd.delaunay(xy)
        assert np.all(got[:, 0] < got[:, 1]) and np.all(got[:, 1] < got[:, 2])
        assert got.tolist() == sorted(got.tolist())
        want = sorted(tuple(sorted(map(int, s))) for s in Delaunay(xy).simplices)
        assert [tuple(r) for r in got.tolist()] == want


def _camera(w=160, h=120, f=140.0):
    K = np.array([[f, 0, (w - 1) / 2], [0, f, (h - 1) / 2], [0, 0, 1.0]])
    th = 0.3
    R = np.array([[np.cos(th), 0, -np.sin(th)], [0, 1, 0], [np.sin(th), 0, np.cos(th)]])
    t = np.array([0.4, -0.2, 1.0])
    return K, R, t, w, h


def _plane_points(K, R, t, w, h, n, rng, n_cam=np.array([0.2, -0.3, -1.0]), d_cam=12.0):
    """World points on the camera-frame plane n.X + d = 0 that project inside the image."""
    n_cam = n_cam / np.linalg.norm(n_cam)
    uv = np.column_stack([rng.uniform(0, w, n), rng.uniform(0, h, n)])
    rays = np.column_stack([(uv[:, 0] - K[0, 2]) / K[0, 0], (uv[:, 1] - K[1, 2]) / K[1, 1], np.ones(n)])
    lam = -d_cam / (rays @ n_cam)
    Xc = rays * lam[:, None]
    return (Xc - t) @ R, n_cam, d_cam       # X_world = R^T (X_cam - t)


def test_points_on_one_plane_give_that_plane(pd):
    K, R, t, w, h = _camera()
    rng = np.random.default_rng(1)
    pts, n_cam, d_cam = _plane_points(K, R, t, w, h, 150, rng)
    planes, masks, ntri = pd.sfm_planar_prior(K, R, t, w, h, 1.0, 100.0, pts)
    m = masks.astype(bool)
    assert ntri > 100 and m.mean() > 0.6
    want = np.append(n_cam, d_cam) * (1 if d_cam >= 0 else -1)
    assert np.allclose(planes[m], want, atol=2e-5)
    assert np.all(planes[~m] == np.array([0, 0, -1, 0], np.float32))
    # depth of the prior = depth of the plane
    ys, xs = np.nonzero(m)
    ray = np.column_stack([(xs - K[0, 2]) / K[0, 0], (ys - K[1, 2]) / K[1, 1], np.ones(len(xs))])
    d = -planes[m][:, 3] / np.sum(planes[m][:, :3] * ray, 1)
    assert np.allclose(d, -d_cam / (ray @ n_cam), rtol=1e-5)


def test_prior_matches_the_oracle(pd):
    K, R, t, w, h = _camera()
    rng = np.random.default_rng(2)
    # two planes at different depths + scattered points
    a, _, _ = _plane_points(K, R, t, w, h, 120, rng)
    b, _, _ = _plane_points(K, R, t, w, h, 60, rng, n_cam=np.array([-0.4, 0.1, -1.0]), d_cam=20.0)
    c = (rng.random((30, 3)) - 0.5) * 20 + np.array([0, 0, 15.0])
    pts = np.concatenate([a, b, c])
    for dmin, dmax in ((1.0, 100.0), (8.0, 18.0)):
        planes, masks, ntri = pd.sfm_planar_prior(K, R, t, w, h, dmin, dmax, pts)
        op, om, ont = po.sfm_planar_prior(K, R, t, w, h, dmin, dmax, pts)
        assert ntri == ont
        assert np.array_equal(masks, om)
        assert np.max(np.abs(planes - op)) < 1e-5


def test_depth_ratio_and_area_gates(pd):
    K, R, t, w, h = _camera()
    rng = np.random.default_rng(3)
    near, _, _ = _plane_points(K, R, t, w, h, 40, rng, d_cam=10.0)
    far, _, _ = _plane_points(K, R, t, w, h, 40, rng, d_cam=30.0)   # 3x deeper: mixed triangles fail 1.5
    pts = np.concatenate([near, far])
    planes, masks, ntri = pd.sfm_planar_prior(K, R, t, w, h, 1.0, 100.0, pts)
    tris = pd.delaunay(np.array([[(K @ (R @ p + t))[0] / (R @ p + t)[2], (K @ (R @ p + t))[1] / (R @ p + t)[2]]
                                 for p in pts]))
    mixed = sum(1 for tri in tris if len({i < 40 for i in tri}) == 2)
    assert mixed > 0 and ntri <= len(tris) - mixed
    assert set(np.round(planes[masks.astype(bool)][:, 3], 3)) <= {10.0, 30.0}
    # four points spanning the image: two triangles of more than 20000 px^2 -> none kept
    Kb, Rb, tb, wb, hb = _camera(400, 300, 300.0)
    corners = np.array([[5, 5], [395, 5], [395, 295], [5, 295]], float)
    rays = np.column_stack([(corners[:, 0] - Kb[0, 2]) / Kb[0, 0], (corners[:, 1] - Kb[1, 2]) / Kb[1, 1],
                            np.ones(4)]) * 10.0
    big = (rays - tb) @ Rb
    _, m2, n2 = pd.sfm_planar_prior(Kb, Rb, tb, wb, hb, 1.0, 100.0, big)
    assert n2 == 0 and m2.sum() == 0
    # fewer than 4 points in the image: no prior
    _, m3, n3 = pd.sfm_planar_prior(K, R, t, w, h, 1.0, 100.0, near[:3])
    assert n3 == 0 and m3.sum() == 0


def _walk_config(rng, n_shots=12, n_sp=400):
    ang = np.linspace(0, 1.5, n_shots)
    origins = np.column_stack([20 * np.cos(ang), 20 * np.sin(ang), np.full(n_shots, 5.0)])
    coords = rng.uniform(-5, 5, (n_sp, 3))
    vis = []
    for p in coords:
        d = np.linalg.norm(origins - p, axis=1)
        k = rng.integers(2, 7)
        vis.append(sorted(np.argsort(d + rng.normal(0, 3, n_shots))[:k].tolist()))
    counts = rng.integers(1, 8, n_sp).tolist()
    return origins, coords, vis, counts


def test_select_neighbors_matches_the_restatement(pd):
    rng = np.random.default_rng(4)
    origins, coords, vis, counts = _walk_config(rng)
    for args in ((5, 3, 1, 1.0, 30.0), (3, 10, 20, 2.0, 10.0), (8, 0, 0, 0.5, 60.0)):
        got = pd.select_neighbors(origins, coords, vis, counts, *args)
        want = po.select_neighbors(origins, coords, vis, counts, *args)
        assert got[0] == want[0] and got[1] == want[1]
        for i, b in enumerate(got[0]):
            assert b[0] == i and len(b) <= 1 + args[0]


def test_select_neighbors_known_angles(pd):
    # shot 0 at the origin looking at points 10 m away; shots 1..3 at 1, 3
    # and 6 m to the side: viewing angles ~5.7, ~16.7 and ~31 degrees
    origins = np.array([[0, 0, 0], [1, 0, 0], [3, 0, 0], [6, 0, 0]], float)
    coords = np.array([[0, 0, 10.0]] * 5)
    vis = [[0, 1, 2, 3]] * 5
    counts = [10] * 5
    best, allv = pd.select_neighbors(origins, coords, vis, counts, 2, 5, 5, 3.0, 20.0)
    assert best[0] == [0, 1, 2] or best[0] == [0, 2, 1]
    assert 3 not in best[0] and allv[0][-1] == 3           # 31 degrees: outside, but shares tracks
    best, _ = pd.select_neighbors(origins, coords, vis, counts, 2, 5, 5, 10.0, 40.0)
    assert best[0] == [0, 2, 3]


@pytest.mark.gpu
def test_sfm_points_equal_the_explicit_prior(pydense, scene_small):
    from tests.helpers import gpu_estimator, params_for
    s = scene_small
    p = params_for(s)
    h, w = s.gt_depth.shape
    K, R, t = s.Ks[0], s.Rs[0], s.ts[0]
    rng = np.random.default_rng(6)
    ys, xs = rng.integers(0, h, 400), rng.integers(0, w, 400)
    d = s.gt_depth[ys, xs]
    ok = d > 0
    Xc = np.column_stack([(xs - K[0, 2]) / K[0, 0], (ys - K[1, 2]) / K[1, 1], np.ones(len(xs))]) * d[:, None]
    pts = ((Xc - t) @ R)[ok]
    a = gpu_estimator(pydense, s, p)
    a.set_sfm_points(pts)
    ra = a.compute()
    planes, masks, ntri = pydense.sfm_planar_prior(K, R, t, w, h, p.depth_min, p.depth_max, pts)
    assert ntri > 0 and masks.sum() > 0
    b = gpu_estimator(pydense, s, p)
    b.set_prior(planes, masks)
    rb = b.compute()
    assert np.array_equal(ra["depth"], rb["depth"]) and np.array_equal(ra["cost"], rb["cost"])
    c = gpu_estimator(pydense, s, p).compute()
    gt = s.gt_depth
    acc = [float(np.mean(np.abs(r["depth"] - gt) < 0.01 * gt)) for r in (ra, c)]
    print(f"SfM prior: {ntri} triangles, {masks.mean():.3f} masked, GT within 1% {acc[0]:.4f} "
          f"(without prior {acc[1]:.4f})", flush=True)
    assert not np.array_equal(ra["depth"], c["depth"])
