"""DepthmapClusterEstimator: several references run level by level, with the
previous level's depth maps of linked references as geometric consistency
inputs (cluster_estimator.cc:82-156; call sites depthmaps.py:208-286)."""
import numpy as np
import pytest

from oracle import pm_oracle as po
from tests.helpers import params_for


def _refs(scene, dmin, dmax):
    """Reference 0 = view 0 with views 1..N; reference 1 = view 1 with view 0
    and views 2..N.  Each takes the other's depth into its slot 0."""
    v = list(range(scene.num_src + 1))
    orders = [v, [1, 0] + v[2:]]
    refs = []
    for i, order in enumerate(orders):
        refs.append(dict(images=[scene.images[j] for j in order], Ks=[scene.Ks[j] for j in order],
                         Rs=[scene.Rs[j] for j in order], ts=[scene.ts[j] for j in order],
                         depth_min=dmin, depth_max=dmax, links=[(0, 1 - i)]))
    return refs


def test_cluster_without_geometry_equals_independent_runs(scene_tiny):
    s = scene_tiny
    p = params_for(s, patch_size=5)
    refs = _refs(s, s.depth_min, s.depth_max)
    out = po.estimate_cluster(refs, p, 2, geom_weight=0.0)
    for r, (d, n, c, conf) in zip(refs, out):
        d1, _, _, _, _ = po.estimate_levels(r["images"], r["Ks"], r["Rs"], r["ts"], p, 2)
        assert np.array_equal(d, d1)


def test_cluster_geometry_changes_only_later_levels(scene_tiny):
    s = scene_tiny
    p = params_for(s, patch_size=5)
    refs = _refs(s, s.depth_min, s.depth_max)
    one = po.estimate_cluster(refs, p, 1, geom_weight=0.5)      # one level: no previous depths
    ind = po.estimate_cluster(refs, p, 1, geom_weight=0.0)
    assert all(np.array_equal(a[0], b[0]) for a, b in zip(one, ind))
    two = po.estimate_cluster(refs, p, 2, geom_weight=0.5)
    ind2 = po.estimate_cluster(refs, p, 2, geom_weight=0.0)
    assert any(not np.array_equal(a[0], b[0]) for a, b in zip(two, ind2))


@pytest.mark.gpu
def test_gpu_cluster_matches_oracle(pydense, scene_small):
    s = scene_small
    p = params_for(s, patch_size=5)
    refs = _refs(s, s.depth_min, s.depth_max)
    c = pydense.DepthmapClusterEstimator()
    c.set_max_iterations(p.max_iterations)
    c.set_patch_size(p.patch_size)
    c.set_hierarchy_levels(2)
    c.set_sigma_spatial(p.sigma_spatial)
    c.set_sigma_color(p.sigma_color)
    c.set_top_k(p.top_k)
    c.set_use_census(p.use_census)
    c.set_smooth_weight(p.smooth_weight)
    c.set_anchor_views(p.anchor_views)
    c.set_checkerboard_filter(False)
    c.set_speckle_min_size(0)
    c.set_gap_max_size(0)
    c.set_geom_consistency_weight(0.05)
    for i, r in enumerate(refs):
        idx = c.begin_ref_view(r["Ks"][0], r["Rs"][0], r["ts"][0], r["images"][0], r["depth_min"], r["depth_max"])
        assert idx == i
        for K, R, t, im in zip(r["Ks"][1:], r["Rs"][1:], r["ts"][1:], r["images"][1:]):
            c.add_source_view(K, R, t, im)
    for i, r in enumerate(refs):
        for slot, frm in r["links"]:
            c.add_geom_link(i, slot, frm)
    got = c.run()
    want = po.estimate_cluster(refs, p, 2, geom_weight=0.05)
    assert len(got) == 2
    for i, ((dg, ng, cg, confg), (do, no, co, confo)) in enumerate(zip(got, want)):
        valid = do > 0
        within = np.mean(np.abs(dg[valid] - do[valid]) < 0.01 * do[valid])
        print(f"cluster ref {i}: bit-equal {np.mean(dg == do):.6f} within1% {within:.6f}", flush=True)
        assert dg.shape == do.shape and ng.shape == do.shape + (3,)
        assert within >= 0.99


@pytest.mark.gpu
def test_gpu_cluster_errors(pydense):
    c = pydense.DepthmapClusterEstimator()
    with pytest.raises(RuntimeError, match="no reference views"):
        c.run()
    with pytest.raises(RuntimeError, match="no ref view started"):
        c.add_source_view(np.eye(3), np.eye(3), np.zeros(3), np.zeros((8, 8), np.uint8))
    with pytest.raises(RuntimeError, match="invalid ref_idx"):
        c.add_geom_link(3, 0, 0)
    assert pydense.DepthmapClusterEstimator.num_devices() >= 1
    assert pydense.DepthmapClusterEstimator.device_memory_bytes(0) > 0
