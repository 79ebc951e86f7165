"""Batched runs (pm_batch_*, pydense.DepthmapBatch): many references of
different sizes, source counts and depth ranges in one launch per kernel,
captured in a HIP graph.  DepthmapClusterEstimator::Run runs the references
one after the other (cluster_estimator.cc:117-151); a batched reference must
give bit-for-bit the result of running it alone, whose parity with the
oracle is pinned elsewhere -- here checked again directly for one reference.
"""
import numpy as np
import pytest

from oracle import pm_oracle as po
from tests.helpers import gpu_estimator, params_for


@pytest.fixture(scope="module")
def scenes():
    from opensfm_amd import synthetic
    return [
        synthetic.three_planes(11, width=160, height=120, num_src=4, base_cell=0.3),
        synthetic.three_planes(12, width=200, height=150, num_src=10, base_cell=0.3),   # > 8 sources
        synthetic.facet_heightfield(13, width=240, height=180, num_src=6),
        synthetic.three_planes(14, width=97, height=61, num_src=3, base_cell=0.3),      # odd sizes
    ]


def _params(s):
    return params_for(s, patch_size=7)


def _alone(pydense, s, storage):
    est = gpu_estimator(pydense, s, _params(s), storage=storage)
    return est.compute()


def _batched(pydense, scenes, storage, config=None):
    ests = [gpu_estimator(pydense, s, _params(s), storage=storage) for s in scenes]
    b = pydense.DepthmapBatch()
    for e in ests:
        b.add(e)
    if config is not None:
        b.set_config(*config)
    ms = b.run()
    return b, ests, [e.download() for e in ests], ms


def _same(a, b):
    for k in ("depth", "plane", "cost", "views", "confidence"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_batch_equals_single_runs(pydense, scenes, storage):
    pydense.DepthmapBatch.reset_tuning()
    alone = [_alone(pydense, s, storage) for s in scenes]
    b, ests, got, ms = _batched(pydense, scenes, storage)
    cfg = b.config()
    print(f"batch {storage}: {len(scenes)} refs, {ms:.2f} ms, config {cfg}", flush=True)
    assert cfg["cost_tile_x"] in (4, 8, 16) and cfg["pixel_tile_x"] in (16, 32)
    log = b.tune_log()
    assert log.shape[0] >= 6 and np.all(log[:, 2:] > 0)
    for a, g in zip(alone, got):
        _same(a, g)
    # a second run relaunches the captured graph: same result
    b.run()
    for e, g in zip(ests, got):
        _same(e.download(), g)


@pytest.mark.gpu
@pytest.mark.parametrize("config", [(16, 16, False), (8, 32, True), (4, 16, True)])
def test_batch_configurations_agree(pydense, scenes, config):
    alone = [_alone(pydense, s, "f16") for s in scenes]
    _, _, got, _ = _batched(pydense, scenes, "f16", config)
    for a, g in zip(alone, got):
        _same(a, g)


@pytest.mark.gpu
def test_single_estimator_tiles_agree(pydense, scenes):
    s = scenes[1]
    base = _alone(pydense, s, "f32")
    for cx, px in ((16, 16), (8, 32), (4, 16)):
        est = gpu_estimator(pydense, s, _params(s))
        est.set_tiles(cx, px)
        assert est.tiles() == (cx, px)
        _same(base, est.compute())


@pytest.mark.gpu
def test_batch_follows_changed_references(pydense, scenes):
    b, ests, got, _ = _batched(pydense, scenes[:2], "f32")
    ests[1].set_depth_range(3.0, 30.0, 0)
    b.run()
    s = scenes[1]
    p = _params(s)
    p.depth_min, p.depth_max = 3.0, 30.0
    ref = gpu_estimator(pydense, s, p).compute()
    _same(ref, ests[1].download())
    _same(got[0], ests[0].download())


@pytest.mark.gpu
def test_batched_reference_matches_oracle(pydense, scenes):
    s = scenes[3]
    p = _params(s)
    _, _, got, _ = _batched(pydense, scenes, "f32")
    d_o, _, c_o, _, _ = po.estimate(s.images, s.Ks, s.Rs, s.ts, p)
    d_g = got[3]["depth"]
    valid = d_o > 0
    within = np.mean(np.abs(d_g[valid] - d_o[valid]) < 0.01 * d_o[valid])
    print(f"batched ref vs oracle: bit-equal {np.mean(d_g == d_o):.6f} within1% {within:.6f}", flush=True)
    assert within >= 0.99


@pytest.mark.gpu
def test_batch_errors(pydense, scenes):
    b = pydense.DepthmapBatch()
    with pytest.raises(RuntimeError, match="no references"):
        b.run()
    e1 = gpu_estimator(pydense, scenes[0], _params(scenes[0]))
    e2 = gpu_estimator(pydense, scenes[3], params_for(scenes[3], patch_size=5))
    b.add(e1)
    with pytest.raises(RuntimeError, match="added twice"):
        b.add(e1)
    b.add(e2)
    with pytest.raises(RuntimeError, match="patch size"):
        b.run()
    with pytest.raises(RuntimeError, match="cost tile"):
        b.set_config(5, 0, True)


@pytest.mark.gpu
def test_cluster_one_level_runs_batched(pydense, scenes):
    # one pyramid level: the cluster estimator runs its references as one batch
    chosen = (scenes[0], scenes[2])
    c = pydense.DepthmapClusterEstimator()
    c.set_max_iterations(3)
    c.set_patch_size(7)
    c.set_hierarchy_levels(1)
    c.set_speckle_min_size(0)
    c.set_gap_max_size(0)
    c.set_checkerboard_filter(False)
    for s in chosen:
        c.begin_ref_view(s.Ks[0], s.Rs[0], s.ts[0], s.images[0], s.depth_min, s.depth_max)
        for K, R, t, im in zip(s.Ks[1:], s.Rs[1:], s.ts[1:], s.images[1:]):
            c.add_source_view(K, R, t, im)
    got = c.run()
    for s, (d, n, cost, conf) in zip(chosen, got):
        est = pydense.DepthmapEstimator()   # the same defaults as the cluster estimator
        for im, K, R, t in zip(s.images, s.Ks, s.Rs, s.ts):
            est.add_view(K, R, t, im)
        est.set_depth_range(s.depth_min, s.depth_max, 0)
        est.set_patchmatch_iterations(3)
        est.set_patch_size(7)
        est.set_hierarchy_levels(1)
        est.set_speckle_min_size(0)
        est.set_gap_max_size(0)
        est.set_checkerboard_filter(False)
        r = est.compute()
        assert np.array_equal(d, r["depth"])
        assert np.array_equal(cost, r["cost"])
        assert np.array_equal(n, r["normal"])


SMALL_SIZES = ((80, 60), (128, 96), (200, 150), (250, 188))


@pytest.fixture(scope="module")
def walk():
    from opensfm_amd import synthetic
    return synthetic.camera_walk(5, num_views=10, sizes=SMALL_SIZES)


def test_camera_walk_layout(walk):
    # every view is a reference with 4..12 sources drawn per reference, sizes
    # from the given set, and a depth range that holds the scene it sees
    assert len(walk.refs) == 10
    for v, srcs, dmin, dmax in walk.refs:
        assert 4 <= len(srcs) <= 12 or len(srcs) == len(walk.refs) - 1
        assert v not in srcs and len(set(srcs)) == len(srcs)
        h, w = walk.images[v].shape
        assert (w, h) in SMALL_SIZES
        gt = walk.gt_depth[v]
        assert np.mean(gt > 0) > 0.5
        assert dmin < gt[gt > 0].min() and gt.max() < dmax
    ref = walk.reference(3)
    assert ref.num_src == len(walk.refs[3][1]) and ref.images[0] is walk.images[walk.refs[3][0]]
    again = __import__("opensfm_amd.synthetic", fromlist=["x"]).camera_walk(5, num_views=10, sizes=SMALL_SIZES)
    assert all(np.array_equal(a, b) for a, b in zip(walk.images, again.images))


@pytest.mark.gpu
def test_camera_walk_batch_equals_single_runs(pydense, walk):
    refs = [walk.reference(i) for i in range(len(walk.refs))]
    alone = [_alone(pydense, s, "f16") for s in refs]
    _, _, got, ms = _batched(pydense, refs, "f16")
    print(f"camera walk: {len(refs)} refs batched in {ms:.2f} ms", flush=True)
    for a, g in zip(alone, got):
        _same(a, g)


@pytest.mark.gpu
def test_batch_levels_equal_single_multilevel_runs(pydense, scenes):
    # every reference with a 3-level pyramid, run level by level through one
    # batch, equals running each reference's pyramid on its own
    refs = scenes[:3]

    def make(s):
        e = gpu_estimator(pydense, s, _params(s), storage="f16")
        e.set_hierarchy_levels(3)
        return e

    alone = [make(s).compute() for s in refs]
    ests = [make(s) for s in refs]
    b = pydense.DepthmapBatch()
    for e in ests:
        b.add(e)
    for level in range(3):
        b.set_level(level)
        b.run()
    for a, e in zip(alone, ests):
        _same(a, e.download())
    with pytest.raises(RuntimeError, match="level"):
        b.set_level(3)
        b.run()


@pytest.mark.gpu
def test_batch_levels_run_in_order_once(pydense, scenes):
    # a finer level starts from the level just below it: running level 1
    # twice (the graph would replay an upsample of buffers replaced by the
    # first run) or skipping a level is refused before any launch
    refs = scenes[:2]
    ests = []
    for s in refs:
        e = gpu_estimator(pydense, s, _params(s), storage="f32")
        e.set_hierarchy_levels(3)
        ests.append(e)
    b = pydense.DepthmapBatch()
    for e in ests:
        b.add(e)
    b.set_level(1)
    with pytest.raises(RuntimeError, match="levels run in order"):
        b.run()                      # level 0 has not run
    b.set_level(0)
    b.run()
    b.set_level(1)
    b.run()
    with pytest.raises(RuntimeError, match="levels run in order"):
        b.run()                      # level 1 again
    b.set_level(2)
    b.run()
    assert ests[0].download(["depth"])["depth"].shape == refs[0].gt_depth.shape


@pytest.mark.gpu
def test_download_has_the_staged_level_size(pydense, scenes):
    # after a coarse level only, the decoded maps have that level's size
    s = scenes[0]
    e = gpu_estimator(pydense, s, _params(s), storage="f32")
    e.set_hierarchy_levels(2)
    b = pydense.DepthmapBatch()
    b.add(e)
    b.run()
    h, w = s.gt_depth.shape
    d = e.download(["depth", "normal"])
    assert d["depth"].shape == (h // 2, w // 2) and d["normal"].shape == (h // 2, w // 2, 3)


@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_batch_with_tiny_references_equals_single_runs(pydense, storage):
    """References of a few pixels batched with ordinary ones: each reference
    owns whole groups of 8 workgroups, most of them past its last tile; the
    batch must still equal the references run alone, bit for bit."""
    from opensfm_amd import synthetic
    scenes = [synthetic.three_planes(21, width=1, height=1, num_src=1, base_cell=0.3, supersample=1),
              synthetic.three_planes(22, width=160, height=120, num_src=4, base_cell=0.3),
              synthetic.three_planes(23, width=5, height=3, num_src=2, base_cell=0.3, supersample=1),
              synthetic.three_planes(24, width=70, height=4, num_src=3, base_cell=0.3, supersample=1),
              synthetic.three_planes(25, width=2, height=2, num_src=2, base_cell=0.3, supersample=1)]
    alone = [_alone(pydense, s, storage) for s in scenes]
    for config in ((4, 32, True), (16, 16, False)):
        b, ests, got, ms = _batched(pydense, scenes, storage, config)
        print(f"tiny batch {storage} {config}: {ms:.2f} ms", flush=True)
        for a, g in zip(alone, got):
            _same(a, g)
