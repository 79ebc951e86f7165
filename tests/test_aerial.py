"""Aerial survey workload: nadir views of a fractal terrain, run in groups
admitted by device memory, with Gauss-Newton refinement.

The scene generator is pinned by an independent ray march (numpy, fine
steps) and the numpy value noise at the hit points, and by the flight
geometry (strip and photo spacing from the overlaps).  Admission is pinned by
its grouping rule; on the GPU, references run in admitted groups must equal
the same references run alone (and refined alone) bit for bit, and one
aerial reference must equal the oracle.
"""
import time

import numpy as np
import pytest

from oracle import pm_oracle as po
from tests.helpers import gpu_estimator, half_round, params_for


def test_admit_groups():
    from opensfm_amd.admission import admit
    assert admit([3, 3, 3, 3], 7) == [[0, 1], [2, 3]]
    assert admit([3, 3, 3, 3], 6) == [[0, 1], [2, 3]]
    assert admit([3, 3, 3, 3], 5) == [[0], [1], [2], [3]]
    assert admit([10, 1, 1, 10, 1], 4) == [[0], [1, 2], [3], [4]]      # too large: alone
    assert admit([], 4) == []
    assert admit([0, 0, 5], 5) == [[0, 1, 2]]
    with pytest.raises(ValueError):
        admit([1], 0)
    with pytest.raises(ValueError):
        admit([-1], 3)


def test_fractal_dem():
    from opensfm_amd.synthetic import fractal_dem
    z = fractal_dem(5)
    assert z.shape == (2001, 2001)
    assert z.min() == 0.0 and abs(z.max() - 80.0) < 1e-9
    assert max(np.abs(np.diff(z, axis=0)).max(), np.abs(np.diff(z, axis=1)).max()) < 0.75
    assert np.array_equal(z, fractal_dem(5)) and not np.array_equal(z, fractal_dem(6))


W, H = 200, 150     # same footprint as 4000x3000 (focal 0.9 x width)


@pytest.fixture(scope="module")
def survey():
    from opensfm_amd import synthetic
    return synthetic.aerial_survey(seed=5, num_views=12, width=W, height=H, neighbours=5)


def _bilinear(z, x, y):
    j, i = int(np.floor(x)), int(np.floor(y))
    fx, fy = x - j, y - i
    return ((1 - fy) * ((1 - fx) * z[i, j] + fx * z[i, j + 1])
            + fy * ((1 - fx) * z[i + 1, j] + fx * z[i + 1, j + 1]))


def test_survey_layout(survey):
    from opensfm_amd.synthetic import fractal_dem
    s = survey
    assert len(s.images) == 12 and all(im.dtype == np.float16 and im.shape == (H, W) for im in s.images)
    C = np.array([-R.T @ t for R, t in zip(s.Rs, s.ts)])
    # 2 strips of 6: photos 100 * 150/180 * 0.2 m apart along a strip, strips
    # 100 * 200/180 * 0.3 m apart (jitter 0.3 m), at 140 m
    along, across = 100 * H / (0.9 * W) * 0.2, 100 * W / (0.9 * W) * 0.3
    assert np.allclose(np.diff(C[:6, 1]), along, atol=1.5)
    assert np.allclose(np.diff(C[6:, 1]), -along, atol=1.5)          # flown back
    assert abs(np.mean(C[6:, 0]) - np.mean(C[:6, 0]) - across) < 1.0
    assert np.allclose(C[:, 2], 140.0, atol=1.5)
    for R in s.Rs:
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-12)
        assert np.degrees(np.arccos(-R[2, 2])) < 2.0                   # looking down
    assert s.Rs[0][0, 0] > 0.99 and s.Rs[6][0, 0] < -0.99              # second strip rotated 180 degrees
    for v, srcs, dmin, dmax in s.refs:
        assert (dmin, dmax) == (60.0, 140.0) and len(srcs) == 5 and v not in srcs
        d = np.linalg.norm(C[srcs] - C[v], axis=1)
        others = np.delete(np.linalg.norm(C - C[v], axis=1), [v] + srcs)
        assert np.all(np.diff(d) >= 0) and d.max() <= others.min()     # the nearest, in order
        g = s.gt_depth[v]
        assert np.all(g > 60.0) and np.all(g < 140.0)
    # forward overlap with the next photo of the strip, near 80 %
    v, u = 2, 3
    ys, xs = np.mgrid[0:H, 0:W]
    d = s.gt_depth[v].astype(np.float64)
    X = s.Rs[v].T @ (np.linalg.inv(s.Ks[v]) @ np.stack([xs.ravel(), ys.ravel(), np.ones(xs.size)]) * d.ravel()
                     - s.ts[v][:, None])
    p = s.Ks[u] @ (s.Rs[u] @ X + s.ts[u][:, None])
    inside = (p[0] / p[2] >= 0) & (p[0] / p[2] <= W - 1) & (p[1] / p[2] >= 0) & (p[1] / p[2] <= H - 1)
    assert 0.7 < inside.mean() < 0.9
    # the same seed gives the same survey
    from opensfm_amd import synthetic
    again = synthetic.aerial_survey(seed=5, num_views=12, width=W, height=H, neighbours=5)
    assert all(np.array_equal(a, b) for a, b in zip(s.images, again.images))
    # ground truth against a fine march along random rays, and the image
    # against the value noise at the hit point
    z = fractal_dem(5)
    noise = synthetic.ValueNoise3D(5)
    rng = np.random.default_rng(0)
    errs, texd = [], []
    for v in (0, 7):
        K, R, t = s.Ks[v], s.Rs[v], s.ts[v]
        Cv = -R.T @ t
        for _ in range(40):
            x, y = rng.integers(0, W), rng.integers(0, H)
            dvec = R.T @ np.linalg.solve(K, [x, y, 1.0])
            lam = np.arange(60.0, 150.0, 0.01)
            P = Cv[:, None] + dvec[:, None] * lam
            g = P[2] - np.array([_bilinear(z, px, py) for px, py in zip(P[0], P[1])])
            k = int(np.argmax(g <= 0))
            l0, l1 = lam[k - 1], lam[k]
            for _ in range(60):                     # bisection inside the fine step
                m = 0.5 * (l0 + l1)
                Pm = Cv + dvec * m
                if Pm[2] - _bilinear(z, Pm[0], Pm[1]) > 0:
                    l0 = m
                else:
                    l1 = m
            errs.append(abs(s.gt_depth[v][y, x] - l0))
            hit = Cv + dvec * l0
            texd.append(abs(float(s.images[v][y, x]) - noise.fbm(hit[None, :], 0.8)[0]))
    assert max(errs) < 2e-4                          # float32 depth of ~100 m
    assert np.median(texd) < 0.012 and max(texd) < 0.05   # sensor noise sigma 0.01 + float16


@pytest.mark.gpu
def test_admitted_groups_equal_single_runs(pydense, survey):
    from opensfm_amd.admission import run_admitted
    s = survey
    refs = [s.reference(i) for i in range(6)]

    def build(i):
        return gpu_estimator(pydense, refs[i], params_for(refs[i]), storage="f16")

    sizes = [build(i).device_bytes() for i in range(len(refs))]
    for steps, overlap in ((0, False), (3, False), (3, True)):
        # overlapping groups get half the budget each: the same pairs either way
        budget = int(2.5 * max(sizes)) * (2 if overlap else 1)
        got, groups = run_admitted(pydense, len(refs), build, budget=budget, refine_steps=steps,
                                   on_result=lambda i, e: e.download(["depth", "normal", "cost"]),
                                   overlap_staging=overlap)
        assert [g.refs for g in groups] == [[0, 1], [2, 3], [4, 5]]
        assert pydense.DepthmapEstimator.buffer_cache_bytes(0) == 0       # emptied at the end
        assert all(g.device_bytes <= budget and g.run_ms > 0 for g in groups)
        assert all((g.refine_ms > 0) == (steps > 0) for g in groups)
        print(f"aerial groups (refine {steps}, overlap {overlap}): " + ", ".join(f"{g.refs} {g.run_ms:.2f}+{g.refine_ms:.2f} ms"
                                                             for g in groups), flush=True)
        for i in range(len(refs)):
            est = build(i)
            est.compute()
            if steps:
                est.gauss_newton(steps, 1e-3)
            want = est.download(["depth", "normal", "cost"])
            assert set(got[i]) == {"depth", "normal", "cost"}
            for k in want:
                assert np.array_equal(got[i][k], want[k]), (steps, i, k)


@pytest.mark.gpu
def test_aerial_reference_matches_oracle(pydense, survey):
    sc = survey.reference(4)
    p = params_for(sc)
    d = gpu_estimator(pydense, sc, p, storage="f16").compute()["depth"]
    d_o, _, _, _, _ = po.estimate(half_round(sc.images), sc.Ks, sc.Rs, sc.ts, p)
    assert np.array_equal(d, d_o)
    gt = sc.gt_depth
    print(f"aerial reference 4: GT within 1% {np.mean(np.abs(d - gt) < 0.01 * gt):.4f}", flush=True)


@pytest.mark.gpu
def test_buffer_cache_reuses_released_blocks(pydense, survey):
    E = pydense.DepthmapEstimator
    refs = [survey.reference(i) for i in (1, 2)]
    fresh = [gpu_estimator(pydense, r, params_for(r), storage="f16").compute()["depth"] for r in refs]
    try:
        E.set_buffer_cache(0, 1 << 40)
        a = gpu_estimator(pydense, refs[0], params_for(refs[0]), storage="f16")
        assert np.array_equal(a.compute()["depth"], fresh[0])
        assert E.buffer_cache_bytes(0) == 0
        a.release_device()
        held = E.buffer_cache_bytes(0)
        assert held >= 0.9 * a.device_bytes()
        b = gpu_estimator(pydense, refs[1], params_for(refs[1]), storage="f16")
        assert np.array_equal(b.compute()["depth"], fresh[1])            # on the recycled blocks
        assert E.buffer_cache_bytes(0) < 0.1 * held
        assert np.array_equal(a.compute()["depth"], fresh[0])            # uploads again
        E.set_buffer_cache(0, held // 2)
        assert E.buffer_cache_bytes(0) <= held // 2
    finally:
        E.set_buffer_cache(0, 0)
    assert E.buffer_cache_bytes(0) == 0


class _StubEstimator:
    def __init__(self, log, i, nbytes, levels=1, delay=0.0):
        self.log, self.i, self.nbytes, self.levels, self.delay = log, i, nbytes, levels, delay

    def device_bytes(self):
        return self.nbytes

    def params(self):
        return {"hierarchy_levels": self.levels}

    def gauss_newton(self, steps, damping):
        self.log.append(("gn", self.i, steps))

    def last_refine_ms(self):
        return 1.0

    def download(self, maps=None):
        if self.delay:   # a download that takes a while (a copy behind running kernels)
            time.sleep(self.delay)
        return {"depth": np.full((2, 2), float(self.i), np.float32)}

    def release_device(self):
        self.log.append(("release", self.i))


class _StubLibrary:
    """The parts of pydense that run_admitted uses, recording the calls."""

    def __init__(self):
        log = self.log = []

        class Est:
            @staticmethod
            def set_buffer_cache(device, max_bytes):
                log.append(("cache", max_bytes))

            @staticmethod
            def device_free_memory(device):
                return 1000

        class Batch:
            def __init__(self, device):
                self.refs = []
                self.level = 0

            def set_level(self, level):
                self.level = level
                log.append(("level", tuple(self.refs), level))

            def add(self, e):
                self.refs.append(e.i)

            def set_config(self, **kw):
                log.append(("config", kw))

            def prepare(self):
                log.append(("prepare", tuple(self.refs)))

            def run(self):
                log.append(("run", tuple(self.refs)) if self.level == 0 else ("run", tuple(self.refs), self.level))
                return 5.0

            def stage_timing(self):
                return {"candidates": (2.0, 3), "fallback": (1.0, 6)}

        self.DepthmapEstimator = Est
        self.DepthmapBatch = Batch


@pytest.mark.parametrize("overlap", [False, True])
def test_run_admitted_control_flow(overlap):
    from opensfm_amd.admission import run_admitted
    lib = _StubLibrary()
    sizes = [30, 30, 30, 50, 10, 90, 20]
    built = []

    def build(i):
        built.append(i)
        return _StubEstimator(lib.log, i, sizes[i])

    budget = 200 if overlap else 100          # groups of at most 100 either way
    res, groups = run_admitted(lib, len(sizes), build, budget=budget, refine_steps=2, overlap_staging=overlap,
                               batch_config={"use_graph": True})
    assert built == list(range(len(sizes)))
    assert [g.refs for g in groups] == [[0, 1, 2], [3, 4], [5], [6]]
    assert all(g.device_bytes <= 100 for g in groups)
    assert sorted(res) == list(range(len(sizes))) and all(res[i]["depth"][0, 0] == i for i in res)
    runs = [e[1] for e in lib.log if e[0] == "run"]
    assert runs == [(0, 1, 2), (3, 4), (5,), (6,)]
    assert [e[1] for e in lib.log if e[0] == "prepare"] == runs
    for g in runs:       # every reference refined and released after its group ran
        for i in g:
            k = lib.log.index(("run", g))
            assert lib.log.index(("gn", i, 2)) > k and lib.log.index(("release", i)) > k
    assert lib.log[0] == ("cache", budget) and lib.log[-1] == ("cache", 0)
    assert all(g.run_ms == 5.0 and g.refine_ms == len(g.refs) for g in groups)


def test_overlapped_groups_hold_at_most_two_groups_of_memory():
    """With overlap_staging the worker collects and releases a finished group
    while the next one runs, and stages the group after that only then: the
    bytes of groups staged and not yet released never exceed the budget."""
    from opensfm_amd.admission import run_admitted
    lib = _StubLibrary()
    sizes = [40, 40, 30, 60, 50, 50, 20, 80, 10]
    res, groups = run_admitted(lib, len(sizes), lambda i: _StubEstimator(lib.log, i, sizes[i]), budget=200,
                               refine_steps=1, overlap_staging=True)
    assert [g.refs for g in groups] == [[0, 1], [2, 3], [4, 5], [6, 7], [8]]
    held = peak = 0
    for e in lib.log:
        if e[0] == "prepare":
            held += sum(sizes[i] for i in e[1])
            peak = max(peak, held)
        elif e[0] == "release":
            held -= sizes[e[1]]
    assert held == 0 and 100 < peak <= 200
    assert sorted(res) == list(range(len(sizes))) and all(res[i]["depth"][0, 0] == i for i in res)
    assert lib.log[-1] == ("cache", 0)


@pytest.mark.parametrize("seed", range(12))
def test_admission_bounds_and_order_on_random_sizes(seed):
    """Random reference sizes and budgets, staging overlapped or not: groups
    take references in order, no group exceeds its share of the budget
    unless it is a single reference larger than that share, the bytes of
    groups staged and not yet released never exceed the budget (two
    references too large for a share aside), and every result is collected
    exactly once."""
    from opensfm_amd.admission import run_admitted
    rng = np.random.default_rng(seed)
    count = int(rng.integers(1, 30))
    sizes = [int(v) for v in rng.integers(1, 120, count)]
    budget = int(rng.integers(60, 400))
    overlap = bool(seed % 2)
    lib = _StubLibrary()
    res, groups = run_admitted(lib, count, lambda i: _StubEstimator(lib.log, i, sizes[i], delay=0.002),
                               budget=budget, refine_steps=int(seed % 3), overlap_staging=overlap)
    share = budget // 2 if overlap else budget
    assert [i for g in groups for i in g.refs] == list(range(count))
    for g in groups:
        assert g.device_bytes <= share or len(g.refs) == 1
        assert g.device_bytes == sum(sizes[i] for i in g.refs)
    held = peak = 0
    for e in lib.log:
        if e[0] == "prepare":
            held += sum(sizes[i] for i in e[1])
            peak = max(peak, held)
        elif e[0] == "release":
            held -= sizes[e[1]]
    assert held == 0
    if all(sz <= share for sz in sizes):
        assert peak <= (budget if overlap else share)
    assert sorted(res) == list(range(count)) and all(res[i]["depth"][0, 0] == i for i in res)
    assert lib.log[-1] == ("cache", 0)


def test_overlapped_result_errors_end_the_run():
    from opensfm_amd.admission import run_admitted
    lib = _StubLibrary()

    def on_result(i, e):
        if i == 3:
            raise RuntimeError("disk full")
        return i

    with pytest.raises(RuntimeError, match="disk full"):
        run_admitted(lib, 8, lambda i: _StubEstimator(lib.log, i, 30), budget=200, on_result=on_result,
                     overlap_staging=True)
    assert lib.log[-1] == ("cache", 0)


def test_run_admitted_runs_every_pyramid_level_in_order():
    # a batch starts at level 0 of each reference's pyramid; the finer levels
    # run after it, in order, before the refinement and the results
    from opensfm_amd.admission import run_admitted
    lib = _StubLibrary()
    res, groups = run_admitted(lib, 3, lambda i: _StubEstimator(lib.log, i, 40, levels=3), budget=100,
                               refine_steps=1)
    runs = [e for e in lib.log if e[0] in ("run", "level")]
    assert runs == [("run", (0, 1)), ("level", (0, 1), 1), ("run", (0, 1), 1), ("level", (0, 1), 2),
                    ("run", (0, 1), 2), ("run", (2,)), ("level", (2,), 1), ("run", (2,), 1),
                    ("level", (2,), 2), ("run", (2,), 2)]
    last = lib.log.index(("run", (0, 1), 2))
    assert lib.log.index(("gn", 0, 1)) > last and lib.log.index(("release", 1)) > last
    assert [g.run_ms for g in groups] == [15.0, 15.0]
    assert groups[0].stage_ms == {"candidates": [6.0, 9], "fallback": [3.0, 18]}


def test_run_admitted_errors_reset_the_cache():
    from opensfm_amd.admission import run_admitted
    lib = _StubLibrary()

    def build(i):
        if i == 2:
            raise RuntimeError("no such image")
        return _StubEstimator(lib.log, i, 10)

    with pytest.raises(RuntimeError, match="no such image"):
        run_admitted(lib, 4, build, budget=100)
    assert lib.log[-1] == ("cache", 0)
    with pytest.raises(ValueError):
        run_admitted(lib, 1, build, budget=0)
    res, groups = run_admitted(lib, 0, build, budget=100)
    assert res == {} and groups == []


@pytest.mark.gpu
def test_aerial_full_size_band_matches_oracle(pydense):
    """The aerial workload at its own size: a middle reference of a 4000x3000
    survey (3 strips of 80 %/70 % overlap, the same flight as the 200-view
    job) with its 10 nearest views as sources, patch 7, 3 iterations, f16
    storage; a band of 64 reference rows through the middle, on the GPU and
    on the oracle, bit for bit."""
    from opensfm_amd import synthetic
    s = synthetic.aerial_survey(seed=5, num_views=22, width=4000, height=3000, neighbours=10, device="cuda")
    ref = s.reference(12)
    assert ref.num_src == 10 and ref.gt_depth.shape == (3000, 4000)
    band = ref.crop_rows(1468, 1532)
    p = params_for(band, patch_size=7, max_iterations=3)
    print(f"aerial 4000x3000 reference 12, rows 1468..1532, 10 sources", flush=True)
    d = gpu_estimator(pydense, band, p, storage="f16").compute()["depth"]
    d_o, _, _, _, _ = po.estimate(half_round(band.images), band.Ks, band.Rs, band.ts, p)
    assert d.shape == (64, 4000)
    assert np.array_equal(d, d_o)
    gt = band.gt_depth
    print(f"aerial 4000x3000 band: GT within 1% {np.mean(np.abs(d - gt) < 0.01 * gt):.4f}", flush=True)
