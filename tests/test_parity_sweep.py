"""Seeded sweep of estimator configurations: the HIP path against the oracle,
bit for bit, on small scenes.

The fixed-scene parity tests cover the default configuration and one
parameter change at a time; this sweep draws many at once from a seeded
generator. It includes:
- odd image sizes;
- 1 to 16 source views, 16 being the most a handle takes;
- every patch size;
- both image storages;
- census on and off;
- several top-k, smoothness and iteration settings.
"""
import numpy as np
import pytest

from oracle import pm_oracle as po
from tests.helpers import gpu_estimator, half_round, params_for


def _configs(n=14, seed=2024):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        out.append(dict(
            seed=int(rng.integers(1, 10_000)),
            width=int(rng.integers(33, 141)),
            height=int(rng.integers(25, 101)),
            num_src=int([1, 2, 3, 5, 8, 11, 16][i % 7]),
            patch_size=int(rng.choice([3, 5, 7])),
            storage=str(rng.choice(["f32", "f16"])),
            max_iterations=int(rng.integers(1, 4)),
            use_census=bool(rng.integers(0, 2)),
            top_k=int(rng.choice([1, 2, 4])),
            smooth_weight=float(rng.choice([0.0, 0.2])),
        ))
    return out


CONFIGS = _configs()


def test_sweep_covers_the_edges():
    assert {c["num_src"] for c in CONFIGS} >= {1, 16}
    assert {c["patch_size"] for c in CONFIGS} == {3, 5, 7}
    assert {c["storage"] for c in CONFIGS} == {"f32", "f16"}
    assert any(c["width"] % 2 for c in CONFIGS) and any(c["height"] % 2 for c in CONFIGS)


@pytest.mark.gpu
@pytest.mark.parametrize("cfg", CONFIGS, ids=[f"c{i}" for i in range(len(CONFIGS))])
def test_config_matches_oracle(pydense, cfg):
    from opensfm_amd import synthetic
    s = synthetic.three_planes(cfg["seed"], width=cfg["width"], height=cfg["height"], num_src=cfg["num_src"],
                               base_cell=0.3, supersample=1)
    p = params_for(s, patch_size=cfg["patch_size"], max_iterations=cfg["max_iterations"],
                   use_census=cfg["use_census"], top_k=cfg["top_k"], smooth_weight=cfg["smooth_weight"])
    print(f"config {cfg}", flush=True)
    out = gpu_estimator(pydense, s, p, storage=cfg["storage"]).compute()
    imgs = half_round(s.images) if cfg["storage"] == "f16" else s.images
    d_o, n_o, c_o, _, _ = po.estimate(imgs, s.Ks, s.Rs, s.ts, p)
    assert np.array_equal(out["depth"], d_o)
    assert np.array_equal(out["cost"], c_o)
    assert np.array_equal(out["normal"], n_o)


TINY = [(1, 1, 1, 5, "f32", 1), (2, 2, 2, 3, "f16", 1), (3, 5, 2, 7, "f32", 1), (8, 3, 3, 5, "f16", 1),
        (17, 9, 2, 7, "f32", 1), (4, 70, 2, 5, "f32", 1), (70, 4, 3, 7, "f16", 1),
        (20, 13, 2, 5, "f32", 4), (5, 3, 2, 5, "f16", 3), (40, 9, 3, 3, "f32", 5)]


@pytest.mark.gpu
@pytest.mark.parametrize("w,h,n,patch,storage,levels", TINY, ids=[f"{t[0]}x{t[1]}-l{t[5]}" for t in TINY])
def test_tiny_images_match_oracle(pydense, w, h, n, patch, storage, levels):
    """Images smaller than a patch or a colour tile, one-pixel-thin strips,
    and pyramids whose coarsest levels shrink to one or two pixels (levels
    stop at 1 x 1, as in BuildImagePyramid, estimator.cc:728-735): partial
    tiles, empty colours and clamped taps everywhere, bit for bit against the
    oracle."""
    from opensfm_amd import synthetic
    s = synthetic.three_planes(3, width=w, height=h, num_src=n, base_cell=0.3, supersample=1)
    p = params_for(s, patch_size=patch)
    print(f"tiny {w}x{h}, {n} sources, patch {patch}, {storage}, {levels} levels", flush=True)
    est = gpu_estimator(pydense, s, p, storage=storage)
    est.set_hierarchy_levels(levels)
    out = est.compute()
    imgs = half_round(s.images) if storage == "f16" else s.images
    d_o, _, c_o, _, _ = po.estimate_levels(imgs, s.Ks, s.Rs, s.ts, p, levels)
    assert out["depth"].shape == (h, w)
    assert np.array_equal(out["depth"], d_o)
    assert np.array_equal(out["cost"], c_o)
