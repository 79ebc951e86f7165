"""The HIP path against the oracle on the benchmark workloads' own scenes.

* Plain single view: the 64-facet heightfield at 1600x1200 with 8 sources
  (seed 2), patch 7, 3 iterations; a full-width band of 96 reference rows
  (all sources at full size) runs on both sides, with float32 and with
  float16 image storage (the oracle on the same half-rounded texels).
* Batched, mixed sizes: a seeded camera walk around the box scene with
  references of four sizes and 4-12 sources each, float16 storage, run as
  one batch (one graph launch); three references of different size and
  source count are run by the oracle.
The bars are those of the other GPU parity tests: depth within 1 % on >= 99 %
of the oracle's valid pixels (bit-equality expected and reported).
"""
import numpy as np
import pytest

from oracle import pm_oracle as po
from tests.helpers import gpu_estimator, half_round, params_for

pytestmark = pytest.mark.gpu


def _compare(tag, d_gpu, d_o):
    valid = d_o > 0
    within = float(np.mean(np.abs(d_gpu[valid] - d_o[valid]) < 0.01 * d_o[valid]))
    exact = float(np.mean(d_gpu == d_o))
    print(f"[{tag}] {d_o.shape[1]}x{d_o.shape[0]}: within 1% of the oracle {within:.6f}, bit-equal {exact:.6f}",
          flush=True)
    return within


@pytest.fixture(scope="module")
def facet_band():
    from opensfm_amd import synthetic
    print("rendering the 1600x1200 facet heightfield (8 sources) ...", flush=True)
    scene = synthetic.facet_heightfield(seed=2, width=1600, height=1200, num_src=8)
    return scene.crop_rows(552, 648)


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_facet_heightfield_band_matches_oracle(pydense, facet_band, storage):
    s = facet_band
    assert s.num_src == 8 and s.images[0].shape == (96, 1600) and s.images[1].shape == (1200, 1600)
    p = params_for(s, patch_size=7, max_iterations=3)
    imgs = s.images if storage == "f32" else half_round(s.images)
    print(f"facet band, {storage} storage: GPU ...", flush=True)
    d_gpu = gpu_estimator(pydense, s, p, storage, images=imgs).compute()["depth"]
    print("facet band: oracle ...", flush=True)
    d_o, _, _, _, _ = po.estimate(imgs, s.Ks, s.Rs, s.ts, p)
    assert _compare(f"facet band {storage}", d_gpu, d_o) >= 0.99


@pytest.fixture(scope="module")
def walk():
    from opensfm_amd import synthetic
    return synthetic.camera_walk(7, num_views=12, sizes=((160, 120), (256, 192), (320, 240), (400, 300)))


def test_camera_walk_batch_matches_oracle(pydense, walk):
    scenes = [walk.reference(i) for i in range(len(walk.refs))]
    ests = [gpu_estimator(pydense, s, params_for(s, patch_size=7), "f16") for s in scenes]
    b = pydense.DepthmapBatch()
    for e in ests:
        b.add(e)
    print(f"camera walk batch of {len(scenes)} references: {b.run():.1f} ms, config {b.config()}", flush=True)
    # three references of different size and source count
    picked, seen = [], set()
    for i, s in sorted(enumerate(scenes), key=lambda t: -t[1].images[0].size):
        key = (s.images[0].shape, s.num_src)
        if s.images[0].shape not in {k[0] for k in seen} and s.num_src not in {k[1] for k in seen}:
            picked.append(i)
            seen.add(key)
        if len(picked) == 3:
            break
    assert len(picked) == 3
    for i in picked:
        s = scenes[i]
        imgs = half_round(s.images)
        print(f"camera walk reference {i}: oracle ({s.images[0].shape[1]}x{s.images[0].shape[0]}, "
              f"{s.num_src} sources) ...", flush=True)
        d_o, _, _, _, _ = po.estimate(imgs, s.Ks, s.Rs, s.ts, params_for(s, patch_size=7))
        d_gpu = ests[i].download(["depth"])["depth"]
        assert _compare(f"walk ref {i}, {s.num_src} sources", d_gpu, d_o) >= 0.99
