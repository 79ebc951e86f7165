"""The reference's u8 upload on the GPU against the oracle
(estimator.cc:313-361): a u8 view becomes value * (1/255) in float32, and
with set_quantize_half that value is rounded to binary16 and back before
anything else sees it.  The oracle runs on the same host-side restatement
(oracle.pm_oracle.quantize_u8_half, pinned in test_oracle.py); the depth
maps must agree to the parity bar (within 1 % on >= 99 % of valid pixels;
bit-equality expected and reported), for float32 and float16 image storage,
from host and from device memory.
"""
import numpy as np
import pytest

from oracle import pm_oracle as po
from tests.helpers import gpu_estimator, half_round, params_for

pytestmark = pytest.mark.gpu


def _u8(scene):
    return [np.clip(np.rint(np.asarray(im, np.float64) * 255.0), 0, 255).astype(np.uint8) for im in scene.images]


def _compare(tag, d_gpu, d_o):
    valid = d_o > 0
    within = float(np.mean(np.abs(d_gpu[valid] - d_o[valid]) < 0.01 * d_o[valid]))
    print(f"[{tag}] within 1% of the oracle {within:.6f}, bit-equal {np.mean(d_gpu == d_o):.6f}", flush=True)
    assert within >= 0.99


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_u8_quantize_half_matches_oracle(pydense, scene_tiny, storage):
    s = scene_tiny
    u8 = _u8(s)
    p = params_for(s, patch_size=5)
    e = gpu_estimator(pydense, s, p, storage, images=u8)
    e.set_quantize_half(True)
    d_gpu = e.compute()["depth"]
    q = [po.quantize_u8_half(im) for im in u8]
    d_o, _, _, _, _ = po.estimate(q, s.Ks, s.Rs, s.ts, p)
    _compare(f"u8 -> /255 -> half, {storage} storage", d_gpu, d_o)


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_u8_without_quantisation_matches_oracle(pydense, scene_tiny, storage):
    s = scene_tiny
    u8 = _u8(s)
    p = params_for(s, patch_size=5)
    d_gpu = gpu_estimator(pydense, s, p, storage, images=u8).compute()["depth"]
    f = [im.astype(np.float32) * np.float32(1.0 / 255.0) for im in u8]
    if storage == "f16":
        f = half_round(f)
    d_o, _, _, _, _ = po.estimate(f, s.Ks, s.Rs, s.ts, p)
    _compare(f"u8 -> /255, {storage} storage", d_gpu, d_o)


def test_u8_device_views_quantize_half_match_oracle(pydense, scene_tiny):
    import torch
    s = scene_tiny
    u8 = _u8(s)
    p = params_for(s, patch_size=5)
    dev = [torch.from_numpy(im).cuda() for im in u8]
    torch.cuda.synchronize()
    e = gpu_estimator(pydense, s, p, "f32", images=[])
    for t, K, R, tt in zip(dev, s.Ks, s.Rs, s.ts):
        e.add_view_device(K, R, tt, t.data_ptr(), "u8", t.shape[1], t.shape[0], t.stride(0))
    e.set_quantize_half(True)
    d_gpu = e.compute()["depth"]
    q = [po.quantize_u8_half(im) for im in u8]
    d_o, _, _, _, _ = po.estimate(q, s.Ks, s.Rs, s.ts, p)
    _compare("device u8 -> /255 -> half", d_gpu, d_o)
