"""Views added from device memory (pm_add_view_device) against the same
views added from host memory: the padding and the u8/f16/f32 conversion run
on the device instead of the host and must store the same texels, so every
decoded map is bit-identical, single and batched.  Host pointers and
several pyramid levels are refused."""
import numpy as np
import pytest

from tests.helpers import gpu_estimator, params_for

pytestmark = pytest.mark.gpu


def _device_estimator(pydense, scene, params, storage, tensors, pitch_pad=0):
    """The estimator of gpu_estimator with every view given by device pointer."""
    est = gpu_estimator(pydense, scene, params, storage, images=[])
    for t, K, R, tt in zip(tensors, scene.Ks, scene.Rs, scene.ts):
        h, w = t.shape[0], t.shape[1] - pitch_pad
        dtype = {1: "u8", 2: "f16", 4: "f32"}[t.element_size()]
        est.add_view_device(K, R, tt, t.data_ptr(), dtype, w, h, t.stride(0) * t.element_size())
    return est


def _same(a, b):
    for k in ("depth", "plane", "normal", "cost", "views", "confidence"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("fmt", ["f32", "f16", "u8"])
@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_device_views_equal_host_views(pydense, scene_tiny, fmt, storage):
    import torch
    s = scene_tiny
    p = params_for(s, patch_size=5)
    if fmt == "u8":
        host = [np.clip(np.rint(im * 255.0), 0, 255).astype(np.uint8) for im in s.images]
    elif fmt == "f16":
        host = [im.astype(np.float16) for im in s.images]
    else:
        host = [np.asarray(im, np.float32) for im in s.images]
    ref = gpu_estimator(pydense, s, p, storage, images=host).compute()
    # rows padded by 5 elements: the row stride is honoured
    pad = 5
    dev = []
    for im in host:
        t = torch.zeros((im.shape[0], im.shape[1] + pad), dtype=torch.from_numpy(im).dtype, device="cuda")
        t[:, :im.shape[1]] = torch.from_numpy(im).cuda()
        dev.append(t)
    torch.cuda.synchronize()
    got = _device_estimator(pydense, s, p, storage, dev, pitch_pad=pad).compute()
    print(f"device views {fmt} -> {storage}: bit-equal {np.mean(got['depth'] == ref['depth']):.6f}", flush=True)
    _same(ref, got)


def test_device_views_quantize_half_and_batch(pydense, scene_tiny):
    import torch
    s = scene_tiny
    p = params_for(s, patch_size=5)
    e_host = gpu_estimator(pydense, s, p, "f32")
    e_host.set_quantize_half(True)
    dev = [torch.from_numpy(np.asarray(im, np.float32)).cuda() for im in s.images]
    torch.cuda.synchronize()
    e_dev = _device_estimator(pydense, s, p, "f32", dev)
    e_dev.set_quantize_half(True)
    _same(e_host.compute(), e_dev.compute())
    # one batch of a device-view and a host-view reference equals their single runs
    a = gpu_estimator(pydense, s, p, "f16")
    b = _device_estimator(pydense, s, p, "f16", dev)
    alone = gpu_estimator(pydense, s, p, "f16").compute()
    batch = pydense.DepthmapBatch()
    batch.add(a)
    batch.add(b)
    batch.run()
    _same(alone, a.download())
    _same(alone, b.download())


def test_device_views_refusals(pydense, scene_tiny):
    import torch
    s = scene_tiny
    est = pydense.DepthmapEstimator()
    host = np.ascontiguousarray(s.images[0], np.float32)
    h, w = host.shape
    with pytest.raises(RuntimeError, match="not device memory"):
        est.add_view_device(s.Ks[0], s.Rs[0], s.ts[0], host.ctypes.data, "f32", w, h)
    with pytest.raises(ValueError, match="dtype"):
        est.add_view_device(s.Ks[0], s.Rs[0], s.ts[0], 0, "f64", w, h)
    dev = [torch.from_numpy(np.asarray(im, np.float32)).cuda() for im in s.images]
    torch.cuda.synchronize()
    p = params_for(s, patch_size=5)
    e = _device_estimator(pydense, s, p, "f32", dev)
    e.set_hierarchy_levels(2)
    with pytest.raises(RuntimeError, match="one pyramid level"):
        e.compute()
