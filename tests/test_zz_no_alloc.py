"""pm_score_planes_noalloc (the allocation-free scoring entry point) against
pm_score_planes, both through the C ABI (ctypes on libopensfm_pm.so).

Two calls in a row on one handle, the second with fewer pairs (the shared
device array keeps the first call's contents), must each equal
pm_score_planes on the same pairs bit for bit; more pairs than
PM_NOALLOC_MAX_PAIRS is refused with an error.
"""
import ctypes
import os

import numpy as np
import pytest

from tests.helpers import random_planes
from tests.test_capi import PmParams

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "opensfm_amd", "libopensfm_pm.so")
MAX_PAIRS = 65536   # PM_NOALLOC_MAX_PAIRS

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import torch  # noqa: F401  (the HIP runtime the library shares, as in the other tests)
    L = ctypes.CDLL(LIB)
    L.pm_last_error.restype = ctypes.c_char_p
    n = ctypes.c_int(0)
    if L.pm_device_count(ctypes.byref(n)) != 0 or n.value < 1:
        pytest.fail("no HIP device visible to libopensfm_pm.so")
    return L


def _ok(L, rc):
    assert rc == 0, L.pm_last_error().decode()


def _handle(L, scene, patch):
    h = ctypes.c_void_p()
    _ok(L, L.pm_create(0, ctypes.byref(h)))
    for im, K, R, t in zip(scene.images, scene.Ks, scene.Rs, scene.ts):
        img = np.ascontiguousarray(im, np.float32)
        k, r, tt = (np.ascontiguousarray(a, np.float64) for a in (K, R, t))
        _ok(L, L.pm_add_view(h, k.ctypes.data_as(ctypes.c_void_p), r.ctypes.data_as(ctypes.c_void_p),
                             tt.ctypes.data_as(ctypes.c_void_p), img.ctypes.data_as(ctypes.c_void_p), 1,
                             img.shape[1], img.shape[0], ctypes.c_int64(0), None))
    p = PmParams()
    L.pm_pipeline_params(ctypes.byref(p))
    p.patch_size, p.hierarchy_levels, p.max_iterations = patch, 1, 3
    p.depth_min, p.depth_max = scene.depth_min, scene.depth_max
    _ok(L, L.pm_set_params(h, ctypes.byref(p)))
    _ok(L, L.pm_upload(h))
    return h


def _score(L, fn, h, xy, planes, n_src):
    n = len(xy)
    xy = np.ascontiguousarray(xy, np.int32)
    planes = np.ascontiguousarray(planes, np.float32)
    costs = np.full((n, n_src), -7.0, np.float32)
    init = np.full(n, -7.0, np.float32)
    sel = np.full(n, 0xDEAD, np.uint32)
    rc = fn(h, n, xy.ctypes.data_as(ctypes.c_void_p), planes.ctypes.data_as(ctypes.c_void_p),
            costs.ctypes.data_as(ctypes.c_void_p), init.ctypes.data_as(ctypes.c_void_p),
            sel.ctypes.data_as(ctypes.c_void_p))
    return rc, costs, init, sel


@pytest.mark.parametrize("patch", [5, 7])
def test_noalloc_scores_equal_the_allocating_entry_point(lib, scene_small, patch):
    L = lib
    s = scene_small
    h = _handle(L, s, patch)
    try:
        for n, seed in ((20000, patch), (1500, patch + 10)):   # the second call is smaller
            print(f"no-alloc scoring: patch {patch}, {n} pairs", flush=True)
            xy, planes = random_planes(s, n, seed=seed)
            rc_a, c_a, i_a, s_a = _score(L, L.pm_score_planes, h, xy, planes, s.num_src)
            _ok(L, rc_a)
            rc_b, c_b, i_b, s_b = _score(L, L.pm_score_planes_noalloc, h, xy, planes, s.num_src)
            _ok(L, rc_b)
            assert np.array_equal(c_a, c_b)
            assert np.array_equal(i_a, i_b)
            assert np.array_equal(s_a, s_b)
            assert np.all(c_b >= 0.0) and np.all(c_b <= 2.0)
        # more pairs than the fixed array holds: refused before any device work
        xy, planes = random_planes(s, MAX_PAIRS + 1, seed=3)
        rc, *_ = _score(L, L.pm_score_planes_noalloc, h, xy, planes, s.num_src)
        assert rc == -1 and b"at most 65536" in L.pm_last_error()
    finally:
        L.pm_destroy(h)
