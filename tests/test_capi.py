"""The C ABI library: exports, error reporting and host-side validation.

CPU part: the library loads and exports every function include/opensfm_pm.h
declares, and reports a missing device as an error (no compute calls).
GPU part: call-order and argument errors come back as status codes with
the reference's messages (estimator.cc:142-149, 175-178).
"""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "opensfm_pm.h")
LIB = os.path.join(ROOT, "opensfm_amd", "libopensfm_pm.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(pm_\w+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libopensfm_pm.so is not built (python -m opensfm_amd.build)")
    L = ctypes.CDLL(LIB)
    L.pm_last_error.restype = ctypes.c_char_p
    return L


def test_header_declares_the_surface():
    names = declared_functions()
    for n in ("pm_create", "pm_destroy", "pm_add_view", "pm_set_depth_range", "pm_set_params",
              "pm_compute", "pm_score_planes", "pm_last_error", "pm_upload", "pm_run", "pm_download"):
        assert n in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_default_params_are_the_pipeline_profile(lib):
    buf = (ctypes.c_char * 256)()
    lib.pm_default_params(buf)
    ints = np.frombuffer(bytes(buf)[:8], np.int32)
    floats = np.frombuffer(bytes(buf)[8:16], np.float32)
    assert ints.tolist() == [3, 7]                      # iterations, patch size
    assert np.allclose(floats, [1.5, 25 / 255])         # config.py:447-449


def test_null_handle_is_an_error(lib):
    assert lib.pm_destroy(None) == 0
    assert lib.pm_upload(None) == -1
    assert b"null" in lib.pm_last_error()
    assert lib.pm_create(0, None) == -1


def test_no_device_is_reported_without_a_gpu(lib):
    n = ctypes.c_int(-1)
    assert lib.pm_device_count(ctypes.byref(n)) == 0
    if n.value > 0:
        pytest.skip("a HIP device is visible; the no-device path is not reachable here")
    h = ctypes.c_void_p()
    assert lib.pm_create(0, ctypes.byref(h)) == -2
    assert b"no HIP device" in lib.pm_last_error()
    assert not h.value


def test_batch_without_a_gpu(lib):
    n = ctypes.c_int(-1)
    assert lib.pm_device_count(ctypes.byref(n)) == 0
    if n.value > 0:
        pytest.skip("a HIP device is visible; the no-device path is not reachable here")
    b = ctypes.c_void_p()
    assert lib.pm_batch_create(0, ctypes.byref(b)) == -2
    assert not b.value
    assert lib.pm_batch_destroy(None) == 0
    assert lib.pm_batch_run(None) == -1
    assert b"null pm_batch" in lib.pm_last_error()
    assert lib.pm_batch_add(None, None) == -1


def test_pydense_surface_names():
    from opensfm_amd import native
    pd = native()
    cls = pd.DepthmapEstimator
    for m in ("add_view", "set_depth_range", "set_patchmatch_iterations", "set_patch_size",
              "set_min_patch_sd", "compute_patch_match", "compute_patch_match_sample", "compute_brute_force",
              "score_planes", "is_available", "num_devices"):
        assert hasattr(cls, m), m
    for m in ("add", "set_config", "config", "tune_log", "prepare", "run", "reset_tuning"):
        assert hasattr(pd.DepthmapBatch, m), m
    for m in ("set_same_depth_threshold", "set_min_consistent_views", "set_device", "set_carving_threshold",
              "set_max_carved_views", "set_grazing_cos_threshold", "add_view", "add_view_with_normal", "clean",
              "clear", "is_available"):     # python/pybind.cc:125-143
        assert hasattr(pd.GPUDepthmapCleaner, m), m
    for f in ("sfm_planar_prior", "delaunay", "select_neighbors", "lz4_frame_compress", "lz4_frame_decompress"):
        assert hasattr(pd, f), f
    assert hasattr(pd.DepthmapEstimator, "set_sfm_points") and hasattr(pd.DepthmapClusterEstimator, "set_sfm_points")
    for m in ("device_bytes", "release_device", "device_free_memory", "gauss_newton", "download"):
        assert hasattr(cls, m), m


@pytest.mark.gpu
def test_call_order_and_argument_errors(pydense, scene_tiny):
    est = pydense.DepthmapEstimator()
    with pytest.raises(RuntimeError, match="no views added"):
        est.compute_patch_match()
    s = scene_tiny
    est.add_view(s.Ks[0], s.Rs[0], s.ts[0], s.images[0])
    with pytest.raises(RuntimeError, match="source view"):
        est.upload()
    est.add_view(s.Ks[1], s.Rs[1], s.ts[1], s.images[1])
    est.set_depth_range(5.0, 4.0, 0)
    with pytest.raises(RuntimeError, match="depth range"):
        est.upload()
    est.set_depth_range(s.depth_min, s.depth_max, 0)
    est.set_patch_size(9)
    with pytest.raises(RuntimeError, match="patch_size"):
        est.upload()
    est.set_patch_size(5)
    with pytest.raises(RuntimeError, match="not initialised"):
        est.upload(); est.iterate(0, 0)
    with pytest.raises(ValueError):
        est.score_planes(np.zeros((3, 2), np.int32), np.zeros((2, 4), np.float32))
    with pytest.raises(RuntimeError, match="outside"):
        est.score_planes(np.array([[10_000, 0]], np.int32), np.zeros((1, 4), np.float32))
    depth, *_ = est.compute_patch_match()
    assert depth.shape == s.gt_depth.shape


@pytest.mark.gpu
def test_too_many_sources(pydense, scene_tiny):
    est = pydense.DepthmapEstimator()
    s = scene_tiny
    for _ in range(pydense.MAX_SOURCES + 2):
        est.add_view(s.Ks[0], s.Rs[0], s.ts[0], s.images[0])
    with pytest.raises(RuntimeError, match="at most"):
        est.upload()
