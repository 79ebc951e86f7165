"""ctypes front-end of the CPU oracle (oracle/pm_oracle.cc).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline of bench.py, never by the product package.

Host-side steps of the reference estimator that the oracle needs are
restated here in numpy:
  * u8 -> /255 -> half input quantisation (estimator.cc:313-361),
  * the float64 -> float32 camera cast (opencl_utils.h:266-280),
  * per-source angle weights (estimator.cc:380-407),
  * the pmo_params mapping of DepthmapParams (estimator.h:16-36).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libpm_oracle.so")
_lib = None


class PmoCamera(ctypes.Structure):
    _fields_ = [("K", ctypes.c_float * 9), ("R", ctypes.c_float * 9), ("t", ctypes.c_float * 3),
                ("width", ctypes.c_int), ("height", ctypes.c_int)]


class PmoScene(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int), ("num_src", ctypes.c_int),
                ("images", ctypes.c_void_p), ("offsets", ctypes.c_void_p), ("widths", ctypes.c_void_p),
                ("heights", ctypes.c_void_p), ("cams", ctypes.c_void_p), ("angle_weights", ctypes.c_void_p)]


class PmoParams(ctypes.Structure):
    _fields_ = [("half_patch", ctypes.c_int), ("sigma_spatial", ctypes.c_float),
                ("sigma_color", ctypes.c_float), ("top_k", ctypes.c_int),
                ("census_weight", ctypes.c_float), ("depth_min", ctypes.c_float),
                ("depth_max", ctypes.c_float), ("max_iterations", ctypes.c_int),
                ("smooth_weight", ctypes.c_float), ("anchor_views", ctypes.c_int),
                ("geom_weight", ctypes.c_float), ("var_threshold", ctypes.c_float),
                ("checkerboard_filter", ctypes.c_int)]


class PmoState(ctypes.Structure):
    _fields_ = [("planes", ctypes.c_void_p), ("costs", ctypes.c_void_p), ("rng", ctypes.c_void_p),
                ("sel", ctypes.c_void_p), ("prior_planes", ctypes.c_void_p),
                ("plane_masks", ctypes.c_void_p), ("prev_depths", ctypes.c_void_p),
                ("prev_depth_mask", ctypes.c_uint32)]


def build() -> str:
    """Compile the oracle with its Makefile (g++, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        f, i, u, vp = ctypes.c_float, ctypes.c_int, ctypes.c_uint32, ctypes.c_void_p
        L.pmo_expf.restype = f; L.pmo_expf.argtypes = [f]
        L.pmo_logf.restype = f; L.pmo_logf.argtypes = [f]
        L.pmo_rand_float.restype = f; L.pmo_rand_float.argtypes = [vp, u]
        L.pmo_bilinear.restype = f; L.pmo_bilinear.argtypes = [vp, i, i, f, f]
        L.pmo_project.argtypes = [vp, vp, vp]
        L.pmo_backproject.argtypes = [vp, f, f, f, vp]
        L.pmo_depth_from_plane.restype = f; L.pmo_depth_from_plane.argtypes = [vp, f, f, vp]
        L.pmo_plane_from_depth_normal.argtypes = [vp, f, f, f, vp, vp]
        L.pmo_ref_patch.argtypes = [vp, vp, i, i, vp]
        L.pmo_score_planes.argtypes = [vp, vp, i, vp, vp, vp, vp, vp]
        for name in ("pmo_random_init", "pmo_prior_reinit", "pmo_run"):
            getattr(L, name).argtypes = [vp, vp, vp]
        L.pmo_sweep.argtypes = [vp, vp, vp, i, i, f]
        L.pmo_checkerboard.argtypes = [vp, vp, vp, i]
        L.pmo_smooth_weight.restype = f; L.pmo_smooth_weight.argtypes = [vp, i]
        L.pmo_decode.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.pmo_set_threads.argtypes = [i]
        L.pmo_max_threads.restype = i
        _lib = L
    return _lib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


# ---------------------------------------------------------------------------
# Host-side restatements
# ---------------------------------------------------------------------------
def quantize_u8_half(img_u8: np.ndarray) -> np.ndarray:
    """u8 -> float32 /255 -> half -> float32 (estimator.cc:313-318, 346-353)."""
    return (img_u8.astype(np.float32) * np.float32(1.0 / 255.0)).astype(np.float16).astype(np.float32)


def make_camera(K, R, t, w, h) -> PmoCamera:
    """float64 -> float32 cast of K, R (row-major) and t (opencl_utils.h:266-280)."""
    c = PmoCamera()
    c.K[:] = np.asarray(K, np.float64).astype(np.float32).ravel().tolist()
    c.R[:] = np.asarray(R, np.float64).astype(np.float32).ravel().tolist()
    c.t[:] = np.asarray(t, np.float64).astype(np.float32).ravel().tolist()
    c.width, c.height = int(w), int(h)
    return c


def angle_weights(Rs: Sequence[np.ndarray], ts: Sequence[np.ndarray]) -> np.ndarray:
    """max(sin(angle(baseline, reference optical axis)), 0.01) per source (estimator.cc:380-407)."""
    c0 = -np.asarray(Rs[0]).T @ np.asarray(ts[0])
    z0 = np.asarray(Rs[0])[2]
    out = []
    for R, t in zip(Rs[1:], ts[1:]):
        ci = -np.asarray(R).T @ np.asarray(t)
        b = ci - c0
        bl = np.linalg.norm(b)
        if bl < 1e-12:
            out.append(0.01)
            continue
        ca = abs(z0 @ (b / bl))
        out.append(max(np.sqrt(1.0 - ca * ca), 0.01))
    return np.asarray(out, np.float64).astype(np.float32)


@dataclass
class Params:
    """DepthmapParams (estimator.h:16-36) plus the variance gate of set_min_patch_sd.

    Defaults are the reference pipeline profile (config.py:407-466) with a
    single hierarchy level and no geometric consistency.
    """
    patch_size: int = 7
    sigma_spatial: float = 1.5
    sigma_color: float = 25.0 / 255.0
    top_k: int = 8
    use_census: bool = True
    depth_min: float = 0.0
    depth_max: float = 1.0
    max_iterations: int = 3
    smooth_weight: float = 0.2
    anchor_views: int = 2
    geom_weight: float = 0.0
    var_threshold: float = 1.5e-6
    checkerboard_filter: bool = False

    def c(self) -> PmoParams:
        p = PmoParams()
        p.half_patch = self.patch_size // 2
        p.sigma_spatial = self.sigma_spatial
        p.sigma_color = self.sigma_color
        p.top_k = self.top_k
        p.census_weight = 1.0 if self.use_census else 0.0
        p.depth_min = self.depth_min
        p.depth_max = self.depth_max
        p.max_iterations = self.max_iterations
        p.smooth_weight = self.smooth_weight
        p.anchor_views = self.anchor_views
        p.geom_weight = self.geom_weight
        p.var_threshold = self.var_threshold
        p.checkerboard_filter = 1 if self.checkerboard_filter else 0
        return p


class OracleScene:
    """Views packed for the oracle.  images: float32 arrays (already quantised as wanted)."""

    def __init__(self, images: Sequence[np.ndarray], Ks, Rs, ts):
        imgs = [np.ascontiguousarray(im, np.float32) for im in images]
        self.n_views = len(imgs)
        sizes = np.array([im.size for im in imgs], np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        self.images = np.concatenate([im.ravel() for im in imgs]).astype(np.float32)
        self.widths = np.array([im.shape[1] for im in imgs], np.int32)
        self.heights = np.array([im.shape[0] for im in imgs], np.int32)
        cams = (PmoCamera * self.n_views)()
        for i in range(self.n_views):
            cams[i] = make_camera(Ks[i], Rs[i], ts[i], self.widths[i], self.heights[i])
        self.cams = cams
        self.aw = angle_weights(Rs, ts)
        if self.aw.size == 0:
            self.aw = np.zeros(1, np.float32)
        self.c = PmoScene()
        self.c.width = int(self.widths[0]); self.c.height = int(self.heights[0])
        self.c.num_src = self.n_views - 1
        self.c.images = self.images.ctypes.data
        self.c.offsets = self.offsets.ctypes.data
        self.c.widths = self.widths.ctypes.data
        self.c.heights = self.heights.ctypes.data
        self.c.cams = ctypes.addressof(self.cams)
        self.c.angle_weights = self.aw.ctypes.data

    @property
    def width(self):
        return int(self.widths[0])

    @property
    def height(self):
        return int(self.heights[0])


class OracleState:
    def __init__(self, w: int, h: int, prior_planes=None, plane_masks=None, prev_depths=None,
                 prev_depth_mask: int = 0):
        n = w * h
        self.planes = np.zeros((h, w, 4), np.float32)
        self.costs = np.zeros((h, w), np.float32)
        self.rng = np.zeros((h, w, 2), np.uint32)
        self.sel = np.zeros((h, w), np.uint32)
        self.prior_planes = None if prior_planes is None else np.ascontiguousarray(prior_planes, np.float32)
        self.plane_masks = None if plane_masks is None else np.ascontiguousarray(plane_masks, np.uint32)
        self.prev_depths = None if prev_depths is None else np.ascontiguousarray(prev_depths, np.float32)
        self.c = PmoState()
        self.c.planes = self.planes.ctypes.data
        self.c.costs = self.costs.ctypes.data
        self.c.rng = self.rng.ctypes.data
        self.c.sel = self.sel.ctypes.data
        self.c.prior_planes = _ptr(self.prior_planes)
        self.c.plane_masks = _ptr(self.plane_masks)
        self.c.prev_depths = _ptr(self.prev_depths)
        self.c.prev_depth_mask = prev_depth_mask
        assert n == self.costs.size


def set_threads(n: int) -> None:
    lib().pmo_set_threads(int(n))


def max_threads() -> int:
    return int(lib().pmo_max_threads())


def score_planes(scene: OracleScene, params: Params, xy: np.ndarray, planes: np.ndarray):
    """Per-view costs (n, num_src), top-k initial cost (n,) and view bitmask (n,)."""
    xy = np.ascontiguousarray(xy, np.int32)
    planes = np.ascontiguousarray(planes, np.float32)
    n = xy.shape[0]
    costs = np.zeros((n, scene.n_views - 1), np.float32)
    init = np.zeros(n, np.float32)
    sel = np.zeros(n, np.uint32)
    p = params.c()
    lib().pmo_score_planes(ctypes.byref(scene.c), ctypes.byref(p), n, _ptr(xy), _ptr(planes),
                           _ptr(costs), _ptr(init), _ptr(sel))
    return costs, init, sel


def ref_patch(scene: OracleScene, params: Params, x: int, y: int) -> np.ndarray:
    P = params.patch_size // 2 * 2 + 1
    out = np.zeros(4 + 2 * P * P, np.float32)
    p = params.c()
    lib().pmo_ref_patch(ctypes.byref(scene.c), ctypes.byref(p), x, y, _ptr(out))
    return out


def random_init(scene: OracleScene, params: Params, state: OracleState) -> None:
    p = params.c()
    lib().pmo_random_init(ctypes.byref(scene.c), ctypes.byref(p), ctypes.byref(state.c))


def prior_reinit(scene: OracleScene, params: Params, state: OracleState) -> None:
    p = params.c()
    lib().pmo_prior_reinit(ctypes.byref(scene.c), ctypes.byref(p), ctypes.byref(state.c))


def sweep(scene: OracleScene, params: Params, state: OracleState, color: int, iteration: int) -> None:
    p = params.c()
    sw = lib().pmo_smooth_weight(ctypes.byref(p), iteration)
    lib().pmo_sweep(ctypes.byref(scene.c), ctypes.byref(p), ctypes.byref(state.c), color, iteration, sw)


def checkerboard(scene: OracleScene, params: Params, state: OracleState, color: int) -> None:
    p = params.c()
    lib().pmo_checkerboard(ctypes.byref(scene.c), ctypes.byref(p), ctypes.byref(state.c), color)


def run(scene: OracleScene, params: Params, state: OracleState) -> None:
    p = params.c()
    lib().pmo_run(ctypes.byref(scene.c), ctypes.byref(p), ctypes.byref(state.c))


def decode(scene: OracleScene, params: Params, state: OracleState):
    h, w = scene.height, scene.width
    depth = np.zeros((h, w), np.float32)
    normal = np.zeros((h, w, 3), np.float32)
    cost = np.zeros((h, w), np.float32)
    conf = np.zeros((h, w), np.float32)
    p = params.c()
    lib().pmo_decode(ctypes.byref(scene.c), ctypes.byref(p), ctypes.byref(state.c), _ptr(depth),
                     _ptr(normal), _ptr(cost), _ptr(conf))
    return depth, normal, cost, conf


def estimate(images: Sequence[np.ndarray], Ks, Rs, ts, params: Params):
    """Full single-level run + decode; returns (depth, normal, cost, confidence, state)."""
    sc = OracleScene(images, Ks, Rs, ts)
    st = OracleState(sc.width, sc.height)
    run(sc, params, st)
    d, n, c, conf = decode(sc, params, st)
    return d, n, c, conf, st


def expf(x: float) -> float:
    return float(lib().pmo_expf(float(x)))


def logf(x: float) -> float:
    return float(lib().pmo_logf(float(x)))


def rand_float(state: np.ndarray, key: int) -> float:
    """One draw; ``state`` (uint32[2]) is advanced in place."""
    assert state.dtype == np.uint32 and state.size == 2
    return float(lib().pmo_rand_float(state.ctypes.data, ctypes.c_uint32(key & 0xFFFFFFFF)))
