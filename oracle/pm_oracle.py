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
# PM_ORACLE_LIB: load another build of pm_oracle.cc instead (the mutation
# tests point it at copies with planted mistakes).
_LIB_PATH = os.environ.get("PM_ORACLE_LIB") or os.path.join(_HERE, "libpm_oracle.so")
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
        L.pmo_warp.restype = i; L.pmo_warp.argtypes = [vp, vp, i, i, vp, i, vp]
        L.pmo_ncc_at_stride.restype = f; L.pmo_ncc_at_stride.argtypes = [vp, vp, i, i, i, f, vp]
        for name in ("pmo_random_init", "pmo_prior_reinit", "pmo_run"):
            getattr(L, name).argtypes = [vp, vp, vp]
        L.pmo_sweep.argtypes = [vp, vp, vp, i, i, f]
        L.pmo_plane_sweep.argtypes = [vp, vp, vp, i]
        L.pmo_sweep_depth.restype = f; L.pmo_sweep_depth.argtypes = [vp, i, i]
        L.pmo_checkerboard.argtypes = [vp, vp, vp, i]
        L.pmo_smooth_weight.restype = f; L.pmo_smooth_weight.argtypes = [vp, i]
        L.pmo_decode.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.pmo_smoothness_cost.restype = f; L.pmo_smoothness_cost.argtypes = [vp, vp, vp, vp, i, i]
        L.pmo_surface_normal.argtypes = [vp, vp, vp, i, i, f, vp]
        L.pmo_geom_cost.restype = f; L.pmo_geom_cost.argtypes = [vp, vp, vp, i, vp, i, i]
        L.pmo_weighted_cost_geom.restype = f
        L.pmo_weighted_cost_geom.argtypes = [vp, vp, vp, vp, vp, f, vp, i, i]
        L.pmo_mc_cost.restype = f; L.pmo_mc_cost.argtypes = [vp, vp, vp, vp, vp, f, vp, i, i, vp, u]
        L.pmo_resize_area.argtypes = [vp, i, i, i, i, i, vp]
        L.pmo_postprocess.argtypes = [vp, vp, vp, vp, i, i, i, i]
        L.pmo_ncc_energy_f64.restype = ctypes.c_double
        L.pmo_ncc_energy_f64.argtypes = [vp, vp, i, i, ctypes.c_double, ctypes.c_double, ctypes.c_double, u, vp]
        L.pmo_view_weights.restype = f
        L.pmo_view_weights.argtypes = [vp, vp, vp, u, u, vp, i, i, vp, vp]
        L.pmo_refinement_hypotheses.restype = i
        L.pmo_refinement_hypotheses.argtypes = [vp, vp, vp, i, i, vp, vp, vp, u, vp]
        L.pmo_resize_linear.argtypes = [vp, i, i, i, i, vp]
        L.pmo_upsample.argtypes = [vp, vp, vp, vp, i, i, vp]
        L.pmo_clean_depthmap.argtypes = [vp, vp, i, i, vp, vp, i, f, i, f, i, f, vp]
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


def angle_weights(Rs: Sequence[np.ndarray], ts: Sequence[np.ndarray], dtype=np.float32) -> np.ndarray:
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
    return np.asarray(out, np.float64).astype(dtype)


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
    speckle_min_size: int = 0
    gap_max_size: int = 0

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


def warp(scene: OracleScene, params: Params, x: int, y: int, plane, src: int):
    """The cost vector's local affine warp into source ``src`` (0-based):
    (code, (scx, scy, jxx, jxy, jyx, jyy)); code -1 = a plane depth <= 0,
    0 = valid, 1 = behind the source camera, 2 = determinant gate."""
    pl = np.ascontiguousarray(plane, np.float32)
    out = np.zeros(6, np.float32)
    p = params.c()
    code = lib().pmo_warp(ctypes.byref(scene.c), ctypes.byref(p), int(x), int(y), _ptr(pl), int(src), _ptr(out))
    return code, out


def ncc_at_stride(scene: OracleScene, params: Params, x: int, y: int, src: int, stride: float, w) -> float:
    """The textureless fallback's strided NCC at reference pixel (x, y) on the
    warp w = (scx, scy, jxx, jxy, jyx, jyy) of source ``src``."""
    wa = np.ascontiguousarray(w, np.float32)
    p = params.c()
    return float(lib().pmo_ncc_at_stride(ctypes.byref(scene.c), ctypes.byref(p), int(x), int(y), int(src),
                                         float(stride), _ptr(wa)))


def ref_patch(scene: OracleScene, params: Params, x: int, y: int) -> np.ndarray:
    P = params.patch_size // 2 * 2 + 1
    out = np.zeros(4 + 2 * P * P, np.float32)
    p = params.c()
    lib().pmo_ref_patch(ctypes.byref(scene.c), ctypes.byref(p), x, y, _ptr(out))
    return out


def random_init(scene: OracleScene, params: Params, state: OracleState) -> None:
    p = params.c()
    lib().pmo_random_init(ctypes.byref(scene.c), ctypes.byref(p), ctypes.byref(state.c))


def plane_sweep(scene: OracleScene, params: Params, state: OracleState, num_planes: int) -> None:
    """Fronto-parallel plane sweep over ``num_planes`` log-spaced depths
    (the state a brute-force run decodes)."""
    p = params.c()
    lib().pmo_plane_sweep(ctypes.byref(scene.c), ctypes.byref(p), ctypes.byref(state.c), int(num_planes))


def sweep_depth(params: Params, i: int, n: int) -> float:
    """Depth of plane i of an n-plane sweep (float32 value)."""
    p = params.c()
    return lib().pmo_sweep_depth(ctypes.byref(p), int(i), int(n))


def estimate_plane_sweep(images: Sequence[np.ndarray], Ks, Rs, ts, params: Params, num_planes: int):
    """Plane sweep + decode; returns (depth, normal, cost, confidence, state)."""
    sc = OracleScene(images, Ks, Rs, ts)
    st = OracleState(sc.width, sc.height)
    plane_sweep(sc, params, st, num_planes)
    d, n, c, conf = decode(sc, params, st)
    return d, n, c, conf, st


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
    if params.speckle_min_size > 0 or params.gap_max_size > 0:
        postprocess(depth, normal, cost, conf, params.speckle_min_size, params.gap_max_size)
    return depth, normal, cost, conf


def postprocess(depth: np.ndarray, normal: np.ndarray, cost: np.ndarray, conf: np.ndarray, speckle_min: int,
                gap_max: int) -> None:
    """Speckle removal, gap filling and confidence, in place (estimator.cc:1194-1222)."""
    for a in (depth, normal, cost, conf):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    h, w = depth.shape
    lib().pmo_postprocess(_ptr(depth), _ptr(normal), _ptr(cost), _ptr(conf), w, h, int(speckle_min), int(gap_max))


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


def _f32(a, n=None):
    a = np.ascontiguousarray(a, np.float32)
    assert n is None or a.size == n
    return a


def smoothness_cost(scene: OracleScene, params: Params, planes: np.ndarray, plane, x: int, y: int) -> float:
    """4-neighbour smoothness of ``plane`` at (x, y) against ``planes`` (h, w, 4) (opencl_kernels.h:649-707)."""
    p = params.c()
    planes = _f32(planes, scene.width * scene.height * 4)
    pl = _f32(plane, 4)
    return float(lib().pmo_smoothness_cost(ctypes.byref(scene.c), ctypes.byref(p), _ptr(planes), _ptr(pl), x, y))


def surface_normal(scene: OracleScene, params: Params, planes: np.ndarray, x: int, y: int,
                   center_depth: float) -> np.ndarray:
    """Normal from the central differences of the 4-neighbour depths (opencl_kernels.h:725-772)."""
    p = params.c()
    planes = _f32(planes, scene.width * scene.height * 4)
    out = np.zeros(3, np.float32)
    lib().pmo_surface_normal(ctypes.byref(scene.c), ctypes.byref(p), _ptr(planes), x, y, float(center_depth),
                             _ptr(out))
    return out


def geom_cost(scene: OracleScene, params: Params, state: OracleState, sidx: int, plane, x: int, y: int) -> float:
    """Forward-backward reprojection cost against source sidx's previous depth (opencl_kernels.h:779-861)."""
    p = params.c()
    pl = _f32(plane, 4)
    return float(lib().pmo_geom_cost(ctypes.byref(scene.c), ctypes.byref(p), ctypes.byref(state.c), sidx,
                                     _ptr(pl), x, y))


def weighted_cost_geom(scene: OracleScene, params: Params, state: OracleState, cv, vw, wnorm: float, plane,
                       x: int, y: int) -> float:
    p = params.c()
    n = scene.n_views - 1
    cv, vw, pl = _f32(cv, n), _f32(vw, n), _f32(plane, 4)
    return float(lib().pmo_weighted_cost_geom(ctypes.byref(scene.c), ctypes.byref(p), ctypes.byref(state.c),
                                              _ptr(cv), _ptr(vw), float(wnorm), _ptr(pl), x, y))


def view_weights(scene: OracleScene, params: Params, cand_costs, flags: int, prev_sel: int, plane, x: int, y: int):
    """View weights of acmmp_patchmatch (opencl_kernels.h:1452-1563) for candidate costs
    (8, num_src); returns (weights, selection mask, weight sum)."""
    p = params.c()
    n = scene.n_views - 1
    cc = _f32(cand_costs, 8 * n)
    pl = _f32(plane, 4)
    w = np.zeros(n, np.float32)
    sel = np.zeros(1, np.uint32)
    wn = lib().pmo_view_weights(ctypes.byref(scene.c), ctypes.byref(p), _ptr(cc), ctypes.c_uint32(flags),
                                ctypes.c_uint32(prev_sel), _ptr(pl), x, y, _ptr(w), _ptr(sel))
    return w, int(sel[0]), float(wn)


def refinement_hypotheses(scene: OracleScene, params: Params, planes: np.ndarray, x: int, y: int, plane_now,
                          rng: np.ndarray, key: int, prior_plane=None) -> np.ndarray:
    """The refinement hypotheses of acmmp_patchmatch at (x, y) (opencl_kernels.h:1669-1742):
    (n, 4) planes, n = 6 or 5; ``rng`` advances in place."""
    assert rng.dtype == np.uint32 and rng.size == 2
    p = params.c()
    planes = _f32(planes, scene.width * scene.height * 4)
    out = np.zeros((6, 4), np.float32)
    pp = None if prior_plane is None else _f32(prior_plane, 4)
    pn = _f32(plane_now, 4)
    n = lib().pmo_refinement_hypotheses(ctypes.byref(scene.c), ctypes.byref(p), _ptr(planes), x, y, _ptr(pn),
                                        _ptr(pp), rng.ctypes.data, ctypes.c_uint32(key & 0xFFFFFFFF), _ptr(out))
    return out[:n]


def mc_cost(scene: OracleScene, params: Params, state: OracleState, cv, vw, wnorm: float, plane, x: int, y: int,
            rng: np.ndarray, key: int) -> float:
    """32-draw view-sampled cost (opencl_kernels.h:910-969); ``rng`` (uint32[2]) advances in place."""
    assert rng.dtype == np.uint32 and rng.size == 2
    p = params.c()
    n = scene.n_views - 1
    cv, vw, pl = _f32(cv, n), _f32(vw, n), _f32(plane, 4)
    return float(lib().pmo_mc_cost(ctypes.byref(scene.c), ctypes.byref(p), ctypes.byref(state.c), _ptr(cv),
                                   _ptr(vw), float(wnorm), _ptr(pl), x, y, rng.ctypes.data,
                                   ctypes.c_uint32(key & 0xFFFFFFFF)))


# ---------------------------------------------------------------------------
# Hierarchy levels (estimator.cc:134-255, 724-821)
# ---------------------------------------------------------------------------
def resize_area(img: np.ndarray, nw: int, nh: int, round_int: bool = False) -> np.ndarray:
    src = np.ascontiguousarray(img, np.float32)
    out = np.zeros((nh, nw), np.float32)
    lib().pmo_resize_area(_ptr(src), src.shape[1], src.shape[0], nw, nh, 1 if round_int else 0, _ptr(out))
    return out


def resize_linear(img: np.ndarray, nw: int, nh: int) -> np.ndarray:
    src = np.ascontiguousarray(img, np.float32)
    out = np.zeros((nh, nw), np.float32)
    lib().pmo_resize_linear(_ptr(src), src.shape[1], src.shape[0], nw, nh, _ptr(out))
    return out


def level_views(images: Sequence[np.ndarray], Ks, scale: float, quantize_half: bool = False):
    """One pyramid level (BuildImagePyramid, estimator.cc:727-756).

    ``images`` as given to the estimator: uint8 (resampled as u8, rounded,
    then /255) or float in [0, 1].  Returns float32 images and scaled K."""
    sc32 = np.float32(scale)
    out_imgs, out_Ks = [], []
    for im, K in zip(images, Ks):
        h, w = im.shape
        if scale < 0.999:
            nw = max(int(np.float32(w) * sc32), 1)
            nh = max(int(np.float32(h) * sc32), 1)
        else:
            nw, nh = w, h
        if im.dtype == np.uint8:
            u = im.astype(np.float32) if (nw, nh) == (w, h) else resize_area(im.astype(np.float32), nw, nh, True)
            lv = u * np.float32(1.0 / 255.0)
        else:
            lv = np.asarray(im, np.float32) if (nw, nh) == (w, h) else resize_area(im, nw, nh, False)
        if quantize_half:
            lv = lv.astype(np.float16).astype(np.float32)
        K2 = np.array(K, np.float64).copy()
        if (nw, nh) != (w, h):
            fx = float(np.float32(nw) / np.float32(w))
            fy = float(np.float32(nh) / np.float32(h))
            K2[0, 0] *= fx; K2[0, 2] *= fx; K2[1, 1] *= fy; K2[1, 2] *= fy
        out_imgs.append(lv)
        out_Ks.append(K2)
    return out_imgs, out_Ks


def upsample(scene: OracleScene, params: Params, state: OracleState, coarse_depth: np.ndarray,
             coarse_normal: np.ndarray) -> None:
    ch, cw = coarse_depth.shape
    cd = np.ascontiguousarray(coarse_depth, np.float32)
    cn = np.ascontiguousarray(coarse_normal, np.float32)
    p = params.c()
    lib().pmo_upsample(ctypes.byref(scene.c), ctypes.byref(p), _ptr(cd), _ptr(cn), cw, ch, ctypes.byref(state.c))


def run_level(images: Sequence[np.ndarray], Ks, Rs, ts, params: Params, level: int, levels: int, coarse=None,
              prev_depths=None, quantize_half: bool = False):
    """One pyramid level (RunLevel, estimator.cc:172-255).

    ``coarse``: (depth, normal) of the previous level of the same view, or
    None at level 0.  ``prev_depths``: {source slot: depth map of any size},
    resized to the level bilinearly (UploadPreviousDepths, estimator.cc:
    66-116).  Returns (depth, normal, cost, confidence, state)."""
    scale = 1.0 / 2.0 ** (levels - 1 - level)
    lv_images, lv_Ks = level_views(images, Ks, scale, quantize_half)
    sc = OracleScene(lv_images, lv_Ks, Rs, ts)
    h, w = sc.height, sc.width
    stack, mask = None, 0
    if prev_depths and params.geom_weight > 0:
        stack = np.zeros((sc.n_views - 1, h, w), np.float32)
        for slot, d in prev_depths.items():
            if d is None or not (0 <= slot < sc.n_views - 1):
                continue
            d = np.asarray(d, np.float32)
            stack[slot] = d if d.shape == (h, w) else resize_linear(d, w, h)
            mask |= 1 << slot
    st = OracleState(w, h, prev_depths=stack, prev_depth_mask=mask)
    if coarse is None:
        random_init(sc, params, st)
    else:
        upsample(sc, params, st, coarse[0], coarse[1])
    prior_reinit(sc, params, st)
    for it in range(params.max_iterations):
        sweep(sc, params, st, 0, it)
        sweep(sc, params, st, 1, it)
    if params.checkerboard_filter:
        checkerboard(sc, params, st, 0)
        checkerboard(sc, params, st, 1)
    d, nrm, c, conf = decode(sc, params, st)
    return d, nrm, c, conf, st


def estimate_levels(images: Sequence[np.ndarray], Ks, Rs, ts, params: Params, levels: int,
                    quantize_half: bool = False):
    """Coarse-to-fine run over ``levels`` pyramid levels (Run / RunLevel,
    estimator.cc:134-255); returns (depth, normal, cost, confidence, state)
    of the full-resolution level."""
    n = max(1, int(levels))
    coarse = None
    for level in range(n):
        d, nrm, c, conf, st = run_level(images, Ks, Rs, ts, params, level, n, coarse,
                                        quantize_half=quantize_half)
        coarse = (d, nrm)
    return d, nrm, c, conf, st


def estimate_cluster(refs, params: Params, levels: int, geom_weight: float):
    """DepthmapClusterEstimator::Run (cluster_estimator.cc:82-156).

    ``refs``: list of dicts with images, Ks, Rs, ts (reference first),
    depth_min, depth_max and links [(source slot, from reference)].  Level
    by level, every reference runs with the previous level's depth maps of
    the references its links name.  Returns [(depth, normal, cost,
    confidence)] per reference."""
    import dataclasses
    n = max(1, int(levels))
    k = len(refs)
    coarse = [None] * k
    prev = [None] * k
    out = [None] * k
    for level in range(n):
        cur = [None] * k
        for i, r in enumerate(refs):
            p = dataclasses.replace(params, depth_min=r["depth_min"], depth_max=r["depth_max"],
                                    geom_weight=geom_weight)
            pd = {}
            if level > 0 and geom_weight > 0:
                for slot, frm in r.get("links", []):
                    if 0 <= frm < k and prev[frm] is not None:
                        pd[slot] = prev[frm]
            d, nrm, c, conf, _ = run_level(r["images"], r["Ks"], r["Rs"], r["ts"], p, level, n, coarse[i], pd)
            coarse[i] = (d, nrm)
            cur[i] = d
            out[i] = (d, nrm, c, conf)
        prev = cur
    return out


def ncc_energy_f64(scene: OracleScene, params: Params, x: int, y: int, d: float, a: float, b: float,
                   views: int):
    """Float64 photometric cost of one pixel (the model of pm_ncc_cost_grad); returns (cost, views used)."""
    p = params.c()
    used = ctypes.c_int(0)
    e = lib().pmo_ncc_energy_f64(ctypes.byref(scene.c), ctypes.byref(p), int(x), int(y), float(d), float(a), float(b),
                                 ctypes.c_uint32(int(views)), ctypes.byref(used))
    return float(e), int(used.value)


def clean_depthmap(ref_depth: np.ndarray, ref_normal, views, same_depth_threshold: float = 0.01,
                   min_consistent_views: int = 3, carving_threshold: float = 0.05, max_carved_views: int = 2,
                   grazing_cos_threshold: float = 0.2) -> np.ndarray:
    """GPUDepthmapCleaner::Clean for one reference (cleaner.cc:95-242).

    ``views``: [(K, R, t)] of the reference then [(K, R, t, depth)] of each
    neighbour -- given as a list whose first entry is (K, R, t) and the rest
    (K, R, t, depth).  ``ref_normal``: (H, W, 3) camera-frame normals or None.
    Neighbour depth maps (and the normal) of another size are resized to the
    reference size with the INTER_LINEAR restatement, as Clean does.
    """
    ref = np.ascontiguousarray(ref_depth, np.float32)
    h, w = ref.shape
    normal = None
    if ref_normal is not None:
        nrm = np.asarray(ref_normal, np.float32).reshape(ref_normal.shape[0], ref_normal.shape[1], 3)
        if nrm.shape[:2] != (h, w):
            nrm = np.stack([resize_linear(nrm[..., c], w, h) for c in range(3)], -1)
        normal = np.ascontiguousarray(nrm, np.float32)
    K0, R0, t0 = views[0][:3]
    cams = (PmoCamera * len(views))()
    cams[0] = make_camera(K0, R0, t0, w, h)
    srcs = []
    for j, (K, R, t, d) in enumerate(views[1:]):
        cams[j + 1] = make_camera(K, R, t, w, h)
        d = np.asarray(d, np.float32)
        srcs.append(np.ascontiguousarray(d) if d.shape == (h, w) else resize_linear(d, w, h))
    ptrs = (ctypes.c_void_p * max(len(srcs), 1))(*[a.ctypes.data for a in srcs])
    n = len(srcs)
    out = np.zeros((h, w), np.float32)
    lib().pmo_clean_depthmap(_ptr(ref), _ptr(normal), w, h, ptrs, cams, n, float(same_depth_threshold),
                             int(min(min_consistent_views, n + 1)), float(carving_threshold), int(max_carved_views),
                             float(grazing_cos_threshold), _ptr(out))
    return out


def sfm_planar_prior(K, R, t, w: int, h: int, depth_min: float, depth_max: float, points):
    """ComputeSfMPlanarPrior (estimator.cc:838-1107) restated with numpy:
    Delaunay by Qhull (scipy.spatial), planes from the float64 SVD null
    vector of [X 1], the parametric rasteriser in float32.  Triangles are
    rasterised in canonical order (vertex indices ascending, sorted) --
    cv::Subdiv2D's order is not reproducible.  Returns (planes (h, w, 4),
    masks (h, w) uint32, triangles kept)."""
    from scipy.spatial import Delaunay
    K = np.asarray(K, np.float64); R = np.asarray(R, np.float64); t = np.asarray(t, np.float64)
    f32 = np.float32
    fx, fy, cx, cy = f32(K[0, 0]), f32(K[1, 1]), f32(K[0, 2]), f32(K[1, 2])
    planes = np.zeros((h, w, 4), np.float32)
    planes[..., 2] = -1.0
    masks = np.zeros((h, w), np.uint32)
    proj = []
    for X in np.asarray(points, np.float64).reshape(-1, 3):
        xc, yc, zc = R @ X + t
        if zc <= 0.0:
            continue
        depth = f32(zc)
        if depth < depth_min or depth > depth_max:
            continue
        u = f32(float(fx) * (xc / zc) + float(cx))
        v = f32(float(fy) * (yc / zc) + float(cy))
        if u < 0 or u >= w or v < 0 or v >= h:
            continue
        proj.append((u, v, depth, np.array([xc, yc, zc], np.float32)))
    if len(proj) < 4:
        return planes, masks, 0
    uv = np.array([(p[0], p[1]) for p in proj], np.float64)
    key = {int(p[1]) * w + int(p[0]): i for i, p in enumerate(proj)}
    simp = Delaunay(uv).simplices
    tris = sorted(tuple(sorted(int(i) for i in s)) for s in simp)
    kept = []
    for tri in tris:
        pts = [(int(uv[i, 0]), int(uv[i, 1])) for i in tri]
        if any(not (0 <= x < w and 0 <= y < h) for x, y in pts):
            continue
        if any(y * w + x not in key for x, y in pts):
            continue
        pi = [key[y * w + x] for x, y in pts]
        ds = [proj[i][2] for i in pi]
        if min(ds) > 0 and f32(max(ds)) / f32(min(ds)) > f32(1.5):
            continue
        ax, ay = f32(pts[1][0] - pts[0][0]), f32(pts[1][1] - pts[0][1])
        bx, by = f32(pts[2][0] - pts[0][0]), f32(pts[2][1] - pts[0][1])
        if f32(0.5) * abs(ax * by - ay * bx) > f32(20000.0):
            continue
        if len(set(pi)) < 3:
            # two vertices in one pixel map to one SfM point (the key map
            # keeps the last point of a pixel, :914-917): the reference takes
            # an arbitrary null vector of a rank-2 matrix there; skipped here
            # and in the library
            continue
        A = np.array([[*proj[i][3].astype(np.float64), 1.0] for i in pi])
        B = np.linalg.svd(A)[2][3]
        nrm = np.linalg.norm(B[:3])
        if nrm < 1e-8:
            continue
        B = B / (-nrm if B[3] < 0 else nrm)
        kept.append((pts, B.astype(np.float32)))
    owner = np.zeros((h, w), np.int64)
    for ti, (pts, _) in enumerate(kept):
        (x1, y1), (x2, y2), (x3, y3) = [(f32(x), f32(y)) for x, y in pts]
        L = [np.hypot(f32(pts[a][0] - pts[b][0]), f32(pts[a][1] - pts[b][1])).astype(np.float32)
             for a, b in ((0, 1), (0, 2), (1, 2))]
        max_edge = max(L)
        if max_edge < 1.0:
            continue
        step = f32(1.0) / f32(max_edge)
        p = f32(0.0)
        qs = [f32(0.0)]
        while qs[-1] < 1.0:
            qs.append(f32(qs[-1] + step))
        qs = np.array(qs, np.float32)
        while p < f32(1.0):
            q = qs[qs < f32(f32(1.0) - p)]
            r = (f32(1.0) - p) - q
            px = ((p * x1 + q * x2) + r * x3).astype(np.float32).astype(np.int64)
            py = ((p * y1 + q * y2) + r * y3).astype(np.float32).astype(np.int64)
            ok = (px >= 0) & (px < w) & (py >= 0) & (py < h)
            owner[py[ok], px[ok]] = ti + 1
            p = f32(p + step)
    ys, xs = np.nonzero(owner)
    for y, x in zip(ys, xs):
        pl = kept[owner[y, x] - 1][1]
        rx = (f32(x) - cx) / fx
        ry = (f32(y) - cy) / fy
        denom = f32(f32(pl[0] * rx) + f32(pl[1] * ry)) + f32(pl[2] * f32(1.0))
        d = -pl[3] / denom if abs(denom) > 1e-6 else f32(0.0)
        if depth_min <= d <= depth_max:
            planes[y, x] = pl
            masks[y, x] = 1
    return planes, masks, len(kept)


def select_neighbors(origins, coords, visibility, track_counts, num_neighbors: int, min_point_best: int,
                     min_point_all: int, theta_min_deg: float, theta_max_deg: float):
    """SelectNeighbors (dense_helpers.cc:331-568) in plain Python; ties of
    equal score ordered by shot index, the 'all' extras by index."""
    import math
    N = len(origins)
    tmin, tmax = math.radians(theta_min_deg), math.radians(theta_max_deg)
    cos_lo, cos_hi = math.cos(tmax), math.cos(tmin)
    per = [dict() for _ in range(N)]
    for P, vis, nt in zip(coords, visibility, track_counts):
        cams = [c for c in vis if 0 <= c < N]
        if len(cams) < 2:
            continue
        vs = [np.asarray(origins[c], float) - np.asarray(P, float) for c in cams]
        ns = [float(np.linalg.norm(v)) for v in vs]
        for i in range(len(cams)):
            if ns[i] < 1e-12:
                continue
            for j in range(i + 1, len(cams)):
                if ns[j] < 1e-12:
                    continue
                c = min(1.0, max(-1.0, float(vs[i] @ vs[j]) / (ns[i] * ns[j])))
                good = cos_lo < c < cos_hi
                for a, b in ((cams[i], cams[j]), (cams[j], cams[i])):
                    s = per[a].setdefault(b, [0, 0, 0.0])
                    s[0] += nt if good else 0
                    s[1] += nt
                    s[2] += c * nt
    best, allv = [], []
    for i in range(N):
        nb, na = [], []
        for o, (sb, cnt, sc) in per[i].items():
            if cnt == 0:
                continue
            if cnt > min_point_all:
                na.append(o)
            if cnt <= min_point_best:
                continue
            avg = math.acos(min(1.0, max(-1.0, sc / cnt)))
            if avg < tmin or avg > tmax:
                continue
            if sb > min_point_best:
                nb.append((-sb, o))
        nb.sort()
        b = [i] + [o for _, o in nb[:num_neighbors]]
        best.append(b)
        allv.append(b + [o for o in sorted(na) if o not in set(b)])
    return best, allv


# ---------------------------------------------------------------------------
# FP64 evaluation: libpm_oracle64.so is pm_oracle.cc built with -DPMO_F64,
# the same algorithm (same random draws) with every value, array and
# interface field in double and libm's exp / log.  Views are passed as their
# float32 values widened exactly; cameras keep their float64 K, R, t.
# ---------------------------------------------------------------------------
_LIB64_PATH = os.path.join(_HERE, "libpm_oracle64.so")
_lib64 = None


class PmoCamera64(ctypes.Structure):
    _fields_ = [("K", ctypes.c_double * 9), ("R", ctypes.c_double * 9), ("t", ctypes.c_double * 3),
                ("width", ctypes.c_int), ("height", ctypes.c_int)]


class PmoParams64(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double if t is ctypes.c_float else t) for n, t in PmoParams._fields_]


def lib64():
    global _lib64
    if _lib64 is None:
        if not os.path.exists(_LIB64_PATH):
            build()
        L = ctypes.CDLL(_LIB64_PATH)
        i, vp = ctypes.c_int, ctypes.c_void_p
        L.pmo_run.argtypes = [vp, vp, vp]
        L.pmo_decode.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.pmo_postprocess.argtypes = [vp, vp, vp, vp, i, i, i, i]
        L.pmo_score_planes.argtypes = [vp, vp, i, vp, vp, vp, vp, vp]
        L.pmo_set_threads.argtypes = [i]
        L.pmo_max_threads.restype = i
        _lib64 = L
    return _lib64


def params64(params: Params) -> PmoParams64:
    """The pmo_params of ``params`` with its real fields in double."""
    p = PmoParams64()
    p.half_patch = params.patch_size // 2
    p.sigma_spatial = params.sigma_spatial
    p.sigma_color = params.sigma_color
    p.top_k = params.top_k
    p.census_weight = 1.0 if params.use_census else 0.0
    p.depth_min = params.depth_min
    p.depth_max = params.depth_max
    p.max_iterations = params.max_iterations
    p.smooth_weight = params.smooth_weight
    p.anchor_views = params.anchor_views
    p.geom_weight = params.geom_weight
    p.var_threshold = params.var_threshold
    p.checkerboard_filter = 1 if params.checkerboard_filter else 0
    return p


class OracleScene64(OracleScene):
    """Views and cameras packed for the FP64 build."""

    def __init__(self, images: Sequence[np.ndarray], Ks, Rs, ts):
        super().__init__(images, Ks, Rs, ts)
        self.images = self.images.astype(np.float64)
        cams = (PmoCamera64 * self.n_views)()
        for k in range(self.n_views):
            c = cams[k]
            c.K[:] = np.asarray(Ks[k], np.float64).ravel().tolist()
            c.R[:] = np.asarray(Rs[k], np.float64).ravel().tolist()
            c.t[:] = np.asarray(ts[k], np.float64).ravel().tolist()
            c.width, c.height = int(self.widths[k]), int(self.heights[k])
        self.cams = cams
        self.aw = angle_weights(Rs, ts, np.float64) if self.n_views > 1 else np.zeros(1, np.float64)
        self.c.images = self.images.ctypes.data
        self.c.cams = ctypes.addressof(self.cams)
        self.c.angle_weights = self.aw.ctypes.data


class OracleState64(OracleState):
    def __init__(self, w: int, h: int):
        super().__init__(w, h)
        self.planes = np.zeros((h, w, 4), np.float64)
        self.costs = np.zeros((h, w), np.float64)
        self.c.planes = self.planes.ctypes.data
        self.c.costs = self.costs.ctypes.data


def estimate_f64(images: Sequence[np.ndarray], Ks, Rs, ts, params: Params):
    """Full single-level run + decode in double; returns (depth, normal, cost, confidence), float64."""
    sc = OracleScene64(images, Ks, Rs, ts)
    st = OracleState64(sc.width, sc.height)
    p = params64(params)
    L = lib64()
    L.pmo_run(ctypes.byref(sc.c), ctypes.byref(p), ctypes.byref(st.c))
    h, w = sc.height, sc.width
    depth = np.zeros((h, w), np.float64)
    normal = np.zeros((h, w, 3), np.float64)
    cost = np.zeros((h, w), np.float64)
    conf = np.zeros((h, w), np.float64)
    L.pmo_decode(ctypes.byref(sc.c), ctypes.byref(p), ctypes.byref(st.c), _ptr(depth), _ptr(normal), _ptr(cost),
                 _ptr(conf))
    if params.speckle_min_size > 0 or params.gap_max_size > 0:
        L.pmo_postprocess(_ptr(depth), _ptr(normal), _ptr(cost), _ptr(conf), w, h, int(params.speckle_min_size),
                          int(params.gap_max_size))
    return depth, normal, cost, conf


def score_planes_f64(images: Sequence[np.ndarray], Ks, Rs, ts, params: Params, xy: np.ndarray, planes: np.ndarray):
    """score_planes of the FP64 build: per-view costs (n, num_src), float64."""
    sc = OracleScene64(images, Ks, Rs, ts)
    xy = np.ascontiguousarray(xy, np.int32)
    planes = np.ascontiguousarray(planes, np.float64)
    n = xy.shape[0]
    costs = np.zeros((n, sc.n_views - 1), np.float64)
    init = np.zeros(n, np.float64)
    sel = np.zeros(n, np.uint32)
    p = params64(params)
    lib64().pmo_score_planes(ctypes.byref(sc.c), ctypes.byref(p), n, _ptr(xy), _ptr(planes), _ptr(costs),
                             _ptr(init), _ptr(sel))
    return costs, init, sel

