"""Depth-map files of the dense pipeline and the caller contract around them.

``<image>.raw.npz`` and ``<image>.clean.npz`` under ``<dataset>/depthmaps``
(UndistortedDataSet.depthmap_file, opensfm/dataset.py:1120-1125) are an
``np.savez`` archive compressed as one LZ4 frame (dataset.py:1416-1452,
1464-1488): raw = depth, plane (the unit normals) and an optional
confidence; clean = depth, plane, score and an optional confidence.  The
raw writer accepts score / nghbr / nghbrs for call-site compatibility and
does not store them; the readers return None for absent fields, so files
written by older versions (with them) load as well.

The LZ4 frame codec is the library's own (pydense.lz4_frame_compress /
lz4_frame_decompress); the reference uses python-lz4 (>= 4.4.5).
"""
from __future__ import annotations

import io
import os
from typing import List, Optional, Sequence, Tuple

import numpy as np


def _codec():
    from . import native
    return native()


def depthmap_file(data_path: str, image: str, suffix: str) -> str:
    """dataset.py:1120-1125: <data_path>/depthmaps/<image>.<suffix>."""
    return os.path.join(data_path, "depthmaps", image + "." + suffix)


def _write(path: str, arrays: dict) -> None:
    buf = io.BytesIO()
    np.savez(buf, **arrays)
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "wb") as f:
        f.write(_codec().lz4_frame_compress(buf.getbuffer()))


def _read(path: str):
    with open(path, "rb") as f:
        raw = _codec().lz4_frame_decompress(f.read())
    return np.load(io.BytesIO(raw), allow_pickle=False)


def save_raw_depthmap(path: str, depth: np.ndarray, plane: np.ndarray, score=None, nghbr=None,
                      nghbrs: Optional[Sequence[str]] = None, confidence: Optional[np.ndarray] = None) -> None:
    """UndistortedDataSet._write_raw_depthmap (dataset.py:1416-1435)."""
    arrays = dict(depth=depth, plane=plane)
    if confidence is not None:
        arrays["confidence"] = confidence
    _write(path, arrays)


def load_raw_depthmap(path: str) -> Tuple[np.ndarray, np.ndarray, Optional[np.ndarray], Optional[np.ndarray],
                                           Optional[np.ndarray], Optional[np.ndarray]]:
    """UndistortedDataSet.load_raw_depthmap (dataset.py:1437-1452):
    (depth, plane, score, nghbr, nghbrs, confidence)."""
    o = _read(path)
    try:
        get = lambda k: o[k] if k in o else None  # noqa: E731
        return o["depth"], o["plane"], get("score"), get("nghbr"), get("nghbrs"), get("confidence")
    finally:
        o.close()


def save_clean_depthmap(path: str, depth: np.ndarray, plane: np.ndarray, score: np.ndarray,
                        confidence: Optional[np.ndarray] = None) -> None:
    """UndistortedDataSet._write_clean_depthmap (dataset.py:1464-1476)."""
    arrays = dict(depth=depth, plane=plane, score=score)
    if confidence is not None:
        arrays["confidence"] = confidence
    _write(path, arrays)


def load_clean_depthmap(path: str) -> Tuple[np.ndarray, np.ndarray, np.ndarray, Optional[np.ndarray]]:
    """UndistortedDataSet.load_clean_depthmap (dataset.py:1478-1488)."""
    o = _read(path)
    try:
        conf = o["confidence"] if "confidence" in o else None
        return o["depth"], o["plane"], o["score"], conf
    finally:
        o.close()


def raw_outputs(depth, normal, cost, confidence, max_cost: float):
    """What _run_cluster_raw does to an estimator result before saving it
    (opensfm/dense/depthmaps.py:297-319): float32 copies, NaN/Inf depths
    zeroed with their normals, and depths of pixels with cost > max_cost
    zeroed when max_cost > 0.  Returns (depth, normal, score, confidence)."""
    depth = np.array(depth, dtype=np.float32)
    normal = np.array(normal, dtype=np.float32)
    cost = np.array(cost, dtype=np.float32)
    confidence = np.array(confidence, dtype=np.float32)
    bad = ~np.isfinite(depth)
    if np.any(bad):
        depth[bad] = 0.0
        normal[np.broadcast_to(bad[:, :, None], normal.shape)] = 0.0
    if max_cost > 0:
        depth[cost > max_cost] = 0.0
    return depth, normal, cost, confidence


def save_cluster_results(data_path: str, ref_ids: List[str], results, max_cost: float,
                         neighbours: Optional[dict] = None) -> List[str]:
    """Save DepthmapClusterEstimator.run() results as raw depth maps, as
    _run_cluster_raw does (depthmaps.py:288-323); returns the saved ids."""
    done = []
    for ref_id, (depth, normal, cost, confidence) in zip(ref_ids, results):
        d, n, s, c = raw_outputs(depth, normal, cost, confidence, max_cost)
        nghbrs = list((neighbours or {}).get(ref_id, []))
        save_raw_depthmap(depthmap_file(data_path, ref_id, "raw.npz"), d, n, s, np.zeros((0,), np.int32), nghbrs, c)
        done.append(ref_id)
    return done
