"""Seeded synthetic multi-view scenes for the PatchMatch depth path.

The reference has no image renderer (its synthetic_data package produces
points and tracks only), so scenes are ray-cast here: convex planar quads,
textured by 4-octave 3-D value noise in [0, 1] plus per-view Gaussian sensor
noise, seen by pinhole cameras.  Conventions follow the reference estimator:
x_cam = R X + t, pixel = K x_cam with pixel centres at integer coordinates,
depth = z in the reference camera frame (opencl_kernels.h:102-173).

Every scene returns a ``Scene`` holding float32 images in [0, 1] (plus
noise), float64 K/R/t per view, the ground-truth reference depth, and the
depth range.  View 0 is the reference.
"""
from __future__ import annotations

import dataclasses
from typing import List, Sequence, Tuple

import numpy as np


@dataclasses.dataclass
class Scene:
    images: List[np.ndarray]          # float32 (H, W) per view, view 0 = reference
    Ks: List[np.ndarray]              # float64 (3, 3)
    Rs: List[np.ndarray]              # float64 (3, 3)
    ts: List[np.ndarray]              # float64 (3,)
    gt_depth: np.ndarray              # float32 (H, W), 0 where the ray hits nothing
    depth_min: float
    depth_max: float
    name: str = ""

    @property
    def num_src(self) -> int:
        return len(self.images) - 1

    def crop_rows(self, y0: int, y1: int) -> "Scene":
        """Reference rows [y0, y1) with the principal point shifted; sources unchanged."""
        K0 = self.Ks[0].copy()
        K0[1, 2] -= y0
        return Scene([self.images[0][y0:y1].copy()] + [im for im in self.images[1:]],
                     [K0] + list(self.Ks[1:]), list(self.Rs), list(self.ts),
                     self.gt_depth[y0:y1].copy(), self.depth_min, self.depth_max,
                     self.name + f"[rows {y0}:{y1}]")


# ---------------------------------------------------------------------------
# Value noise
# ---------------------------------------------------------------------------
class ValueNoise3D:
    """Lattice value noise with smoothstep-weighted trilinear interpolation."""

    def __init__(self, seed: int, table: int = 64):
        rng = np.random.default_rng(seed)
        self.n = table
        self.values = rng.random((table, table, table)).astype(np.float64)

    def _octave(self, p: np.ndarray) -> np.ndarray:
        n = self.n
        fl = np.floor(p)
        f = p - fl
        i = fl.astype(np.int64) % n
        f = f * f * (3.0 - 2.0 * f)
        out = np.zeros(p.shape[0])
        for dx in (0, 1):
            wx = f[:, 0] if dx else 1.0 - f[:, 0]
            ix = (i[:, 0] + dx) % n
            for dy in (0, 1):
                wy = f[:, 1] if dy else 1.0 - f[:, 1]
                iy = (i[:, 1] + dy) % n
                for dz in (0, 1):
                    wz = f[:, 2] if dz else 1.0 - f[:, 2]
                    iz = (i[:, 2] + dz) % n
                    out += wx * wy * wz * self.values[ix, iy, iz]
        return out

    def fbm(self, p: np.ndarray, base_cell: float, octaves: int = 4) -> np.ndarray:
        """Sum of ``octaves`` octaves (cell size halves, amplitude halves), in [0, 1]."""
        total = np.zeros(p.shape[0])
        norm = 0.0
        amp = 1.0
        cell = base_cell
        for o in range(octaves):
            total += amp * self._octave(p / cell + 17.31 * o)
            norm += amp
            amp *= 0.5
            cell *= 0.5
        return total / norm


# ---------------------------------------------------------------------------
# Cameras
# ---------------------------------------------------------------------------
def look_at(center: np.ndarray, target: np.ndarray, up=(0.0, -1.0, 0.0)) -> Tuple[np.ndarray, np.ndarray]:
    """R, t of a camera at ``center`` looking at ``target`` (image y down)."""
    z = np.asarray(target, float) - np.asarray(center, float)
    z /= np.linalg.norm(z)
    x = np.cross(np.asarray(up, float), z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    R = np.stack([x, y, z])
    t = -R @ np.asarray(center, float)
    return R, t


def intrinsics(width: int, height: int, focal: float) -> np.ndarray:
    return np.array([[focal, 0.0, (width - 1) / 2.0],
                     [0.0, focal, (height - 1) / 2.0],
                     [0.0, 0.0, 1.0]])


# ---------------------------------------------------------------------------
# Ray casting of convex planar quads
# ---------------------------------------------------------------------------
def _quad_plane(quad: np.ndarray) -> Tuple[np.ndarray, float]:
    n = np.cross(quad[1] - quad[0], quad[2] - quad[0])
    n /= np.linalg.norm(n)
    return n, -float(n @ quad[0])


def render(quads: Sequence[np.ndarray], noise: ValueNoise3D, base_cell: float,
           K: np.ndarray, R: np.ndarray, t: np.ndarray, width: int, height: int,
           sensor_sigma: float, rng: np.random.Generator,
           background: float = 0.5, supersample: int = 1) -> Tuple[np.ndarray, np.ndarray]:
    """Ray-cast ``quads`` (each 4x3 world corners, convex, coplanar).

    With ``supersample`` s > 1 the radiance is the mean of s x s rays per
    pixel (box filter); depth is the centre ray's.
    Returns (image float32, depth float32 = z in this camera, 0 = no hit).
    """
    if supersample > 1:
        s = supersample
        Ks = K.copy()
        Ks[0, 0] *= s; Ks[1, 1] *= s
        Ks[0, 2] = (K[0, 2] + 0.5) * s - 0.5
        Ks[1, 2] = (K[1, 2] + 0.5) * s - 0.5
        hi, _ = render(quads, noise, base_cell, Ks, R, t, width * s, height * s, 0.0, rng, background, 1)
        img = hi.astype(np.float64).reshape(height, s, width, s).mean(axis=(1, 3))
        _, depth = render(quads, noise, base_cell, K, R, t, width, height, 0.0, rng, background, 1)
        img = img + rng.normal(0.0, sensor_sigma, img.shape)
        return img.astype(np.float32), depth
    C = -R.T @ t
    Kinv = np.linalg.inv(K)
    zbuf = np.full((height, width), np.inf)
    hit = np.zeros((height, width, 3))
    for quad in quads:
        quad = np.asarray(quad, float)
        n, d = _quad_plane(quad)
        qc = (R @ quad.T).T + t
        if np.all(qc[:, 2] > 1e-3):
            pr = (K @ qc.T).T
            uv = pr[:, :2] / pr[:, 2:3]
            x0 = int(max(0, np.floor(uv[:, 0].min()) - 1))
            x1 = int(min(width, np.ceil(uv[:, 0].max()) + 2))
            y0 = int(max(0, np.floor(uv[:, 1].min()) - 1))
            y1 = int(min(height, np.ceil(uv[:, 1].max()) + 2))
        else:
            x0, x1, y0, y1 = 0, width, 0, height
        if x0 >= x1 or y0 >= y1:
            continue
        ys, xs = np.mgrid[y0:y1, x0:x1]
        pix = np.stack([xs.ravel(), ys.ravel(), np.ones(xs.size)], 0).astype(float)
        dirs = R.T @ (Kinv @ pix)                      # world ray directions (z_cam = 1)
        denom = n @ dirs
        with np.errstate(divide="ignore", invalid="ignore"):
            lam = -(n @ C + d) / denom                 # = depth along the optical axis
        X = C[:, None] + dirs * lam[None, :]
        inside = (lam > 1e-6) & np.isfinite(lam)
        for k in range(4):
            a, b = quad[k], quad[(k + 1) % 4]
            e = np.cross(b - a, X.T - a) @ n
            inside &= e >= -1e-9
        zb = zbuf[y0:y1, x0:x1].ravel()
        take = inside & (lam < zb)
        if not np.any(take):
            continue
        zb[take] = lam[take]
        zbuf[y0:y1, x0:x1] = zb.reshape(y1 - y0, x1 - x0)
        hb = hit[y0:y1, x0:x1].reshape(-1, 3)
        hb[take] = X[:, take].T
        hit[y0:y1, x0:x1] = hb.reshape(y1 - y0, x1 - x0, 3)
    valid = np.isfinite(zbuf)
    img = np.full((height, width), background)
    if np.any(valid):
        img[valid] = noise.fbm(hit[valid], base_cell)
    img = img + rng.normal(0.0, sensor_sigma, img.shape)
    depth = np.where(valid, zbuf, 0.0)
    return img.astype(np.float32), depth.astype(np.float32)


def _rect(center, u_axis, v_axis, half_u, half_v) -> np.ndarray:
    c = np.asarray(center, float)
    u = np.asarray(u_axis, float) / np.linalg.norm(u_axis)
    v = np.asarray(v_axis, float) / np.linalg.norm(v_axis)
    return np.array([c - half_u * u - half_v * v, c + half_u * u - half_v * v,
                     c + half_u * u + half_v * v, c - half_u * u + half_v * v])


def _render_views(quads, noise, base_cell, Ks, Rs, ts, sizes, sigma, seed, supersample=1):
    rng = np.random.default_rng(seed + 1000)
    images, gt = [], None
    for i, (K, R, t, (w, h)) in enumerate(zip(Ks, Rs, ts, sizes)):
        img, depth = render(quads, noise, base_cell, K, R, t, w, h, sigma, rng, supersample=supersample)
        images.append(img)
        if i == 0:
            gt = depth
    return images, gt


# ---------------------------------------------------------------------------
# Workloads
# ---------------------------------------------------------------------------
def three_planes(seed: int = 1, width: int = 320, height: int = 240, num_src: int = 4,
                 sigma: float = 0.01, base_cell: float = 0.4, supersample: int = 2) -> Scene:
    """One reference + ``num_src`` neighbours on a 1 m arc 10 m from a 3-plane scene.

    Focal 0.85*width, depth range [5, 20] m.
    """
    rng = np.random.default_rng(seed)
    f = 0.85 * width
    K = intrinsics(width, height, f)
    target = np.array([0.0, 0.0, 10.0])
    # arc offsets (metres of arc) for the neighbours, reference at 0:
    # +-0.5, +-1.0 for four neighbours (1 m baseline at most)
    half = (num_src + 1) // 2
    offs = [0.0] + [(1.0 if i % 2 == 0 else -1.0) * (i // 2 + 1) / half for i in range(num_src)]
    Ks, Rs, ts = [], [], []
    for o in offs:
        th = o / 10.0
        C = np.array([10.0 * np.sin(th), 0.05 * o, 10.0 - 10.0 * np.cos(th)])
        R, t = look_at(C, target)
        Ks.append(K.copy()); Rs.append(R); ts.append(t)
    j = rng.uniform(-1, 1, 6)
    back = _rect([0.0, 0.0, 16.0 + j[0]], [1.0, 0.0, 0.12 + 0.03 * j[1]], [0.0, 1.0, 0.05], 14.0, 11.0)
    mid = _rect([-2.2 + 0.3 * j[2], 0.3, 11.5], [1.0, 0.0, 0.45], [0.0, 1.0, -0.1], 2.6, 3.2)
    front = _rect([1.6 + 0.3 * j[3], -0.4 + 0.2 * j[4], 7.5 + 0.3 * j[5]], [1.0, 0.0, -0.2],
                  [0.0, 1.0, 0.35], 1.3, 1.1)
    noise = ValueNoise3D(seed)
    images, gt = _render_views([back, mid, front], noise, base_cell, Ks, Rs, ts,
                               [(width, height)] * len(Ks), sigma, seed, supersample)
    return Scene(images, Ks, Rs, ts, gt, 5.0, 20.0, f"three_planes(seed={seed},{width}x{height},N={num_src})")


def facet_heightfield(seed: int = 2, width: int = 1600, height: int = 1200, num_src: int = 8,
                      grid: int = 8, sigma: float = 0.01, max_tilt_deg: float = 40.0,
                      focal_frac: float = 0.85) -> Scene:
    """Piecewise-planar heightfield of grid x grid tilted facets seen from a 3x3 camera grid.

    Facet normals lie within ``max_tilt_deg`` of the view axis; facet depths
    in [10, 30] m; cameras on a jittered 3x3 grid with 0.5-2 m baselines;
    depth range [4, 40] m.
    """
    rng = np.random.default_rng(seed)
    f = focal_frac * width
    K = intrinsics(width, height, f)
    zc = 20.0
    target = np.array([0.0, 0.0, zc])
    sx, sy = rng.uniform(0.5, 1.4, 2)
    centers = [np.zeros(3)]
    for gy in (-1, 0, 1):
        for gx in (-1, 0, 1):
            if gx == 0 and gy == 0:
                continue
            centers.append(np.array([gx * sx, gy * sy, 0.0]) + np.append(rng.uniform(-0.1, 0.1, 2), 0.0))
    centers = centers[: num_src + 1]
    Ks, Rs, ts = [], [], []
    for i, C in enumerate(centers):
        if i == 0:
            R, t = np.eye(3), np.zeros(3)
        else:
            R, t = look_at(C, target + np.append(rng.uniform(-0.5, 0.5, 2), 0.0))
        Ks.append(K.copy()); Rs.append(R); ts.append(t)
    # facets: cells in normalised reference coordinates, 1.25x the field of view
    hu = 1.25 * (width / 2.0) / f
    hv = 1.25 * (height / 2.0) / f
    us = np.linspace(-hu, hu, grid + 1)
    vs = np.linspace(-hv, hv, grid + 1)
    quads = []
    tilt = np.deg2rad(max_tilt_deg)
    for gy in range(grid):
        for gx in range(grid):
            z0 = rng.uniform(10.0, 30.0)
            a = np.arccos(rng.uniform(np.cos(tilt), 1.0))
            phi = rng.uniform(0, 2 * np.pi)
            n = np.array([np.sin(a) * np.cos(phi), np.sin(a) * np.sin(phi), -np.cos(a)])
            uc, vc = 0.5 * (us[gx] + us[gx + 1]), 0.5 * (vs[gy] + vs[gy + 1])
            P = z0 * np.array([uc, vc, 1.0])
            d = -n @ P
            corners = []
            for (u, v) in ((us[gx], vs[gy]), (us[gx + 1], vs[gy]), (us[gx + 1], vs[gy + 1]), (us[gx], vs[gy + 1])):
                ray = np.array([u, v, 1.0])
                lam = -d / (n @ ray)
                corners.append(lam * ray)
            quads.append(np.array(corners))
    quads.append(_rect([0.0, 0.0, 45.0], [1, 0, 0], [0, 1, 0], 200.0, 200.0))  # background beyond range
    noise = ValueNoise3D(seed)
    images, gt = _render_views(quads, noise, 0.5, Ks, Rs, ts, [(width, height)] * len(Ks), sigma, seed)
    return Scene(images, Ks, Rs, ts, gt, 4.0, 40.0, f"facet_heightfield(seed={seed},{width}x{height},N={num_src})")


def _render_surface(zf, K, R, t, width, height, noise, base_cell, sigma, rng, supersample=1, iters=40):
    """Ray-cast the graph z = zf(x, y) (world frame; returns (z, dz/dx, dz/dy))
    by Newton's method along each pixel ray, started on the plane z = 10.
    Returns (image, depth along the camera's optical axis)."""
    s = max(1, int(supersample))
    Ks = K.copy()
    if s > 1:
        Ks[0, 0] *= s; Ks[1, 1] *= s
        Ks[0, 2] = (K[0, 2] + 0.5) * s - 0.5
        Ks[1, 2] = (K[1, 2] + 0.5) * s - 0.5

    def cast(Kc, w, h):
        C = -R.T @ t
        ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
        pix = np.stack([xs.ravel(), ys.ravel(), np.ones(xs.size)], 0)
        dirs = R.T @ (np.linalg.inv(Kc) @ pix)       # world directions with z_cam = 1
        lam = (10.0 - C[2]) / dirs[2]
        for _ in range(iters):
            X = C[:, None] + dirs * lam
            zz, gx, gy = zf(X[0], X[1])
            g = X[2] - zz
            dg = dirs[2] - gx * dirs[0] - gy * dirs[1]
            lam = lam - g / np.where(np.abs(dg) > 1e-9, dg, 1e-9)
            if np.max(np.abs(g)) < 1e-10:
                break
        X = C[:, None] + dirs * lam
        ok = (lam > 0) & (np.abs(X[2] - zf(X[0], X[1])[0]) < 1e-6)
        return X, lam, ok

    X, lam, ok = cast(Ks, width * s, height * s)
    img = np.full(X.shape[1], 0.5)
    img[ok] = noise.fbm(X[:, ok].T, base_cell)
    img = img.reshape(height * s, width * s)
    if s > 1:
        img = img.reshape(height, s, width, s).mean(axis=(1, 3))
    img = img + rng.normal(0.0, sigma, img.shape)
    _, lam1, ok1 = cast(K, width, height)
    depth = np.where(ok1, lam1, 0.0).reshape(height, width)
    return img.astype(np.float32), depth.astype(np.float32)


def sinusoidal_bumps(seed: int = 4, width: int = 1024, height: int = 768, num_src: int = 6,
                     amplitude: float = 0.5, sigma: float = 0.01, supersample: int = 2) -> Scene:
    """A plane 10 m from the reference camera with smooth sinusoidal bumps of
    ``amplitude`` m (z = 10 + A sin(2 pi x / lx + p1) sin(2 pi y / ly + p2),
    seeded wavelengths 2-4 m, i.e. bumps over a 10 m span), seen by the
    reference and ``num_src`` neighbours on a 0.8 m ring.  Exact ray casting of
    the smooth surface; focal 0.9 * width; depth range [6, 14] m."""
    rng = np.random.default_rng(seed)
    f = 0.9 * width
    K = intrinsics(width, height, f)
    lx, ly = rng.uniform(2.0, 4.0, 2)
    p1, p2 = rng.uniform(0, 2 * np.pi, 2)
    kx, ky = 2 * np.pi / lx, 2 * np.pi / ly

    def zf(x, y):
        sx, sy = np.sin(kx * x + p1), np.sin(ky * y + p2)
        return (10.0 + amplitude * sx * sy, amplitude * kx * np.cos(kx * x + p1) * sy,
                amplitude * ky * sx * np.cos(ky * y + p2))

    target = np.array([0.0, 0.0, 10.0])
    Ks, Rs, ts = [K.copy()], [np.eye(3)], [np.zeros(3)]
    for k in range(num_src):
        ang = 2 * np.pi * k / num_src + rng.uniform(-0.2, 0.2)
        C = np.array([0.8 * np.cos(ang), 0.8 * np.sin(ang), rng.uniform(-0.1, 0.1)])
        R, t = look_at(C, target + np.append(rng.uniform(-0.3, 0.3, 2), 0.0))
        Ks.append(K.copy()); Rs.append(R); ts.append(t)
    noise = ValueNoise3D(seed)
    rr = np.random.default_rng(seed + 1000)
    images, gt = [], None
    for i, (Kc, R, t) in enumerate(zip(Ks, Rs, ts)):
        img, depth = _render_surface(zf, Kc, R, t, width, height, noise, 0.25, sigma, rr, supersample)
        images.append(img)
        if i == 0:
            gt = depth
    return Scene(images, Ks, Rs, ts, gt, 6.0, 14.0,
                 f"sinusoidal_bumps(seed={seed},{width}x{height},N={num_src},A={amplitude})")
