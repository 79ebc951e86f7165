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


# ---------------------------------------------------------------------------
# Many references: a camera walk around a piecewise-planar scene
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class WalkScene:
    """Views of a camera walk and the references built on them.

    ``refs[i] = (view index, [source view indices], depth_min, depth_max)``;
    ``gt_depth[v]`` is the ground-truth depth of view v (0 = no hit).
    """
    images: List[np.ndarray]
    Ks: List[np.ndarray]
    Rs: List[np.ndarray]
    ts: List[np.ndarray]
    gt_depth: List[np.ndarray]
    refs: List[Tuple[int, List[int], float, float]]
    name: str = ""

    def reference(self, i: int) -> Scene:
        """Reference i as a single-reference Scene (view 0 = the reference)."""
        v, srcs, dmin, dmax = self.refs[i]
        order = [v] + list(srcs)
        return Scene([self.images[j] for j in order], [self.Ks[j] for j in order], [self.Rs[j] for j in order],
                     [self.ts[j] for j in order], self.gt_depth[v], dmin, dmax, f"{self.name}[ref {i}]")

    def units(self, iterations: int = 3) -> float:
        """Mpix x source views x iterations of all references."""
        return sum(self.images[v].size / 1e6 * len(s) * iterations for v, s, _, _ in self.refs)


def _box_quads(cx, cy, hx, hy, height, angle):
    """Top and four sides of a box standing on z = 0 (world z up)."""
    c, s = np.cos(angle), np.sin(angle)
    ax, ay = np.array([c, s, 0.0]) * hx, np.array([-s, c, 0.0]) * hy
    base = np.array([cx, cy, 0.0])
    up = np.array([0.0, 0.0, height])
    corners = [base - ax - ay, base + ax - ay, base + ax + ay, base - ax + ay]
    quads = [np.array([p + up for p in corners])]
    for k in range(4):
        a, b = corners[k], corners[(k + 1) % 4]
        quads.append(np.array([a, b, b + up, a + up]))
    return quads


def _fbm_torch(noise: ValueNoise3D, p, base_cell: float, octaves: int = 4):
    """ValueNoise3D.fbm of the (n, 3) float64 tensor ``p``, on its device."""
    import torch

    table = torch.as_tensor(noise.values, dtype=p.dtype, device=p.device)
    nt = noise.n
    total = torch.zeros(p.shape[0], dtype=p.dtype, device=p.device)
    amp, cell, norm = 1.0, base_cell, 0.0
    for o in range(octaves):
        q = p / cell + 17.31 * o
        fl = torch.floor(q)
        f = q - fl
        i = fl.to(torch.int64) % nt
        f = f * f * (3.0 - 2.0 * f)
        acc = torch.zeros_like(total)
        for dx in (0, 1):
            wx = f[:, 0] if dx else 1.0 - f[:, 0]
            for dy in (0, 1):
                wy = f[:, 1] if dy else 1.0 - f[:, 1]
                for dz in (0, 1):
                    wz = f[:, 2] if dz else 1.0 - f[:, 2]
                    acc += wx * wy * wz * table[(i[:, 0] + dx) % nt, (i[:, 1] + dy) % nt, (i[:, 2] + dz) % nt]
        total += amp * acc
        norm += amp
        amp *= 0.5
        cell *= 0.5
    return total / norm


def _render_torch(quads, noise, base_cell, K, R, t, width, height, sigma, gen, device):
    """Ray-cast convex planar quads with torch (float64) on ``device``; the
    same model as ``render`` (z-buffer, 4-octave value noise, sensor noise).
    Returns (image float32, depth float32) as numpy arrays."""
    import torch

    dt = torch.float64
    Kt = torch.as_tensor(K, dtype=dt, device=device)
    Rt = torch.as_tensor(R, dtype=dt, device=device)
    tt = torch.as_tensor(t, dtype=dt, device=device)
    C = -Rt.T @ tt
    ys, xs = torch.meshgrid(torch.arange(height, dtype=dt, device=device),
                            torch.arange(width, dtype=dt, device=device), indexing="ij")
    pix = torch.stack([xs.reshape(-1), ys.reshape(-1), torch.ones(width * height, dtype=dt, device=device)], 0)
    dirs = Rt.T @ (torch.linalg.inv(Kt) @ pix)         # z_cam = 1 per ray
    zbuf = torch.full((width * height,), float("inf"), dtype=dt, device=device)
    hit = torch.zeros((3, width * height), dtype=dt, device=device)
    for quad in quads:
        q = torch.as_tensor(np.asarray(quad, float), dtype=dt, device=device)
        n = torch.linalg.cross(q[1] - q[0], q[2] - q[0])
        n = n / torch.linalg.norm(n)
        d = -(n @ q[0])
        denom = n @ dirs
        lam = -(n @ C + d) / denom
        X = C[:, None] + dirs * lam[None, :]
        inside = (lam > 1e-6) & torch.isfinite(lam)
        for k in range(4):
            a, b = q[k], q[(k + 1) % 4]
            e = torch.linalg.cross((b - a).expand(X.shape[1], 3), (X.T - a), dim=1) @ n
            inside &= e >= -1e-9
        take = inside & (lam < zbuf)
        zbuf = torch.where(take, lam, zbuf)
        hit = torch.where(take[None, :], X, hit)
    valid = torch.isfinite(zbuf)
    total = _fbm_torch(noise, hit.T, base_cell)
    img = torch.where(valid, total, torch.full_like(total, 0.5))
    img = img + sigma * torch.randn(img.shape, generator=gen, dtype=dt, device=device)
    depth = torch.where(valid, zbuf, torch.zeros_like(zbuf))
    return (img.reshape(height, width).to(torch.float32).cpu().numpy(),
            depth.reshape(height, width).to(torch.float32).cpu().numpy())


WALK_SIZES = ((640, 480), (1024, 768), (1600, 1200), (2000, 1500))


def camera_walk(seed: int = 3, num_views: int = 64, sizes: Sequence[Tuple[int, int]] = WALK_SIZES,
                neighbours: Tuple[int, int] = (4, 12), num_boxes: int = 14, sigma: float = 0.01,
                focal_frac: float = 0.85, device: str = "cpu") -> WalkScene:
    """A seeded random camera walk around a 20 x 20 m piecewise-planar scene.

    The scene (world z up) is a 20 x 20 m ground plane with ``num_boxes``
    rotated boxes 0.5-4 m tall.  ``num_views`` cameras walk an arc 12-15 m
    from the centre at 10-13 m height, each looking at the centre with some
    jitter (so that the scene fills the views); each view draws its resolution from ``sizes`` (focal
    ``focal_frac`` x width).  Every view is a reference; its sources are
    4-12 (``neighbours``, drawn per reference) views nearest along the walk
    whose viewing directions are within 35 degrees and whose baseline is
    0.2-8 m.  The depth range of a reference is the depth span of the
    scene's bounding box in its camera, widened by 10 %.
    """
    import torch

    rng = np.random.default_rng(seed)
    quads = [_rect([0.0, 0.0, 0.0], [1, 0, 0], [0, 1, 0], 10.0, 10.0)]
    for _ in range(num_boxes):
        cx, cy = rng.uniform(-7.5, 7.5, 2)
        hx, hy = rng.uniform(0.4, 1.8, 2)
        quads += _box_quads(cx, cy, hx, hy, rng.uniform(0.5, 4.0), rng.uniform(0, np.pi / 2))
    bbox = np.array([[x, y, z] for x in (-10, 10) for y in (-10, 10) for z in (0.0, 4.0)])

    phi = rng.uniform(0, 2 * np.pi)
    radius, height = 13.5, 11.5
    Ks, Rs, ts, centres, axes, sz = [], [], [], [], [], []
    for _ in range(num_views):
        phi += rng.uniform(0.03, 0.08)
        radius = float(np.clip(radius + rng.normal(0, 0.3), 12.0, 15.0))
        height = float(np.clip(height + rng.normal(0, 0.3), 10.0, 13.0))
        C = np.array([radius * np.cos(phi), radius * np.sin(phi), height])
        target = np.array([rng.normal(0, 0.8), rng.normal(0, 0.8), 0.5])
        R, t = look_at(C, target, up=(0.0, 0.0, -1.0))
        w, h = sizes[rng.integers(len(sizes))]
        Ks.append(intrinsics(w, h, focal_frac * w)); Rs.append(R); ts.append(t)
        centres.append(C); axes.append(R[2]); sz.append((w, h))

    refs = []
    for i in range(num_views):
        want = int(rng.integers(neighbours[0], neighbours[1] + 1))
        cand = []
        for j in range(num_views):
            if j == i:
                continue
            base = np.linalg.norm(centres[j] - centres[i])
            ang = np.degrees(np.arccos(np.clip(axes[i] @ axes[j], -1.0, 1.0)))
            if 0.2 <= base <= 8.0 and ang <= 35.0:
                cand.append((abs(j - i), j))
        srcs = [j for _, j in sorted(cand)[:want]]
        z = (Rs[i] @ bbox.T).T[:, 2] + ts[i][2]
        refs.append((i, srcs, float(max(0.5, 0.9 * z.min())), float(1.1 * z.max())))

    noise = ValueNoise3D(seed)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed + 1000)
    images, gts = [], []
    for K, R, t, (w, h) in zip(Ks, Rs, ts, sz):
        img, depth = _render_torch(quads, noise, 0.4, K, R, t, w, h, sigma, gen, device)
        images.append(img)
        gts.append(depth)
    return WalkScene(images, Ks, Rs, ts, gts, refs,
                     f"camera_walk(seed={seed},{num_views} views,{len(quads)} quads)")


# ---------------------------------------------------------------------------
# Aerial survey: nadir views over a fractal terrain
# ---------------------------------------------------------------------------
def fractal_dem(seed: int, extent: float = 2000.0, grid: float = 1.0, relief: float = 80.0,
                beta: float = 3.8) -> np.ndarray:
    """Seeded fractal elevation grid ((n, n) float64, n = extent / grid + 1,
    heights in [0, relief] m): spectral synthesis with power spectrum
    f^-beta on a power-of-two periodic grid, cropped.  beta = 3.8 keeps the
    steepest 1 m step at about 0.7."""
    n = int(round(extent / grid)) + 1
    m = 1 << int(np.ceil(np.log2(n)))
    rng = np.random.default_rng(seed)
    fx = np.fft.fftfreq(m)[:, None]
    fy = np.fft.rfftfreq(m)[None, :]
    f = np.hypot(fx, fy)
    f[0, 0] = 1.0
    amp = f ** (-beta / 2.0)
    amp[0, 0] = 0.0
    spec = amp * (rng.standard_normal(f.shape) + 1j * rng.standard_normal(f.shape))
    z = np.fft.irfft2(spec, s=(m, m))[:n, :n]
    return (z - z.min()) / (z.max() - z.min()) * relief


def _rot_xyz(rx: float, ry: float, rz: float) -> np.ndarray:
    cx, sx, cy, sy, cz, sz = np.cos(rx), np.sin(rx), np.cos(ry), np.sin(ry), np.cos(rz), np.sin(rz)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def _render_dem_torch(dem, grid: float, noise: ValueNoise3D, base_cell: float, K, R, t, width: int, height: int,
                      sigma: float, gen, device, iters: int = 48, keep_image: bool = False):
    """Ray-cast the bilinear surface of ``dem`` (heights at x = j * grid,
    y = i * grid for dem[i, j]) by bracketed Newton along each pixel ray.
    The caller guarantees that the height along every ray is monotone (one
    crossing), so the root is the first hit.  Returns (image float16, depth
    float32) numpy arrays (``keep_image``: the image as a torch tensor on
    ``device``); rays leaving the grid get depth 0 and grey 0.5."""
    import torch

    dt = torch.float64
    Z = torch.as_tensor(dem, dtype=dt, device=device)
    ny, nx = Z.shape
    Kt = torch.as_tensor(K, dtype=dt, device=device)
    Rt = torch.as_tensor(R, dtype=dt, device=device)
    tt = torch.as_tensor(t, dtype=dt, device=device)
    C = -Rt.T @ tt
    ys, xs = torch.meshgrid(torch.arange(height, dtype=dt, device=device),
                            torch.arange(width, dtype=dt, device=device), indexing="ij")
    pix = torch.stack([xs.reshape(-1), ys.reshape(-1), torch.ones(width * height, dtype=dt, device=device)], 0)
    dirs = Rt.T @ (torch.linalg.inv(Kt) @ pix)         # z_cam = 1 per ray
    del pix, xs, ys
    dx, dy, dz = dirs[0], dirs[1], dirs[2]

    def surface(lam):
        gx = ((C[0] + lam * dx) / grid).clamp(0.0, nx - 1.0)
        gy = ((C[1] + lam * dy) / grid).clamp(0.0, ny - 1.0)
        ix = torch.floor(gx).clamp(max=nx - 2)
        iy = torch.floor(gy).clamp(max=ny - 2)
        fx, fy = gx - ix, gy - iy
        k = iy.to(torch.int64) * nx + ix.to(torch.int64)
        zf = Z.reshape(-1)
        z00, z10, z01, z11 = zf[k], zf[k + 1], zf[k + nx], zf[k + nx + 1]
        h = (1 - fy) * ((1 - fx) * z00 + fx * z10) + fy * ((1 - fx) * z01 + fx * z11)
        hx = ((1 - fy) * (z10 - z00) + fy * (z11 - z01)) / grid
        hy = ((1 - fx) * (z01 - z00) + fx * (z11 - z10)) / grid
        return h, hx, hy

    zmin, zmax = float(Z.min()), float(Z.max())
    lo = (zmax + 1e-3 - C[2]) / dz                     # above the highest point
    hi = (zmin - 1e-3 - C[2]) / dz                     # below the lowest point
    lam = 0.5 * (lo + hi)
    g_prev = torch.full_like(lam, float("inf"))
    for _ in range(iters):
        h, hx, hy = surface(lam)
        g = C[2] + lam * dz - h
        above = g > 0
        lo = torch.where(above, lam, lo)
        hi = torch.where(above, hi, lam)
        dg = dz - hx * dx - hy * dy
        nl = lam - g / dg
        # Newton inside the bracket while |g| at least halves; bisection
        # otherwise (Newton can cycle across the kinks between grid cells)
        newton = (nl > lo) & (nl < hi) & (torch.abs(g) <= 0.5 * g_prev)
        g_prev = torch.abs(g)
        lam = torch.where(newton, nl, 0.5 * (lo + hi))
    h, _, _ = surface(lam)
    X = C[:, None] + dirs * lam[None, :]
    inside = (X[0] >= 0) & (X[0] <= (nx - 1) * grid) & (X[1] >= 0) & (X[1] <= (ny - 1) * grid)
    ok = inside & (torch.abs(X[2] - h) < 1e-6)
    tex = _fbm_torch(noise, X.T, base_cell)
    img = torch.where(ok, tex, torch.full_like(tex, 0.5))
    img = img + sigma * torch.randn(img.shape, generator=gen, dtype=dt, device=device)
    depth = torch.where(ok, lam, torch.zeros_like(lam))
    img16 = img.reshape(height, width).to(torch.float16).contiguous()
    return (img16 if keep_image else img16.cpu().numpy(),
            depth.reshape(height, width).to(torch.float32).cpu().numpy())


def aerial_survey(seed: int = 5, num_views: int = 200, width: int = 4000, height: int = 3000,
                  neighbours: int = 10, extent: float = 2000.0, grid: float = 1.0, relief: float = 80.0,
                  flight_height: float = 100.0, forward_overlap: float = 0.8, side_overlap: float = 0.7,
                  focal_frac: float = 0.9, depth_range: Tuple[float, float] = (60.0, 140.0),
                  sigma: float = 0.01, device: str = "cpu", images_on_device: bool = False) -> WalkScene:
    """Nadir views of a seeded fractal terrain, flown in a lawn-mower pattern.

    Terrain: ``fractal_dem`` over ``extent`` x ``extent`` m at ``grid`` m,
    heights 0..``relief`` m, textured by 4-octave value noise (0.8 m down to
    0.1 m cells) in world coordinates.  Cameras (focal ``focal_frac`` x
    width) fly ``flight_height`` m above the mid-relief level, looking
    straight down with small seeded attitude (0.3 degree) and position
    (0.3 m) jitter, along strips parallel to y: the image height lies along
    the strip, consecutive photos overlap by ``forward_overlap`` and
    neighbouring strips by ``side_overlap`` (at the mid-relief level), and
    every other strip is flown in the opposite direction (images rotated by
    180 degrees).  Strips: about sqrt(num_views / 2), the survey centred on
    the terrain.  Every view is a reference; its sources are the
    ``neighbours`` views with the nearest camera centres (ties by index),
    and its depth range is ``depth_range``.  Images are float16: numpy
    arrays, or with ``images_on_device`` the rendering device's torch
    tensors (contiguous (H, W)), which stay there.
    """
    import torch

    rng = np.random.default_rng(seed)
    dem = fractal_dem(seed, extent, grid, relief)
    f = focal_frac * width
    K = intrinsics(width, height, f)
    along = height / f * flight_height * (1.0 - forward_overlap)
    across = width / f * flight_height * (1.0 - side_overlap)
    strips = max(1, int(round(np.sqrt(num_views / 2.0))))
    per = int(np.ceil(num_views / strips))
    z0 = relief / 2.0 + flight_height
    base = np.array([[1.0, 0.0, 0.0], [0.0, -1.0, 0.0], [0.0, 0.0, -1.0]])   # image up = +y, looking down
    flip = np.diag([-1.0, -1.0, 1.0])                                         # 180 degrees about the axis
    Ks, Rs, ts, centres = [], [], [], []
    for s in range(strips):
        for k in range(per):
            if len(Ks) == num_views:
                break
            kk = k if s % 2 == 0 else per - 1 - k
            C = np.array([extent / 2 + (s - (strips - 1) / 2) * across,
                          extent / 2 + (kk - (per - 1) / 2) * along, z0])
            C = C + np.append(rng.normal(0.0, 0.3, 2), rng.normal(0.0, 0.3))
            R0 = base if s % 2 == 0 else flip @ base
            R = _rot_xyz(*np.radians(rng.normal(0.0, 0.3, 3))) @ R0
            Ks.append(K.copy()); Rs.append(R); ts.append(-R @ C); centres.append(C)
    centres = np.array(centres)

    # one crossing per ray: the height change along a ray's horizontal step
    # stays below its vertical step (bilinear cells: per-axis step bounds)
    sx = np.abs(np.diff(dem, axis=1)).max() / grid
    sy = np.abs(np.diff(dem, axis=0)).max() / grid
    corners = np.linalg.inv(K) @ np.array([[0, width - 1, 0, width - 1], [0, 0, height - 1, height - 1], [1, 1, 1, 1.0]])
    for R in Rs:
        d = R.T @ corners
        if np.any(sx * np.abs(d[0]) + sy * np.abs(d[1]) >= 0.95 * np.abs(d[2])):
            raise ValueError("aerial_survey: terrain too steep for single-crossing rays")

    refs = []
    for i in range(num_views):
        dist = np.linalg.norm(centres - centres[i], axis=1)
        order = [j for j in np.lexsort((np.arange(num_views), dist)) if j != i]
        refs.append((i, [int(j) for j in order[:neighbours]], float(depth_range[0]), float(depth_range[1])))

    noise = ValueNoise3D(seed)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed + 1000)
    images, gts = [], []
    for K_, R, t in zip(Ks, Rs, ts):
        img, depth = _render_dem_torch(dem, grid, noise, 0.8, K_, R, t, width, height, sigma, gen, device,
                                       keep_image=images_on_device)
        images.append(img)
        gts.append(depth)
    return WalkScene(images, Ks, Rs, ts, gts, refs,
                     f"aerial_survey(seed={seed},{num_views} views,{width}x{height},{strips} strips)")
