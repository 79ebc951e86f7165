"""Pins of the oracle's textureless fallback (strided NCC) and of its warp.

The strided NCC (``compute_ncc_at_stride``, opencl_kernels.h:188-258, called
at :531-553) and the local affine warp (:401-461) are checked against what
the mathematics fixes, not against a retyped copy of the oracle:

* Decimation: on a reference whose stride-1 window (and, for stride 3, the
  stride-2 window) is flat, the stride-k NCC of the full images equals the
  stride-1 NCC of the k-times decimated images with sigma_spatial / k.  The
  second is a different code path (the cached reference taps of the fused
  stride-1 loop), so a wrong sampling position, weight, stride or moment of
  the strided path fails.  The warps are rolled and zoomed (J = s R(theta)),
  so a transposed Jacobian term fails too.  The source is the bilinear
  upsampling of a coarse image, so fine and coarse bilinear samples agree.
* Parallel stereo: a source that is the reference shifted by an integer
  disparity gives NCC = 1 (cost 0) at the true depth through the strided
  path, where census would give 1.0.
* Warp: the centre equals the plane-induced homography
  H = K2 [Q12 + a12 v^T] K1^-1 (doc/dense_matching.md:40-53) and the
  Jacobian equals its forward difference, within the second-order remainder
  of its differential Df (:55-73).
* A float64 evaluation of the documented strided formula on random patches
  and warps, with the invariances it implies.
"""
import numpy as np
import pytest

from oracle import pm_oracle as po

SIGMA_S = 1.5
SIGMA_C = 25.0 / 255.0


def _K(f, cx, cy):
    return np.array([[f, 0.0, cx], [0.0, f, cy], [0.0, 0.0, 1.0]])


def _rz(theta):
    c, s = np.cos(theta), np.sin(theta)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def _bilinear64(img, x, y):
    """Bilinear sample with texel centres at integers and clamp-to-edge, float64."""
    h, w = img.shape
    x = np.clip(np.asarray(x, np.float64), 0.0, w - 1.0)
    y = np.clip(np.asarray(y, np.float64), 0.0, h - 1.0)
    x0 = np.floor(x).astype(int)
    y0 = np.floor(y).astype(int)
    x1 = np.minimum(x0 + 1, w - 1)
    y1 = np.minimum(y0 + 1, h - 1)
    ax, ay = x - x0, y - y0
    im = img.astype(np.float64)
    top = im[y0, x0] * (1 - ax) + im[y0, x1] * ax
    bot = im[y1, x0] * (1 - ax) + im[y1, x1] * ax
    return top * (1 - ay) + bot * ay


def _upsampled(coarse, k, ox, oy, w, h):
    """Fine image whose pixel (ox + k*u, oy + k*v) is coarse[v, u], bilinear in between:
    its bilinear samples equal the coarse image's at the mapped coordinates."""
    xs = (np.arange(w) - ox) / k
    ys = (np.arange(h) - oy) / k
    X, Y = np.meshgrid(xs, ys)
    return _bilinear64(coarse, X, Y).astype(np.float32)


# --------------------------------------------------------------------------
# Decimation
# --------------------------------------------------------------------------
def _decimation_case(seed, k, patch, noise):
    """Reference/source pair with test pixels whose stride-1 (and stride-2
    for k = 3) windows are flat and whose stride-k lattice is textured; the
    source sees the lattice through an affine warp.  Returns the fine and
    the decimated problems."""
    rng = np.random.default_rng(seed)
    hp = patch // 2
    spacing = 16 if k == 2 else 15
    ncx, ncy = 5, 4
    W, H = spacing * (ncx + 1) + 6, spacing * (ncy + 1) + 6
    ox, oy = 1, 2 if k == 3 else 0
    f, Z = 80.0, 5.0
    K1 = _K(f, (W - 1) / 2.0, (H - 1) / 2.0)
    theta = rng.uniform(0.2, 0.5) * rng.choice([-1, 1])
    zoom = rng.uniform(0.8, 1.25)
    f2 = f * zoom
    K2 = _K(f2, (W - 1) / 2.0 + rng.uniform(-3, 3), (H - 1) / 2.0 + rng.uniform(-3, 3))
    R2 = _rz(theta)
    t2 = np.array([rng.uniform(-0.2, 0.2), rng.uniform(-0.2, 0.2), 0.0])
    # fronto-parallel plane z = Z in the reference frame: n = (0, 0, -1), d = Z;
    # t2_z = 0 and a roll keep H affine: x2 = A x1 + c
    Kf = [po.make_camera(K, np.eye(3), np.zeros(3), W, H) for K in (K1, K2)]
    A = (f2 / f) * R2[:2, :2]
    c = K2[:2, 2] + f2 * t2[:2] / Z - A @ K1[:2, 2]
    # coarse source, smooth random texture; the fine source is its bilinear upsampling
    cw, chh = W // k + 4, H // k + 4
    coarse = rng.uniform(0.15, 0.85, (chh, cw))
    coarse = 0.5 * coarse + 0.25 * (np.roll(coarse, 1, 0) + np.roll(coarse, 1, 1))
    src = _upsampled(coarse, k, ox, oy, W, H)
    c0 = 0.5
    ref = np.full((H, W), c0, np.float32)
    centres = []
    for j in range(ncy):
        for i in range(ncx):
            xc = spacing * (i + 1) + ox
            yc = spacing * (j + 1) + oy
            xc -= (xc - ox) % k
            yc -= (yc - oy) % k
            # keep the pixels whose warped strided footprint stays inside the
            # source (clamp-to-edge differs between the fine and coarse grids)
            reach = np.array([[xc + k * hp * sx, yc + k * hp * sy] for sx in (-1, 1) for sy in (-1, 1)])
            warped = reach @ A.T + c
            if warped.min() < 2 or warped[:, 0].max() > W - 3 or warped[:, 1].max() > H - 3:
                continue
            centres.append((xc, yc))
            for iy in range(-hp, hp + 1):
                for ix in range(-hp, hp + 1):
                    if max(abs(k * ix), abs(k * iy)) <= hp:
                        continue  # the stride-1 window stays flat
                    if k == 3 and (k * ix) % 2 == 0 and (k * iy) % 2 == 0:
                        continue  # never reached: only the centre is on both lattices
                    x, y = xc + k * ix, yc + k * iy
                    sx, sy = A @ np.array([x, y]) + c
                    v = _bilinear64(src, sx, sy) + noise * rng.standard_normal()
                    ref[y, x] = np.float32(v)
    xy = np.array(centres, np.int32)
    plane = np.array([0.0, 0.0, -1.0, Z], np.float32)
    fine = dict(images=[ref, src], Ks=[K1, K2], Rs=[np.eye(3), R2], ts=[np.zeros(3), t2], xy=xy,
                planes=np.tile(plane, (len(xy), 1)), sigma_spatial=SIGMA_S)
    # decimated problem: every k-th pixel from (ox, oy), intrinsics scaled
    D1 = _K(f / k, (K1[0, 2] - ox) / k, (K1[1, 2] - oy) / k)
    D2 = _K(f2 / k, (K2[0, 2] - ox) / k, (K2[1, 2] - oy) / k)
    dec = dict(images=[ref[oy::k, ox::k].copy(), coarse.astype(np.float32)], Ks=[D1, D2], Rs=fine["Rs"],
               ts=fine["ts"], xy=np.stack([(xy[:, 0] - ox) // k, (xy[:, 1] - oy) // k], 1).astype(np.int32),
               planes=fine["planes"], sigma_spatial=SIGMA_S / k)
    return fine, dec, (A, c, Kf)


def _costs(problem, patch, census=True):
    sc = po.OracleScene(problem["images"], problem["Ks"], problem["Rs"], problem["ts"])
    p = po.Params(patch_size=patch, depth_min=1.0, depth_max=20.0, sigma_spatial=problem["sigma_spatial"],
                  sigma_color=SIGMA_C, use_census=census)
    return po.score_planes(sc, p, problem["xy"], problem["planes"])[0][:, 0], sc, p


@pytest.mark.parametrize("k,patch", [(2, 7), (2, 5), (3, 5)])
def test_strided_ncc_equals_stride1_on_decimated_images(k, patch):
    hp = patch // 2
    checked = 0
    for seed in range(8):
        for noise in (0.0, 0.03, 0.08):
            fine, dec, (A, c, _) = _decimation_case(100 * k + seed, k, patch, noise)
            cf, scf, pf = _costs(fine, patch)
            cd, _, _ = _costs(dec, patch)
            ref, src = fine["images"]
            for n, (x, y) in enumerate(fine["xy"]):
                # construction: the stride-1 window is flat, so the fine cost
                # can only come from a strided NCC (or census, which gives 1.0)
                assert np.all(ref[y - hp:y + hp + 1, x - hp:x + hp + 1] == ref[y, x])
                if k == 3:
                    assert np.all(ref[y - 2 * hp:y + 2 * hp + 1:2, x - 2 * hp:x + 2 * hp + 1:2] == ref[y, x])
                # Both sides return 1 - NCC only when that NCC exceeds 0.1
                # (otherwise each goes on to further strides of its own); the
                # branch is read off the float64 formula, away from the threshold.
                wa = A @ np.array([x, y]) + c
                ncc, _ = _ncc64(ref, src, x, y, np.array([wa[0], wa[1], A[0, 0], A[1, 0], A[0, 1], A[1, 1]]), k, hp)
                if ncc < 0.11:
                    continue
                assert abs(cf[n] - cd[n]) < 2e-5, (k, patch, seed, noise, n, cf[n], cd[n])
                assert abs(cf[n] - (1.0 - ncc)) < 1e-4
                checked += 1
            # the warp the oracle uses is the affine map that built the scene
            code, w = po.warp(scf, pf, int(fine["xy"][0, 0]), int(fine["xy"][0, 1]), fine["planes"][0], 0)
            x, y = fine["xy"][0]
            assert code == 0
            assert np.allclose(w[:2], A @ np.array([x, y]) + c, atol=2e-4)
            assert np.allclose([[w[2], w[4]], [w[3], w[5]]], A, atol=2e-4)
    print(f"decimation k={k} patch={patch}: {checked} pixels")
    assert checked >= 60


def test_transposed_jacobian_would_fail_the_decimation_pin():
    # the same scene with the warp's off-diagonal terms swapped gives a
    # different strided NCC, so the pin above detects a transposition
    fine, dec, (A, c, _) = _decimation_case(7, 2, 7, 0.0)
    assert abs(A[0, 1] - A[1, 0]) > 0.2
    sc = po.OracleScene(fine["images"], fine["Ks"], fine["Rs"], fine["ts"])
    p = po.Params(patch_size=7, depth_min=1.0, depth_max=20.0, sigma_spatial=SIGMA_S, sigma_color=SIGMA_C)
    diffs = []
    for x, y in fine["xy"]:
        code, w = po.warp(sc, p, int(x), int(y), fine["planes"][0], 0)
        wt = w.copy()
        wt[3], wt[4] = w[4], w[3]
        diffs.append(abs(po.ncc_at_stride(sc, p, x, y, 0, 2.0, w) - po.ncc_at_stride(sc, p, x, y, 0, 2.0, wt)))
    assert np.median(diffs) > 0.05


# --------------------------------------------------------------------------
# Parallel stereo at an integer disparity
# --------------------------------------------------------------------------
@pytest.mark.parametrize("k,patch", [(2, 7), (3, 5)])
def test_strided_ncc_is_one_at_the_true_integer_disparity(k, patch):
    fine, _, _ = _decimation_case(11 * k, k, patch, 0.05)
    ref = fine["images"][0]
    H, W = ref.shape
    f, b, disp = 80.0, 0.25, 4
    Z = f * b / disp
    K = _K(f, (W - 1) / 2.0, (H - 1) / 2.0)
    src = np.full_like(ref, ref[0, 0])
    src[:, :W - disp] = ref[:, disp:]            # x2 = x1 - disp
    t2 = np.array([-b, 0.0, 0.0])                # source centre at +b on the x axis
    sc = po.OracleScene([ref, src], [K, K], [np.eye(3)] * 2, [np.zeros(3), t2])
    xy = fine["xy"]
    for census in (True, False):
        p = po.Params(patch_size=patch, depth_min=1.0, depth_max=20.0, sigma_spatial=SIGMA_S,
                      sigma_color=SIGMA_C, use_census=census)
        at_truth = po.score_planes(sc, p, xy, np.tile(np.array([0, 0, -1.0, Z], np.float32), (len(xy), 1)))[0]
        # census of the flat window would give 1.0; the strided NCC gives 1 - 1 = 0
        assert np.all(at_truth < 1e-5), at_truth.ravel()
        for wrong in (disp - 1, disp + 1):
            off = po.score_planes(sc, p, xy, np.tile(np.array([0, 0, -1.0, f * b / wrong], np.float32),
                                                     (len(xy), 1)))[0]
            assert np.median(off) > 0.05
        # the warp is the translation by -disp
        code, w = po.warp(sc, p, int(xy[0, 0]), int(xy[0, 1]), np.array([0, 0, -1.0, Z], np.float32), 0)
        assert code == 0 and np.allclose(w, [xy[0, 0] - disp, xy[0, 1], 1, 0, 0, 1], atol=1e-4)


# --------------------------------------------------------------------------
# The warp against the plane-induced homography
# --------------------------------------------------------------------------
def _rot(axis, angle):
    axis = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    Kx = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(angle) * Kx + (1 - np.cos(angle)) * Kx @ Kx


def _f32(a):
    return np.asarray(a, np.float64).astype(np.float32).astype(np.float64)


def test_warp_is_the_plane_induced_homography():
    rng = np.random.default_rng(21)
    W, H = 640, 480
    img = np.zeros((H, W), np.float32)
    n_ok = n_asym = 0
    worst_c = worst_fd = 0.0
    for case in range(300):
        K1 = _K(rng.uniform(450, 700), rng.uniform(300, 340), rng.uniform(220, 260))
        K2 = _K(rng.uniform(450, 700), rng.uniform(300, 340), rng.uniform(220, 260))
        R1 = _rot(rng.standard_normal(3), rng.uniform(0, 0.3))
        t1 = rng.uniform(-1, 1, 3)
        R2 = _rot(rng.standard_normal(3), rng.uniform(0.05, 0.35)) @ R1
        C1 = -R1.T @ t1
        C2 = C1 + rng.uniform(-1.2, 1.2, 3)
        t2 = -R2 @ C2
        sc = po.OracleScene([img, img], [K1, K2], [R1, R2], [t1, t2])
        p = po.Params(patch_size=5, depth_min=0.5, depth_max=100.0)
        x, y = int(rng.integers(40, W - 40)), int(rng.integers(40, H - 40))
        # plane through the pixel's point at depth Z with a normal within 40 deg of -z
        Z = rng.uniform(4, 20)
        nrm = np.array([rng.uniform(-0.6, 0.6), rng.uniform(-0.6, 0.6), -1.0])
        nrm /= np.linalg.norm(nrm)
        P = Z * np.linalg.solve(K1, [x, y, 1.0])
        plane = np.array([*nrm, -nrm @ P], np.float32)
        code, w = po.warp(sc, p, x, y, plane, 0)
        # float64 homography on the float32 values the kernels see
        k1, k2, r1, r2, tt1, tt2 = (_f32(a) for a in (K1, K2, R1, R2, t1, t2))
        v = plane[:3].astype(np.float64) / np.float64(plane[3])
        Q = r2 @ r1.T
        a = Q @ tt1 - tt2
        Hm = k2 @ (Q + np.outer(a, v)) @ np.linalg.inv(k1)

        def f(px, py):
            u = Hm @ np.array([px, py, 1.0])
            return u[:2] / u[2], u

        c0, u0 = f(x, y)
        if u0[2] < 1e-3 or code == -1:
            continue
        Df = np.array([[Hm[0, 0] * u0[2] - Hm[2, 0] * u0[0], Hm[0, 1] * u0[2] - Hm[2, 1] * u0[0]],
                       [Hm[1, 0] * u0[2] - Hm[2, 0] * u0[1], Hm[1, 1] * u0[2] - Hm[2, 1] * u0[1]]]) / u0[2] ** 2
        fd = np.stack([f(x + 1, y)[0] - c0, f(x, y + 1)[0] - c0], 1)   # columns: d/dx, d/dy
        det = np.linalg.det(fd)
        if not (0.03 < det < 45):
            assert code in (1, 2) or 0.019 < det < 50.5
            continue
        assert code == 0, (case, det)
        J = np.array([[w[2], w[4]], [w[3], w[5]]])     # sx = scx + jxx dx + jyx dy
        worst_c = max(worst_c, float(np.max(np.abs(w[:2] - c0))))
        worst_fd = max(worst_fd, float(np.max(np.abs(J - fd))))
        assert np.allclose(w[:2], c0, atol=5e-4), (case, w[:2], c0)
        assert np.allclose(J, fd, atol=5e-4), (case, J, fd)
        # forward difference vs the differential: the second-order remainder
        d2 = np.stack([f(x + 1, y)[0] - 2 * c0 + f(x - 1, y)[0], f(x, y + 1)[0] - 2 * c0 + f(x, y - 1)[0]], 1)
        assert np.all(np.abs(J - Df) <= 0.5 * np.abs(d2) * 1.01 + 5e-4), (case, J, Df)
        n_ok += 1
        n_asym += abs(Df[0, 1] - Df[1, 0]) > 0.05
    print(f"warp: {n_ok} cases, centre error <= {worst_c:.2e} px, Jacobian vs forward difference <= {worst_fd:.2e}")
    assert n_ok >= 200
    assert n_asym >= n_ok // 2   # a transposed Jacobian would fail most cases


# --------------------------------------------------------------------------
# The documented strided formula in float64 on random patches and warps
# --------------------------------------------------------------------------
def _ncc64(ref, src, x, y, w, stride, hp, var_thr=1.5e-6):
    """opencl_kernels.h:216-257 in float64 (exact bilinear, clamp-to-edge)."""
    sf = -1.0 / (2.0 * SIGMA_S ** 2)
    cf = -1.0 / (2.0 * SIGMA_C ** 2)
    iy, ix = np.mgrid[-hp:hp + 1, -hp:hp + 1]
    fdx, fdy = stride * ix, stride * iy
    ry = np.clip(y + fdy, 0, ref.shape[0] - 1).astype(int)
    rx = np.clip(x + fdx, 0, ref.shape[1] - 1).astype(int)
    rv = ref.astype(np.float64)[ry, rx]
    sv = _bilinear64(src, w[0] + w[2] * fdx + w[4] * fdy, w[1] + w[3] * fdx + w[5] * fdy)
    r = rv - float(ref[y, x])
    s = sv - _bilinear64(src, w[0], w[1])
    wt = np.exp(fdx * fdx * sf + fdy * fdy * sf + r * r * cf)
    sw = wt.sum()
    mr, ms = (wt * r).sum() / sw, (wt * s).sum() / sw
    vr = (wt * r * r).sum() / sw - mr * mr
    vs = (wt * s * s).sum() / sw - ms * ms
    if vr < var_thr or vs < var_thr:
        return -1.0, min(vr, vs)
    return ((wt * r * s).sum() / sw - mr * ms) / np.sqrt(vr * vs), min(vr, vs)


def _noise_image(rng, h, w):
    base = rng.uniform(0, 1, (h // 4 + 2, w // 4 + 2))
    img = _bilinear64(base, *np.meshgrid(np.arange(w) / 4.0, np.arange(h) / 4.0))
    return (0.7 * img + 0.3 * rng.uniform(0, 1, (h, w))).astype(np.float32)


def test_strided_ncc_matches_the_float64_formula():
    rng = np.random.default_rng(5)
    H, W = 72, 96
    ref = _noise_image(rng, H, W)
    src = _noise_image(rng, H, W)
    K = _K(80.0, (W - 1) / 2, (H - 1) / 2)
    sc = po.OracleScene([ref, src], [K, K], [np.eye(3)] * 2, [np.zeros(3)] * 2)
    worst = 0.0
    n = 0
    for patch in (3, 5, 7):
        p = po.Params(patch_size=patch, depth_min=1, depth_max=10, sigma_spatial=SIGMA_S, sigma_color=SIGMA_C)
        for _ in range(150):
            x, y = int(rng.integers(12, W - 12)), int(rng.integers(12, H - 12))
            theta, s = rng.uniform(-0.6, 0.6), rng.uniform(0.7, 1.4)
            A = s * _rz(theta)[:2, :2] + rng.uniform(-0.15, 0.15, (2, 2))
            w = np.array([x + rng.uniform(-6, 6), y + rng.uniform(-6, 6), A[0, 0], A[1, 0], A[0, 1], A[1, 1]],
                         np.float32)
            for stride in (2.0, 3.0):
                got = po.ncc_at_stride(sc, p, x, y, 0, stride, w)
                want, v = _ncc64(ref, src, x, y, w.astype(np.float64), stride, patch // 2)
                if 1.2e-6 < v < 1.8e-6:
                    continue  # at the variance gate: float32 and float64 may branch apart
                worst = max(worst, abs(got - want))
                assert abs(got - want) < 1e-5, (patch, stride, got, want)
                n += 1
    print(f"strided NCC vs float64: {n} cases, max |diff| {worst:.2e}")
    assert n > 800


def test_strided_ncc_invariances():
    rng = np.random.default_rng(9)
    H, W = 64, 80
    ref = _noise_image(rng, H, W)
    K = _K(80.0, (W - 1) / 2, (H - 1) / 2)
    p = po.Params(patch_size=7, depth_min=1, depth_max=10, sigma_spatial=SIGMA_S, sigma_color=SIGMA_C)
    scenes = {a: po.OracleScene([ref, (a * ref + (0.5 if a < 0 else 0.1)).astype(np.float32)], [K, K],
                                [np.eye(3)] * 2, [np.zeros(3)] * 2) for a in (1.0, 0.6, -0.8)}
    for _ in range(60):
        x, y = int(rng.integers(10, W - 10)), int(rng.integers(10, H - 10))
        ident = np.array([x, y, 1, 0, 0, 1], np.float32)
        A = _rz(rng.uniform(-0.5, 0.5))[:2, :2] * rng.uniform(0.8, 1.2)
        gen = np.array([x + rng.uniform(-3, 3), y + rng.uniform(-3, 3), A[0, 0], A[1, 0], A[0, 1], A[1, 1]],
                       np.float32)
        for stride in (2.0, 3.0):
            one = po.ncc_at_stride(scenes[1.0], p, x, y, 0, stride, ident)
            assert abs(one - 1.0) < 1e-5                      # identical patches
            base = po.ncc_at_stride(scenes[1.0], p, x, y, 0, stride, gen)
            aff = po.ncc_at_stride(scenes[0.6], p, x, y, 0, stride, gen)
            neg = po.ncc_at_stride(scenes[-0.8], p, x, y, 0, stride, gen)
            assert abs(aff - base) < 2e-5                      # s -> a s + b, a > 0
            assert abs(neg + base) < 2e-5                      # a < 0 flips the sign
            assert -1.0 - 1e-6 <= base <= 1.0 + 1e-6
