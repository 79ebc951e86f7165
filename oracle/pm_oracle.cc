// CPU oracle for the PatchMatch multi-view depth path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity checker and the timed
// CPU baseline; it is never linked into, loaded by, or called from the
// product library.  Only tests/, the smoke check and the CPU-baseline leg of
// the benchmark may load it.
//
// It restates, in plain C++17 + OpenMP, the ACMMP-style OpenCL program of the
// reference (opensfm/src/lib/dense/opencl_kernels.h:22-1856) and the host
// schedule of opensfm/src/lib/dense/src/estimator.cc:134-255, 515-720,
// 1112-1224.  Every function cites the lines it follows.
//
// Numerics contract (the reference's own GPU numerics are not reproducible:
// hardware bilinear filtering, -cl-fast-relaxed-math, -cl-mad-enable at
// opencl_utils.h:84-85).  This oracle fixes them explicitly and the HIP path
// follows the same contract, so the two can be compared bit for bit:
//   * compiled with -ffp-contract=off: a*b+c is two roundings unless written
//     as fmaf(); fmaf() is used exactly where the device code uses it (the
//     per-tap bilinear lerps, warp coordinates and NCC accumulations -- the
//     fusions -cl-mad-enable licenses in the reference);
//   * exp/log are pm_expf/pm_logf below (Cody-Waite reduction + Taylor /
//     atanh series, written from the textbook method), not libm;
//   * sqrt and '/' are IEEE correctly rounded;
//   * bilinear sampling is exact software bilinear with texel centres at
//     integer coordinates and clamp-to-edge addressing; the reference's
//     "+0.5f" (opencl_kernels.h:224,229,459,491) only converts pixel-centre
//     coordinates to OpenCL's texel-corner convention, so the oracle samples
//     at the pixel-centre coordinate directly.
//
// Every function here is pinned by tests/ against closed forms, special
// cases, brute force or the reference's documented rules (test_oracle.py,
// test_oracle_pins.py, test_levels.py): geometry, RNG, bilinear, cost
// vectors, top-k, reference patch, surface normal, smoothness, geometric
// consistency, view weights (validation sum, anchoring, angle weight,
// incident-angle factor, all-fail fallback), weighted and Monte-Carlo costs
// (draw frequencies), candidate search pattern (one-pass propagation of a
// seed), refinement hypotheses (structure and draw distributions),
// checkerboard median, prior re-init bounds, decode, pyramid resampling,
// coarse-to-fine transfer, speckle removal and gap filling.  The whole
// algorithm is pinned by the plane-sweep argmin and ground-truth tests.
// The fronto-parallel plane sweep (pmo_plane_sweep) is pinned by recovering
// the exact sweep depth of a rendered fronto-parallel plane and by the
// one-plane case (the range's geometric mean).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <vector>
#include <omp.h>

// Arithmetic type.  The default build computes in float under the numerics
// contract above.  The FP64 build (-DPMO_F64, libpm_oracle64.so) runs the
// same algorithm with every value, array and interface field in double and
// libm's exp and log: the float64 evaluation that float32 and
// float16-storage results are measured against.  The random generator and
// the contract's bit-level helpers stay float in both.
#ifdef PMO_F64
typedef double real;
#else
typedef float real;
#endif

namespace {

constexpr int kMaxSrc = 16;   // reference: MAX_SOURCES 8 (opencl_kernels.h:26), lifted
constexpr int kMcSamples = 32;  // MC_NUM_SAMPLES (opencl_kernels.h:910)
constexpr real kPiF = 3.14159274101257324219f;  // M_PI_F

}  // namespace

extern "C" {

// Camera layout of the reference's CLCamera (opencl_utils.h:258-280; kernel
// struct opencl_kernels.h:62-68): K and R row-major, t, size.  92 bytes.
typedef struct {
  real K[9];
  real R[9];
  real t[3];
  int width;
  int height;
} pmo_camera;

typedef struct {
  int width, height;          // reference view size
  int num_src;                // number of source views (<= 16)
  const real* images;        // all views, row-major, concatenated (ref first)
  const int64_t* offsets;     // num_src+1 element offsets into images
  const int* widths;          // num_src+1
  const int* heights;         // num_src+1
  const pmo_camera* cams;     // num_src+1 (index 0 = reference)
  const real* angle_weights; // num_src (estimator.cc:380-407)
} pmo_scene;

typedef struct {
  int half_patch;
  real sigma_spatial;
  real sigma_color;
  int top_k;
  real census_weight;   // use_census ? 1 : 0 (estimator.cc:536)
  real depth_min, depth_max;
  int max_iterations;
  real smooth_weight;
  int anchor_views;
  real geom_weight;
  real var_threshold;   // 1.5e-6 in the reference (opencl_kernels.h:254,516,519)
  int checkerboard_filter;
} pmo_params;

typedef struct {
  real* planes;               // 4 * npix (nx, ny, nz, d)
  real* costs;                // npix
  uint32_t* rng;               // 2 * npix
  uint32_t* sel;               // npix
  const real* prior_planes;   // 4 * npix or null
  const uint32_t* plane_masks; // npix or null
  const real* prev_depths;    // num_src * npix (ref resolution) or null
  uint32_t prev_depth_mask;
} pmo_state;

}  // extern "C"

namespace {

// ---------------------------------------------------------------------------
// Elementary functions of the numerics contract.
// ---------------------------------------------------------------------------
inline float bits_to_float(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline uint32_t float_to_bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

// e^x: k = rint(x/ln2), r = x - k*ln2 in two Cody-Waite steps, e^r by its
// degree-7 Taylor polynomial (|r| <= 0.35, truncation < 6e-9 relative),
// scaled by 2^k through the exponent field.  x < -86 returns 0 (no
// subnormal results), x > 88 returns +inf.
inline float pm_expf(float x) {
  if (x != x) return x;
  if (x < -86.0f) return 0.0f;
  if (x > 88.0f) return INFINITY;
  const float kf = std::rint(x * 1.44269502162933349609f);
  float r = std::fmaf(-kf, 0.693145751953125f, x);
  r = std::fmaf(-kf, 1.42860682030941723212e-6f, r);
  float p = 1.98412698412698412698e-4f;           // 1/7!
  p = std::fmaf(p, r, 1.38888888888888888889e-3f);  // 1/6!
  p = std::fmaf(p, r, 8.33333333333333333333e-3f);  // 1/5!
  p = std::fmaf(p, r, 4.16666666666666666667e-2f);  // 1/4!
  p = std::fmaf(p, r, 1.66666666666666666667e-1f);  // 1/3!
  p = std::fmaf(p, r, 0.5f);
  p = std::fmaf(p, r, 1.0f);
  p = std::fmaf(p, r, 1.0f);
  const int k = static_cast<int>(kf);
  return p * bits_to_float(static_cast<uint32_t>(k + 127) << 23);
}

// ln x for finite x > 0: x = m * 2^e with m in [sqrt(1/2), sqrt(2)),
// ln m = 2 atanh(s), s = (m-1)/(m+1), series to s^11 (|s| < 0.172).
inline float pm_logf(float x) {
  if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
  if (x == INFINITY) return x;
  uint32_t u = float_to_bits(x);
  int e = static_cast<int>((u >> 23) & 0xffu) - 127;
  if (e == -127) {  // subnormal: renormalise
    x = x * 8388608.0f;
    u = float_to_bits(x);
    e = static_cast<int>((u >> 23) & 0xffu) - 127 - 23;
  }
  float m = bits_to_float((u & 0x7fffffu) | 0x3f800000u);
  if (m > 1.41421353816986083984f) { m = m * 0.5f; e += 1; }
  const float f = m - 1.0f;
  const float s = f / (2.0f + f);
  const float s2 = s * s;
  float p = 9.09090909090909090909e-2f;            // 1/11
  p = std::fmaf(p, s2, 1.11111111111111111111e-1f); // 1/9
  p = std::fmaf(p, s2, 1.42857142857142857143e-1f); // 1/7
  p = std::fmaf(p, s2, 2.0e-1f);                    // 1/5
  p = std::fmaf(p, s2, 3.33333333333333333333e-1f); // 1/3
  p = std::fmaf(p, s2, 1.0f);
  const float lm = (2.0f * s) * p;
  const float fe = static_cast<float>(e);
  return std::fmaf(fe, 0.693145751953125f, std::fmaf(fe, 1.42860682030941723212e-6f, lm));
}

inline real rfma(real a, real b, real c) { return std::fma(a, b, c); }
#ifdef PMO_F64
inline real rexp(real x) { return std::exp(x); }
inline real rlog(real x) { return std::log(x); }
#else
inline real rexp(real x) { return pm_expf(x); }
inline real rlog(real x) { return pm_logf(x); }
#endif

struct V3 { real x, y, z; };
inline real dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline real len3(V3 a) { return std::sqrt(dot3(a, a)); }
// OpenCL normalize(): all-zero input is returned unchanged.
inline V3 normalize3(V3 a) {
  const real l = len3(a);
  if (l == 0.0f) return a;
  return V3{a.x / l, a.y / l, a.z / l};
}
inline real clampf(real v, real lo, real hi) { return std::fmin(std::fmax(v, lo), hi); }

struct Plane { real x, y, z, w; };

// ---------------------------------------------------------------------------
// RNG: opencl_kernels.h:78-97.
// ---------------------------------------------------------------------------
inline void philox_round(uint32_t st[2], uint32_t key) {
  const uint64_t prod = static_cast<uint64_t>(st[0]) * 0xD256D193ull;
  const uint32_t lo = static_cast<uint32_t>(prod);
  const uint32_t hi = static_cast<uint32_t>(prod >> 32);
  const uint32_t new_x = st[1] ^ key ^ hi;
  st[0] = new_x;
  st[1] = lo;
}

inline float rand_float(uint32_t st[2], uint32_t key) {
  for (uint32_t i = 0; i < 4; ++i) philox_round(st, key + i * 0x9E3779B9u);
  return static_cast<float>(st[0]) / 4294967296.0f;
}

// ---------------------------------------------------------------------------
// Geometry: opencl_kernels.h:102-173.
// ---------------------------------------------------------------------------
inline V3 project(V3 X, const pmo_camera& c) {
  V3 xc;
  xc.x = c.R[0] * X.x + c.R[1] * X.y + c.R[2] * X.z + c.t[0];
  xc.y = c.R[3] * X.x + c.R[4] * X.y + c.R[5] * X.z + c.t[1];
  xc.z = c.R[6] * X.x + c.R[7] * X.y + c.R[8] * X.z + c.t[2];
  V3 p;
  p.x = c.K[0] * xc.x + c.K[1] * xc.y + c.K[2] * xc.z;
  p.y = c.K[3] * xc.x + c.K[4] * xc.y + c.K[5] * xc.z;
  p.z = c.K[6] * xc.x + c.K[7] * xc.y + c.K[8] * xc.z;
  return p;
}

inline V3 backproject(real x, real y, real depth, const pmo_camera& c) {
  const real fx = c.K[0], fy = c.K[4], cx = c.K[2], cy = c.K[5];
  V3 pc{depth * (x - cx) / fx, depth * (y - cy) / fy, depth};
  V3 d{pc.x - c.t[0], pc.y - c.t[1], pc.z - c.t[2]};
  V3 w;
  w.x = c.R[0] * d.x + c.R[3] * d.y + c.R[6] * d.z;
  w.y = c.R[1] * d.x + c.R[4] * d.y + c.R[7] * d.z;
  w.z = c.R[2] * d.x + c.R[5] * d.y + c.R[8] * d.z;
  return w;
}

inline real depth_from_plane(real x, real y, Plane p, const pmo_camera& c) {
  const real fx = c.K[0], fy = c.K[4], cx = c.K[2], cy = c.K[5];
  const real denom = p.x * (x - cx) / fx + p.y * (y - cy) / fy + p.z;
  if (std::fabs(denom) < 1e-6f) return 0.0f;
  return -p.w / denom;
}

inline Plane plane_from_depth_normal(real x, real y, real depth, V3 n, const pmo_camera& c) {
  const real fx = c.K[0], fy = c.K[4], cx = c.K[2], cy = c.K[5];
  V3 pt{depth * (x - cx) / fx, depth * (y - cy) / fy, depth};
  const real d = -dot3(n, pt);
  return Plane{n.x, n.y, n.z, d};
}

// ---------------------------------------------------------------------------
// Image access.
// ---------------------------------------------------------------------------
struct Img {
  const real* data;
  int w, h;
  inline real texel(int ix, int iy) const {
    ix = std::min(std::max(ix, 0), w - 1);
    iy = std::min(std::max(iy, 0), h - 1);
    return data[static_cast<int64_t>(iy) * w + ix];
  }
  // Bilinear at a pixel-centre coordinate, clamp-to-edge.  Coordinates are
  // first clamped to [-1, w-1] x [-1, h-1]: outside that range both taps of
  // an axis hit the same edge texel, so the clamp does not change the value
  // (lerp(a, a, t) == a exactly) and keeps the real->int conversion defined.
  inline real bilinear(real sx, real sy) const {
    sx = clampf(sx, -1.0f, static_cast<real>(w - 1));
    sy = clampf(sy, -1.0f, static_cast<real>(h - 1));
    const real fx0 = std::floor(sx), fy0 = std::floor(sy);
    const real ax = sx - fx0, ay = sy - fy0;
    const int x0 = static_cast<int>(fx0), y0 = static_cast<int>(fy0);
    const real t00 = texel(x0, y0), t01 = texel(x0 + 1, y0);
    const real t10 = texel(x0, y0 + 1), t11 = texel(x0 + 1, y0 + 1);
    const real v0 = rfma(ax, t01 - t00, t00);
    const real v1 = rfma(ax, t11 - t10, t10);
    return rfma(ay, v1 - v0, v0);
  }
};

struct Ctx {
  const pmo_scene* sc;
  const pmo_params* pr;
  Img ref;
  Img src[kMaxSrc];
  const pmo_camera* cams;
  int n_src;
  real spatial_factor, color_factor;
  Ctx(const pmo_scene* s, const pmo_params* p) : sc(s), pr(p) {
    ref = Img{s->images + s->offsets[0], s->widths[0], s->heights[0]};
    n_src = std::min(s->num_src, kMaxSrc);
    for (int i = 0; i < n_src; ++i)
      src[i] = Img{s->images + s->offsets[i + 1], s->widths[i + 1], s->heights[i + 1]};
    cams = s->cams;
    spatial_factor = -1.0f / (2.0f * p->sigma_spatial * p->sigma_spatial);
    color_factor = -1.0f / (2.0f * p->sigma_color * p->sigma_color);
  }
};

// ---------------------------------------------------------------------------
// Reference patch: opencl_kernels.h:42-49, 328-366.
// ---------------------------------------------------------------------------
struct RefPatch {
  real ref_center, sum_w, sum_wr, sum_wrr;
  real w[(2 * 7 + 1) * (2 * 7 + 1)];
  real wr[(2 * 7 + 1) * (2 * 7 + 1)];
};

void compute_ref_patch(const Ctx& c, int x, int y, RefPatch* rp) {
  const int hp = c.pr->half_patch;
  const real center = c.ref.texel(x, y);
  real sum_w = 0.0f, sum_wr = 0.0f, sum_wrr = 0.0f;
  int p = 0;
  for (int iy = -hp; iy <= hp; ++iy) {
    for (int ix = -hp; ix <= hp; ++ix) {
      const real rv = c.ref.texel(x + ix, y + iy);
      const real r = rv - center;
      const real w = rexp(c.spatial_factor * static_cast<real>(ix * ix + iy * iy) +
                              c.color_factor * r * r);
      sum_w += w;
      const real wr = w * r;
      sum_wr += wr;
      sum_wrr = rfma(wr, r, sum_wrr);
      rp->w[p] = w;
      rp->wr[p] = wr;
      ++p;
    }
  }
  rp->ref_center = center;
  rp->sum_w = sum_w;
  rp->sum_wr = sum_wr;
  rp->sum_wrr = sum_wrr;
}

// Strided ZNCC fallback: opencl_kernels.h:188-258.
real ncc_at_stride(const Ctx& c, int s, real ref_center, real src_center, real scx, real scy,
                    int cx, int cy, real stride, real jxx, real jxy, real jyx, real jyy) {
  const int hp = c.pr->half_patch;
  real sum_w = 0.0f, sum_r = 0.0f, sum_s = 0.0f, sum_rr = 0.0f, sum_ss = 0.0f, sum_rs = 0.0f;
  for (int iy = -hp; iy <= hp; ++iy) {
    const real fdy = stride * static_cast<real>(iy);
    // source coordinates and sums contracted as the reference's
    // -cl-fast-relaxed-math build allows (opencl_utils.h:84-85)
    const real bx = rfma(jyx, fdy, scx);
    const real by = rfma(jyy, fdy, scy);
    for (int ix = -hp; ix <= hp; ++ix) {
      const real fdx = stride * static_cast<real>(ix);
      const real rx = static_cast<real>(cx) + fdx;
      const real ry = static_cast<real>(cy) + fdy;
      const real ref_val = c.ref.texel(static_cast<int>(rx), static_cast<int>(ry));
      const real sx = rfma(jxx, fdx, bx);
      const real sy = rfma(jxy, fdx, by);
      const real src_val = c.src[s].bilinear(sx, sy);
      const real r = ref_val - ref_center;
      const real w = rexp(fdx * fdx * c.spatial_factor + fdy * fdy * c.spatial_factor +
                              r * r * c.color_factor);
      const real sv = src_val - src_center;
      const real wsv = w * sv;
      sum_r += w * r;
      sum_s += wsv;
      sum_rr += w * r * r;
      sum_ss = rfma(wsv, sv, sum_ss);
      sum_rs = rfma(w * r, sv, sum_rs);
      sum_w += w;
    }
  }
  if (sum_w < 1e-6f) return -1.0f;
  const real inv_w = 1.0f / sum_w;
  const real mean_r = sum_r * inv_w;
  const real mean_s = sum_s * inv_w;
  const real var_ref = sum_rr * inv_w - mean_r * mean_r;
  const real var_src = sum_ss * inv_w - mean_s * mean_s;
  if (var_ref < c.pr->var_threshold || var_src < c.pr->var_threshold) return -1.0f;
  const real covar = sum_rs * inv_w - mean_r * mean_s;
  return covar / std::sqrt(var_ref * var_src);
}

// Census fallback: opencl_kernels.h:270-315.
real census_inner(const Ctx& c, int s, real ref_center, real src_center, int cx, int cy,
                   real scx, real scy, real jxx, real jxy, real jyx, real jyy) {
  uint32_t rc = 0u, sc = 0u;
  int bit = 0;
  for (int dy = -2; dy <= 2; ++dy) {
    for (int dx = -2; dx <= 2; ++dx) {
      if (dx == 0 && dy == 0) continue;
      const real ref_val = c.ref.texel(cx + dx, cy + dy);
      const real sx = scx + jxx * static_cast<real>(dx) + jyx * static_cast<real>(dy);
      const real sy = scy + jxy * static_cast<real>(dx) + jyy * static_cast<real>(dy);
      const real src_val = c.src[s].bilinear(sx, sy);
      if (ref_val > ref_center) rc |= (1u << bit);
      if (src_val > src_center) sc |= (1u << bit);
      ++bit;
    }
  }
  const int hamming = __builtin_popcount(rc ^ sc);
  if (__builtin_popcount(rc) < 3 && __builtin_popcount(sc) < 3) return 1.0f;
  return static_cast<real>(hamming) / 24.0f;
}

// The plane's points at (x, y), (x+1, y) and (x, y+1) in the world
// (opencl_kernels.h:401-416); false when a depth is <= 0 (cost 2.0).
struct PlanePoints { V3 cw, wdx, wdy; };
bool plane_points(const pmo_camera& c0, int x, int y, Plane plane, PlanePoints* pp) {
  const real depth = depth_from_plane(static_cast<real>(x), static_cast<real>(y), plane, c0);
  if (depth <= 0.0f) return false;
  const real depth_dx = depth_from_plane(static_cast<real>(x + 1), static_cast<real>(y), plane, c0);
  const real depth_dy = depth_from_plane(static_cast<real>(x), static_cast<real>(y + 1), plane, c0);
  if (depth_dx <= 0.0f || depth_dy <= 0.0f) return false;
  pp->cw = backproject(static_cast<real>(x), static_cast<real>(y), depth, c0);
  pp->wdx = backproject(static_cast<real>(x + 1), static_cast<real>(y), depth_dx, c0);
  pp->wdy = backproject(static_cast<real>(x), static_cast<real>(y + 1), depth_dy, c0);
  return true;
}

// Local affine warp into source camera cs (opencl_kernels.h:432-456): the
// projected centre and the forward-difference Jacobian, source pixel
// (sx, sy) = (scx + jxx*dx + jyx*dy, scy + jxy*dx + jyy*dy) for a reference
// offset (dx, dy).  0: valid; 1: a point is behind the source camera;
// 2: det J outside [0.02, 50].  (Both invalid cases cost 2.0.)
struct Warp { real scx, scy, jxx, jxy, jyx, jyy; };
int source_warp(const PlanePoints& pp, const pmo_camera& cs, Warp* w) {
  const V3 pc = project(pp.cw, cs);
  if (pc.z < 1e-6f) return 1;
  w->scx = pc.x / pc.z;
  w->scy = pc.y / pc.z;
  const V3 pdx = project(pp.wdx, cs);
  const V3 pdy = project(pp.wdy, cs);
  if (pdx.z < 1e-6f || pdy.z < 1e-6f) return 1;
  w->jxx = pdx.x / pdx.z - w->scx;
  w->jxy = pdx.y / pdx.z - w->scy;
  w->jyx = pdy.x / pdy.z - w->scx;
  w->jyy = pdy.y / pdy.z - w->scy;
  const real det = w->jxx * w->jyy - w->jxy * w->jyx;
  return (det < 0.02f || det > 50.0f) ? 2 : 0;
}

// Per-view cost vector ("ComputePlaneScore"): opencl_kernels.h:381-557.
// The reference loops taps outer / views inner; each view's sums are
// accumulated over the taps in the same order either way, so the oracle
// loops views outer.
int cost_vector(const Ctx& c, int x, int y, Plane plane, const RefPatch& rp, real* out) {
  const int n_src = c.n_src;
  const int hp = c.pr->half_patch;
  PlanePoints pp;
  if (!plane_points(c.cams[0], x, y, plane, &pp)) {
    for (int i = 0; i < n_src; ++i) out[i] = 2.0f;
    return n_src;
  }

  const bool ref_ok = rp.sum_w >= 1e-6f;
  real inv_w = 0.0f, mean_r = 0.0f, var_ref = 0.0f;
  if (ref_ok) {
    inv_w = 1.0f / rp.sum_w;
    mean_r = rp.sum_wr * inv_w;
    var_ref = rp.sum_wrr * inv_w - mean_r * mean_r;
  }

  for (int s = 0; s < n_src; ++s) {
    Warp wp;
    if (source_warp(pp, c.cams[s + 1], &wp) != 0) { out[s] = 2.0f; continue; }
    const real scx = wp.scx, scy = wp.scy, jxx = wp.jxx, jxy = wp.jxy, jyx = wp.jyx, jyy = wp.jyy;
    const real scen = c.src[s].bilinear(scx, scy);

    // Fused stride-1 ZNCC over the cached reference weights (:471-499).
    real sum_ws = 0.0f, sum_wss = 0.0f, sum_wrs = 0.0f;
    int p = 0;
    for (int iy = -hp; iy <= hp; ++iy) {
      const real fiy = static_cast<real>(iy);
      const real bx = rfma(jyx, fiy, scx);
      const real by = rfma(jyy, fiy, scy);
      for (int ix = -hp; ix <= hp; ++ix) {
        const real fix = static_cast<real>(ix);
        const real w = rp.w[p];
        const real wr = rp.wr[p];
        ++p;
        const real sx = rfma(jxx, fix, bx);
        const real sy = rfma(jxy, fix, by);
        const real sv = c.src[s].bilinear(sx, sy) - scen;
        const real wsv = w * sv;
        sum_ws += wsv;
        sum_wss = rfma(wsv, sv, sum_wss);
        sum_wrs = rfma(wr, sv, sum_wrs);
      }
    }
    if (!ref_ok) { out[s] = 2.0f; continue; }  // :502-505

    real ncc1 = -1.0f;
    if (var_ref >= c.pr->var_threshold) {
      const real mean_s = sum_ws * inv_w;
      const real var_src = sum_wss * inv_w - mean_s * mean_s;
      if (var_src >= c.pr->var_threshold) {
        const real covar = sum_wrs * inv_w - mean_r * mean_s;
        ncc1 = covar / std::sqrt(var_ref * var_src);
      }
    }
    if (ncc1 > 0.1f) { out[s] = 1.0f - ncc1; continue; }
    real best = ncc1;
    const real strides[2] = {2.0f, 3.0f};
    for (int st = 0; st < 2; ++st) {
      const real ncc = ncc_at_stride(c, s, rp.ref_center, scen, scx, scy, x, y, strides[st],
                                      jxx, jxy, jyx, jyy);
      if (ncc > best) best = ncc;
      if (best > 0.1f) break;
    }
    if (best <= -0.5f) {
      out[s] = (c.pr->census_weight > 1e-6f)
                   ? census_inner(c, s, rp.ref_center, scen, x, y, scx, scy, jxx, jxy, jyx, jyy)
                   : 2.0f;
    } else {
      out[s] = 1.0f - best;
    }
  }
  return n_src;
}

// Top-k initial cost + view bitmask: opencl_kernels.h:565-621.
real initial_cost_and_views(const Ctx& c, int x, int y, Plane plane, const RefPatch& rp,
                             uint32_t* out_sel) {
  real cv[kMaxSrc], cc[kMaxSrc];
  for (int i = 0; i < kMaxSrc; ++i) { cv[i] = 2.0f; cc[i] = 2.0f; }
  const int n_src = c.n_src;
  cost_vector(c, x, y, plane, rp, cv);
  int num_valid = 0;
  for (int i = 0; i < n_src; ++i) {
    cc[i] = cv[i];
    if (cv[i] < 2.0f) ++num_valid;
  }
  for (int i = 1; i < n_src; ++i) {
    const real tmp = cv[i];
    int j = i;
    for (; j >= 1 && tmp < cv[j - 1]; --j) cv[j] = cv[j - 1];
    cv[j] = tmp;
  }
  *out_sel = 0u;
  const int k = std::min(std::min(c.pr->top_k, num_valid), n_src);
  if (k <= 0) return 2.0f;
  real total = 0.0f;
  for (int i = 0; i < k; ++i) total += cv[i];
  const real thr = cv[k - 1];
  for (int i = 0; i < n_src; ++i)
    if (cc[i] <= thr) *out_sel |= (1u << i);
  return total / static_cast<real>(k);
}

inline Plane load_plane(const real* planes, int64_t idx) {
  return Plane{planes[4 * idx], planes[4 * idx + 1], planes[4 * idx + 2], planes[4 * idx + 3]};
}
inline void store_plane(real* planes, int64_t idx, Plane p) {
  planes[4 * idx] = p.x; planes[4 * idx + 1] = p.y; planes[4 * idx + 2] = p.z; planes[4 * idx + 3] = p.w;
}

// Smoothness: opencl_kernels.h:649-707.
real smoothness_cost(const Ctx& c, const real* planes, Plane cur, int x, int y) {
  const pmo_camera& c0 = c.cams[0];
  const int W = c.sc->width, H = c.sc->height;
  const real cur_depth = depth_from_plane(static_cast<real>(x), static_cast<real>(y), cur, c0);
  if (cur_depth <= 0.0f) return 0.0f;
  const V3 cur_n = normalize3(V3{cur.x, cur.y, cur.z});
  const real sigma_rel = 0.02f;
  const int dxs[4] = {-1, 1, 0, 0};
  const int dys[4] = {0, 0, -1, 1};
  real tw = 0.0f, tc = 0.0f;
  for (int k = 0; k < 4; ++k) {
    const int nx = x + dxs[k], ny = y + dys[k];
    if (nx < 0 || nx >= W || ny < 0 || ny >= H) continue;
    const Plane nb = load_plane(planes, static_cast<int64_t>(ny) * W + nx);
    const real nb_depth = depth_from_plane(static_cast<real>(nx), static_cast<real>(ny), nb, c0);
    if (nb_depth <= 0.0f) continue;
    const real rel = std::fabs(cur_depth - nb_depth) / cur_depth;
    const real w = rexp(-rel * rel / (2.0f * sigma_rel * sigma_rel));
    const V3 nb_n = normalize3(V3{nb.x, nb.y, nb.z});
    const real normal_cost = 1.0f - dot3(cur_n, nb_n);
    const real predicted = depth_from_plane(static_cast<real>(nx), static_cast<real>(ny), cur, c0);
    const real dev = (predicted > 0.0f) ? std::fabs(nb_depth - predicted) / cur_depth : 0.0f;
    const real cost = 0.5f * clampf(normal_cost, 0.0f, 1.0f) + 0.5f * clampf(dev * 10.0f, 0.0f, 1.0f);
    tc += w * cost;
    tw += w;
  }
  if (tw < 1e-6f) return 0.0f;
  return tc / tw;
}

// Surface normal from neighbour depths: opencl_kernels.h:725-772.
V3 surface_normal(const Ctx& c, const real* planes, int x, int y, real center_depth) {
  const int W = c.sc->width, H = c.sc->height;
  if (center_depth <= 0.0f) return V3{0, 0, 0};
  if (x < 1 || x >= W - 1 || y < 1 || y >= H - 1) return V3{0, 0, 0};
  const pmo_camera& c0 = c.cams[0];
  const real fx = c0.K[0], fy = c0.K[4], cx = c0.K[2], cy = c0.K[5];
  const Plane pl = load_plane(planes, static_cast<int64_t>(y) * W + (x - 1));
  const Plane pr = load_plane(planes, static_cast<int64_t>(y) * W + (x + 1));
  const Plane pu = load_plane(planes, static_cast<int64_t>(y - 1) * W + x);
  const Plane pd = load_plane(planes, static_cast<int64_t>(y + 1) * W + x);
  const real dl = depth_from_plane(static_cast<real>(x - 1), static_cast<real>(y), pl, c0);
  const real dr = depth_from_plane(static_cast<real>(x + 1), static_cast<real>(y), pr, c0);
  const real du = depth_from_plane(static_cast<real>(x), static_cast<real>(y - 1), pu, c0);
  const real dd = depth_from_plane(static_cast<real>(x), static_cast<real>(y + 1), pd, c0);
  if (dl <= 0.0f || dr <= 0.0f || du <= 0.0f || dd <= 0.0f) return V3{0, 0, 0};
  const real dx = dr - dl;
  const real dy = dd - du;
  V3 n{fx * dx, fy * dy, (cx - static_cast<real>(x)) * dx + (cy - static_cast<real>(y)) * dy - center_depth};
  const real l = len3(n);
  if (l < 1e-8f) return V3{0, 0, 0};
  n = V3{n.x / l, n.y / l, n.z / l};
  if (n.z > 0.0f) n = V3{-n.x, -n.y, -n.z};
  return n;
}

// Geometric consistency: opencl_kernels.h:779-861.
real geom_cost(const Ctx& c, const pmo_state* st, int sidx, Plane plane, int px, int py) {
  const int pdw = c.sc->width, pdh = c.sc->height;
  const pmo_camera& rc = c.cams[0];
  const pmo_camera& scm = c.cams[sidx + 1];
  const real bad = 1.0f;
  const real max_dim = static_cast<real>(std::max(pdw, pdh));
  if (max_dim < 1.0f) return bad;
  const real good_thresh = 3.0f / max_dim;
  const real full_thresh = 6.0f / max_dim;
  const real depth = depth_from_plane(static_cast<real>(px), static_cast<real>(py), plane, rc);
  if (depth <= 0.0f) return bad;
  const V3 wp = backproject(static_cast<real>(px), static_cast<real>(py), depth, rc);
  const V3 sp = project(wp, scm);
  if (sp.z < 1e-6f) return bad;
  const real sx = sp.x / sp.z, sy = sp.y / sp.z;
  if (sx < 0.0f || sx >= static_cast<real>(pdw) || sy < 0.0f || sy >= static_cast<real>(pdh)) return bad;
  const int ix = std::min(std::max(static_cast<int>(sx + 0.5f), 0), pdw - 1);
  const int iy = std::min(std::max(static_cast<int>(sy + 0.5f), 0), pdh - 1);
  const real sd = st->prev_depths[static_cast<int64_t>(sidx) * pdw * pdh + static_cast<int64_t>(iy) * pdw + ix];
  if (sd <= 0.0f) return bad;
  const V3 sw = backproject(sx, sy, sd, scm);
  const V3 rp = project(sw, rc);
  if (rp.z < 1e-6f) return bad;
  const real bx = rp.x / rp.z, by = rp.y / rp.z;
  const real dx = static_cast<real>(px) - bx;
  const real dy = static_cast<real>(py) - by;
  const real err = std::sqrt(dx * dx + dy * dy) / max_dim;
  if (err <= good_thresh) return 0.0f;
  if (err >= full_thresh) return 1.0f;
  return (err - good_thresh) / (full_thresh - good_thresh);
}

// View-weighted cost with geometry: opencl_kernels.h:871-900.
real weighted_cost_geom(const Ctx& c, const pmo_state* st, const real* cv, const real* vw,
                         real wnorm, Plane plane, int x, int y) {
  if (wnorm < 1e-6f) return 2.0f;
  real total = 0.0f;
  const real gw = c.pr->geom_weight;
  for (int i = 0; i < c.n_src; ++i) {
    if (vw[i] > 0.0f) {
      const real photo = cv[i];
      real geom = 0.0f;
      if (gw > 0.0f && (st->prev_depth_mask & (1u << i))) geom = geom_cost(c, st, i, plane, x, y);
      total += vw[i] * (photo + gw * geom);
    }
  }
  return total / wnorm;
}

// Monte-Carlo view-sampled cost: opencl_kernels.h:910-969.
real mc_cost(const Ctx& c, const pmo_state* st, const real* cv, const real* vw, real wnorm,
              Plane plane, int x, int y, uint32_t rs[2], uint32_t key) {
  if (wnorm < 1e-6f) return 2.0f;
  const int n_src = c.n_src;
  real cdf[kMaxSrc];
  real cum = 0.0f;
  for (int i = 0; i < n_src; ++i) { cum += vw[i]; cdf[i] = cum; }
  const real gw = c.pr->geom_weight;
  real total = 0.0f;
  int valid = 0;
  for (int s = 0; s < kMcSamples; ++s) {
    const real u = rand_float(rs, key + static_cast<uint32_t>(s) * 7919u);
    const real thr = u * cum;
    int v = n_src - 1;
    for (int i = 0; i < n_src; ++i)
      if (cdf[i] > thr) { v = i; break; }
    const real photo = cv[v];
    real geom = 0.0f;
    if (gw > 0.0f && (st->prev_depth_mask & (1u << v))) geom = geom_cost(c, st, v, plane, x, y);
    total += photo + gw * geom;
    ++valid;
  }
  return valid > 0 ? total / static_cast<real>(valid) : 2.0f;
}

// acmmp_random_init: opencl_kernels.h:979-1039.
void random_init_px(const Ctx& c, pmo_state* st, int x, int y) {
  const int W = c.sc->width;
  const int idx = y * W + x;
  uint32_t rs[2] = {static_cast<uint32_t>(idx) * 1099087573u + 2654435769u,
                    static_cast<uint32_t>(idx) * 2246822519u + 13u};
  const uint32_t key = static_cast<uint32_t>(idx + 7919);
  const real ldmin = rlog(std::fmax(c.pr->depth_min, 1e-6f));
  const real ldmax = rlog(std::fmax(c.pr->depth_max, 1e-5f));
  const real depth = rexp(ldmin + rand_float(rs, key) * (ldmax - ldmin));
  V3 n;
  n.x = rand_float(rs, key + 3u) * 2.0f - 1.0f;
  n.y = rand_float(rs, key + 5u) * 2.0f - 1.0f;
  n.z = -1.0f;
  n = normalize3(n);
  const Plane plane = plane_from_depth_normal(static_cast<real>(x), static_cast<real>(y), depth, n, c.cams[0]);
  store_plane(st->planes, idx, plane);
  st->rng[2 * idx] = rs[0];
  st->rng[2 * idx + 1] = rs[1];
  RefPatch rp;
  compute_ref_patch(c, x, y, &rp);
  uint32_t sel = 0u;
  st->costs[idx] = initial_cost_and_views(c, x, y, plane, rp, &sel);
  st->sel[idx] = sel;
}

// Fronto-parallel plane sweep (compute_brute_force of the pydense stub,
// pydense.pyi:53, which has no implementation in the reference tree).
// Plane i of n has depth exp(log dmin + i * (log dmax - log dmin) / (n - 1))
// (the geometric mean of the range when n == 1) and normal (0, 0, -1); it is
// scored like a random-init plane (stride-1 NCC, fallbacks, top-k mean of
// opencl_kernels.h:565-621), the lowest cost wins and the first plane wins
// ties.  The generator state is the random init's seed, before any draw.
real sweep_depth(const pmo_params& P, int i, int n) {
  const real ldmin = rlog(std::fmax(P.depth_min, 1e-6f));
  const real ldmax = rlog(std::fmax(P.depth_max, 1e-5f));
  if (n <= 1) return rexp(0.5f * (ldmin + ldmax));
  const real step = (ldmax - ldmin) / static_cast<real>(n - 1);
  return rexp(ldmin + static_cast<real>(i) * step);
}

void plane_sweep_px(const Ctx& c, pmo_state* st, int x, int y, int n) {
  const int idx = y * c.sc->width + x;
  st->rng[2 * idx] = static_cast<uint32_t>(idx) * 1099087573u + 2654435769u;
  st->rng[2 * idx + 1] = static_cast<uint32_t>(idx) * 2246822519u + 13u;
  RefPatch rp;
  compute_ref_patch(c, x, y, &rp);
  const V3 nrm{0.0f, 0.0f, -1.0f};
  real best = 0.0f;
  uint32_t best_sel = 0u;
  Plane best_plane{};
  for (int i = 0; i < n; ++i) {
    const real depth = sweep_depth(*c.pr, i, n);
    const Plane plane = plane_from_depth_normal(static_cast<real>(x), static_cast<real>(y), depth, nrm, c.cams[0]);
    uint32_t sel = 0u;
    const real cost = initial_cost_and_views(c, x, y, plane, rp, &sel);
    if (i == 0 || cost < best) {
      best = cost;
      best_sel = sel;
      best_plane = plane;
    }
  }
  store_plane(st->planes, idx, best_plane);
  st->costs[idx] = best;
  st->sel[idx] = best_sel;
}

// acmmp_prior_reinit: opencl_kernels.h:1056-1182.
void prior_reinit_px(const Ctx& c, pmo_state* st, const real* snap, int x, int y) {
  const int W = c.sc->width;
  const int idx = y * W + x;
  if (!st->plane_masks || st->plane_masks[idx] == 0u) return;
  const pmo_camera& c0 = c.cams[0];
  const Plane pp = load_plane(st->prior_planes, idx);
  const real dprior = depth_from_plane(static_cast<real>(x), static_cast<real>(y), pp, c0);
  const real dmin = c.pr->depth_min, dmax = c.pr->depth_max;
  if (dprior <= 0.0f || dprior < dmin || dprior > dmax) return;
  uint32_t rs[2] = {st->rng[2 * idx], st->rng[2 * idx + 1]};
  const uint32_t key = static_cast<uint32_t>(idx + 31337u);
  real dp = dprior * 0.99f + rand_float(rs, key) * dprior * 0.02f;
  dp = clampf(dp, dmin, dmax);
  const V3 pn = normalize3(V3{pp.x, pp.y, pp.z});
  const Plane plane = plane_from_depth_normal(static_cast<real>(x), static_cast<real>(y), dp, pn, c0);
  const int n_src = c.n_src;
  const uint32_t cur_sel = st->sel[idx];
  real vw[kMaxSrc];
  real wnorm = 0.0f;
  for (int i = 0; i < kMaxSrc; ++i) {
    if (i < n_src && (cur_sel & (1u << i))) { vw[i] = 1.0f; wnorm += 1.0f; } else vw[i] = 0.0f;
  }
  RefPatch rp;
  compute_ref_patch(c, x, y, &rp);
  const real sw = c.pr->smooth_weight;
  if (wnorm < 1e-6f) {
    uint32_t sel = 0u;
    real cost = initial_cost_and_views(c, x, y, plane, rp, &sel);
    if (sw > 1e-6f) cost += sw * smoothness_cost(c, snap, plane, x, y);
    if (cost < st->costs[idx]) {
      store_plane(st->planes, idx, plane);
      st->costs[idx] = cost;
      st->sel[idx] = sel;
    }
    st->rng[2 * idx] = rs[0];
    st->rng[2 * idx + 1] = rs[1];
    return;
  }
  real cv[kMaxSrc];
  for (int i = 0; i < kMaxSrc; ++i) cv[i] = 2.0f;
  cost_vector(c, x, y, plane, rp, cv);
  real cost = weighted_cost_geom(c, st, cv, vw, wnorm, plane, x, y);
  if (sw > 1e-6f) cost += sw * smoothness_cost(c, snap, plane, x, y);
  if (cost < st->costs[idx]) {
    store_plane(st->planes, idx, plane);
    st->costs[idx] = cost;
  }
  st->rng[2 * idx] = rs[0];
  st->rng[2 * idx + 1] = rs[1];
}

// Refinement hypotheses of acmmp_patchmatch (opencl_kernels.h:1669-1742):
//   0 (random depth, current normal)   1 (current depth, random normal)
//   2 (random depth, random normal)    3 (current depth, perturbed normal)
//   4 (perturbed depth, current normal) 5 (current depth, surface normal)
// With a prior the random draws sit around the prior.  Returns 6, or 5 when
// no surface normal is available.
int refinement_hypotheses(const Ctx& c, const real* planes, int x, int y, Plane plane_now, real depth_now,
                          bool have_prior, real depth_prior, V3 prior_n, real depth_sigma, real angle_sigma,
                          uint32_t rs[2], uint32_t key, real rd[6], V3 rn[6]) {
  const pmo_params& P = *c.pr;
  const real depth_perturbation = 0.005f;
  const real normal_perturbation = 0.01f * kPiF;
  const V3 cur_normal = normalize3(V3{plane_now.x, plane_now.y, plane_now.z});
  real depth_rand, depth_pert;
  V3 nrand, npert;
  if (have_prior && depth_prior > 0.0f) {
    depth_rand = depth_prior - 3.0f * depth_sigma + rand_float(rs, key + 503u) * 6.0f * depth_sigma;
    depth_rand = clampf(depth_rand, P.depth_min, P.depth_max);
    const real pa1 = (rand_float(rs, key + 504u) - 0.5f) * angle_sigma;
    const real pa2 = (rand_float(rs, key + 505u) - 0.5f) * angle_sigma;
    nrand = normalize3(V3{prior_n.x + pa1, prior_n.y + pa2, prior_n.z});
    if (nrand.z > 0.0f) nrand.z = -nrand.z;
  } else {
    const real ldmin = rlog(std::fmax(P.depth_min, 1e-6f));
    const real ldmax = rlog(std::fmax(P.depth_max, 1e-5f));
    depth_rand = rexp(ldmin + rand_float(rs, key + 503u) * (ldmax - ldmin));
    nrand.x = rand_float(rs, key + 504u) * 2.0f - 1.0f;
    nrand.y = rand_float(rs, key + 505u) * 2.0f - 1.0f;
    nrand.z = -1.0f;
    nrand = normalize3(nrand);
  }
  {
    const real dlo = (1.0f - depth_perturbation) * depth_now;
    const real dhi = (1.0f + depth_perturbation) * depth_now;
    depth_pert = dlo + rand_float(rs, key + 510u) * (dhi - dlo);
    depth_pert = clampf(depth_pert, P.depth_min, P.depth_max);
  }
  {
    const real pa1 = (rand_float(rs, key + 520u) - 0.5f) * normal_perturbation;
    const real pa2 = (rand_float(rs, key + 521u) - 0.5f) * normal_perturbation;
    npert = normalize3(V3{cur_normal.x + pa1, cur_normal.y + pa2, cur_normal.z});
    if (npert.z > 0.0f) npert.z = -npert.z;
  }
  const V3 sn = surface_normal(c, planes, x, y, depth_now);
  const bool have_sn = len3(sn) > 0.5f;
  const real d6[6] = {depth_rand, depth_now, depth_rand, depth_now, depth_pert, depth_now};
  const V3 n6[6] = {cur_normal, nrand, nrand, npert, cur_normal, sn};
  for (int h = 0; h < 6; ++h) { rd[h] = d6[h]; rn[h] = n6[h]; }
  return have_sn ? 6 : 5;
}

// View weights of acmmp_patchmatch (opencl_kernels.h:1452-1563): per view
// the sum of (1 - cost) over the flagged candidates it validates (cost <
// 0.8); up to anchor_views views of the previous selection without support
// get 1; positive weights are scaled by the angle weight and by the
// incident-angle factor exp(-(1 - |n . ray|)^2 / (2 * 0.6^2)) of the
// current plane; if nothing is left, every view with some candidate cost
// below 2 gets 1.  ca[j * kMaxSrc + i] = cost of candidate j in view i.
// Returns the weight sum; writes the weights and the selection mask.
real view_weights(const Ctx& c, const real* ca, const bool* flags, uint32_t prev_sel, Plane cur, int x, int y,
                   real vw[kMaxSrc], uint32_t* sel_out) {
  const pmo_params& P = *c.pr;
  const pmo_camera& c0 = c.cams[0];
  const int n_src = c.n_src;
  const real vthr = 0.8f;
  for (int i = 0; i < kMaxSrc; ++i) vw[i] = 0.0f;
  for (int i = 0; i < n_src; ++i) {
    real vc = 0.0f;
    for (int j = 0; j < 8; ++j) {
      const real cc = ca[j * kMaxSrc + i];
      if (flags[j] && cc < vthr) vc += (1.0f - cc);
    }
    if (vc > 0.0f) vw[i] = vc;
  }
  if (P.anchor_views > 0) {
    int anchored = 0;
    for (int i = 0; i < n_src && anchored < P.anchor_views; ++i) {
      if ((prev_sel & (1u << i)) && vw[i] < 1e-6f) { vw[i] = 1.0f; ++anchored; }
    }
  }
  {
    V3 cur_n{cur.x, cur.y, cur.z};
    const real nl = len3(cur_n);
    const bool have_normal = nl > 0.1f;
    if (have_normal) cur_n = V3{cur_n.x / nl, cur_n.y / nl, cur_n.z / nl};
    const real cur_depth = depth_from_plane(static_cast<real>(x), static_cast<real>(y), cur, c0);
    const V3 p3 = backproject(static_cast<real>(x), static_cast<real>(y), cur_depth, c0);
    const real inc_sigma = 0.6f;
    const real inv2s2 = -0.5f / (inc_sigma * inc_sigma);
    for (int i = 0; i < n_src; ++i) {
      if (vw[i] < 1e-6f) continue;
      vw[i] *= c.sc->angle_weights[i];
      if (have_normal && cur_depth > 0.0f) {
        const pmo_camera& ci = c.cams[i + 1];
        V3 C;
        C.x = -(ci.R[0] * ci.t[0] + ci.R[3] * ci.t[1] + ci.R[6] * ci.t[2]);
        C.y = -(ci.R[1] * ci.t[0] + ci.R[4] * ci.t[1] + ci.R[7] * ci.t[2]);
        C.z = -(ci.R[2] * ci.t[0] + ci.R[5] * ci.t[1] + ci.R[8] * ci.t[2]);
        const V3 ray = normalize3(V3{C.x - p3.x, C.y - p3.y, C.z - p3.z});
        const real cos_inc = std::fabs(dot3(cur_n, ray));
        const real xi = 1.0f - cos_inc;
        vw[i] *= rexp(xi * xi * inv2s2);
      }
    }
  }
  uint32_t new_sel = 0u;
  real wnorm = 0.0f;
  for (int i = 0; i < n_src; ++i)
    if (vw[i] > 0.0f) { new_sel |= (1u << i); wnorm += vw[i]; }
  if (wnorm < 1e-6f) {
    for (int i = 0; i < n_src; ++i) {
      for (int j = 0; j < 8; ++j) {
        if (flags[j] && ca[j * kMaxSrc + i] < 2.0f) {
          vw[i] = 1.0f; new_sel |= (1u << i); wnorm += 1.0f; break;
        }
      }
    }
  }
  *sel_out = new_sel;
  return wnorm;
}

// acmmp_patchmatch, one pixel of one colour: opencl_kernels.h:1197-1778.
// Updates in place: every neighbour the pixel reads has the other colour
// (odd Manhattan offsets), so this equals the reference's snapshot reads.
void patchmatch_px(const Ctx& c, pmo_state* st, int x, int y, int iteration, real smooth_w) {
  const int W = c.sc->width, H = c.sc->height;
  const int idx = y * W + x;
  const pmo_camera& c0 = c.cams[0];
  const pmo_params& P = *c.pr;
  uint32_t rs[2] = {st->rng[2 * idx], st->rng[2 * idx + 1]};
  const uint32_t key = static_cast<uint32_t>(idx + iteration * 7919 + 42);
  const real* planes = st->planes;
  const real* costs = st->costs;

  RefPatch rp;
  compute_ref_patch(c, x, y, &rp);

  // Step 1: candidate positions (:1254-1430).
  int cpx[8], cpy[8];
  bool flags[8];
  real ca[8 * kMaxSrc];
  for (int i = 0; i < 8; ++i) { cpx[i] = x; cpy[i] = y; flags[i] = false; }
  for (int i = 0; i < 8 * kMaxSrc; ++i) ca[i] = 2.0f;
  auto cost_at = [&](int px, int py) { return costs[static_cast<int64_t>(py) * W + px]; };
  auto eval = [&](int slot, int px, int py) {
    cpx[slot] = px; cpy[slot] = py; flags[slot] = true;
    cost_vector(c, x, y, load_plane(planes, static_cast<int64_t>(py) * W + px), rp, &ca[slot * kMaxSrc]);
  };
  if (y > 2) {  // far up (:1262-1278)
    int by = y - 3; real bc = cost_at(x, by);
    for (int i = 1; i < 11; ++i) { const int yy = y - 3 - 2 * i; if (yy < 0) break;
      const real cc = cost_at(x, yy); if (cc < bc) { bc = cc; by = yy; } }
    eval(1, x, by);
  }
  if (y < H - 3) {  // far down (:1281-1296)
    int by = y + 3; real bc = cost_at(x, by);
    for (int i = 1; i < 11; ++i) { const int yy = y + 3 + 2 * i; if (yy >= H) break;
      const real cc = cost_at(x, yy); if (cc < bc) { bc = cc; by = yy; } }
    eval(3, x, by);
  }
  if (x > 2) {  // far left (:1299-1314)
    int bx = x - 3; real bc = cost_at(bx, y);
    for (int i = 1; i < 11; ++i) { const int xx = x - 3 - 2 * i; if (xx < 0) break;
      const real cc = cost_at(xx, y); if (cc < bc) { bc = cc; bx = xx; } }
    eval(5, bx, y);
  }
  if (x < W - 3) {  // far right (:1317-1332)
    int bx = x + 3; real bc = cost_at(bx, y);
    for (int i = 1; i < 11; ++i) { const int xx = x + 3 + 2 * i; if (xx >= W) break;
      const real cc = cost_at(xx, y); if (cc < bc) { bc = cc; bx = xx; } }
    eval(7, bx, y);
  }
  if (y > 0) {  // near up (:1335-1356)
    int bx = x, by = y - 1; real bc = cost_at(bx, by);
    for (int i = 0; i < 3; ++i) {
      if (y > 1 + i && x > i) { const int px = x - i - 1, py = y - 2 - i; const real cc = cost_at(px, py);
        if (cc < bc) { bc = cc; bx = px; by = py; } }
      if (y > 1 + i && x < W - 1 - i) { const int px = x + i + 1, py = y - 2 - i; const real cc = cost_at(px, py);
        if (cc < bc) { bc = cc; bx = px; by = py; } }
    }
    eval(0, bx, by);
  }
  if (y < H - 1) {  // near down (:1361-1382)
    int bx = x, by = y + 1; real bc = cost_at(bx, by);
    for (int i = 0; i < 3; ++i) {
      if (y < H - 2 - i && x > i) { const int px = x - i - 1, py = y + 2 + i; const real cc = cost_at(px, py);
        if (cc < bc) { bc = cc; bx = px; by = py; } }
      if (y < H - 2 - i && x < W - 1 - i) { const int px = x + i + 1, py = y + 2 + i; const real cc = cost_at(px, py);
        if (cc < bc) { bc = cc; bx = px; by = py; } }
    }
    eval(2, bx, by);
  }
  if (x > 0) {  // near left (:1385-1406)
    int bx = x - 1, by = y; real bc = cost_at(bx, by);
    for (int i = 0; i < 3; ++i) {
      if (x > 1 + i && y > i) { const int px = x - 2 - i, py = y - i - 1; const real cc = cost_at(px, py);
        if (cc < bc) { bc = cc; bx = px; by = py; } }
      if (x > 1 + i && y < H - 1 - i) { const int px = x - 2 - i, py = y + i + 1; const real cc = cost_at(px, py);
        if (cc < bc) { bc = cc; bx = px; by = py; } }
    }
    eval(4, bx, by);
  }
  if (x < W - 1) {  // near right (:1409-1430)
    int bx = x + 1, by = y; real bc = cost_at(bx, by);
    for (int i = 0; i < 3; ++i) {
      if (x < W - 2 - i && y > i) { const int px = x + 2 + i, py = y - i - 1; const real cc = cost_at(px, py);
        if (cc < bc) { bc = cc; bx = px; by = py; } }
      if (x < W - 2 - i && y < H - 1 - i) { const int px = x + 2 + i, py = y + i + 1; const real cc = cost_at(px, py);
        if (cc < bc) { bc = cc; bx = px; by = py; } }
    }
    eval(6, bx, by);
  }

  // Step 2: view weights (:1452-1563).
  real vw[kMaxSrc];
  uint32_t new_sel = 0u;
  const real wnorm = view_weights(c, ca, flags, st->sel[idx], load_plane(planes, idx), x, y, vw, &new_sel);

  // Step 3: candidate costs (:1570-1619).
  real fc[8];
  for (int i = 0; i < 8; ++i) {
    if (flags[i]) {
      const Plane cp = load_plane(planes, static_cast<int64_t>(cpy[i]) * W + cpx[i]);
      real ph = mc_cost(c, st, &ca[i * kMaxSrc], vw, wnorm, cp, x, y, rs,
                         key + 600u + static_cast<uint32_t>(i) * 31u);
      if (smooth_w > 1e-6f) ph += smooth_w * smoothness_cost(c, planes, cp, x, y);
      fc[i] = ph;
    } else {
      fc[i] = 2.0f;
    }
  }
  int best = 0;
  real best_cost = fc[0];
  for (int i = 1; i < 8; ++i) if (fc[i] < best_cost) { best_cost = fc[i]; best = i; }

  Plane plane_now = load_plane(planes, idx);
  real cvn[kMaxSrc];
  for (int i = 0; i < kMaxSrc; ++i) cvn[i] = 2.0f;
  cost_vector(c, x, y, plane_now, rp, cvn);
  real cost_now = mc_cost(c, st, cvn, vw, wnorm, plane_now, x, y, rs, key + 700u);
  if (smooth_w > 1e-6f) cost_now += smooth_w * smoothness_cost(c, planes, plane_now, x, y);
  st->costs[idx] = cost_now;
  real depth_now = depth_from_plane(static_cast<real>(x), static_cast<real>(y), plane_now, c0);

  const bool have_prior = st->plane_masks && st->plane_masks[idx] > 0u;
  real depth_prior = 0.0f;
  V3 prior_n{0.0f, 0.0f, -1.0f};
  const real depth_sigma = (P.depth_max - P.depth_min) / 32.0f;
  const real angle_sigma = kPiF * (5.0f / 180.0f);
  if (have_prior) {
    const Plane pp = load_plane(st->prior_planes, idx);
    depth_prior = depth_from_plane(static_cast<real>(x), static_cast<real>(y), pp, c0);
    prior_n = normalize3(V3{pp.x, pp.y, pp.z});
  }

  // Step 3b (:1643-1656).
  if (flags[best] && best_cost < cost_now) {
    const Plane bp = load_plane(planes, static_cast<int64_t>(cpy[best]) * W + cpx[best]);
    const real db = depth_from_plane(static_cast<real>(x), static_cast<real>(y), bp, c0);
    if (db >= P.depth_min && db <= P.depth_max) {
      depth_now = db;
      plane_now = bp;
      cost_now = best_cost;
      st->sel[idx] = new_sel;
    }
  }

  // Step 4: refinement (:1669-1772).
  {
    real rd[6];
    V3 rn[6];
    const int nc = refinement_hypotheses(c, planes, x, y, plane_now, depth_now, have_prior, depth_prior, prior_n,
                                         depth_sigma, angle_sigma, rs, key, rd, rn);
    for (int h = 0; h < nc; ++h) {
      const Plane cand = plane_from_depth_normal(static_cast<real>(x), static_cast<real>(y), rd[h], rn[h], c0);
      real cv[kMaxSrc];
      for (int i = 0; i < kMaxSrc; ++i) cv[i] = 2.0f;
      cost_vector(c, x, y, cand, rp, cv);
      real tc = mc_cost(c, st, cv, vw, wnorm, cand, x, y, rs, key + 800u + static_cast<uint32_t>(h) * 37u);
      if (smooth_w > 1e-6f) tc += smooth_w * smoothness_cost(c, planes, cand, x, y);
      const real db = depth_from_plane(static_cast<real>(x), static_cast<real>(y), cand, c0);
      if (db < P.depth_min || db > P.depth_max) continue;
      if (tc < cost_now) { depth_now = db; plane_now = cand; cost_now = tc; }
    }
  }
  store_plane(st->planes, idx, plane_now);
  st->costs[idx] = cost_now;
  st->rng[2 * idx] = rs[0];
  st->rng[2 * idx + 1] = rs[1];
}

// acmmp_checkerboard_filter, one pixel: opencl_kernels.h:1883-2039.
void checkerboard_px(const Ctx& c, pmo_state* st, int x, int y) {
  const int W = c.sc->width, H = c.sc->height;
  const int center = y * W + x;
  const pmo_camera& c0 = c.cams[0];
  if (st->costs[center] < 0.001f) return;
  real buf[21];
  int n = 0;
  auto dep = [&](int px, int py) {
    return depth_from_plane(static_cast<real>(px), static_cast<real>(py),
                            load_plane(st->planes, static_cast<int64_t>(py) * W + px), c0);
  };
  buf[n++] = dep(x, y);
  if (y > 0) buf[n++] = dep(x, y - 1);
  if (y > 2) buf[n++] = dep(x, y - 3);
  if (y > 4) buf[n++] = dep(x, y - 5);
  if (y < H - 1) buf[n++] = dep(x, y + 1);
  if (y < H - 3) buf[n++] = dep(x, y + 3);
  if (y < H - 5) buf[n++] = dep(x, y + 5);
  if (x > 0) buf[n++] = dep(x - 1, y);
  if (x > 2) buf[n++] = dep(x - 3, y);
  if (x > 4) buf[n++] = dep(x - 5, y);
  if (x < W - 1) buf[n++] = dep(x + 1, y);
  if (x < W - 3) buf[n++] = dep(x + 3, y);
  if (x < W - 5) buf[n++] = dep(x + 5, y);
  if (y > 0 && x < W - 2) buf[n++] = dep(x + 2, y - 1);
  if (y < H - 1 && x < W - 2) buf[n++] = dep(x + 2, y + 1);
  if (y > 0 && x > 1) buf[n++] = dep(x - 2, y - 1);
  if (y < H - 1 && x > 1) buf[n++] = dep(x - 2, y + 1);
  if (x > 0 && y > 2) buf[n++] = dep(x - 1, y - 2);
  if (x < W - 1 && y > 2) buf[n++] = dep(x + 1, y - 2);
  if (x > 0 && y < H - 2) buf[n++] = dep(x - 1, y + 2);
  if (x < W - 1 && y < H - 2) buf[n++] = dep(x + 1, y + 2);
  {
    const real percent = 0.03f;
    const real cd = buf[0];
    if (cd > 0.0f) {
      const real lo = cd * (1 - percent);
      const real hi = cd * (1 + percent);
      int m = 1;
      for (int i = 1; i < n; ++i)
        if (buf[i] > lo && buf[i] < hi) buf[m++] = buf[i];
      n = m;
    }
  }
  for (int i = 1; i < n; ++i) {
    const real tmp = buf[i];
    int j;
    for (j = i; j >= 1 && tmp < buf[j - 1]; --j) buf[j] = buf[j - 1];
    buf[j] = tmp;
  }
  const int mid = n / 2;
  const real med = (n % 2 == 0) ? (buf[mid - 1] + buf[mid]) * 0.5f : buf[mid];
  if (med > 0.0f) {
    const Plane cp = load_plane(st->planes, center);
    const V3 nn = normalize3(V3{cp.x, cp.y, cp.z});
    store_plane(st->planes, center, plane_from_depth_normal(static_cast<real>(x), static_cast<real>(y), med, nn, c0));
  }
}

template <typename F>
void for_colour(int W, int H, int color, F f) {
  const int hw = (W + 1) / 2;
#pragma omp parallel for schedule(dynamic, 4)
  for (int gy = 0; gy < H; ++gy) {
    for (int gx = 0; gx < hw; ++gx) {
      const int x = gx * 2 + ((gy + color) % 2);
      if (x >= W) continue;
      f(x, gy);
    }
  }
}

}  // namespace

extern "C" {

void pmo_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
int pmo_max_threads(void) { return omp_get_max_threads(); }

float pmo_expf(float x) { return pm_expf(x); }
float pmo_logf(float x) { return pm_logf(x); }

float pmo_rand_float(uint32_t* state, uint32_t key) { return rand_float(state, key); }

real pmo_bilinear(const real* img, int w, int h, real sx, real sy) {
  Img im{img, w, h};
  return im.bilinear(sx, sy);
}

void pmo_project(const pmo_camera* cam, const real* X, real* out) {
  const V3 p = project(V3{X[0], X[1], X[2]}, *cam);
  out[0] = p.x; out[1] = p.y; out[2] = p.z;
}
void pmo_backproject(const pmo_camera* cam, real x, real y, real depth, real* out) {
  const V3 p = backproject(x, y, depth, *cam);
  out[0] = p.x; out[1] = p.y; out[2] = p.z;
}
real pmo_depth_from_plane(const pmo_camera* cam, real x, real y, const real* plane) {
  return depth_from_plane(x, y, Plane{plane[0], plane[1], plane[2], plane[3]}, *cam);
}
void pmo_plane_from_depth_normal(const pmo_camera* cam, real x, real y, real depth,
                                 const real* n, real* out) {
  const Plane p = plane_from_depth_normal(x, y, depth, V3{n[0], n[1], n[2]}, *cam);
  out[0] = p.x; out[1] = p.y; out[2] = p.z; out[3] = p.w;
}

// Reference patch moments (centre, sum_w, sum_wr, sum_wrr) then w[P], wr[P].
void pmo_ref_patch(const pmo_scene* sc, const pmo_params* pr, int x, int y, real* out) {
  Ctx c(sc, pr);
  RefPatch rp;
  compute_ref_patch(c, x, y, &rp);
  const int P = (2 * pr->half_patch + 1) * (2 * pr->half_patch + 1);
  out[0] = rp.ref_center; out[1] = rp.sum_w; out[2] = rp.sum_wr; out[3] = rp.sum_wrr;
  for (int i = 0; i < P; ++i) { out[4 + i] = rp.w[i]; out[4 + P + i] = rp.wr[i]; }
}

// Pixel-area downscale (the cv::resize INTER_AREA of BuildImagePyramid,
// estimator.cc:727-756): each output pixel is the mean of the input over its
// footprint, edge pixels weighted by the covered fraction; sums in double,
// optionally rounded to an integer (u8 images).
void pmo_resize_area(const real* src, int w, int h, int nw, int nh, int round_int, real* dst) {
  const double sx = static_cast<double>(w) / nw, sy = static_cast<double>(h) / nh;
  for (int dy = 0; dy < nh; ++dy) {
    const double y0 = dy * sy, y1 = std::min((dy + 1) * sy, static_cast<double>(h));
    const int iy0 = static_cast<int>(std::floor(y0)), iy1 = std::min(h, static_cast<int>(std::ceil(y1)));
    for (int dx = 0; dx < nw; ++dx) {
      const double x0 = dx * sx, x1 = std::min((dx + 1) * sx, static_cast<double>(w));
      const int ix0 = static_cast<int>(std::floor(x0)), ix1 = std::min(w, static_cast<int>(std::ceil(x1)));
      double acc = 0.0;
      for (int yy = iy0; yy < iy1; ++yy) {
        const double wy = std::min(y1, yy + 1.0) - std::max(y0, static_cast<double>(yy));
        double row = 0.0;
        for (int xx = ix0; xx < ix1; ++xx) {
          const double wx = std::min(x1, xx + 1.0) - std::max(x0, static_cast<double>(xx));
          row += wx * static_cast<double>(src[static_cast<int64_t>(yy) * w + xx]);
        }
        acc += wy * row;
      }
      double v = acc / ((x1 - x0) * (y1 - y0));
      if (round_int) v = std::nearbyint(v);
      dst[static_cast<int64_t>(dy) * nw + dx] = static_cast<real>(v);
    }
  }
}

// Bilinear resize with pixel-centre alignment and edge clamping (the
// cv::resize INTER_LINEAR of UploadPreviousDepths, estimator.cc:92-101).
void pmo_resize_linear(const real* src, int w, int h, int nw, int nh, real* dst) {
  const double sx = static_cast<double>(w) / nw, sy = static_cast<double>(h) / nh;
  for (int dy = 0; dy < nh; ++dy) {
    const double fy = (dy + 0.5) * sy - 0.5;
    int y0 = static_cast<int>(std::floor(fy));
    double ay = fy - y0;
    if (y0 < 0) { y0 = 0; ay = 0.0; }
    if (y0 >= h - 1) { y0 = h - 1; ay = 0.0; }
    const int y1 = std::min(y0 + 1, h - 1);
    for (int dx = 0; dx < nw; ++dx) {
      const double fx = (dx + 0.5) * sx - 0.5;
      int x0 = static_cast<int>(std::floor(fx));
      double ax = fx - x0;
      if (x0 < 0) { x0 = 0; ax = 0.0; }
      if (x0 >= w - 1) { x0 = w - 1; ax = 0.0; }
      const int x1 = std::min(x0 + 1, w - 1);
      const double v00 = src[static_cast<int64_t>(y0) * w + x0], v01 = src[static_cast<int64_t>(y0) * w + x1];
      const double v10 = src[static_cast<int64_t>(y1) * w + x0], v11 = src[static_cast<int64_t>(y1) * w + x1];
      const double v = (1.0 - ay) * ((1.0 - ax) * v00 + ax * v01) + ay * ((1.0 - ax) * v10 + ax * v11);
      dst[static_cast<int64_t>(dy) * nw + dx] = static_cast<real>(v);
    }
  }
}

// Coarse-to-fine transfer (acmmp_upsample, opencl_kernels.h:1792-1856) from a
// decoded src_w x src_h level; the coarse plane of each coarse pixel is
// built from its decoded depth and normal as UpsampleDepthNormal does
// (estimator.cc:775-791), with the fine camera's intrinsics scaled by the
// size ratio.  Sets planes, costs (2.0), rng and sel (0) of the fine state.
void pmo_upsample(const pmo_scene* sc, const pmo_params* pr, const real* cdepth, const real* cnormal, int src_w,
                  int src_h, pmo_state* st) {
  Ctx c(sc, pr);
  const int W = sc->width, H = sc->height;
  const pmo_camera& c0 = c.cams[0];
#pragma omp parallel for schedule(static)
  for (int y = 0; y < H; ++y) {
    for (int x = 0; x < W; ++x) {
      const int idx = y * W + x;
      uint32_t rs[2] = {static_cast<uint32_t>(idx) * 1099087573u + 2654435769u,
                        static_cast<uint32_t>(idx) * 2246822519u + 13u};
      const real scale_x = static_cast<real>(src_w) / static_cast<real>(W);
      const real scale_y = static_cast<real>(src_h) / static_cast<real>(H);
      const int sx = std::min(std::max(static_cast<int>(static_cast<real>(x) * scale_x), 0), src_w - 1);
      const int sy = std::min(std::max(static_cast<int>(static_cast<real>(y) * scale_y), 0), src_h - 1);
      const int sidx = sy * src_w + sx;
      const real d = cdepth[sidx];
      const V3 cn{cnormal[3 * sidx], cnormal[3 * sidx + 1], cnormal[3 * sidx + 2]};
      const real fx = c0.K[0] * static_cast<real>(src_w) / static_cast<real>(W);
      const real fy = c0.K[4] * static_cast<real>(src_h) / static_cast<real>(H);
      const real cxs = c0.K[2] * static_cast<real>(src_w) / static_cast<real>(W);
      const real cys = c0.K[5] * static_cast<real>(src_h) / static_cast<real>(H);
      const V3 pt{d * (static_cast<real>(sx) - cxs) / fx, d * (static_cast<real>(sy) - cys) / fy, d};
      const Plane cp{cn.x, cn.y, cn.z, -(cn.x * pt.x + cn.y * pt.y + cn.z * pt.z)};
      const real depth = depth_from_plane(static_cast<real>(x), static_cast<real>(y), cp, c0);
      Plane out;
      if (depth > 0.0f) {
        out = plane_from_depth_normal(static_cast<real>(x), static_cast<real>(y), depth,
                                      normalize3(V3{cp.x, cp.y, cp.z}), c0);
      } else {
        const uint32_t key = static_cast<uint32_t>(idx + 7919);
        const real ldmin = rlog(std::fmax(pr->depth_min, 1e-6f));
        const real ldmax = rlog(std::fmax(pr->depth_max, 1e-5f));
        const real rd = rexp(ldmin + rand_float(rs, key) * (ldmax - ldmin));
        V3 n;
        n.x = rand_float(rs, key + 3u) * 2.0f - 1.0f;
        n.y = rand_float(rs, key + 5u) * 2.0f - 1.0f;
        n.z = -1.0f;
        n = normalize3(n);
        out = plane_from_depth_normal(static_cast<real>(x), static_cast<real>(y), rd, n, c0);
      }
      store_plane(st->planes, idx, out);
      st->costs[idx] = 2.0f;
      st->rng[2 * idx] = rs[0];
      st->rng[2 * idx + 1] = rs[1];
      st->sel[idx] = 0u;
    }
  }
}

// Host post-processing of a decoded map (ReadBackResults, estimator.cc:
// 1194-1222): speckle removal (RemoveSmallSegments, :1231-1312: 4-connected
// segments of depths within 1 % relative of their neighbour; smaller than
// speckle_min pixels are cleared), gap filling (GapInterpolation,
// :1319-1418: runs of at most gap_max empty pixels between two depths within
// 2.5 % are filled linearly, rows first, then still-empty column runs), and
// the confidence of the result.  Each step runs only when its size is > 0.
void pmo_postprocess(real* depth, real* normal, const real* cost, real* confidence, int w, int h,
                     int speckle_min, int gap_max) {
  const int npix = w * h;
  if (speckle_min > 0) {
    const real thr = 0.01f;
    std::vector<char> seen(static_cast<size_t>(npix), 0);
    std::vector<int> comp;
    const int ox[4] = {1, -1, 0, 0}, oy[4] = {0, 0, 1, -1};
    for (int s = 0; s < npix; ++s) {
      if (seen[s] || depth[s] <= 0.0f) { seen[s] = 1; continue; }
      comp.assign(1, s);
      seen[s] = 1;
      for (size_t k = 0; k < comp.size(); ++k) {
        const int c = comp[k], cy = c / w, cx = c % w;
        for (int d = 0; d < 4; ++d) {
          const int nx = cx + ox[d], ny = cy + oy[d];
          if (nx < 0 || ny < 0 || nx >= w || ny >= h) continue;
          const int ni = ny * w + nx;
          if (seen[ni]) continue;
          if (depth[ni] <= 0.0f) { seen[ni] = 1; continue; }
          if (std::fabs(depth[c] - depth[ni]) / std::max(depth[c], depth[ni]) < thr) {
            seen[ni] = 1;
            comp.push_back(ni);
          }
        }
      }
      if (static_cast<int>(comp.size()) < speckle_min)
        for (int c : comp) { depth[c] = 0.0f; normal[3 * c] = normal[3 * c + 1] = normal[3 * c + 2] = 0.0f; }
    }
  }
  if (gap_max > 0) {
    const real thr = 0.025f;
    // fill pixel q from valid a to valid b at fraction t
    auto fill = [&](int a, int b, int q, real t, bool only_empty) {
      if (only_empty && depth[q] > 0.0f) return;
      depth[q] = depth[a] + t * (depth[b] - depth[a]);
      real v0 = (1.0f - t) * normal[3 * a] + t * normal[3 * b];
      real v1 = (1.0f - t) * normal[3 * a + 1] + t * normal[3 * b + 1];
      real v2 = (1.0f - t) * normal[3 * a + 2] + t * normal[3 * b + 2];
      const real len = std::sqrt(v0 * v0 + v1 * v1 + v2 * v2);
      if (len > 1e-6f) { v0 /= len; v1 /= len; v2 /= len; }
      normal[3 * q] = v0; normal[3 * q + 1] = v1; normal[3 * q + 2] = v2;
    };
    // one pass over lines; idx(line, pos) maps to a pixel
    auto pass = [&](int lines, int len, auto idx, bool only_empty) {
      for (int l = 0; l < lines; ++l) {
        int start = -1;
        for (int i = 0; i < len; ++i) {
          const int b = idx(l, i);
          if (depth[b] <= 0.0f) { if (start < 0) start = i; continue; }
          if (start >= 0 && start - 1 >= 0 && i - start <= gap_max) {
            const int a = idx(l, start - 1);
            // depth[a] is read before the run is filled (the run starts after a)
            const real da = depth[a], db = depth[b];
            if (std::fabs(da - db) / std::max(da, db) < thr)
              for (int g = start; g < i; ++g)
                fill(a, b, idx(l, g), static_cast<real>(g - (start - 1)) / static_cast<real>(i - (start - 1)),
                     only_empty);
          }
          start = -1;
        }
      }
    };
    pass(h, w, [w](int l, int i) { return l * w + i; }, false);
    pass(w, h, [w](int l, int i) { return i * w + l; }, true);
  }
  for (int i = 0; i < npix; ++i)
    confidence[i] = depth[i] > 0.0f ? 1.0f / (1.0f + rexp((cost[i] - 0.5f) * 5.0f)) : 0.0f;
}

// Float64 evaluation of the differentiable photometric cost of one pixel
// (the checker of pm_ncc_cost_grad's gradient by finite differences): mean
// over the views of `views` valid at theta of 1 - NCC, stride-1 bilateral
// NCC on the local affine warp of the plane through the pixel's point at
// depth d with normal (a, b, -1).  Returns the cost and sets *used to the
// number of valid views.
double pmo_ncc_energy_f64(const pmo_scene* sc, const pmo_params* pr, int x, int y, double d, double a, double b,
                          uint32_t views, int* used) {
  Ctx c(sc, pr);
  RefPatch rp;
  compute_ref_patch(c, x, y, &rp);
  const pmo_camera& c0 = c.cams[0];
  const double fx = c0.K[0], fy = c0.K[4], cx = c0.K[2], cy = c0.K[5];
  const double r0x = (x - cx) / fx, r0y = (y - cy) / fy, r1x = (x + 1 - cx) / fx, r2y = (y + 1 - cy) / fy;
  const double ur0 = a * r0x + b * r0y - 1.0, ur1 = a * r1x + b * r0y - 1.0, ur2 = a * r0x + b * r2y - 1.0;
  *used = 0;
  if (std::fabs(ur0) < 1e-6 || std::fabs(ur1) < 1e-6 || std::fabs(ur2) < 1e-6) return 0.0;
  const double dk[3] = {d, d * ur0 / ur1, d * ur0 / ur2};
  if (!(dk[0] > 0 && dk[1] > 0 && dk[2] > 0)) return 0.0;
  const double rx[3] = {r0x, r1x, r0x}, ry[3] = {r0y, r0y, r2y};
  const int hp = pr->half_patch, D = 2 * hp + 1;
  const double W = rp.sum_w;
  if (!(W >= 1e-6)) return 0.0;
  const double mean_r = rp.sum_wr / W, var_r = rp.sum_wrr / W - mean_r * mean_r;
  if (var_r < pr->var_threshold) return 0.0;
  double e = 0.0;
  for (int s = 0; s < c.n_src; ++s) {
    if (!((views >> s) & 1u)) continue;
    const pmo_camera& ci = c.cams[s + 1];
    double M[9], T[3], A[9], B[3];
    for (int r = 0; r < 3; ++r)
      for (int q = 0; q < 3; ++q)
        M[3 * r + q] = static_cast<double>(ci.R[3 * r]) * c0.R[3 * q] + static_cast<double>(ci.R[3 * r + 1]) * c0.R[3 * q + 1] +
                       static_cast<double>(ci.R[3 * r + 2]) * c0.R[3 * q + 2];
    for (int r = 0; r < 3; ++r) T[r] = ci.t[r] - (M[3 * r] * c0.t[0] + M[3 * r + 1] * c0.t[1] + M[3 * r + 2] * c0.t[2]);
    for (int r = 0; r < 3; ++r) {
      for (int q = 0; q < 3; ++q)
        A[3 * r + q] = static_cast<double>(ci.K[3 * r]) * M[q] + static_cast<double>(ci.K[3 * r + 1]) * M[3 + q] +
                       static_cast<double>(ci.K[3 * r + 2]) * M[6 + q];
      B[r] = ci.K[3 * r] * T[0] + ci.K[3 * r + 1] * T[1] + ci.K[3 * r + 2] * T[2];
    }
    double u[3], v[3];
    bool ok = true;
    for (int k = 0; k < 3; ++k) {
      const double px = dk[k] * (A[0] * rx[k] + A[1] * ry[k] + A[2]) + B[0];
      const double py = dk[k] * (A[3] * rx[k] + A[4] * ry[k] + A[5]) + B[1];
      const double pz = dk[k] * (A[6] * rx[k] + A[7] * ry[k] + A[8]) + B[2];
      if (!(pz > 1e-6)) { ok = false; break; }
      u[k] = px / pz;
      v[k] = py / pz;
    }
    if (!ok) continue;
    const double jxx = u[1] - u[0], jxy = v[1] - v[0], jyx = u[2] - u[0], jyy = v[2] - v[0];
    const double det = jxx * jyy - jxy * jyx;
    if (det < 0.02 || det > 50.0) continue;
    const Img& im = c.src[s];
    auto bil = [&](double sx, double sy) {
      sx = std::min(std::max(sx, -1.0), static_cast<double>(im.w - 1));
      sy = std::min(std::max(sy, -1.0), static_cast<double>(im.h - 1));
      const double fx0 = std::floor(sx), fy0 = std::floor(sy), ax = sx - fx0, ay = sy - fy0;
      const int x0 = static_cast<int>(fx0), y0 = static_cast<int>(fy0);
      const double t00 = im.texel(x0, y0), t01 = im.texel(x0 + 1, y0);
      const double t10 = im.texel(x0, y0 + 1), t11 = im.texel(x0 + 1, y0 + 1);
      const double v0 = t00 + ax * (t01 - t00), v1 = t10 + ax * (t11 - t10);
      return v0 + ay * (v1 - v0);
    };
    const double scen = bil(u[0], v[0]);
    double ss = 0, sss = 0, srs = 0;
    int p = 0;
    for (int iy = -hp; iy <= hp; ++iy)
      for (int ix = -hp; ix <= hp; ++ix, ++p) {
        const double sd = bil(u[0] + ix * jxx + iy * jyx, v[0] + ix * jxy + iy * jyy) - scen;
        ss += rp.w[p] * sd;
        sss += rp.w[p] * sd * sd;
        srs += rp.wr[p] * sd;
      }
    (void)D;
    const double mean_s = ss / W, var_s = sss / W - mean_s * mean_s;
    if (var_s < pr->var_threshold) continue;
    const double ncc = (srs / W - mean_r * mean_s) / std::sqrt(var_r * var_s);
    e += 1.0 - ncc;
    ++*used;
  }
  return *used ? e / *used : 0.0;
}

// Internal pieces, exposed so the tests can pin them one by one.
real pmo_smoothness_cost(const pmo_scene* sc, const pmo_params* pr, const real* planes, const real* plane,
                          int x, int y) {
  Ctx c(sc, pr);
  return smoothness_cost(c, planes, Plane{plane[0], plane[1], plane[2], plane[3]}, x, y);
}
void pmo_surface_normal(const pmo_scene* sc, const pmo_params* pr, const real* planes, int x, int y,
                        real center_depth, real* out) {
  Ctx c(sc, pr);
  const V3 n = surface_normal(c, planes, x, y, center_depth);
  out[0] = n.x; out[1] = n.y; out[2] = n.z;
}
real pmo_geom_cost(const pmo_scene* sc, const pmo_params* pr, const pmo_state* st, int sidx, const real* plane,
                    int x, int y) {
  Ctx c(sc, pr);
  return geom_cost(c, st, sidx, Plane{plane[0], plane[1], plane[2], plane[3]}, x, y);
}
real pmo_weighted_cost_geom(const pmo_scene* sc, const pmo_params* pr, const pmo_state* st, const real* cv,
                             const real* vw, real wnorm, const real* plane, int x, int y) {
  Ctx c(sc, pr);
  return weighted_cost_geom(c, st, cv, vw, wnorm, Plane{plane[0], plane[1], plane[2], plane[3]}, x, y);
}
real pmo_mc_cost(const pmo_scene* sc, const pmo_params* pr, const pmo_state* st, const real* cv, const real* vw,
                  real wnorm, const real* plane, int x, int y, uint32_t* rng, uint32_t key) {
  Ctx c(sc, pr);
  return mc_cost(c, st, cv, vw, wnorm, Plane{plane[0], plane[1], plane[2], plane[3]}, x, y, rng, key);
}

// View weights at (x, y): cand_costs is (8, num_src) row-major, flag bit j
// marks candidate j present.  Writes num_src weights and the selection mask;
// returns the weight sum.
real pmo_view_weights(const pmo_scene* sc, const pmo_params* pr, const real* cand_costs, uint32_t flag_bits,
                       uint32_t prev_sel, const real* plane, int x, int y, real* weights, uint32_t* sel) {
  Ctx c(sc, pr);
  real ca[8 * kMaxSrc];
  bool flags[8];
  for (int j = 0; j < 8; ++j) {
    flags[j] = (flag_bits >> j) & 1u;
    for (int i = 0; i < kMaxSrc; ++i) ca[j * kMaxSrc + i] = i < c.n_src ? cand_costs[j * c.n_src + i] : 2.0f;
  }
  real vw[kMaxSrc];
  const real wn = view_weights(c, ca, flags, prev_sel, Plane{plane[0], plane[1], plane[2], plane[3]}, x, y, vw, sel);
  for (int i = 0; i < c.n_src; ++i) weights[i] = vw[i];
  return wn;
}

// Refinement hypotheses at (x, y) for the current plane (and an optional
// prior plane); writes n planes (n, d) to out and returns n.  rng advances.
int pmo_refinement_hypotheses(const pmo_scene* sc, const pmo_params* pr, const real* planes, int x, int y,
                              const real* plane_now, const real* prior_plane, uint32_t* rng, uint32_t key,
                              real* out) {
  Ctx c(sc, pr);
  const pmo_camera& c0 = c.cams[0];
  const Plane pn{plane_now[0], plane_now[1], plane_now[2], plane_now[3]};
  const real depth_now = depth_from_plane(static_cast<real>(x), static_cast<real>(y), pn, c0);
  bool have_prior = prior_plane != nullptr;
  real depth_prior = 0.0f;
  V3 prior_n{0.0f, 0.0f, -1.0f};
  if (have_prior) {
    const Plane pp{prior_plane[0], prior_plane[1], prior_plane[2], prior_plane[3]};
    depth_prior = depth_from_plane(static_cast<real>(x), static_cast<real>(y), pp, c0);
    prior_n = normalize3(V3{pp.x, pp.y, pp.z});
  }
  const real depth_sigma = (pr->depth_max - pr->depth_min) / 32.0f;
  const real angle_sigma = kPiF * (5.0f / 180.0f);
  real rd[6];
  V3 rn[6];
  const int n = refinement_hypotheses(c, planes, x, y, pn, depth_now, have_prior, depth_prior, prior_n,
                                      depth_sigma, angle_sigma, rng, key, rd, rn);
  for (int h = 0; h < n; ++h) {
    const Plane q = plane_from_depth_normal(static_cast<real>(x), static_cast<real>(y), rd[h], rn[h], c0);
    out[4 * h] = q.x; out[4 * h + 1] = q.y; out[4 * h + 2] = q.z; out[4 * h + 3] = q.w;
  }
  return n;
}

// Score n (pixel, plane) pairs: per-view costs (n x num_src), top-k initial
// cost and view bitmask.  xy: 2n ints; planes: 4n floats.
// The local affine warp of cost_vector for one source: out = (scx, scy,
// jxx, jxy, jyx, jyy); returns -1 when a plane depth is <= 0, else
// source_warp's code (0 valid, 1 behind the source camera, 2 det gate).
int pmo_warp(const pmo_scene* sc, const pmo_params* pr, int x, int y, const real* plane, int src, real* out) {
  Ctx c(sc, pr);
  PlanePoints pp;
  if (!plane_points(c.cams[0], x, y, load_plane(plane, 0), &pp)) return -1;
  Warp w{};
  const int code = source_warp(pp, c.cams[src + 1], &w);
  const real v[6] = {w.scx, w.scy, w.jxx, w.jxy, w.jyx, w.jyy};
  std::memcpy(out, v, sizeof(v));
  return code;
}

// The strided NCC of the textureless fallback on a given warp w = (scx, scy,
// jxx, jxy, jyx, jyy) of source src at reference pixel (x, y); the source
// centre value is the bilinear sample at (scx, scy), as in cost_vector.
real pmo_ncc_at_stride(const pmo_scene* sc, const pmo_params* pr, int x, int y, int src, real stride,
                        const real* w) {
  Ctx c(sc, pr);
  const real scen = c.src[src].bilinear(w[0], w[1]);
  return ncc_at_stride(c, src, c.ref.texel(x, y), scen, w[0], w[1], x, y, stride, w[2], w[3], w[4], w[5]);
}

void pmo_score_planes(const pmo_scene* sc, const pmo_params* pr, int n, const int* xy,
                      const real* planes, real* costs, real* init_cost, uint32_t* sel) {
  Ctx c(sc, pr);
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < n; ++i) {
    const int x = xy[2 * i], y = xy[2 * i + 1];
    RefPatch rp;
    compute_ref_patch(c, x, y, &rp);
    const Plane p{planes[4 * i], planes[4 * i + 1], planes[4 * i + 2], planes[4 * i + 3]};
    real cv[kMaxSrc];
    for (int k = 0; k < kMaxSrc; ++k) cv[k] = 2.0f;
    cost_vector(c, x, y, p, rp, cv);
    for (int k = 0; k < c.n_src; ++k) costs[static_cast<int64_t>(i) * c.n_src + k] = cv[k];
    uint32_t s = 0u;
    const real ic = initial_cost_and_views(c, x, y, p, rp, &s);
    if (init_cost) init_cost[i] = ic;
    if (sel) sel[i] = s;
  }
}

void pmo_random_init(const pmo_scene* sc, const pmo_params* pr, pmo_state* st) {
  Ctx c(sc, pr);
  const int W = sc->width, H = sc->height;
#pragma omp parallel for schedule(dynamic, 4)
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) random_init_px(c, st, x, y);
}

void pmo_plane_sweep(const pmo_scene* sc, const pmo_params* pr, pmo_state* st, int num_planes) {
  Ctx c(sc, pr);
  const int W = sc->width, H = sc->height;
  if (num_planes < 1) return;
#pragma omp parallel for schedule(dynamic, 4)
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) plane_sweep_px(c, st, x, y, num_planes);
}

real pmo_sweep_depth(const pmo_params* pr, int i, int n) { return sweep_depth(*pr, i, n); }

void pmo_prior_reinit(const pmo_scene* sc, const pmo_params* pr, pmo_state* st) {
  Ctx c(sc, pr);
  const int W = sc->width, H = sc->height;
  if (!st->plane_masks) return;
  // Every pixel runs at once here (no colour split), and the smoothness term
  // reads 4-neighbour planes from the pre-kernel snapshot (planes_img at
  // opencl_kernels.h:1144,1170), so those reads go to a copy.
  real* snap = new real[static_cast<size_t>(W) * H * 4];
  std::memcpy(snap, st->planes, sizeof(real) * 4 * static_cast<size_t>(W) * H);
#pragma omp parallel for schedule(dynamic, 4)
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) prior_reinit_px(c, st, snap, x, y);
  delete[] snap;
}

void pmo_sweep(const pmo_scene* sc, const pmo_params* pr, pmo_state* st, int color, int iteration,
               real smooth_w) {
  Ctx c(sc, pr);
  for_colour(sc->width, sc->height, color, [&](int x, int y) { patchmatch_px(c, st, x, y, iteration, smooth_w); });
}

void pmo_checkerboard(const pmo_scene* sc, const pmo_params* pr, pmo_state* st, int color) {
  Ctx c(sc, pr);
  for_colour(sc->width, sc->height, color, [&](int x, int y) { checkerboard_px(c, st, x, y); });
}

// Smooth weight of iteration `iter` (estimator.cc:646-649).
real pmo_smooth_weight(const pmo_params* pr, int iter) {
  const real decay = 1.0f - 0.9f * static_cast<real>(iter) /
                                 static_cast<real>(std::max(pr->max_iterations - 1, 1));
  return pr->smooth_weight * decay;
}

// Decode planes/costs into depth, normal, cost, confidence
// (estimator.cc:1112-1224; speckle / gap filling off, median off).
void pmo_decode(const pmo_scene* sc, const pmo_params* pr, const pmo_state* st, real* depth,
                real* normal, real* cost, real* confidence) {
  const int W = sc->width, H = sc->height;
  const pmo_camera& c0 = sc->cams[0];
#pragma omp parallel for schedule(static)
  for (int y = 0; y < H; ++y) {
    for (int x = 0; x < W; ++x) {
      const int64_t idx = static_cast<int64_t>(y) * W + x;
      const real* p = &st->planes[4 * idx];
      depth[idx] = 0.0f;
      normal[3 * idx] = 0.0f; normal[3 * idx + 1] = 0.0f; normal[3 * idx + 2] = 0.0f;
      cost[idx] = 0.0f;
      confidence[idx] = 0.0f;
      if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2]) || !std::isfinite(p[3]))
        continue;
      const real rx = (static_cast<real>(x) - c0.K[2]) / c0.K[0];
      const real ry = (static_cast<real>(y) - c0.K[5]) / c0.K[4];
      const real denom = p[0] * rx + p[1] * ry + p[2] * 1.0f;
      real d = (std::fabs(denom) > 1e-6f) ? -p[3] / denom : 0.0f;
      if (!std::isfinite(d)) d = 0.0f;
      if (d > 0.0f && d >= pr->depth_min && d <= pr->depth_max) {
        depth[idx] = d;
        const real len = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        if (len > 1e-6f) {
          normal[3 * idx] = p[0] / len; normal[3 * idx + 1] = p[1] / len; normal[3 * idx + 2] = p[2] / len;
        }
      }
      const real cc = st->costs[idx];
      cost[idx] = std::isfinite(cc) ? cc : 0.0f;
      if (depth[idx] > 0.0f)
        confidence[idx] = 1.0f / (1.0f + rexp((cost[idx] - 0.5f) * 5.0f));
    }
  }
}

// Whole single-level run (estimator.cc:134-255 with hierarchy_levels = 1):
// random init, prior re-init, iterations x {red, black}, optional
// checkerboard filter.  Decode separately with pmo_decode.
void pmo_run(const pmo_scene* sc, const pmo_params* pr, pmo_state* st) {
  pmo_random_init(sc, pr, st);
  pmo_prior_reinit(sc, pr, st);
  for (int it = 0; it < pr->max_iterations; ++it) {
    const real sw = pmo_smooth_weight(pr, it);
    pmo_sweep(sc, pr, st, 0, it, sw);
    pmo_sweep(sc, pr, st, 1, it, sw);
  }
  if (pr->checkerboard_filter) {
    pmo_checkerboard(sc, pr, st, 0);
    pmo_checkerboard(sc, pr, st, 1);
  }
}


// ---------------------------------------------------------------------------
// Depth-map cleaning: acmmp_clean_depthmap (opencl_kernels.h:2116-2279) as
// run by GPUDepthmapCleaner::Clean (cleaner.cc:95-242).  src_depths are the
// neighbours' depth maps already at the reference size (Clean resizes them
// with cv::resize INTER_LINEAR, cleaner.cc:148-158); cams[0] is the
// reference, cams[1..n_src] the neighbours (CLCamera of each view's own K,
// R, t, cleaner.cc:196-201).  Neighbour depths are read bilinearly at the
// projected pixel-centre coordinate (read_src_depth, :2090-2112); the
// reference depth at the pixel (nearest sampler).  normal: the reference's
// camera-frame normals (3 per pixel) or null (has_normal = 0).
// min_consistent_views is the effective minimum Clean passes,
// min(min_consistent_views, num_views) (cleaner.cc:222).
// ---------------------------------------------------------------------------
void pmo_clean_depthmap(const real* ref_depth, const real* normal, int w, int h, const real* const* src_depths,
                        const pmo_camera* cams, int n_src_in, real same_depth_threshold, int min_consistent_views,
                        real carving_threshold, int max_carved_views, real grazing_cos_threshold, real* out) {
  const int n_src = std::min(n_src_in, kMaxSrc);  // MAX_CLEAN_SOURCES 16 (:2057)
  const pmo_camera& rc = cams[0];
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; ++y) {
    for (int x = 0; x < w; ++x) {
      const int idx = y * w + x;
      const real ref_d = ref_depth[idx];
      if (!(ref_d > 0.0f)) {
        out[idx] = 0.0f;
        continue;
      }
      // grazing angle from the normal (:2142-2160)
      int is_grazing = 0;
      if (normal) {
        const real nnx = normal[3 * idx], nny = normal[3 * idx + 1], nnz = normal[3 * idx + 2];
        const real nlen = std::sqrt(nnx * nnx + nny * nny + nnz * nnz);
        if (nlen > 1e-6f) {
          const real ray_x = (static_cast<real>(x) - rc.K[2]) / rc.K[0];
          const real ray_y = (static_cast<real>(y) - rc.K[5]) / rc.K[4];
          const real ray_z = 1.0f;
          const real rlen = std::sqrt(ray_x * ray_x + ray_y * ray_y + ray_z * ray_z);
          const real cos_angle = std::fabs((nnx * ray_x + nny * ray_y + nnz * ray_z) / (nlen * rlen));
          is_grazing = (cos_angle < grazing_cos_threshold) ? 1 : 0;
        }
      }
      const int is_suspicious = is_grazing;
      const real fx = rc.K[0], cx = rc.K[2], fy = rc.K[4], cy = rc.K[5];
      // back-project into the world (:2164-2177)
      const real cam_x = (static_cast<real>(x) - cx) / fx * ref_d;
      const real cam_y = (static_cast<real>(y) - cy) / fy * ref_d;
      const real cam_z = ref_d;
      const real pcx = cam_x - rc.t[0], pcy = cam_y - rc.t[1], pcz = cam_z - rc.t[2];
      const real wx = rc.R[0] * pcx + rc.R[3] * pcy + rc.R[6] * pcz;
      const real wy = rc.R[1] * pcx + rc.R[4] * pcy + rc.R[7] * pcz;
      const real wz = rc.R[2] * pcx + rc.R[5] * pcy + rc.R[8] * pcz;
      int consistent = 1;
      int carved = 0;
      for (int v = 0; v < n_src; ++v) {
        const pmo_camera& sc = cams[v + 1];
        const real sx = sc.R[0] * wx + sc.R[1] * wy + sc.R[2] * wz + sc.t[0];
        const real sy = sc.R[3] * wx + sc.R[4] * wy + sc.R[5] * wz + sc.t[1];
        const real sz = sc.R[6] * wx + sc.R[7] * wy + sc.R[8] * wz + sc.t[2];
        if (sz <= 1e-6f) continue;
        const real su = sc.K[0] * sx / sz + sc.K[2];
        const real sv = sc.K[4] * sy / sz + sc.K[5];
        if (su < 0.0f || sv < 0.0f || su >= static_cast<real>(w - 1) || sv >= static_cast<real>(h - 1)) continue;
        const real src_d = Img{src_depths[v], w, h}.bilinear(su, sv);
        if (!(src_d > 0.0f)) continue;
        // the source surface point back in the world (:2212-2226, 2239-2252)
        const real scx = (su - sc.K[2]) / sc.K[0] * src_d;
        const real scy = (sv - sc.K[5]) / sc.K[4] * src_d;
        const real scz = src_d;
        const real qx = scx - sc.t[0], qy = scy - sc.t[1], qz = scz - sc.t[2];
        const real wx2 = sc.R[0] * qx + sc.R[3] * qy + sc.R[6] * qz;
        const real wy2 = sc.R[1] * qx + sc.R[4] * qy + sc.R[7] * qz;
        const real wz2 = sc.R[2] * qx + sc.R[5] * qy + sc.R[8] * qz;
        // space carving: the source sees further (:2210-2235)
        if (src_d > sz * (1.0f + carving_threshold)) {
          const real rz = rc.R[6] * wx2 + rc.R[7] * wy2 + rc.R[8] * wz2 + rc.t[2];
          if (rz > 1e-6f && rz > ref_d * (1.0f + same_depth_threshold)) carved++;
          continue;
        }
        // forward depth check, then back into the reference (:2238-2267)
        if (std::fabs(src_d - sz) >= sz * same_depth_threshold) continue;
        const real rx = rc.R[0] * wx2 + rc.R[1] * wy2 + rc.R[2] * wz2 + rc.t[0];
        const real ry = rc.R[3] * wx2 + rc.R[4] * wy2 + rc.R[5] * wz2 + rc.t[1];
        const real rz = rc.R[6] * wx2 + rc.R[7] * wy2 + rc.R[8] * wz2 + rc.t[2];
        if (rz <= 1e-6f) continue;
        const real ru = fx * rx / rz + cx;
        const real rv = fy * ry / rz + cy;
        if (std::fabs(ru - static_cast<real>(x)) + std::fabs(rv - static_cast<real>(y)) > 2.0f) continue;
        if (std::fabs(rz - ref_d) >= ref_d * same_depth_threshold) continue;
        consistent++;
      }
      // adaptive decision (:2270-2276)
      const int eff_min = is_suspicious ? (min_consistent_views + 1) : min_consistent_views;
      const int eff_max_carve = is_suspicious ? std::max(1, max_carved_views - 1) : max_carved_views;
      const int carve_budget = eff_max_carve + std::max(0, consistent - eff_min);
      const bool keep = (consistent >= eff_min) && (carved <= carve_budget);
      out[idx] = keep ? ref_d : 0.0f;
    }
  }
}

}  // extern "C"
