// PatchMatch kernels for gfx950 (MI355X).
//
// One lane per reference pixel; the 64 lanes of a wave cover a 16 x 4 tile
// of same-colour pixels (32 x 4 image pixels), so the warped source patches
// of neighbouring lanes overlap and the bilinear gathers of a wave hit the
// same cache lines.  The bilateral reference patch (w, w*r per tap) lives
// in VGPRs for the whole pass (patch size is a template parameter, loops
// are fully unrolled); per-candidate cost vectors live in LDS, one column
// per lane, so the Monte-Carlo view sampling can index them at random.
// Colour passes update the state in place: every neighbour a pixel reads
// has the other colour (odd Manhattan offsets, opencl_kernels.h:1262-1430,
// 664-668, 740-746), which makes an in-place pass equal to the reference's
// snapshot pass.
#include "pm_kernels.h"

namespace pmhip {

namespace {

constexpr int kWave = 64;
constexpr int kTileX = 16;
constexpr int kTileY = 4;
constexpr int kSweepSlots = 9;   // 8 propagation candidates + current/refinement

template <int HP>
struct RefPatch {
  static constexpr int P = (2 * HP + 1) * (2 * HP + 1);
  float center, sum_w, sum_wr, sum_wrr;
  float w[P];
  float wr[P];
};

__device__ __forceinline__ Plane to_plane(float4 v) { return Plane{v.x, v.y, v.z, v.w}; }
__device__ __forceinline__ float4 to_f4(Plane p) { return make_float4(p.x, p.y, p.z, p.w); }

// opencl_kernels.h:328-366
template <int HP, typename T>
__device__ __forceinline__ void compute_ref_patch(const T* __restrict__ img, const PmView& rv, int x, int y,
                                                  float sf, float cf, RefPatch<HP>& rp) {
  const float center = texel(img, rv, x, y);
  float sum_w = 0.0f, sum_wr = 0.0f, sum_wrr = 0.0f;
  int p = 0;
#pragma unroll
  for (int iy = -HP; iy <= HP; ++iy) {
#pragma unroll
    for (int ix = -HP; ix <= HP; ++ix) {
      const float r = texel(img, rv, x + ix, y + iy) - center;
      const float w = pm_expf(sf * static_cast<float>(ix * ix + iy * iy) + cf * r * r);
      sum_w += w;
      const float wr = w * r;
      sum_wr += wr;
      sum_wrr = fmaf(wr, r, sum_wrr);
      rp.w[p] = w;
      rp.wr[p] = wr;
      ++p;
    }
  }
  rp.center = center;
  rp.sum_w = sum_w;
  rp.sum_wr = sum_wr;
  rp.sum_wrr = sum_wrr;
}

// opencl_kernels.h:188-258
template <int HP, typename T>
__device__ __forceinline__ float ncc_at_stride(const PmArgs& a, const T* __restrict__ img, int s,
                                                         float ref_center, float src_center, float scx, float scy,
                                                         int cx, int cy, float stride, float jxx, float jxy,
                                                         float jyx, float jyy) {
  const PmView rv = a.views[0];
  const PmView sv = a.views[s + 1];
  float sum_w = 0.0f, sum_r = 0.0f, sum_s = 0.0f, sum_rr = 0.0f, sum_ss = 0.0f, sum_rs = 0.0f;
  for (int iy = -HP; iy <= HP; ++iy) {
    for (int ix = -HP; ix <= HP; ++ix) {
      const float fdx = stride * static_cast<float>(ix);
      const float fdy = stride * static_cast<float>(iy);
      const float rx = static_cast<float>(cx) + fdx;
      const float ry = static_cast<float>(cy) + fdy;
      const float ref_val = texel(img, rv, static_cast<int>(rx), static_cast<int>(ry));
      const float sx = scx + jxx * fdx + jyx * fdy;
      const float sy = scy + jxy * fdx + jyy * fdy;
      const float src_val = bilinear(img, sv, sx, sy);
      const float r = ref_val - ref_center;
      const float w = pm_expf(fdx * fdx * a.spatial_factor + fdy * fdy * a.spatial_factor +
                              r * r * a.color_factor);
      const float sd = src_val - src_center;
      sum_r += w * r;
      sum_s += w * sd;
      sum_rr += w * r * r;
      sum_ss += w * sd * sd;
      sum_rs += w * r * sd;
      sum_w += w;
    }
  }
  if (sum_w < 1e-6f) return -1.0f;
  const float inv_w = 1.0f / sum_w;
  const float mean_r = sum_r * inv_w;
  const float mean_s = sum_s * inv_w;
  const float var_ref = sum_rr * inv_w - mean_r * mean_r;
  const float var_src = sum_ss * inv_w - mean_s * mean_s;
  if (var_ref < a.var_threshold || var_src < a.var_threshold) return -1.0f;
  const float covar = sum_rs * inv_w - mean_r * mean_s;
  return covar / sqrtf(var_ref * var_src);
}

// opencl_kernels.h:270-315
template <typename T>
__device__ __forceinline__ float census_inner(const PmArgs& a, const T* __restrict__ img, int s,
                                                        float ref_center, float src_center, int cx, int cy,
                                                        float scx, float scy, float jxx, float jxy, float jyx,
                                                        float jyy) {
  const PmView rv = a.views[0];
  const PmView sv = a.views[s + 1];
  uint32_t rc = 0u, sc = 0u;
  int bit = 0;
  for (int dy = -2; dy <= 2; ++dy) {
    for (int dx = -2; dx <= 2; ++dx) {
      if (dx == 0 && dy == 0) continue;
      const float ref_val = texel(img, rv, cx + dx, cy + dy);
      const float sx = scx + jxx * static_cast<float>(dx) + jyx * static_cast<float>(dy);
      const float sy = scy + jxy * static_cast<float>(dx) + jyy * static_cast<float>(dy);
      const float src_val = bilinear(img, sv, sx, sy);
      if (ref_val > ref_center) rc |= (1u << bit);
      if (src_val > src_center) sc |= (1u << bit);
      ++bit;
    }
  }
  const int hamming = __popc(rc ^ sc);
  if (__popc(rc) < 3 && __popc(sc) < 3) return 1.0f;
  return static_cast<float>(hamming) / 24.0f;
}

// Per-view cost vector of one plane at (x, y): opencl_kernels.h:381-557.
// Writes n_src costs to out[s * ostride].  Views are the outer loop; each
// view accumulates its taps in row-major order, as in the reference.
template <int HP, typename T>
__device__ __forceinline__ void cost_vector(const PmArgs& a, const T* __restrict__ img, int x, int y, Plane pl,
                                            const RefPatch<HP>& rp, float* out, int ostride) {
  const int n_src = a.n_src;
  const PmCam& c0 = a.cams[0];
  const float fx = static_cast<float>(x), fy = static_cast<float>(y);
  const float depth = depth_from_plane(fx, fy, pl, c0);
  const float depth_dx = depth_from_plane(static_cast<float>(x + 1), fy, pl, c0);
  const float depth_dy = depth_from_plane(fx, static_cast<float>(y + 1), pl, c0);
  if (depth <= 0.0f || depth_dx <= 0.0f || depth_dy <= 0.0f) {  // :402-412
    for (int s = 0; s < n_src; ++s) out[s * ostride] = 2.0f;
    return;
  }
  const V3 cw = backproject(fx, fy, depth, c0);
  const V3 wdx = backproject(static_cast<float>(x + 1), fy, depth_dx, c0);
  const V3 wdy = backproject(fx, static_cast<float>(y + 1), depth_dy, c0);

  const bool ref_ok = rp.sum_w >= 1e-6f;
  float inv_w = 0.0f, mean_r = 0.0f, var_ref = 0.0f;
  if (ref_ok) {
    inv_w = 1.0f / rp.sum_w;
    mean_r = rp.sum_wr * inv_w;
    var_ref = rp.sum_wrr * inv_w - mean_r * mean_r;
  }

  for (int s = 0; s < n_src; ++s) {
    const PmCam& cs = a.cams[s + 1];
    const PmView sv = a.views[s + 1];
    float cost = 2.0f;
    const V3 pc = project(cw, cs);
    if (!(pc.z < 1e-6f)) {
      const float scx = pc.x / pc.z;
      const float scy = pc.y / pc.z;
      const V3 pdx = project(wdx, cs);
      const V3 pdy = project(wdy, cs);
      if (!(pdx.z < 1e-6f || pdy.z < 1e-6f)) {
        const float jxx = pdx.x / pdx.z - scx;
        const float jxy = pdx.y / pdx.z - scy;
        const float jyx = pdy.x / pdy.z - scx;
        const float jyy = pdy.y / pdy.z - scy;
        const float det = jxx * jyy - jxy * jyx;
        if (!(det < 0.02f || det > 50.0f) && ref_ok) {
          const float scen = bilinear(img, sv, scx, scy);
          // Fused stride-1 ZNCC over the cached reference weights (:471-499).
          float sum_ws = 0.0f, sum_wss = 0.0f, sum_wrs = 0.0f;
          int p = 0;
#pragma unroll
          for (int iy = -HP; iy <= HP; ++iy) {
            const float fiy = static_cast<float>(iy);
            const float bx = fmaf(jyx, fiy, scx);
            const float by = fmaf(jyy, fiy, scy);
#pragma unroll
            for (int ix = -HP; ix <= HP; ++ix) {
              const float fix = static_cast<float>(ix);
              const float sx = fmaf(jxx, fix, bx);
              const float sy = fmaf(jxy, fix, by);
              const float sd = bilinear(img, sv, sx, sy) - scen;
              const float wsv = rp.w[p] * sd;
              sum_ws += wsv;
              sum_wss = fmaf(wsv, sd, sum_wss);
              sum_wrs = fmaf(rp.wr[p], sd, sum_wrs);
              ++p;
            }
          }
          float ncc1 = -1.0f;
          if (var_ref >= a.var_threshold) {
            const float mean_s = sum_ws * inv_w;
            const float var_src = sum_wss * inv_w - mean_s * mean_s;
            if (var_src >= a.var_threshold) {
              const float covar = sum_wrs * inv_w - mean_r * mean_s;
              ncc1 = covar / sqrtf(var_ref * var_src);
            }
          }
          if (ncc1 > 0.1f) {
            cost = 1.0f - ncc1;
          } else {
            // Textureless fallback: strides 2 then 3, stop once NCC > 0.1 (:532-542).
            float best = ncc1;
#pragma unroll 1
            for (int st = 0; st < 2; ++st) {
              const float ncc = ncc_at_stride<HP>(a, img, s, rp.center, scen, scx, scy, x, y,
                                                  st == 0 ? 2.0f : 3.0f, jxx, jxy, jyx, jyy);
              if (ncc > best) best = ncc;
              if (best > 0.1f) break;
            }
            if (best <= -0.5f) {
              cost = (a.census_weight > 1e-6f)
                         ? census_inner(a, img, s, rp.center, scen, x, y, scx, scy, jxx, jxy, jyx, jyy)
                         : 2.0f;
            } else {
              cost = 1.0f - best;
            }
          }
        }
      }
    }
    out[s * ostride] = cost;
  }
}

// Smoothness against the 4 neighbours (opencl_kernels.h:649-707); the
// neighbour planes are preloaded (they have the other colour and cannot
// change during the pass).
struct Nbrs {
  Plane p[4];     // left, right, up, down
  bool ok[4];
};

__device__ __forceinline__ float smoothness_cost(const PmArgs& a, const Nbrs& nb, Plane cur, int x, int y) {
  const PmCam& c0 = a.cams[0];
  const float cur_depth = depth_from_plane(static_cast<float>(x), static_cast<float>(y), cur, c0);
  if (cur_depth <= 0.0f) return 0.0f;
  const V3 cur_n = normalize3(V3{cur.x, cur.y, cur.z});
  const float sigma_rel = 0.02f;
  const int dxs[4] = {-1, 1, 0, 0};
  const int dys[4] = {0, 0, -1, 1};
  float tw = 0.0f, tc = 0.0f;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (!nb.ok[k]) continue;
    const int nx = x + dxs[k], ny = y + dys[k];
    const Plane q = nb.p[k];
    const float nd = depth_from_plane(static_cast<float>(nx), static_cast<float>(ny), q, c0);
    if (nd <= 0.0f) continue;
    const float rel = fabsf(cur_depth - nd) / cur_depth;
    const float w = pm_expf(-rel * rel / (2.0f * sigma_rel * sigma_rel));
    const V3 nn = normalize3(V3{q.x, q.y, q.z});
    const float normal_cost = 1.0f - dot3(cur_n, nn);
    const float predicted = depth_from_plane(static_cast<float>(nx), static_cast<float>(ny), cur, c0);
    const float dev = (predicted > 0.0f) ? fabsf(nd - predicted) / cur_depth : 0.0f;
    const float cost = 0.5f * clampf(normal_cost, 0.0f, 1.0f) + 0.5f * clampf(dev * 10.0f, 0.0f, 1.0f);
    tc += w * cost;
    tw += w;
  }
  if (tw < 1e-6f) return 0.0f;
  return tc / tw;
}

__device__ __forceinline__ void load_nbrs(const PmArgs& a, const float4* __restrict__ planes, int x, int y, Nbrs& nb) {
  const int W = a.width, H = a.height;
  const int dxs[4] = {-1, 1, 0, 0};
  const int dys[4] = {0, 0, -1, 1};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int nx = x + dxs[k], ny = y + dys[k];
    nb.ok[k] = !(nx < 0 || nx >= W || ny < 0 || ny >= H);
    nb.p[k] = nb.ok[k] ? to_plane(gp(planes)[static_cast<int64_t>(ny) * W + nx]) : Plane{0, 0, 0, 0};
  }
}

// opencl_kernels.h:725-772
__device__ __forceinline__ V3 surface_normal(const PmArgs& a, const Nbrs& nb, int x, int y, float center_depth) {
  const int W = a.width, H = a.height;
  if (center_depth <= 0.0f) return V3{0, 0, 0};
  if (x < 1 || x >= W - 1 || y < 1 || y >= H - 1) return V3{0, 0, 0};
  const PmCam& c0 = a.cams[0];
  const float fx = c0.K[0], fy = c0.K[4], cx = c0.K[2], cy = c0.K[5];
  const float dl = depth_from_plane(static_cast<float>(x - 1), static_cast<float>(y), nb.p[0], c0);
  const float dr = depth_from_plane(static_cast<float>(x + 1), static_cast<float>(y), nb.p[1], c0);
  const float du = depth_from_plane(static_cast<float>(x), static_cast<float>(y - 1), nb.p[2], c0);
  const float dd = depth_from_plane(static_cast<float>(x), static_cast<float>(y + 1), nb.p[3], c0);
  if (dl <= 0.0f || dr <= 0.0f || du <= 0.0f || dd <= 0.0f) return V3{0, 0, 0};
  const float dx = dr - dl;
  const float dy = dd - du;
  V3 n{fx * dx, fy * dy, (cx - static_cast<float>(x)) * dx + (cy - static_cast<float>(y)) * dy - center_depth};
  const float l = len3(n);
  if (l < 1e-8f) return V3{0, 0, 0};
  n = V3{n.x / l, n.y / l, n.z / l};
  if (n.z > 0.0f) n = V3{-n.x, -n.y, -n.z};
  return n;
}

// opencl_kernels.h:779-861
__device__ __forceinline__ float geom_cost(const PmArgs& a, int sidx, Plane plane, int px, int py) {
  const int pdw = a.width, pdh = a.height;
  const PmCam& rc = a.cams[0];
  const PmCam& scm = a.cams[sidx + 1];
  const float max_dim = static_cast<float>(max(pdw, pdh));
  if (max_dim < 1.0f) return 1.0f;
  const float good_thresh = 3.0f / max_dim;
  const float full_thresh = 6.0f / max_dim;
  const float depth = depth_from_plane(static_cast<float>(px), static_cast<float>(py), plane, rc);
  if (depth <= 0.0f) return 1.0f;
  const V3 wp = backproject(static_cast<float>(px), static_cast<float>(py), depth, rc);
  const V3 sp = project(wp, scm);
  if (sp.z < 1e-6f) return 1.0f;
  const float sx = sp.x / sp.z, sy = sp.y / sp.z;
  if (sx < 0.0f || sx >= static_cast<float>(pdw) || sy < 0.0f || sy >= static_cast<float>(pdh)) return 1.0f;
  const int ix = min(max(static_cast<int>(sx + 0.5f), 0), pdw - 1);
  const int iy = min(max(static_cast<int>(sy + 0.5f), 0), pdh - 1);
  const float sd = a.prev_depths[static_cast<int64_t>(sidx) * pdw * pdh + static_cast<int64_t>(iy) * pdw + ix];
  if (sd <= 0.0f) return 1.0f;
  const V3 sw = backproject(sx, sy, sd, scm);
  const V3 rp = project(sw, rc);
  if (rp.z < 1e-6f) return 1.0f;
  const float bx = rp.x / rp.z, by = rp.y / rp.z;
  const float dx = static_cast<float>(px) - bx;
  const float dy = static_cast<float>(py) - by;
  const float err = sqrtf(dx * dx + dy * dy) / max_dim;
  if (err <= good_thresh) return 0.0f;
  if (err >= full_thresh) return 1.0f;
  return (err - good_thresh) / (full_thresh - good_thresh);
}

// opencl_kernels.h:910-969; cv is this lane's LDS column (stride kWave).
__device__ __forceinline__ float mc_cost(const PmArgs& a, const float* cv, const float (&vw)[kMaxSrc], float wnorm,
                                         Plane plane, int x, int y, uint32_t& s0, uint32_t& s1, uint32_t key) {
  if (wnorm < 1e-6f) return 2.0f;
  const int n_src = a.n_src;
  float cdf[kMaxSrc];
  float cum = 0.0f;
#pragma unroll
  for (int i = 0; i < kMaxSrc; ++i) {
    if (i < n_src) cum += vw[i];
    cdf[i] = cum;
  }
  const float gw = a.geom_weight;
  const bool use_geom = gw > 0.0f && a.prev_depth_mask != 0u;
  float total = 0.0f;
  for (int smp = 0; smp < kMcSamples; ++smp) {
    const float u = rand_float(s0, s1, key + static_cast<uint32_t>(smp) * 7919u);
    const float thr = u * cum;
    int v = n_src - 1;
#pragma unroll
    for (int i = kMaxSrc - 1; i >= 0; --i)
      if (i < n_src && cdf[i] > thr) v = i;
    const float photo = cv[v * kWave];
    float geom = 0.0f;
    if (use_geom && (a.prev_depth_mask & (1u << v))) geom = geom_cost(a, v, plane, x, y);
    total += photo + gw * geom;
  }
  return total / static_cast<float>(kMcSamples);
}

// opencl_kernels.h:871-900
__device__ __forceinline__ float weighted_cost_geom(const PmArgs& a, const float* cv, const float (&vw)[kMaxSrc],
                                                    float wnorm, Plane plane, int x, int y) {
  if (wnorm < 1e-6f) return 2.0f;
  float total = 0.0f;
  const float gw = a.geom_weight;
#pragma unroll
  for (int i = 0; i < kMaxSrc; ++i) {
    if (i < a.n_src && vw[i] > 0.0f) {
      const float photo = cv[i * kWave];
      float geom = 0.0f;
      if (gw > 0.0f && (a.prev_depth_mask & (1u << i))) geom = geom_cost(a, i, plane, x, y);
      total += vw[i] * (photo + gw * geom);
    }
  }
  return total / wnorm;
}

// opencl_kernels.h:565-621 on an LDS column: cv (original) and tmp (sorted).
__device__ __forceinline__ float initial_cost_and_views(const PmArgs& a, const float* cv, float* tmp,
                                                        uint32_t* out_sel) {
  const int n_src = a.n_src;
  int num_valid = 0;
  for (int i = 0; i < n_src; ++i) {
    tmp[i * kWave] = cv[i * kWave];
    if (cv[i * kWave] < 2.0f) ++num_valid;
  }
  for (int i = 1; i < n_src; ++i) {
    const float t = tmp[i * kWave];
    int j = i;
    for (; j >= 1 && t < tmp[(j - 1) * kWave]; --j) tmp[j * kWave] = tmp[(j - 1) * kWave];
    tmp[j * kWave] = t;
  }
  *out_sel = 0u;
  const int k = min(min(a.top_k, num_valid), n_src);
  if (k <= 0) return 2.0f;
  float total = 0.0f;
  for (int i = 0; i < k; ++i) total += tmp[i * kWave];
  const float thr = tmp[(k - 1) * kWave];
  uint32_t s = 0u;
  for (int i = 0; i < n_src; ++i)
    if (cv[i * kWave] <= thr) s |= (1u << i);
  *out_sel = s;
  return total / static_cast<float>(k);
}

template <typename T>
__device__ __forceinline__ const T* images_of(const PmArgs& a) { return static_cast<const T*>(a.images); }

// ---------------------------------------------------------------------------
// acmmp_random_init (opencl_kernels.h:979-1039)
// ---------------------------------------------------------------------------
template <int HP, typename T>
__global__ __launch_bounds__(kWave) void init_kernel(PmArgs a) {
  extern __shared__ float smem[];
  const int lane = threadIdx.x;
  const int x = blockIdx.x * kTileX + (lane % kTileX);
  const int y = blockIdx.y * kTileY + (lane / kTileX);
  if (x >= a.width || y >= a.height) return;
  const T* img = images_of<T>(a);
  const int idx = y * a.width + x;
  uint32_t s0 = static_cast<uint32_t>(idx) * 1099087573u + 2654435769u;
  uint32_t s1 = static_cast<uint32_t>(idx) * 2246822519u + 13u;
  const uint32_t key = static_cast<uint32_t>(idx + 7919);
  const float ldmin = pm_logf(fmaxf(a.depth_min, 1e-6f));
  const float ldmax = pm_logf(fmaxf(a.depth_max, 1e-5f));
  const float depth = pm_expf(ldmin + rand_float(s0, s1, key) * (ldmax - ldmin));
  V3 n;
  n.x = rand_float(s0, s1, key + 3u) * 2.0f - 1.0f;
  n.y = rand_float(s0, s1, key + 5u) * 2.0f - 1.0f;
  n.z = -1.0f;
  n = normalize3(n);
  const Plane plane = plane_from_depth_normal(static_cast<float>(x), static_cast<float>(y), depth, n, a.cams[0]);
  a.planes[idx] = to_f4(plane);
  a.rng[idx] = make_uint2(s0, s1);
  RefPatch<HP> rp;
  compute_ref_patch<HP>(img, a.views[0], x, y, a.spatial_factor, a.color_factor, rp);
  float* cv = smem + lane;
  float* tmp = smem + a.n_src * kWave + lane;
  cost_vector<HP>(a, img, x, y, plane, rp, cv, kWave);
  uint32_t sel = 0u;
  a.costs[idx] = initial_cost_and_views(a, cv, tmp, &sel);
  a.sel[idx] = sel;
}

// ---------------------------------------------------------------------------
// acmmp_prior_reinit (opencl_kernels.h:1056-1182); neighbour planes of the
// smoothness term come from the pre-launch snapshot, as in the reference.
// ---------------------------------------------------------------------------
template <int HP, typename T>
__global__ __launch_bounds__(kWave) void prior_kernel(PmArgs a) {
  extern __shared__ float smem[];
  const int lane = threadIdx.x;
  const int x = blockIdx.x * kTileX + (lane % kTileX);
  const int y = blockIdx.y * kTileY + (lane / kTileX);
  if (x >= a.width || y >= a.height) return;
  const int idx = y * a.width + x;
  if (a.plane_masks[idx] == 0u) return;
  const T* img = images_of<T>(a);
  const PmCam& c0 = a.cams[0];
  const Plane pp = to_plane(a.prior_planes[idx]);
  const float dprior = depth_from_plane(static_cast<float>(x), static_cast<float>(y), pp, c0);
  if (dprior <= 0.0f || dprior < a.depth_min || dprior > a.depth_max) return;
  uint2 st = a.rng[idx];
  uint32_t s0 = st.x, s1 = st.y;
  const uint32_t key = static_cast<uint32_t>(idx + 31337u);
  float dp = dprior * 0.99f + rand_float(s0, s1, key) * dprior * 0.02f;
  dp = clampf(dp, a.depth_min, a.depth_max);
  const V3 pn = normalize3(V3{pp.x, pp.y, pp.z});
  const Plane plane = plane_from_depth_normal(static_cast<float>(x), static_cast<float>(y), dp, pn, c0);
  const int n_src = a.n_src;
  const uint32_t cur_sel = a.sel[idx];
  float vw[kMaxSrc];
  float wnorm = 0.0f;
#pragma unroll
  for (int i = 0; i < kMaxSrc; ++i) {
    if (i < n_src && (cur_sel & (1u << i))) { vw[i] = 1.0f; wnorm += 1.0f; } else { vw[i] = 0.0f; }
  }
  RefPatch<HP> rp;
  compute_ref_patch<HP>(img, a.views[0], x, y, a.spatial_factor, a.color_factor, rp);
  Nbrs nb;
  load_nbrs(a, a.snapshot, x, y, nb);
  const float sw = a.smooth_w;
  float* cv = smem + lane;
  float* tmp = smem + n_src * kWave + lane;
  cost_vector<HP>(a, img, x, y, plane, rp, cv, kWave);
  if (wnorm < 1e-6f) {
    uint32_t sel = 0u;
    float cost = initial_cost_and_views(a, cv, tmp, &sel);
    if (sw > 1e-6f) cost += sw * smoothness_cost(a, nb, plane, x, y);
    if (cost < a.costs[idx]) {
      a.planes[idx] = to_f4(plane);
      a.costs[idx] = cost;
      a.sel[idx] = sel;
    }
  } else {
    float cost = weighted_cost_geom(a, cv, vw, wnorm, plane, x, y);
    if (sw > 1e-6f) cost += sw * smoothness_cost(a, nb, plane, x, y);
    if (cost < a.costs[idx]) {
      a.planes[idx] = to_f4(plane);
      a.costs[idx] = cost;
    }
  }
  a.rng[idx] = make_uint2(s0, s1);
}

// ---------------------------------------------------------------------------
// acmmp_patchmatch, one colour (opencl_kernels.h:1197-1778)
//
// The 15 plane hypotheses of a pixel (8 propagated candidates, the current
// plane, up to 6 refinements) run through ONE loop with one inlined copy of
// the cost function, which keeps the kernel small enough for the
// instruction cache.  Evaluating the current plane's cost vector together
// with the candidates is equivalent to the reference order: cost vectors
// consume no random numbers, so the RNG stream (candidate MC draws, then
// the current plane's, then the refinement draws) is unchanged.
// ---------------------------------------------------------------------------
// Occupancy target of the colour-pass kernel (waves per SIMD); the register
// allocator trims hoisting to fit.
#ifndef PM_SWEEP_WAVES_PER_EU
#define PM_SWEEP_WAVES_PER_EU 1
#endif

template <int HP, typename T>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(PM_SWEEP_WAVES_PER_EU)))
void sweep_kernel(PmArgs a) {
  extern __shared__ float smem[];
  const int lane = threadIdx.x;
  const int W = a.width, H = a.height;
  const int gx = blockIdx.x * kTileX + (lane % kTileX);
  const int y = blockIdx.y * kTileY + (lane / kTileX);
  const int half_width = (W + 1) / 2;
  if (gx >= half_width || y >= H) return;
  const int x = gx * 2 + ((y + a.color) & 1);
  if (x >= W) return;
  const T* img = images_of<T>(a);
  const PmCam& c0 = a.cams[0];
  const int n_src = a.n_src;
  const int idx = y * W + x;
  const uint2 st = gp(a.rng)[idx];
  uint32_t s0 = st.x, s1 = st.y;
  const uint32_t key = static_cast<uint32_t>(idx + a.iteration * 7919 + 42);
  // Other-colour state is read only; this pixel's own entries are written
  // once at the end.
  const PM_GLOBAL float4* planes = gp(a.planes);
  const PM_GLOBAL float* costs = gp(a.costs);
  float* ca = smem + lane;  // ca[(slot * n_src + view) * kWave]
  int* cpos = reinterpret_cast<int*>(smem + kSweepSlots * n_src * kWave) + lane;  // cpos[slot * kWave]

  RefPatch<HP> rp;
  compute_ref_patch<HP>(img, a.views[0], x, y, a.spatial_factor, a.color_factor, rp);

  // Step 1: candidate positions (:1254-1430), cheapest-cost probes on the
  // other colour.  Slot layout 0=near up, 1=far up, 2=near down, 3=far down,
  // 4=near left, 5=far left, 6=near right, 7=far right.
  uint32_t flags = 0u;
  auto cost_at = [&](int px, int py) { return costs[py * W + px]; };
  auto put = [&](int slot, int px, int py) {
    cpos[slot * kWave] = py * W + px;
    flags |= 1u << slot;
  };
  if (y > 2) {
    int by = y - 3; float bc = cost_at(x, by);
    for (int i = 1; i < 11; ++i) { const int yy = y - 3 - 2 * i; if (yy < 0) break;
      const float c = cost_at(x, yy); if (c < bc) { bc = c; by = yy; } }
    put(1, x, by);
  }
  if (y < H - 3) {
    int by = y + 3; float bc = cost_at(x, by);
    for (int i = 1; i < 11; ++i) { const int yy = y + 3 + 2 * i; if (yy >= H) break;
      const float c = cost_at(x, yy); if (c < bc) { bc = c; by = yy; } }
    put(3, x, by);
  }
  if (x > 2) {
    int bx = x - 3; float bc = cost_at(bx, y);
    for (int i = 1; i < 11; ++i) { const int xx = x - 3 - 2 * i; if (xx < 0) break;
      const float c = cost_at(xx, y); if (c < bc) { bc = c; bx = xx; } }
    put(5, bx, y);
  }
  if (x < W - 3) {
    int bx = x + 3; float bc = cost_at(bx, y);
    for (int i = 1; i < 11; ++i) { const int xx = x + 3 + 2 * i; if (xx >= W) break;
      const float c = cost_at(xx, y); if (c < bc) { bc = c; bx = xx; } }
    put(7, bx, y);
  }
  if (y > 0) {
    int bx = x, by = y - 1; float bc = cost_at(bx, by);
    for (int i = 0; i < 3; ++i) {
      if (y > 1 + i && x > i) { const int px = x - i - 1, py = y - 2 - i; const float c = cost_at(px, py);
        if (c < bc) { bc = c; bx = px; by = py; } }
      if (y > 1 + i && x < W - 1 - i) { const int px = x + i + 1, py = y - 2 - i; const float c = cost_at(px, py);
        if (c < bc) { bc = c; bx = px; by = py; } }
    }
    put(0, bx, by);
  }
  if (y < H - 1) {
    int bx = x, by = y + 1; float bc = cost_at(bx, by);
    for (int i = 0; i < 3; ++i) {
      if (y < H - 2 - i && x > i) { const int px = x - i - 1, py = y + 2 + i; const float c = cost_at(px, py);
        if (c < bc) { bc = c; bx = px; by = py; } }
      if (y < H - 2 - i && x < W - 1 - i) { const int px = x + i + 1, py = y + 2 + i; const float c = cost_at(px, py);
        if (c < bc) { bc = c; bx = px; by = py; } }
    }
    put(2, bx, by);
  }
  if (x > 0) {
    int bx = x - 1, by = y; float bc = cost_at(bx, by);
    for (int i = 0; i < 3; ++i) {
      if (x > 1 + i && y > i) { const int px = x - 2 - i, py = y - i - 1; const float c = cost_at(px, py);
        if (c < bc) { bc = c; bx = px; by = py; } }
      if (x > 1 + i && y < H - 1 - i) { const int px = x - 2 - i, py = y + i + 1; const float c = cost_at(px, py);
        if (c < bc) { bc = c; bx = px; by = py; } }
    }
    put(4, bx, by);
  }
  if (x < W - 1) {
    int bx = x + 1, by = y; float bc = cost_at(bx, by);
    for (int i = 0; i < 3; ++i) {
      if (x < W - 2 - i && y > i) { const int px = x + 2 + i, py = y - i - 1; const float c = cost_at(px, py);
        if (c < bc) { bc = c; bx = px; by = py; } }
      if (x < W - 2 - i && y < H - 1 - i) { const int px = x + 2 + i, py = y + i + 1; const float c = cost_at(px, py);
        if (c < bc) { bc = c; bx = px; by = py; } }
    }
    put(6, bx, by);
  }

  Nbrs nb;
  load_nbrs(a, a.planes, x, y, nb);
  const float smooth_w = a.smooth_w;
  const uint32_t prev_sel = gp(a.sel)[idx];
  Plane plane_now = to_plane(planes[idx]);
  float cost_now = 2.0f, depth_now = 0.0f;
  uint32_t sel_out = prev_sel;

  float vw[kMaxSrc];
#pragma unroll
  for (int i = 0; i < kMaxSrc; ++i) vw[i] = 0.0f;
  float wnorm = 0.0f;
  // refinement hypotheses (filled after the current plane is scored)
  float depth_rand = 0.0f, depth_pert = 0.0f, depth0 = 0.0f;
  V3 cur_normal{0, 0, 0}, nrand{0, 0, 0}, npert{0, 0, 0}, sn{0, 0, 0};
  int n_ref = 0;

#pragma unroll 1
  for (int h = 0; h < 9 + 6; ++h) {
    Plane hyp;
    bool active;
    int slot;
    if (h < 8) {
      active = (flags >> h) & 1u;
      slot = h;
      hyp = active ? to_plane(planes[cpos[h * kWave]]) : Plane{0, 0, 0, 0};
    } else if (h == 8) {
      active = true;
      slot = 8;
      hyp = plane_now;
    } else {
      const int r = h - 9;
      active = r < n_ref;
      slot = 8;
      float rd;
      V3 rn;
      switch (r) {  // :1733-1741
        case 0: rd = depth_rand; rn = cur_normal; break;
        case 1: rd = depth0; rn = nrand; break;
        case 2: rd = depth_rand; rn = nrand; break;
        case 3: rd = depth0; rn = npert; break;
        case 4: rd = depth_pert; rn = cur_normal; break;
        default: rd = depth0; rn = sn; break;
      }
      hyp = plane_from_depth_normal(static_cast<float>(x), static_cast<float>(y), rd, rn, c0);
    }
    if (active) cost_vector<HP>(a, img, x, y, hyp, rp, ca + slot * n_src * kWave, kWave);

    if (h == 8) {
      // Step 2: view weights (:1452-1563).
      const float vthr = 0.8f;
#pragma unroll
      for (int i = 0; i < kMaxSrc; ++i) {
        if (i < n_src) {
          float vc = 0.0f;
          for (int j = 0; j < 8; ++j) {
            const float c = ca[(j * n_src + i) * kWave];
            if (((flags >> j) & 1u) && c < vthr) vc += (1.0f - c);
          }
          if (vc > 0.0f) vw[i] = vc;
        }
      }
      if (a.anchor_views > 0) {
        int anchored = 0;
#pragma unroll
        for (int i = 0; i < kMaxSrc; ++i) {
          if (i < n_src && anchored < a.anchor_views) {
            if ((prev_sel & (1u << i)) && vw[i] < 1e-6f) { vw[i] = 1.0f; ++anchored; }
          }
        }
      }
      {
        V3 cur_n{plane_now.x, plane_now.y, plane_now.z};
        const float nl = len3(cur_n);
        const bool have_normal = nl > 0.1f;
        if (have_normal) cur_n = V3{cur_n.x / nl, cur_n.y / nl, cur_n.z / nl};
        const float cur_depth = depth_from_plane(static_cast<float>(x), static_cast<float>(y), plane_now, c0);
        const V3 p3 = backproject(static_cast<float>(x), static_cast<float>(y), cur_depth, c0);
        const float inc_sigma = 0.6f;
        const float inv2s2 = -0.5f / (inc_sigma * inc_sigma);
#pragma unroll
        for (int i = 0; i < kMaxSrc; ++i) {
          if (i >= n_src || vw[i] < 1e-6f) continue;
          const PmView& v = a.views[i + 1];
          vw[i] *= v.angle_w;
          if (have_normal && cur_depth > 0.0f) {
            const V3 ray = normalize3(V3{v.cx - p3.x, v.cy - p3.y, v.cz - p3.z});
            const float cos_inc = fabsf(dot3(cur_n, ray));
            const float xi = 1.0f - cos_inc;
            vw[i] *= pm_expf(xi * xi * inv2s2);
          }
        }
      }
      uint32_t new_sel = 0u;
#pragma unroll
      for (int i = 0; i < kMaxSrc; ++i)
        if (i < n_src && vw[i] > 0.0f) { new_sel |= (1u << i); wnorm += vw[i]; }
      if (wnorm < 1e-6f) {
#pragma unroll
        for (int i = 0; i < kMaxSrc; ++i) {
          if (i >= n_src) continue;
          for (int j = 0; j < 8; ++j) {
            if (((flags >> j) & 1u) && ca[(j * n_src + i) * kWave] < 2.0f) {
              vw[i] = 1.0f; new_sel |= (1u << i); wnorm += 1.0f; break;
            }
          }
        }
      }

      // Step 3: Monte-Carlo costs of the candidates, then of the current
      // plane (:1570-1619).
      float best_cost = 2.0f;
      int best = 0;
#pragma unroll 1
      for (int i = 0; i < 9; ++i) {
        const bool flagged = i == 8 || ((flags >> i) & 1u);
        float fc = 2.0f;
        if (flagged) {
          const Plane cp = i == 8 ? plane_now : to_plane(planes[cpos[i * kWave]]);
          const uint32_t k2 = i == 8 ? key + 700u : key + 600u + static_cast<uint32_t>(i) * 31u;
          fc = mc_cost(a, ca + i * n_src * kWave, vw, wnorm, cp, x, y, s0, s1, k2);
          if (smooth_w > 1e-6f) fc += smooth_w * smoothness_cost(a, nb, cp, x, y);
        }
        if (i == 8) {
          cost_now = fc;
        } else if (i == 0) {
          best_cost = fc;
        } else if (fc < best_cost) {
          best_cost = fc;
          best = i;
        }
      }
      depth_now = depth_from_plane(static_cast<float>(x), static_cast<float>(y), plane_now, c0);

      const bool have_prior = a.plane_masks != nullptr && gp(a.plane_masks)[idx] > 0u;
      float depth_prior = 0.0f;
      V3 prior_n{0.0f, 0.0f, -1.0f};
      const float depth_sigma = (a.depth_max - a.depth_min) / 32.0f;
      const float angle_sigma = kPiF * (5.0f / 180.0f);
      if (have_prior) {
        const Plane pp = to_plane(gp(a.prior_planes)[idx]);
        depth_prior = depth_from_plane(static_cast<float>(x), static_cast<float>(y), pp, c0);
        prior_n = normalize3(V3{pp.x, pp.y, pp.z});
      }

      // Step 3b (:1643-1656).
      if (((flags >> best) & 1u) && best_cost < cost_now) {
        const Plane bp = to_plane(planes[cpos[best * kWave]]);
        const float db = depth_from_plane(static_cast<float>(x), static_cast<float>(y), bp, c0);
        if (db >= a.depth_min && db <= a.depth_max) {
          depth_now = db;
          plane_now = bp;
          cost_now = best_cost;
          sel_out = new_sel;
        }
      }

      // Step 4 set-up: the refinement hypotheses (:1669-1743).
      const float depth_perturbation = 0.005f;
      const float normal_perturbation = 0.01f * kPiF;
      cur_normal = normalize3(V3{plane_now.x, plane_now.y, plane_now.z});
      if (have_prior && depth_prior > 0.0f) {
        depth_rand = depth_prior - 3.0f * depth_sigma + rand_float(s0, s1, key + 503u) * 6.0f * depth_sigma;
        depth_rand = clampf(depth_rand, a.depth_min, a.depth_max);
        const float pa1 = (rand_float(s0, s1, key + 504u) - 0.5f) * angle_sigma;
        const float pa2 = (rand_float(s0, s1, key + 505u) - 0.5f) * angle_sigma;
        nrand = normalize3(V3{prior_n.x + pa1, prior_n.y + pa2, prior_n.z});
        if (nrand.z > 0.0f) nrand.z = -nrand.z;
      } else {
        const float ldmin = pm_logf(fmaxf(a.depth_min, 1e-6f));
        const float ldmax = pm_logf(fmaxf(a.depth_max, 1e-5f));
        depth_rand = pm_expf(ldmin + rand_float(s0, s1, key + 503u) * (ldmax - ldmin));
        nrand.x = rand_float(s0, s1, key + 504u) * 2.0f - 1.0f;
        nrand.y = rand_float(s0, s1, key + 505u) * 2.0f - 1.0f;
        nrand.z = -1.0f;
        nrand = normalize3(nrand);
      }
      {
        const float dlo = (1.0f - depth_perturbation) * depth_now;
        const float dhi = (1.0f + depth_perturbation) * depth_now;
        depth_pert = dlo + rand_float(s0, s1, key + 510u) * (dhi - dlo);
        depth_pert = clampf(depth_pert, a.depth_min, a.depth_max);
      }
      {
        const float pa1 = (rand_float(s0, s1, key + 520u) - 0.5f) * normal_perturbation;
        const float pa2 = (rand_float(s0, s1, key + 521u) - 0.5f) * normal_perturbation;
        npert = normalize3(V3{cur_normal.x + pa1, cur_normal.y + pa2, cur_normal.z});
        if (npert.z > 0.0f) npert.z = -npert.z;
      }
      sn = surface_normal(a, nb, x, y, depth_now);
      n_ref = len3(sn) > 0.5f ? 6 : 5;
      depth0 = depth_now;
    } else if (h > 8 && active) {
      // Step 4: score one refinement hypothesis (:1744-1771).
      float tc = mc_cost(a, ca + 8 * n_src * kWave, vw, wnorm, hyp, x, y, s0, s1,
                         key + 800u + static_cast<uint32_t>(h - 9) * 37u);
      if (smooth_w > 1e-6f) tc += smooth_w * smoothness_cost(a, nb, hyp, x, y);
      const float db = depth_from_plane(static_cast<float>(x), static_cast<float>(y), hyp, c0);
      if (!(db < a.depth_min || db > a.depth_max) && tc < cost_now) {
        depth_now = db;
        plane_now = hyp;
        cost_now = tc;
      }
    }
  }
  gpw(a.planes)[idx] = to_f4(plane_now);
  gpw(a.costs)[idx] = cost_now;
  gpw(a.rng)[idx] = make_uint2(s0, s1);
  gpw(a.sel)[idx] = sel_out;
}

// ---------------------------------------------------------------------------
// acmmp_checkerboard_filter, one colour (opencl_kernels.h:1883-2039)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kWave) void filter_kernel(PmArgs a) {
  const int lane = threadIdx.x;
  const int W = a.width, H = a.height;
  const int gx = blockIdx.x * kTileX + (lane % kTileX);
  const int y = blockIdx.y * kTileY + (lane / kTileX);
  const int half_width = (W + 1) / 2;
  if (gx >= half_width || y >= H) return;
  const int x = gx * 2 + ((y + a.color) & 1);
  if (x >= W) return;
  const int center = y * W + x;
  if (a.costs[center] < 0.001f) return;
  const PmCam& c0 = a.cams[0];
  float buf[21];
  int n = 0;
  auto dep = [&](int px, int py) {
    return depth_from_plane(static_cast<float>(px), static_cast<float>(py), to_plane(a.planes[py * W + px]), c0);
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
    const float percent = 0.03f;
    const float cd = buf[0];
    if (cd > 0.0f) {
      const float lo = cd * (1 - percent);
      const float hi = cd * (1 + percent);
      int m = 1;
      for (int i = 1; i < n; ++i)
        if (buf[i] > lo && buf[i] < hi) buf[m++] = buf[i];
      n = m;
    }
  }
  for (int i = 1; i < n; ++i) {
    const float tmp = buf[i];
    int j;
    for (j = i; j >= 1 && tmp < buf[j - 1]; --j) buf[j] = buf[j - 1];
    buf[j] = tmp;
  }
  const int mid = n / 2;
  const float med = (n % 2 == 0) ? (buf[mid - 1] + buf[mid]) * 0.5f : buf[mid];
  if (med > 0.0f) {
    const Plane cp = to_plane(a.planes[center]);
    const V3 nn = normalize3(V3{cp.x, cp.y, cp.z});
    a.planes[center] = to_f4(plane_from_depth_normal(static_cast<float>(x), static_cast<float>(y), med, nn, c0));
  }
}

// ---------------------------------------------------------------------------
// Decode (estimator.cc:1112-1224; median, speckle and gap filling off).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void decode_kernel(PmArgs a, PmDecodeOut o) {
  const int W = a.width, H = a.height;
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<int64_t>(W) * H) return;
  const int x = static_cast<int>(idx % W), y = static_cast<int>(idx / W);
  const PmCam& c0 = a.cams[0];
  const float4 p = a.planes[idx];
  float depth = 0.0f, nx = 0.0f, ny = 0.0f, nz = 0.0f, cost = 0.0f, conf = 0.0f;
  if (isfinite(p.x) && isfinite(p.y) && isfinite(p.z) && isfinite(p.w)) {
    const float rx = (static_cast<float>(x) - c0.K[2]) / c0.K[0];
    const float ry = (static_cast<float>(y) - c0.K[5]) / c0.K[4];
    const float denom = p.x * rx + p.y * ry + p.z * 1.0f;
    float d = (fabsf(denom) > 1e-6f) ? -p.w / denom : 0.0f;
    if (!isfinite(d)) d = 0.0f;
    if (d > 0.0f && d >= a.depth_min && d <= a.depth_max) {
      depth = d;
      const float len = sqrtf(p.x * p.x + p.y * p.y + p.z * p.z);
      if (len > 1e-6f) { nx = p.x / len; ny = p.y / len; nz = p.z / len; }
    }
    const float cc = a.costs[idx];
    cost = isfinite(cc) ? cc : 0.0f;
    if (depth > 0.0f) conf = 1.0f / (1.0f + pm_expf((cost - 0.5f) * 5.0f));
  }
  if (o.depth) o.depth[idx] = depth;
  if (o.normal) { o.normal[3 * idx] = nx; o.normal[3 * idx + 1] = ny; o.normal[3 * idx + 2] = nz; }
  if (o.cost) o.cost[idx] = cost;
  if (o.confidence) o.confidence[idx] = conf;
}

// ---------------------------------------------------------------------------
// Score caller-given (pixel, plane) pairs.
// ---------------------------------------------------------------------------
template <int HP, typename T>
__global__ __launch_bounds__(kWave) void score_kernel(PmArgs a, int n, const int32_t* __restrict__ xy,
                                                      const float4* __restrict__ pl, float* __restrict__ costs,
                                                      float* __restrict__ init_cost, uint32_t* __restrict__ sel) {
  extern __shared__ float smem[];
  const int lane = threadIdx.x;
  const int i = blockIdx.x * kWave + lane;
  if (i >= n) return;
  const T* img = images_of<T>(a);
  const int x = xy[2 * i], y = xy[2 * i + 1];
  RefPatch<HP> rp;
  compute_ref_patch<HP>(img, a.views[0], x, y, a.spatial_factor, a.color_factor, rp);
  float* cv = smem + lane;
  float* tmp = smem + a.n_src * kWave + lane;
  cost_vector<HP>(a, img, x, y, to_plane(pl[i]), rp, cv, kWave);
  for (int s = 0; s < a.n_src; ++s) costs[static_cast<int64_t>(i) * a.n_src + s] = cv[s * kWave];
  uint32_t sv = 0u;
  const float ic = initial_cost_and_views(a, cv, tmp, &sv);
  if (init_cost) init_cost[i] = ic;
  if (sel) sel[i] = sv;
}

dim3 full_grid(const PmArgs& a) {
  return dim3((a.width + kTileX - 1) / kTileX, (a.height + kTileY - 1) / kTileY);
}
dim3 colour_grid(const PmArgs& a) {
  const int hw = (a.width + 1) / 2;
  return dim3((hw + kTileX - 1) / kTileX, (a.height + kTileY - 1) / kTileY);
}

}  // namespace

#define PM_DISPATCH(KERNEL, HP, F16, GRID, LDS, STREAM, ...)                                         \
  do {                                                                                            \
    if (F16) {                                                                                    \
      switch (HP) {                                                                               \
        case 1: hipLaunchKernelGGL((KERNEL<1, _Float16>), GRID, dim3(kWave), LDS, STREAM, __VA_ARGS__); break; \
        case 2: hipLaunchKernelGGL((KERNEL<2, _Float16>), GRID, dim3(kWave), LDS, STREAM, __VA_ARGS__); break; \
        case 3: hipLaunchKernelGGL((KERNEL<3, _Float16>), GRID, dim3(kWave), LDS, STREAM, __VA_ARGS__); break; \
        default: return hipErrorInvalidValue;                                                     \
      }                                                                                           \
    } else {                                                                                      \
      switch (HP) {                                                                               \
        case 1: hipLaunchKernelGGL((KERNEL<1, float>), GRID, dim3(kWave), LDS, STREAM, __VA_ARGS__); break; \
        case 2: hipLaunchKernelGGL((KERNEL<2, float>), GRID, dim3(kWave), LDS, STREAM, __VA_ARGS__); break; \
        case 3: hipLaunchKernelGGL((KERNEL<3, float>), GRID, dim3(kWave), LDS, STREAM, __VA_ARGS__); break; \
        default: return hipErrorInvalidValue;                                                     \
      }                                                                                           \
    }                                                                                             \
    return hipGetLastError();                                                                     \
  } while (0)

hipError_t launch_init(const PmArgs& a, int hp, bool f16, hipStream_t stream) {
  const size_t lds = sizeof(float) * 2 * a.n_src * kWave;
  PM_DISPATCH(init_kernel, hp, f16, full_grid(a), lds, stream, a);
}

hipError_t launch_prior(const PmArgs& a, int hp, bool f16, hipStream_t stream) {
  const size_t lds = sizeof(float) * 2 * a.n_src * kWave;
  PM_DISPATCH(prior_kernel, hp, f16, full_grid(a), lds, stream, a);
}

hipError_t launch_sweep(const PmArgs& a, int hp, bool f16, hipStream_t stream) {
  const size_t lds = sizeof(float) * (kSweepSlots * a.n_src + 8) * kWave;
  PM_DISPATCH(sweep_kernel, hp, f16, colour_grid(a), lds, stream, a);
}

hipError_t launch_filter(const PmArgs& a, hipStream_t stream) {
  hipLaunchKernelGGL(filter_kernel, colour_grid(a), dim3(kWave), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_decode(const PmArgs& a, const PmDecodeOut& o, hipStream_t stream) {
  const int64_t n = static_cast<int64_t>(a.width) * a.height;
  hipLaunchKernelGGL(decode_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, stream, a, o);
  return hipGetLastError();
}

hipError_t launch_score(const PmArgs& a, int hp, bool f16, int n, const int32_t* xy, const float4* planes,
                        float* costs, float* init_cost, uint32_t* sel, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  const size_t lds = sizeof(float) * 2 * a.n_src * kWave;
  const dim3 grid((n + kWave - 1) / kWave);
  PM_DISPATCH(score_kernel, hp, f16, grid, lds, stream, a, n, xy, planes, costs, init_cost, sel);
}

}  // namespace pmhip
