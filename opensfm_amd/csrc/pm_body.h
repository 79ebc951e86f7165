// Device code of the PatchMatch kernels for gfx950 (MI355X): every kernel
// is a body function of (arguments, workgroup index), instantiated by the
// single-reference and the batched wrappers at the end of this file.
//
// One lane per reference pixel; the 64 lanes of a wave cover a TX x (64/TX)
// tile of same-colour pixels (2*TX x 64/TX image pixels), so the warped source patches
// of neighbouring lanes overlap and the bilinear gathers of a wave hit the
// same cache lines.  The bilateral reference patch (w, w*r per tap) lives
// in VGPRs for the whole pass (patch size is a template parameter, loops
// are fully unrolled); per-candidate cost vectors live in LDS, one column
// per lane, so the Monte-Carlo view sampling can index them at random.
// Colour passes update the state in place: every neighbour a pixel reads
// has the other colour (odd Manhattan offsets, opencl_kernels.h:1262-1430,
// 664-668, 740-746), which makes an in-place pass equal to the reference's
// snapshot pass.
#pragma once
#include <type_traits>

#include "pm_kernels.h"

namespace pmhip {

namespace {

constexpr int kWave = 64;
constexpr int kFullTileX = 16;   // full-resolution tiles (refpatch, init, prior): 16 x 4 pixels
constexpr int kFullTileY = kWave / kFullTileX;
constexpr int kSweepSlots = 9;   // 8 propagation candidates + current/refinement
constexpr int kXcds = 8;         // MI355X: 8 XCDs, each with its own L2

// Tile of this workgroup.  The dispatcher hands workgroups to the XCDs
// round-robin (linear id mod 8), so consecutive ids land on different L2s.
// The tiles are ordered in vertical strips of about kBandStripPx pixels,
// each strip top to bottom and row by row, strip after strip; every XCD
// takes one contiguous eighth of that order (equal tile counts), so the
// tiles in flight on one XCD are a few rows of one strip.  Their gathers then
// stay within that strip's epipolar segments whichever way a view's baseline
// runs, and neighbouring tiles still share cache lines of the
// reference-side data.  The grid is 1-D with kXcds * ceil(tiles / kXcds)
// workgroups; b is the workgroup's index in the grid of its reference (a
// multiple of kXcds apart from blockIdx.x, so b mod kXcds is still the XCD).
#ifndef PM_BAND_STRIP_PX
#define PM_BAND_STRIP_PX 64
#endif
#ifndef PM_BAND_SLAB_PX
#define PM_BAND_SLAB_PX 384
#endif
constexpr int kBandStripPx = PM_BAND_STRIP_PX;  // 0: whole rows
constexpr int kBandSlabPx = PM_BAND_SLAB_PX;    // 0: the whole height

// Tile t of the strip order: horizontal slabs of `slab` tile rows (0: the
// whole height) one after the other; inside a slab, vertical strips of
// `strip` tiles (0: whole rows) left to right; inside a strip, row by row.
__device__ __forceinline__ void strip_tile(int tiles_x, int tiles_y, int t, int strip, int slab, int& tx, int& ty) {
  const int sh = slab > 0 ? min(slab, tiles_y) : tiles_y;
  const int k = t / (sh * tiles_x);
  const int rows = min(sh, tiles_y - k * sh);  // the last slab may be shorter
  const int r0 = t - k * sh * tiles_x;
  const int sw = strip > 0 ? min(strip, tiles_x) : tiles_x;
  const int s = r0 / (sw * rows);
  const int rem = r0 - s * sw * rows;
  const int ws = min(sw, tiles_x - s * sw);  // the last strip may be narrower
  ty = k * sh + rem / ws;
  tx = s * sw + rem % ws;
}

__device__ __forceinline__ bool tile_of(int tiles_x, int tiles_y, int b, int strip, int slab, int& tx, int& ty) {
  const int ntiles = tiles_x * tiles_y;
  const int per = (ntiles + kXcds - 1) / kXcds;
  const int t = (b % kXcds) * per + b / kXcds;
  if (t >= ntiles) return false;
  strip_tile(tiles_x, tiles_y, t, strip, slab, tx, ty);
  return true;
}

__device__ __forceinline__ bool full_tile_pixel(int width, int height, int b, int& x, int& y) {
  int tx, ty;
  if (!tile_of((width + kFullTileX - 1) / kFullTileX, (height + kFullTileY - 1) / kFullTileY, b,
               kBandStripPx / kFullTileX, kBandSlabPx / kFullTileY, tx, ty))
    return false;
  x = tx * kFullTileX + static_cast<int>(threadIdx.x % kFullTileX);
  y = ty * kFullTileY + static_cast<int>(threadIdx.x / kFullTileX);
  return x < width && y < height;
}

// Colour tiles are TX x (kWave / TX) same-colour pixels.
template <int TX>
__device__ __forceinline__ int colour_tiles_x(const PmArgs& a) { return ((a.width + 1) / 2 + TX - 1) / TX; }
template <int TX>
__device__ __forceinline__ int colour_tiles_y(const PmArgs& a) { return (a.height + kWave / TX - 1) / (kWave / TX); }

// Colour tile of workgroup b (false: past the last tile, for the whole
// workgroup).
template <int TX>
__device__ __forceinline__ bool colour_tile(const PmArgs& a, int b, int& tx, int& ty) {
  return tile_of(colour_tiles_x<TX>(a), colour_tiles_y<TX>(a), b, kBandStripPx / (2 * TX), kBandSlabPx / (kWave / TX),
                 tx, ty);
}

__device__ __forceinline__ Plane to_plane(float4 v) { return Plane{v.x, v.y, v.z, v.w}; }
__device__ __forceinline__ float4 to_f4(Plane p) { return make_float4(p.x, p.y, p.z, p.w); }
template <typename T>
__device__ __forceinline__ Sampler make_sampler(const PmArgs& a, const PmView& v) {
  return Sampler{reinterpret_cast<const PM_GLOBAL char*>(gp(static_cast<const char*>(a.quads))) +
                     v.qbase * QuadOf<T>::bytes,
                 static_cast<float>(v.pitch), static_cast<float>(v.pitch + 1), static_cast<float>(v.w - 1),
                 static_cast<float>(v.h - 1)};
}

// Camera / view i for a wave-uniform i (scalar loads from the constant space).
__device__ __forceinline__ PmCam kcam(const PmArgs& a, int i) {
  return ((const PM_CONST PmCam*)(a.cams))[i];
}
__device__ __forceinline__ PmView kview(const PmArgs& a, int i) {
  return ((const PM_CONST PmView*)(a.views))[i];
}
template <typename T>
__device__ __forceinline__ const T* images_of(const PmArgs& a) { return static_cast<const T*>(a.images); }

// ---------------------------------------------------------------------------
// Reference-side patch data.  The bilateral weights w and the products w*r
// of every tap, and the reference moments, depend only on the reference
// image, the pixel and the sampling stride, so they are computed once per
// upload and kept in HBM: the stride-1 taps (392 B per pixel at patch 7)
// and the moments of strides 1, 2 and 3:
//   ref_taps[p * ref_plane + cidx] = (w, w*r)
//   ref_sums[s * ref_plane + cidx] = (centre, sum_w, sum_wr, sum_wrr)   s = stride - 1
// cidx is the colour-major pixel index, so the 64 lanes of a colour-pass
// wave read consecutive entries.  The strided taps are recomputed where
// they are used with this same arithmetic.  Values are exactly those the
// reference recomputes for every hypothesis (opencl_kernels.h:328-366 for
// stride 1, the reference half of :216-244 for strides 2 and 3).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int colour_index(const PmArgs& a, int x, int y) {
  const int hw = (a.width + 1) / 2;
  return ((x + y) & 1) * (a.height * hw) + y * hw + (x >> 1);
}

template <int HP, typename T>
__device__ __forceinline__ void refpatch_body(const PmArgs& a, int b) {
  int x, y;
  if (!full_tile_pixel(a.width, a.height, b, x, y)) return;
  const T* img = images_of<T>(a);
  const PmView rv = kview(a, 0);
  const int cidx = colour_index(a, x, y);
  const float sf = a.spatial_factor, cf = a.color_factor;
  const float center = texel(img, rv, x, y);
  PM_GLOBAL float2* taps = gpw(a.ref_taps);
  PM_GLOBAL float4* sums = gpw(a.ref_sums);
  {  // stride 1 (opencl_kernels.h:340-365)
    float sum_w = 0.0f, sum_wr = 0.0f, sum_wrr = 0.0f;
    int p = 0;
    for (int iy = -HP; iy <= HP; ++iy) {
      for (int ix = -HP; ix <= HP; ++ix) {
        const float r = texel(img, rv, x + ix, y + iy) - center;
        const float w = pm_expf(sf * static_cast<float>(ix * ix + iy * iy) + cf * r * r);
        sum_w += w;
        const float wr = w * r;
        sum_wr += wr;
        sum_wrr = fmaf(wr, r, sum_wrr);
        taps[static_cast<int64_t>(p) * a.ref_plane + cidx] = make_float2(w, wr);
        ++p;
      }
    }
    sums[cidx] = make_float4(center, sum_w, sum_wr, sum_wrr);
  }
  for (int si = 1; si < 3; ++si) {  // strides 2 and 3 (opencl_kernels.h:216-244)
    const float stride = static_cast<float>(si + 1);
    float sum_w = 0.0f, sum_r = 0.0f, sum_rr = 0.0f;
    for (int iy = -HP; iy <= HP; ++iy) {
      for (int ix = -HP; ix <= HP; ++ix) {
        const float fdx = stride * static_cast<float>(ix);
        const float fdy = stride * static_cast<float>(iy);
        const float rx = static_cast<float>(x) + fdx;
        const float ry = static_cast<float>(y) + fdy;
        const float r = texel(img, rv, static_cast<int>(rx), static_cast<int>(ry)) - center;
        const float w = pm_expf(fdx * fdx * sf + fdy * fdy * sf + r * r * cf);
        const float wr = w * r;
        sum_r += wr;
        sum_rr += wr * r;
        sum_w += w;
      }
    }
    sums[static_cast<int64_t>(si) * a.ref_plane + cidx] = make_float4(center, sum_w, sum_r, sum_rr);
  }
}

// Camera / view of a per-lane (divergent) source index: read from the
// device copies instead of the kernel arguments.
__device__ __forceinline__ PmView lane_view(const PmArgs& a, int s) { return gp(a.views)[s + 1]; }
__device__ __forceinline__ PmCam lane_cam(const PmArgs& a, int s) { return gp(a.cams)[s + 1]; }

// Warp of one source view for a plane (opencl_kernels.h:432-461):
// false = invalid (cost 2.0).
struct Warp { float scx, scy, jxx, jxy, jyx, jyy; };

__device__ __forceinline__ bool warp_of(V3 cw, V3 wdx, V3 wdy, const PmCam& cs, Warp& w) {
  const V3 pc = project(cw, cs);
  if (pc.z < 1e-6f) return false;
  w.scx = pc.x / pc.z;
  w.scy = pc.y / pc.z;
  const V3 pdx = project(wdx, cs);
  const V3 pdy = project(wdy, cs);
  if (pdx.z < 1e-6f || pdy.z < 1e-6f) return false;
  w.jxx = pdx.x / pdx.z - w.scx;
  w.jxy = pdx.y / pdx.z - w.scy;
  w.jyx = pdy.x / pdy.z - w.scx;
  w.jyy = pdy.y / pdy.z - w.scy;
  const float det = w.jxx * w.jyy - w.jxy * w.jyx;
  return !(det < 0.02f || det > 50.0f);
}

// The three reference points of a plane at (x, y) (opencl_kernels.h:401-416).
struct PlanePts { V3 cw, wdx, wdy; };

__device__ __forceinline__ bool plane_points(const PmCam& c0, int x, int y, Plane pl, PlanePts& pp) {
  const float fx = static_cast<float>(x), fy = static_cast<float>(y);
  const float depth = depth_from_plane(fx, fy, pl, c0);
  const float depth_dx = depth_from_plane(static_cast<float>(x + 1), fy, pl, c0);
  const float depth_dy = depth_from_plane(fx, static_cast<float>(y + 1), pl, c0);
  if (depth <= 0.0f || depth_dx <= 0.0f || depth_dy <= 0.0f) return false;
  pp.cw = backproject(fx, fy, depth, c0);
  pp.wdx = backproject(static_cast<float>(x + 1), fy, depth_dx, c0);
  pp.wdy = backproject(fx, static_cast<float>(y + 1), depth_dy, c0);
  return true;
}

// Stride-1 per-view costs of one plane (opencl_kernels.h:381-557 without
// the textureless fallback).  Views run in chunks of kViewChunk inside the
// tap loop, so each tap's (w, w*r) is read once per chunk; each view still
// accumulates its taps in row-major order.  With the taps in LDS, one view
// per chunk is fastest on MI355X (fewer live registers, more waves: 21.7 vs
// 31.5 ms per candidates launch at 1600x1200 with 8 sources).  out[s * os] gets the
// cost, or -- for views whose NCC is <= 0.1 -- the stride-1 NCC, with bit s
// set in the returned pending mask for resolve_fallbacks().
#ifndef PM_VIEW_CHUNK
#define PM_VIEW_CHUNK 1
#endif
constexpr int kViewChunk = PM_VIEW_CHUNK;

// True when every tap of a patch of half-size HP (taps spaced `stride`
// apart) warped by w lies in [-1, w-1] x [-1, h-1] of the view, where
// sample() needs no clamp.  The corners bound the affine tap positions; the
// margin covers the rounding of the taps' two FMAs (views are at most 65536
// pixels wide); NaN fails.
template <int HP>
__device__ __forceinline__ bool taps_in_view(const Warp& w, const Sampler& sm, float stride = 1.0f) {
  const float ex = (fabsf(w.jxx) + fabsf(w.jyx)) * (stride * static_cast<float>(HP)) + 0.05f;
  const float ey = (fabsf(w.jxy) + fabsf(w.jyy)) * (stride * static_cast<float>(HP)) + 0.05f;
  return w.scx - ex >= -1.0f && w.scx + ex <= sm.xmax && w.scy - ey >= -1.0f && w.scy + ey <= sm.ymax;
}

// kLdsTaps: the pixel's stride-1 (w, w*r) are read from the workgroup's LDS
// copy (lt[tap * kWave], see stage_tile_taps) instead of ref_taps.
// Only the views [v_lo, v_hi) are scored (all by default).
template <int HP, typename T, bool kLdsTaps = false>
__device__ __forceinline__ uint32_t cost_vector(const PmArgs& a, const T* __restrict__ img, int x, int y, int cidx,
                                                Plane pl, float* out, int64_t os, const float2* lt = nullptr,
                                                int v_lo = 0, int v_hi = kMaxSrc) {
  constexpr int D = 2 * HP + 1;
  const int n_src = min(a.n_src, v_hi);
  PlanePts pp;
  if (!plane_points(kcam(a, 0), x, y, pl, pp)) {
    for (int s = v_lo; s < n_src; ++s) out[s * os] = 2.0f;
    return 0u;
  }
  const float4 rs = gp(a.ref_sums)[cidx];
  const bool ref_ok = rs.y >= 1e-6f;
  float inv_w = 0.0f, mean_r = 0.0f, var_ref = 0.0f;
  if (ref_ok) {
    inv_w = 1.0f / rs.y;
    mean_r = rs.z * inv_w;
    var_ref = rs.w * inv_w - mean_r * mean_r;
  }
  const int64_t plane_stride = a.ref_plane;
  const uint32_t ucidx = static_cast<uint32_t>(cidx);
  uint32_t pending = 0u;
#pragma unroll 1
  for (int c0v = v_lo; c0v < n_src; c0v += kViewChunk) {
    // Views of the chunk past n_src, or with an invalid warp, are sampled
    // anyway (coordinates are clamped, so every access stays inside the
    // view) and their sums discarded: the tap loop stays branch-free and
    // the loads of all views and taps of a row can be in flight together.
    Warp wp[kViewChunk];
    Sampler sm[kViewChunk];
    float scen[kViewChunk], s_ws[kViewChunk], s_wss[kViewChunk], s_wrs[kViewChunk];
    bool ok[kViewChunk];
#pragma unroll
    for (int k = 0; k < kViewChunk; ++k) {
      const int s = c0v + k;
      const int sv = s < n_src ? s : n_src - 1;
      ok[k] = s < n_src && warp_of(pp.cw, pp.wdx, pp.wdy, kcam(a, sv + 1), wp[k]) && ref_ok;
      if (!ok[k]) wp[k] = Warp{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
      sm[k] = make_sampler<T>(a, kview(a, sv + 1));
      scen[k] = sample<T>(sm[k], wp[k].scx, wp[k].scy);
      s_ws[k] = 0.0f; s_wss[k] = 0.0f; s_wrs[k] = 0.0f;
    }
    // The tap loop, with or without the coordinate clamp: when every lane's
    // patch lies inside its view the clamp is the identity and is left out.
    auto taps = [&](auto clamp) {
      constexpr bool kClamp = decltype(clamp)::value;
#pragma unroll 1
      for (int iy = -HP; iy <= HP; ++iy) {
        const float fiy = static_cast<float>(iy);
        float bx[kViewChunk], by[kViewChunk];
#pragma unroll
        for (int k = 0; k < kViewChunk; ++k) {
          bx[k] = fmaf(wp[k].jyx, fiy, wp[k].scx);
          by[k] = fmaf(wp[k].jyy, fiy, wp[k].scy);
        }
        const PM_GLOBAL float2* trow = gp(a.ref_taps) + static_cast<int64_t>((iy + HP) * D) * plane_stride;
#pragma unroll
        for (int ix = -HP; ix <= HP; ++ix) {
          const float fix = static_cast<float>(ix);
          float2 wv;
          if constexpr (kLdsTaps) {
            wv = lt[((iy + HP) * D + (ix + HP)) * kWave];
          } else {
            const PM_GLOBAL char* tp =
                reinterpret_cast<const PM_GLOBAL char*>(trow + static_cast<int64_t>(ix + HP) * plane_stride);
            wv = *reinterpret_cast<const PM_GLOBAL float2*>(tp + (ucidx << 3));
          }
#pragma unroll
          for (int k = 0; k < kViewChunk; ++k) {
            const float sx = fmaf(wp[k].jxx, fix, bx[k]);
            const float sy = fmaf(wp[k].jxy, fix, by[k]);
            const float sd = sample<T, kClamp>(sm[k], sx, sy) - scen[k];
            const float wsv = wv.x * sd;
            s_ws[k] += wsv;
            s_wss[k] = fmaf(wsv, sd, s_wss[k]);
            s_wrs[k] = fmaf(wv.y, sd, s_wrs[k]);
          }
        }
      }
    };
    bool inside = true;
#pragma unroll
    for (int k = 0; k < kViewChunk; ++k) inside = inside && taps_in_view<HP>(wp[k], sm[k]);
    if (__all(inside)) taps(std::false_type{});
    else taps(std::true_type{});
#pragma unroll
    for (int k = 0; k < kViewChunk; ++k) {
      const int s = c0v + k;
      if (s >= n_src) continue;
      float cost = 2.0f;
      if (ok[k]) {
        float ncc1 = -1.0f;
        if (var_ref >= a.var_threshold) {
          const float mean_s = s_ws[k] * inv_w;
          const float var_src = s_wss[k] * inv_w - mean_s * mean_s;
          if (var_src >= a.var_threshold) {
            const float covar = s_wrs[k] * inv_w - mean_r * mean_s;
            ncc1 = covar / sqrtf(var_ref * var_src);
          }
        }
        if (ncc1 > 0.1f) {
          cost = 1.0f - ncc1;
        } else {
          cost = ncc1;
          pending |= 1u << s;
        }
      }
      out[s * os] = cost;
    }
  }
  return pending;
}

// Reference texels for the fallbacks: straight from the padded image, or
// from a workgroup's LDS copy of the clamped texels around its tile.
template <typename T>
struct GlobalRef {
  const T* img;
  PmView rv;
  __device__ __forceinline__ float operator()(int x, int y) const { return texel(img, rv, x, y); }
};
struct LdsRef {
  const float* t;
  int x0, y0, pitch;
  __device__ __forceinline__ float operator()(int x, int y) const { return t[(y - y0) * pitch + (x - x0)]; }
};

// Census fallback (opencl_kernels.h:270-315).
template <typename T, typename Ref>
__device__ __forceinline__ float census_inner(const Ref& ref, const Sampler& sm, float ref_center, float src_center,
                                              int cx, int cy, const Warp& w) {
  uint32_t rc = 0u, sc = 0u;
  int bit = 0;
  // rows one at a time: a rare path that must not set the kernel's registers
#pragma unroll 1
  for (int dy = -2; dy <= 2; ++dy) {
#pragma unroll
    for (int dx = -2; dx <= 2; ++dx) {
      if (dx == 0 && dy == 0) continue;
      const float ref_val = ref(cx + dx, cy + dy);
      const float sx = w.scx + w.jxx * static_cast<float>(dx) + w.jyx * static_cast<float>(dy);
      const float sy = w.scy + w.jxy * static_cast<float>(dx) + w.jyy * static_cast<float>(dy);
      const float src_val = sample<T>(sm, sx, sy);
      if (ref_val > ref_center) rc |= (1u << bit);
      if (src_val > src_center) sc |= (1u << bit);
      ++bit;
    }
  }
  const int hamming = __popc(rc ^ sc);
  if (__popc(rc) < 3 && __popc(sc) < 3) return 1.0f;
  return static_cast<float>(hamming) / 24.0f;
}

// Textureless fallback of the views left pending by cost_vector()
// (opencl_kernels.h:531-553): strides 2 then 3 (stop once NCC > 0.1), then
// census or 2.0.  The wave walks the views in order and the lanes with that
// view pending work on it together: the view is wave-uniform (scalar
// descriptors, one source image per gather instruction, neighbouring
// pixels), which keeps the gathers coherent.
// Strided NCC of the textureless fallback (opencl_kernels.h:188-258) at
// stride SS, where the fallback runs inline (random init, plane sweep,
// caller-given planes).  The reference-side (w, w*r) of each tap are
// recomputed from the reference texels with the arithmetic of
// refpatch_kernel (bit-identical to the stored strides 2 and 3); the moments
// come from ref_sums.  The tap at (ix, iy) reads the reference texel
// (x + SS*ix, y + SS*iy) -- the int conversion of x + float(SS*ix) is exact.
// The weight's argument is <= 0.
template <int HP, typename T, typename Ref>
__device__ __forceinline__ float ncc_at_stride(const PmArgs& a, const Ref& ref, const Sampler& sm, int x, int y,
                                               int cidx, float src_center, const Warp& w, int SS) {
  const float stride = static_cast<float>(SS);
  const float4 rs = gp(a.ref_sums)[static_cast<int64_t>(SS - 1) * a.ref_plane + cidx];  // (c, sum_w, sum_r, sum_rr)
  const float sf = a.spatial_factor, cf = a.color_factor;
  float sum_s = 0.0f, sum_ss = 0.0f, sum_rs = 0.0f;
#pragma unroll 1
  for (int iy = -HP; iy <= HP; ++iy) {
    const float fdy = stride * static_cast<float>(iy);
    const float bx = fmaf(w.jyx, fdy, w.scx);
    const float by = fmaf(w.jyy, fdy, w.scy);
#pragma unroll
    for (int ix = -HP; ix <= HP; ++ix) {
      const float fdx = stride * static_cast<float>(ix);
      const float sx = fmaf(w.jxx, fdx, bx);
      const float sy = fmaf(w.jxy, fdx, by);
      const float sd = sample<T>(sm, sx, sy) - src_center;
      const float rx = static_cast<float>(x) + fdx;
      const float ry = static_cast<float>(y) + fdy;
      const float r = ref(static_cast<int>(rx), static_cast<int>(ry)) - rs.x;
      const float wt = pm_expf_nonpos(fdx * fdx * sf + fdy * fdy * sf + r * r * cf);
      const float wr = wt * r;
      const float wsd = wt * sd;
      sum_s += wsd;
      sum_ss = fmaf(wsd, sd, sum_ss);
      sum_rs = fmaf(wr, sd, sum_rs);
    }
  }
  if (rs.y < 1e-6f) return -1.0f;
  const float inv_w = 1.0f / rs.y;
  const float mean_r = rs.z * inv_w;
  const float mean_s = sum_s * inv_w;
  const float var_ref = rs.w * inv_w - mean_r * mean_r;
  const float var_src = sum_ss * inv_w - mean_s * mean_s;
  if (var_ref < a.var_threshold || var_src < a.var_threshold) return -1.0f;
  const float covar = sum_rs * inv_w - mean_r * mean_s;
  return covar / sqrtf(var_ref * var_src);
}

// The fallback of one (pixel, plane, view) whose stride-1 NCC was <= 0.1
// (opencl_kernels.h:531-553): strides 2 then 3 (stop once NCC > 0.1), then
// census or 2.0.  Returns the cost.
template <int HP, typename T, typename Ref>
__device__ __forceinline__ float fallback_cost(const PmArgs& a, const Ref& ref, const Sampler& sm, int x, int y,
                                               int cidx, float ref_center, const Warp& w, float ncc1) {
  const float scen = sample<T>(sm, w.scx, w.scy);
  float best = ncc1;
#pragma unroll 1
  for (int s = 2; s <= 3; ++s) {  // strides 2 then 3 (:533-542)
    const float ncc = ncc_at_stride<HP, T>(a, ref, sm, x, y, cidx, scen, w, s);
    if (ncc > best) best = ncc;
    if (best > 0.1f) break;
  }
  if (best <= -0.5f) return (a.census_weight > 1e-6f) ? census_inner<T>(ref, sm, ref_center, scen, x, y, w) : 2.0f;
  return 1.0f - best;
}

template <int HP, typename T>
__device__ __forceinline__ void resolve_fallbacks(const PmArgs& a, const T* __restrict__ img, int x, int y, int cidx,
                                                  Plane pl, float* out, int64_t os, uint32_t pending) {
  PlanePts pp{};
  if (pending) plane_points(kcam(a, 0), x, y, pl, pp);
  const float ref_center = gp(a.ref_sums)[cidx].x;
  const GlobalRef<T> ref{img, kview(a, 0)};
  const int n_src = a.n_src;
#pragma unroll 1
  for (int s = 0; s < n_src; ++s) {
    const bool mine = (pending >> s) & 1u;
    if (!__any(mine)) continue;
    if (mine) {
      const Sampler sm = make_sampler<T>(a, kview(a, s + 1));
      Warp w;
      warp_of(pp.cw, pp.wdx, pp.wdy, kcam(a, s + 1), w);  // valid: it was valid at stride 1
      out[s * os] = fallback_cost<HP, T>(a, ref, sm, x, y, cidx, ref_center, w, out[s * os]);
    }
  }
}

// ---------------------------------------------------------------------------
// Textureless fallbacks of the colour-pass cost kernels are resolved by a
// separate kernel with every lane busy.  Each hypothesis wave records, per
// source view, the ballot of its lanes left pending:
//   fb_masks[(tile * 9 + slot) * n_src + view]   (bit = lane of the tile)
// and fallback_kernel gives each (view, tile) one wave that walks the
// pending (slot, lane) pairs 64 at a time.  The lanes of a wave are then
// neighbouring pixels warped into one source view -- coherent gathers --
// and the reference texels around the tile sit in LDS.
// Every hypothesis wave writes its masks on every pass (zero when nothing
// is pending), so no clearing is needed.  The stride-1 NCC is already in
// the cost slot.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void write_fb_masks(const PmArgs& a, int tile, int slot, uint32_t pending) {
  const int lane = static_cast<int>(threadIdx.x);
  uint64_t mine = 0ull;
#pragma unroll 1
  for (int s = 0; s < a.n_src; ++s) {
    const uint64_t b = __ballot((pending >> s) & 1u);
    if (lane == s) mine = b;
  }
  if (lane < a.n_src) gpw(a.fb_masks)[(static_cast<int64_t>(tile) * 9 + slot) * a.n_src + lane] = mine;
}

// Index of the n-th (0-based) set bit of m; m has more than n set bits.
__device__ __forceinline__ int nth_set_bit(uint64_t m, int n) {
  int pos = 0;
#pragma unroll
  for (int width = 32; width >= 1; width >>= 1) {
    const uint64_t lo = m & ((1ull << width) - 1ull);
    const int c = __popcll(lo);
    if (n >= c) {
      n -= c;
      m >>= width;
      pos += width;
    } else {
      m = lo;
    }
  }
  return pos;
}

// ---------------------------------------------------------------------------
// Tile-workgroup fallback: one workgroup of kFbWaves waves per colour tile.
// The bilateral weights of the strided taps depend only on the reference
// pixel, the stride and the tap, not on the plane or the view, so the
// workgroup computes them once for the tile's pixels that have pending jobs
// (strides 2 and 3, into LDS; bit-identical to computing them per job) and
// every (slot, view) job of those pixels reads them back: about half of a
// strided tap's instructions are the weight's argument and exp.  The waves
// then take (view, 64-job chunk) items from an LDS counter in view order,
// so the lanes of a wave share a source view (scalar camera and sampler,
// neighbouring pixels) and the workgroup's waves stay busy when views have
// uneven job counts.
// ---------------------------------------------------------------------------
#ifndef PM_FB_WAVES
#define PM_FB_WAVES 4
#endif
constexpr int kFbWaves = PM_FB_WAVES;

// Strided NCC of one job with the tile's LDS weights wt[tap * kWave] of
// the job's pixel (opencl_kernels.h:188-258): the same sums, in the same
// order, as ncc_at_stride.  When the reference moments already fail the
// variance gate the result is -1 whatever the source, so the taps are skipped.
template <typename T>
struct GlobalSampler {
  Sampler sm;
  __device__ __forceinline__ float operator()(float sx, float sy) const { return sample<T>(sm, sx, sy); }
};

// Strided NCC of one job with the tile's LDS weights wt[tap * kWave] of
// the job's pixel (opencl_kernels.h:188-258): the same sums, in the same
// order, as ncc_at_stride.  When the reference moments already fail the
// variance gate the result is -1 whatever the source, so the taps are skipped.
template <int HP, int S, typename Smp>
__device__ __forceinline__ float ncc_at_stride_w(const PmArgs& a, const LdsRef& ref, const float* wt,
                                                 const Smp& smp, int x, int y, int cidx, float src_center,
                                                 const Warp& w) {
  constexpr float stride = static_cast<float>(S);
  const float4 rs = gp(a.ref_sums)[static_cast<int64_t>(S - 1) * a.ref_plane + cidx];  // (c, sum_w, sum_r, sum_rr)
  if (rs.y < 1e-6f) return -1.0f;
  const float inv_w = 1.0f / rs.y;
  const float mean_r = rs.z * inv_w;
  const float var_ref = rs.w * inv_w - mean_r * mean_r;
  if (var_ref < a.var_threshold) return -1.0f;
  float sum_s = 0.0f, sum_ss = 0.0f, sum_rs = 0.0f;
#pragma unroll 1
  for (int iy = -HP; iy <= HP; ++iy) {
    const float fdy = stride * static_cast<float>(iy);
    const float bx = fmaf(w.jyx, fdy, w.scx);
    const float by = fmaf(w.jyy, fdy, w.scy);
    const float* rrow = ref.t + (y + S * iy - ref.y0) * ref.pitch + (x - ref.x0);
    const float* wrow = wt + (iy + HP) * (2 * HP + 1) * kWave;
#pragma unroll
    for (int ix = -HP; ix <= HP; ++ix) {
      const float fdx = stride * static_cast<float>(ix);
      const float sx = fmaf(w.jxx, fdx, bx);
      const float sy = fmaf(w.jxy, fdx, by);
      const float sd = smp(sx, sy) - src_center;
      const float r = rrow[S * ix] - rs.x;
      const float wgt = wrow[(ix + HP) * kWave];
      const float wr = wgt * r;
      const float wsd = wgt * sd;
      sum_s += wsd;
      sum_ss = fmaf(wsd, sd, sum_ss);
      sum_rs = fmaf(wr, sd, sum_rs);
    }
  }
  const float mean_s = sum_s * inv_w;
  const float var_src = sum_ss * inv_w - mean_s * mean_s;
  if (var_src < a.var_threshold) return -1.0f;
  const float covar = sum_rs * inv_w - mean_r * mean_s;
  return covar / sqrtf(var_ref * var_src);
}

// Census on the job's sampler (opencl_kernels.h:270-315), as census_inner.
template <typename Smp>
__device__ __forceinline__ float census_w(const LdsRef& ref, const Smp& smp, float ref_center, float src_center,
                                          int cx, int cy, const Warp& w) {
  uint32_t rc = 0u, sc = 0u;
  int bit = 0;
#pragma unroll 1
  for (int dy = -2; dy <= 2; ++dy) {
#pragma unroll
    for (int dx = -2; dx <= 2; ++dx) {
      if (dx == 0 && dy == 0) continue;
      const float ref_val = ref(cx + dx, cy + dy);
      const float sx = w.scx + w.jxx * static_cast<float>(dx) + w.jyx * static_cast<float>(dy);
      const float sy = w.scy + w.jxy * static_cast<float>(dx) + w.jyy * static_cast<float>(dy);
      const float src_val = smp(sx, sy);
      if (ref_val > ref_center) rc |= (1u << bit);
      if (src_val > src_center) sc |= (1u << bit);
      ++bit;
    }
  }
  const int hamming = __popc(rc ^ sc);
  if (__popc(rc) < 3 && __popc(sc) < 3) return 1.0f;
  return static_cast<float>(hamming) / 24.0f;
}

// Strides 2 then 3 (stop once NCC > 0.1), then census or 2.0
// (opencl_kernels.h:531-553) on one sampler.
template <int HP, typename Smp>
__device__ __forceinline__ float fallback_w(const PmArgs& a, const LdsRef& ref, const float* wt, const Smp& smp,
                                            int x, int y, int cidx, float scen, const Warp& w, float ncc1) {
  constexpr int P = (2 * HP + 1) * (2 * HP + 1);
  float best = ncc1;
  const float ncc2 = ncc_at_stride_w<HP, 2>(a, ref, wt, smp, x, y, cidx, scen, w);
  if (ncc2 > best) best = ncc2;
  if (!(best > 0.1f)) {
    const float ncc3 = ncc_at_stride_w<HP, 3>(a, ref, wt + P * kWave, smp, x, y, cidx, scen, w);
    if (ncc3 > best) best = ncc3;
  }
  if (best <= -0.5f) return (a.census_weight > 1e-6f) ? census_w(ref, smp, ref(x, y), scen, x, y, w) : 2.0f;
  return 1.0f - best;
}

template <int HP, typename T, int TX>
__device__ __forceinline__ void fallback_tile_body(const PmArgs& a, int b) {
  constexpr int D = 2 * HP + 1;
  constexpr int P = D * D;
  constexpr int TY = kWave / TX;
  constexpr int R = (3 * HP > 2) ? 3 * HP : 2;       // reach of stride 3 and of the 5 x 5 census
  constexpr int RW = 2 * TX + 2 * R, RH = TY + 2 * R;
  __shared__ float ref_tile[RW * RH];
  __shared__ float wts[2 * P * kWave];               // [stride - 2][tap][tile lane]
  __shared__ int next_item;
  int tx, ty;
  if (!colour_tile<TX>(a, b, tx, ty)) return;        // uniform over the workgroup
  const int tile = ty * colour_tiles_x<TX>(a) + tx;
  const int n_src = a.n_src;
  const int nslots = a.phase == 0 ? 9 : 6;
  const int lane = static_cast<int>(threadIdx.x);
  const int t = static_cast<int>(threadIdx.y) * kWave + lane;
  // written by the cost kernel before this launch: read-only here (scalar loads)
  const PM_CONST uint64_t* mk = (const PM_CONST uint64_t*)(a.fb_masks) + static_cast<int64_t>(tile) * 9 * n_src;
  uint64_t any = 0ull;
  int chunks = 0;
#pragma unroll 1
  for (int s = 0; s < n_src; ++s) {
    int tot = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const uint64_t m = k < nslots ? mk[k * n_src + s] : 0ull;
      any |= m;
      tot += __popcll(m);
    }
    chunks += (tot + kWave - 1) / kWave;
  }
  if (chunks == 0) return;                           // uniform over the workgroup
  const int x0 = tx * 2 * TX - R, y0 = ty * TY - R;
  {
    const T* img = images_of<T>(a);
    const PmView rv = kview(a, 0);
    for (int i = t; i < RW * RH; i += kWave * kFbWaves) ref_tile[i] = texel(img, rv, x0 + i % RW, y0 + i / RW);
  }
  if (t == 0) next_item = 0;
  __syncthreads();
  // weights of the strided taps of the tile's pixels with pending jobs
  {
    const float sf = a.spatial_factor, cf = a.color_factor;
    for (int i = t; i < 2 * P * kWave; i += kWave * kFbWaves) {
      const int l = i % kWave;
      if (!((any >> l) & 1ull)) continue;
      const int tap = (i / kWave) % P;
      const int SS = 2 + i / (P * kWave);
      const int iy = tap / D - HP, ix = tap % D - HP;
      const int yy = ty * TY + l / TX;
      const int xx = (tx * TX + l % TX) * 2 + ((yy + a.color) & 1);
      const float fdx = static_cast<float>(SS) * static_cast<float>(ix);
      const float fdy = static_cast<float>(SS) * static_cast<float>(iy);
      const float r = ref_tile[(yy + SS * iy - y0) * RW + (xx + SS * ix - x0)] - ref_tile[(yy - y0) * RW + (xx - x0)];
      wts[i] = pm_expf_nonpos(fdx * fdx * sf + fdy * fdy * sf + r * r * cf);
    }
  }
  __syncthreads();
  const LdsRef ref{ref_tile, x0, y0, RW};
  const int64_t pl = a.ref_plane;
  const PmCam c0 = kcam(a, 0);
#pragma unroll 1
  for (;;) {
    int item = 0;
    if (lane == 0) item = atomicAdd(&next_item, 1);
    item = __builtin_amdgcn_readfirstlane(__shfl(item, 0));
    if (item >= chunks) break;
    // view s and chunk c of this item (views in order)
    int s = 0, c = item;
    uint64_t m[9];
    int cum[9];
    int total = 0;
#pragma unroll 1
    for (;; ++s) {
      total = 0;
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        m[k] = k < nslots ? mk[k * n_src + s] : 0ull;
        total += __popcll(m[k]);
        cum[k] = total;
      }
      const int nc = (total + kWave - 1) / kWave;
      if (c < nc) break;
      c -= nc;
    }
    const int q = c * kWave + lane;
    if (q >= total) continue;
    int slot = 0, prev = 0;
    uint64_t mm = m[0];
#pragma unroll
    for (int k = 1; k < 9; ++k) {
      if (q >= cum[k - 1]) { slot = k; prev = cum[k - 1]; mm = m[k]; }
    }
    const int l = nth_set_bit(mm, q - prev);
    const int gx = tx * TX + (l % TX);
    const int y = ty * TY + (l / TX);
    const int x = gx * 2 + ((y + a.color) & 1);
    const int cidx = colour_index(a, x, y);
    Plane pl4;
    if (a.phase == 0) {
      const int pos = slot < 8 ? gp(a.cpos)[slot * pl + cidx] : y * a.width + x;
      pl4 = to_plane(gp(a.planes)[pos]);
    } else {
      pl4 = to_plane(gp(a.refp)[slot * pl + cidx]);
    }
    float* out = a.ca + (static_cast<int64_t>(slot) * n_src + s) * pl + cidx;
    const GlobalSampler<T> smp{make_sampler<T>(a, kview(a, s + 1))};
    PlanePts pp;
    plane_points(c0, x, y, pl4, pp);
    Warp w;
    warp_of(pp.cw, pp.wdx, pp.wdy, kcam(a, s + 1), w);  // valid: it was valid at stride 1
    const float scen = smp(w.scx, w.scy);
    *out = fallback_w<HP>(a, ref, wts + l, smp, x, y, cidx, scen, w, *out);
  }
}

// Full cost vector: stride 1 for all views, then the fallbacks.
template <int HP, typename T>
__device__ __forceinline__ void full_cost_vector(const PmArgs& a, const T* __restrict__ img, int x, int y, int cidx,
                                                 Plane pl, float* out, int64_t os = kWave) {
  const uint32_t pending = cost_vector<HP>(a, img, x, y, cidx, pl, out, os);
  resolve_fallbacks<HP>(a, img, x, y, cidx, pl, out, os, pending);
}

// Smoothness against the 4 neighbours (opencl_kernels.h:649-707); the
// neighbour planes are preloaded (they have the other colour and cannot
// change during the pass).
// The 4-neighbourhood of a pixel (left, right, up, down): whether each
// neighbour is inside the image, its depth on its own plane at its own
// pixel, and its unit normal.  These do not depend on the hypothesis being
// scored, so they are computed once per pixel (0 depth and a zero normal
// outside the image, as the zero plane gives).
struct Nbrs {
  float nd[4];
  V3 nn[4];
  bool ok[4];
};

__device__ __forceinline__ float smoothness_cost(const PmArgs& a, const Nbrs& nb, Plane cur, int x, int y) {
  const PmCam c0 = kcam(a, 0);
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
    const float nd = nb.nd[k];
    if (nd <= 0.0f) continue;
    const float rel = fabsf(cur_depth - nd) / cur_depth;
    const float w = pm_expf(-rel * rel / (2.0f * sigma_rel * sigma_rel));
    const float normal_cost = 1.0f - dot3(cur_n, nb.nn[k]);
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
  const PmCam c0 = kcam(a, 0);
  const int dxs[4] = {-1, 1, 0, 0};
  const int dys[4] = {0, 0, -1, 1};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int nx = x + dxs[k], ny = y + dys[k];
    nb.ok[k] = !(nx < 0 || nx >= W || ny < 0 || ny >= H);
    const Plane q = nb.ok[k] ? to_plane(gp(planes)[static_cast<int64_t>(ny) * W + nx]) : Plane{0, 0, 0, 0};
    nb.nd[k] = depth_from_plane(static_cast<float>(nx), static_cast<float>(ny), q, c0);
    nb.nn[k] = normalize3(V3{q.x, q.y, q.z});
  }
}

// opencl_kernels.h:725-772
__device__ __forceinline__ V3 surface_normal(const PmArgs& a, const Nbrs& nb, int x, int y, float center_depth) {
  const int W = a.width, H = a.height;
  if (center_depth <= 0.0f) return V3{0, 0, 0};
  if (x < 1 || x >= W - 1 || y < 1 || y >= H - 1) return V3{0, 0, 0};
  const PmCam c0 = kcam(a, 0);
  const float fx = c0.K[0], fy = c0.K[4], cx = c0.K[2], cy = c0.K[5];
  const float dl = nb.nd[0], dr = nb.nd[1], du = nb.nd[2], dd = nb.nd[3];
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
  const PmCam rc = kcam(a, 0);
  const PmCam scm = lane_cam(a, sidx);
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
template <int NV>
__device__ __forceinline__ float mc_cost(const PmArgs& a, const float* cv, int64_t cs, const float (&vw)[NV],
                                         float wnorm, Plane plane, int x, int y, uint32_t& s0, uint32_t& s1,
                                         uint32_t key) {
  if (wnorm < 1e-6f) return 2.0f;
  const int n_src = a.n_src;
  float cdf[NV];
  float cum = 0.0f;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    if (i < n_src) cum += vw[i];
    cdf[i] = cum;
  }
  const float gw = a.geom_weight;
  const bool use_geom = gw > 0.0f && a.prev_depth_mask != 0u;
  float total = 0.0f;
  if (!use_geom) {
    // The hypothesis's view costs are loaded once, before the draws; each
    // draw takes its view's cost by the same comparisons that pick the view,
    // so no draw waits on a load of its own.
    float c[NV];
    float last = 0.0f;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      c[i] = i < n_src ? cv[i * cs] : 0.0f;
      if (i == n_src - 1) last = c[i];
    }
    const float geom_term = gw * 0.0f;
    for (int smp = 0; smp < kMcSamples; ++smp) {
      const float u = rand_float(s0, s1, key + static_cast<uint32_t>(smp) * 7919u);
      const float thr = u * cum;
      float photo = last;
#pragma unroll
      for (int i = NV - 1; i >= 0; --i)
        if (i < n_src && cdf[i] > thr) photo = c[i];
      total += photo + geom_term;
    }
    return total / static_cast<float>(kMcSamples);
  }
  for (int smp = 0; smp < kMcSamples; ++smp) {
    const float u = rand_float(s0, s1, key + static_cast<uint32_t>(smp) * 7919u);
    const float thr = u * cum;
    int v = n_src - 1;
#pragma unroll
    for (int i = NV - 1; i >= 0; --i)
      if (i < n_src && cdf[i] > thr) v = i;
    const float photo = cv[v * cs];
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

// ---------------------------------------------------------------------------
// acmmp_random_init (opencl_kernels.h:979-1039)
// ---------------------------------------------------------------------------
// Random plane of pixel (x, y) from its seeded stream (opencl_kernels.h:
// 1003-1021): log-uniform depth in [depth_min, depth_max], normal
// (u, v, -1) normalised.
__device__ __forceinline__ Plane random_plane(const PmArgs& a, int x, int y, uint32_t& s0, uint32_t& s1,
                                              uint32_t key) {
  const float ldmin = pm_logf(fmaxf(a.depth_min, 1e-6f));
  const float ldmax = pm_logf(fmaxf(a.depth_max, 1e-5f));
  const float depth = pm_expf(ldmin + rand_float(s0, s1, key) * (ldmax - ldmin));
  V3 n;
  n.x = rand_float(s0, s1, key + 3u) * 2.0f - 1.0f;
  n.y = rand_float(s0, s1, key + 5u) * 2.0f - 1.0f;
  n.z = -1.0f;
  n = normalize3(n);
  return plane_from_depth_normal(static_cast<float>(x), static_cast<float>(y), depth, n, kcam(a, 0));
}

// The init runs in two kernels.  init_views: one wave per (source view,
// 16x4 tile) scores the tile's random planes in that view (stride-1 NCC,
// then the textureless fallback inline) into ca[view][cidx].  Each XCD takes
// its band of tiles view by view, so the gathers in flight on one L2 are
// those of a single source view: random depths scatter a band's patches
// along that view's epipolar band, which then stays resident, instead of
// along all views' bands at once.  init_select then reduces each pixel's
// column of costs (top-k, selected views).  Every pixel regenerates its
// plane from its seeded stream in every view's wave; the view-0 wave stores
// it and the generator state.
#ifndef PM_INIT_STRIP
#define PM_INIT_STRIP 16
#endif
constexpr int kInitStrip = PM_INIT_STRIP;  // tiles per strip (0: whole rows)

template <int HP, typename T>
__device__ __forceinline__ void init_views_body(const PmArgs& a, int b) {
  const int tiles_x = (a.width + kFullTileX - 1) / kFullTileX;
  const int ntiles = tiles_x * ((a.height + kFullTileY - 1) / kFullTileY);
  const int per = (ntiles + kXcds - 1) / kXcds;  // slots of one view in this XCD's sequence
  const int q = b / kXcds;                       // position in this XCD's sequence
  const int v = q / per;
  const int t = (b % kXcds) * per + (q - v * per);
  if (v >= a.n_src || t >= ntiles) return;
  // Within a view the tiles go in the strip order of tile_of, strips of
  // kInitStrip tiles: a random depth moves a patch along its epipolar
  // segment, and a strip's segments fit one L2 whichever way they run.
  int tx, ty;
  strip_tile(tiles_x, (a.height + kFullTileY - 1) / kFullTileY, t, kInitStrip, 0, tx, ty);
  const int x = tx * kFullTileX + static_cast<int>(threadIdx.x % kFullTileX);
  const int y = ty * kFullTileY + static_cast<int>(threadIdx.x / kFullTileX);
  if (x >= a.width || y >= a.height) return;
  const T* img = images_of<T>(a);
  const int idx = y * a.width + x;
  uint32_t s0 = static_cast<uint32_t>(idx) * 1099087573u + 2654435769u;
  uint32_t s1 = static_cast<uint32_t>(idx) * 2246822519u + 13u;
  const Plane plane = random_plane(a, x, y, s0, s1, static_cast<uint32_t>(idx + 7919));
  if (v == 0) {
    a.planes[idx] = to_f4(plane);
    a.rng[idx] = make_uint2(s0, s1);
  }
  const int cidx = colour_index(a, x, y);
  float* out = a.ca + cidx;
  const uint32_t pending = cost_vector<HP>(a, img, x, y, cidx, plane, out, a.ref_plane, nullptr, v, v + 1);
  resolve_fallbacks<HP>(a, img, x, y, cidx, plane, out, a.ref_plane, pending);
}

__device__ __forceinline__ void init_select_body(const PmArgs& a, int b) {
  extern __shared__ float smem[];
  const int lane = threadIdx.x;
  int x, y;
  if (!full_tile_pixel(a.width, a.height, b, x, y)) return;
  const int idx = y * a.width + x;
  const int cidx = colour_index(a, x, y);
  float* cv = smem + lane;
  float* tmp = smem + a.n_src * kWave + lane;
  // The random plane is the pixel's current plane: its costs also go to
  // slot 8, where the first candidates pass finds them.
  float* cur = a.ca + static_cast<int64_t>(8) * a.n_src * a.ref_plane + cidx;
  for (int s = 0; s < a.n_src; ++s) {
    const float c = a.ca[s * a.ref_plane + cidx];
    cv[s * kWave] = c;
    cur[s * a.ref_plane] = c;
  }
  uint32_t sel = 0u;
  a.costs[idx] = initial_cost_and_views(a, cv, tmp, &sel);
  a.sel[idx] = sel;
}

// ---------------------------------------------------------------------------
// Fronto-parallel plane sweep (compute_brute_force of the pydense stub,
// pydense.pyi:53, which has no implementation in the reference tree), one
// launch per plane: plane a.sweep_index of a.sweep_count has depth
// exp(log dmin + i * (log dmax - log dmin) / (n - 1)) (the geometric mean of
// the range for one plane) and normal (0, 0, -1), and is scored like a
// random-init plane.  The first launch writes every pixel; later ones keep
// the lower cost, so the first plane wins ties.  The generator state is the
// random init's seed, before any draw.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float sweep_depth(const PmArgs& a) {
  const float ldmin = pm_logf(fmaxf(a.depth_min, 1e-6f));
  const float ldmax = pm_logf(fmaxf(a.depth_max, 1e-5f));
  if (a.sweep_count <= 1) return pm_expf(0.5f * (ldmin + ldmax));
  const float step = (ldmax - ldmin) / static_cast<float>(a.sweep_count - 1);
  return pm_expf(ldmin + static_cast<float>(a.sweep_index) * step);
}

template <int HP, typename T>
__device__ __forceinline__ void plane_sweep_body(const PmArgs& a, int b) {
  extern __shared__ float smem[];
  const int lane = threadIdx.x;
  int x, y;
  if (!full_tile_pixel(a.width, a.height, b, x, y)) return;
  const int idx = y * a.width + x;
  const Plane plane = plane_from_depth_normal(static_cast<float>(x), static_cast<float>(y), sweep_depth(a),
                                              V3{0.0f, 0.0f, -1.0f}, kcam(a, 0));
  float* cv = smem + lane;
  float* tmp = smem + a.n_src * kWave + lane;
  full_cost_vector<HP>(a, images_of<T>(a), x, y, colour_index(a, x, y), plane, cv);
  uint32_t sel = 0u;
  const float cost = initial_cost_and_views(a, cv, tmp, &sel);
  const bool first = a.sweep_index == 0;
  if (first) {
    a.rng[idx] = make_uint2(static_cast<uint32_t>(idx) * 1099087573u + 2654435769u,
                            static_cast<uint32_t>(idx) * 2246822519u + 13u);
  }
  if (first || cost < a.costs[idx]) {
    a.planes[idx] = to_f4(plane);
    a.costs[idx] = cost;
    a.sel[idx] = sel;
  }
}

// ---------------------------------------------------------------------------
// acmmp_prior_reinit (opencl_kernels.h:1056-1182); neighbour planes of the
// smoothness term come from the pre-launch snapshot, as in the reference.
// ---------------------------------------------------------------------------
template <int HP, typename T>
__device__ __forceinline__ void prior_body(const PmArgs& a, int b) {
  extern __shared__ float smem[];
  const int lane = threadIdx.x;
  int x, y;
  if (!full_tile_pixel(a.width, a.height, b, x, y)) return;
  const int idx = y * a.width + x;
  if (a.plane_masks[idx] == 0u) return;
  const T* img = images_of<T>(a);
  const PmCam c0 = kcam(a, 0);
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
  Nbrs nb;
  load_nbrs(a, a.snapshot, x, y, nb);
  const float sw = a.smooth_w;
  float* cv = smem + lane;
  float* tmp = smem + n_src * kWave + lane;
  full_cost_vector<HP>(a, img, x, y, colour_index(a, x, y), plane, cv);
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
// acmmp_patchmatch, one colour (opencl_kernels.h:1197-1778), as four kernels:
//   pass_candidates  cost vectors of the 8 propagated candidates and of the
//                    current plane, one lane per (pixel, hypothesis);
//   pass_select      view weights, Monte-Carlo costs, acceptance of the best
//                    candidate, refinement hypotheses (one lane per pixel);
//   pass_refine_cost cost vectors of the refinement hypotheses, one lane per
//                    (pixel, hypothesis);
//   pass_refine      their Monte-Carlo costs and acceptance (lane per pixel).
// Cost vectors draw no random numbers, so computing them ahead of the MC
// steps keeps the reference's RNG stream: candidate MC draws, current-plane
// MC draws, refinement set-up draws, refinement MC draws.  Per-pixel
// intermediate state lives in HBM, indexed by the colour-major index.
// ---------------------------------------------------------------------------
struct ColourPixel {
  int x, y, idx, cidx;
};

// Pixel of this lane in colour tile (tx, ty).
template <int TX>
__device__ __forceinline__ bool colour_lane_pixel(const PmArgs& a, int tx, int ty, ColourPixel& cp) {
  const int hw = (a.width + 1) / 2;
  const int gx = tx * TX + static_cast<int>(threadIdx.x % TX);
  const int y = ty * (kWave / TX) + static_cast<int>(threadIdx.x / TX);
  if (gx >= hw || y >= a.height) return false;
  const int x = gx * 2 + ((y + a.color) & 1);
  if (x >= a.width) return false;
  cp.x = x;
  cp.y = y;
  cp.idx = y * a.width + x;
  cp.cidx = colour_index(a, x, y);
  return true;
}

// Copies the stride-1 reference taps of the 64 pixels of colour tile
// (tx, ty) to LDS, lt[tap * kWave + lane], with every thread of the
// workgroup; the hypothesis waves then read them with LDS loads, which
// takes a third of the vector-memory instructions off the texture path.
template <int P, int TX>
__device__ __forceinline__ void stage_tile_taps(const PmArgs& a, int tx, int ty, float2* lt) {
  const int hw = (a.width + 1) / 2;
  const int nthreads = static_cast<int>(blockDim.x * blockDim.y);
  const int t = static_cast<int>(threadIdx.y * blockDim.x + threadIdx.x);
  const int64_t cbase = static_cast<int64_t>(a.color) * a.height * hw;
  const PM_GLOBAL float2* taps = gp(a.ref_taps);
  for (int k = t; k < P * kWave; k += nthreads) {
    const int tap = k / kWave, l = k % kWave;
    const int gx = tx * TX + l % TX, y = ty * (kWave / TX) + l / TX;
    float2 v = make_float2(0.0f, 0.0f);
    if (gx < hw && y < a.height) v = taps[static_cast<int64_t>(tap) * a.ref_plane + cbase + y * hw + gx];
    lt[k] = v;
  }
  __syncthreads();
}

// Pixel of this lane in the colour tile of workgroup b.
template <int TX>
__device__ __forceinline__ bool colour_pixel(const PmArgs& a, int b, ColourPixel& cp) {
  int tx, ty;
  if (!colour_tile<TX>(a, b, tx, ty)) return false;
  return colour_lane_pixel<TX>(a, tx, ty, cp);
}

// Candidate slot h of pixel (x, y) (opencl_kernels.h:1254-1430): the
// cheapest-cost probe of that search pattern, or -1 when the pattern does
// not fit in the image.  0 near up, 1 far up, 2 near down, 3 far down,
// 4 near left, 5 far left, 6 near right, 7 far right.
__device__ __forceinline__ int candidate_position(const PmArgs& a, int h, int x, int y) {
  const int W = a.width, H = a.height;
  const PM_GLOBAL float* costs = gp(a.costs);
  auto cost_at = [&](int px, int py) { return costs[py * W + px]; };
  switch (h) {
    case 1:
      if (y > 2) {
        int by = y - 3; float bc = cost_at(x, by);
        for (int i = 1; i < 11; ++i) { const int yy = y - 3 - 2 * i; if (yy < 0) break;
          const float c = cost_at(x, yy); if (c < bc) { bc = c; by = yy; } }
        return by * W + x;
      }
      return -1;
    case 3:
      if (y < H - 3) {
        int by = y + 3; float bc = cost_at(x, by);
        for (int i = 1; i < 11; ++i) { const int yy = y + 3 + 2 * i; if (yy >= H) break;
          const float c = cost_at(x, yy); if (c < bc) { bc = c; by = yy; } }
        return by * W + x;
      }
      return -1;
    case 5:
      if (x > 2) {
        int bx = x - 3; float bc = cost_at(bx, y);
        for (int i = 1; i < 11; ++i) { const int xx = x - 3 - 2 * i; if (xx < 0) break;
          const float c = cost_at(xx, y); if (c < bc) { bc = c; bx = xx; } }
        return y * W + bx;
      }
      return -1;
    case 7:
      if (x < W - 3) {
        int bx = x + 3; float bc = cost_at(bx, y);
        for (int i = 1; i < 11; ++i) { const int xx = x + 3 + 2 * i; if (xx >= W) break;
          const float c = cost_at(xx, y); if (c < bc) { bc = c; bx = xx; } }
        return y * W + bx;
      }
      return -1;
    case 0:
      if (y > 0) {
        int bx = x, by = y - 1; float bc = cost_at(bx, by);
        for (int i = 0; i < 3; ++i) {
          if (y > 1 + i && x > i) { const int px = x - i - 1, py = y - 2 - i; const float c = cost_at(px, py);
            if (c < bc) { bc = c; bx = px; by = py; } }
          if (y > 1 + i && x < W - 1 - i) { const int px = x + i + 1, py = y - 2 - i; const float c = cost_at(px, py);
            if (c < bc) { bc = c; bx = px; by = py; } }
        }
        return by * W + bx;
      }
      return -1;
    case 2:
      if (y < H - 1) {
        int bx = x, by = y + 1; float bc = cost_at(bx, by);
        for (int i = 0; i < 3; ++i) {
          if (y < H - 2 - i && x > i) { const int px = x - i - 1, py = y + 2 + i; const float c = cost_at(px, py);
            if (c < bc) { bc = c; bx = px; by = py; } }
          if (y < H - 2 - i && x < W - 1 - i) { const int px = x + i + 1, py = y + 2 + i; const float c = cost_at(px, py);
            if (c < bc) { bc = c; bx = px; by = py; } }
        }
        return by * W + bx;
      }
      return -1;
    case 4:
      if (x > 0) {
        int bx = x - 1, by = y; float bc = cost_at(bx, by);
        for (int i = 0; i < 3; ++i) {
          if (x > 1 + i && y > i) { const int px = x - 2 - i, py = y - i - 1; const float c = cost_at(px, py);
            if (c < bc) { bc = c; bx = px; by = py; } }
          if (x > 1 + i && y < H - 1 - i) { const int px = x - 2 - i, py = y + i + 1; const float c = cost_at(px, py);
            if (c < bc) { bc = c; bx = px; by = py; } }
        }
        return by * W + bx;
      }
      return -1;
    default:  // 6
      if (x < W - 1) {
        int bx = x + 1, by = y; float bc = cost_at(bx, by);
        for (int i = 0; i < 3; ++i) {
          if (x < W - 2 - i && y > i) { const int px = x + 2 + i, py = y - i - 1; const float c = cost_at(px, py);
            if (c < bc) { bc = c; bx = px; by = py; } }
          if (x < W - 2 - i && y < H - 1 - i) { const int px = x + 2 + i, py = y + i + 1; const float c = cost_at(px, py);
            if (c < bc) { bc = c; bx = px; by = py; } }
        }
        return by * W + bx;
      }
      return -1;
  }
}

template <int HP, typename T, int TX>
__device__ __forceinline__ void candidates_body(const PmArgs& a, int b) {
  constexpr int P = (2 * HP + 1) * (2 * HP + 1);
  __shared__ float2 lt[P * kWave];
  int tx, ty;
  if (!colour_tile<TX>(a, b, tx, ty)) return;
  stage_tile_taps<P, TX>(a, tx, ty, lt);
  const int h = threadIdx.y;  // 0..7 candidates, 8 = current plane (one wave each)
  ColourPixel cp;
  uint32_t pending = 0u;
  // the current plane's costs are already in slot 8 when cur_cached
  if (!(h == 8 && a.cur_cached) && colour_lane_pixel<TX>(a, tx, ty, cp)) {
    const int pos = h == 8 ? cp.idx : candidate_position(a, h, cp.x, cp.y);
    if (h < 8) gpw(a.cpos)[static_cast<int64_t>(h) * a.ref_plane + cp.cidx] = pos;
    if (pos >= 0) {
      const Plane hyp = to_plane(gp(a.planes)[pos]);
      float* out = a.ca + static_cast<int64_t>(h) * a.n_src * a.ref_plane + cp.cidx;
      pending = cost_vector<HP, T, true>(a, images_of<T>(a), cp.x, cp.y, cp.cidx, hyp, out, a.ref_plane,
                                         lt + threadIdx.x);
    }
  }
  write_fb_masks(a, ty * colour_tiles_x<TX>(a) + tx, h, pending);
}

// The per-view costs of hypothesis `slot` become the costs of the pixel's
// current plane (slot 8), which the next candidates pass of this colour
// then need not score again.
__device__ __forceinline__ void copy_costs_to_current(const PmArgs& a, int cidx, int slot) {
  const int64_t pl = a.ref_plane;
  const float* src = a.ca + static_cast<int64_t>(slot) * a.n_src * pl + cidx;
  float* dst = a.ca + static_cast<int64_t>(8) * a.n_src * pl + cidx;
#pragma unroll 1
  for (int s = 0; s < a.n_src; ++s) dst[s * pl] = src[s * pl];
}

template <int NV, int TX>
__device__ __forceinline__ void select_body(const PmArgs& a, int b) {
  ColourPixel cp;
  if (!colour_pixel<TX>(a, b, cp)) return;
  const int x = cp.x, y = cp.y, idx = cp.idx;
  const int64_t pl = a.ref_plane;
  const PmCam c0 = kcam(a, 0);
  const int n_src = a.n_src;
  const uint2 st = gp(a.rng)[idx];
  uint32_t s0 = st.x, s1 = st.y;
  const uint32_t key = static_cast<uint32_t>(idx + a.iteration * 7919 + 42);
  const PM_GLOBAL float4* planes = gp(a.planes);
  const float* ca = a.ca + cp.cidx;  // ca[(slot * n_src + view) * pl]
  int cpos[8];
  uint32_t flags = 0u;
#pragma unroll
  for (int h = 0; h < 8; ++h) {
    cpos[h] = gp(a.cpos)[h * pl + cp.cidx];
    if (cpos[h] >= 0) flags |= 1u << h;
  }
  const uint32_t prev_sel = gp(a.sel)[idx];
  Plane plane_now = to_plane(planes[idx]);

  // Step 2: view weights (:1452-1563).
  float vw[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) vw[i] = 0.0f;
  const float vthr = 0.8f;
  for (int j = 0; j < 8; ++j) {  // per view the sum runs over j = 0..7, as in the reference
    if (!((flags >> j) & 1u)) continue;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      if (i < n_src) {
        const float c = ca[(j * n_src + i) * pl];
        if (c < vthr) vw[i] += (1.0f - c);
      }
    }
  }
  if (a.anchor_views > 0) {
    int anchored = 0;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
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
    for (int i = 0; i < NV; ++i) {
      if (i >= n_src || vw[i] < 1e-6f) continue;
      const PmView v = kview(a, i + 1);
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
  float wnorm = 0.0f;
#pragma unroll
  for (int i = 0; i < NV; ++i)
    if (i < n_src && vw[i] > 0.0f) { new_sel |= (1u << i); wnorm += vw[i]; }
  if (wnorm < 1e-6f) {  // :1552-1563
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      if (i >= n_src) continue;
      bool any = false;
      for (int j = 0; j < 8; ++j)
        any |= ((flags >> j) & 1u) && ca[(j * n_src + i) * pl] < 2.0f;
      if (any) { vw[i] = 1.0f; new_sel |= (1u << i); wnorm += 1.0f; }
    }
  }

  Nbrs nb;
  load_nbrs(a, a.planes, x, y, nb);
  const float smooth_w = a.smooth_w;

  // Step 3: Monte-Carlo costs of the candidates, then of the current plane
  // (:1570-1619).
  float best_cost = 2.0f, cost_now = 2.0f;
  int best = 0;
  for (int i = 0; i < 9; ++i) {
    const bool flagged = i == 8 || ((flags >> i) & 1u);
    float fc = 2.0f;
    if (flagged) {
      const Plane cand = i == 8 ? plane_now : to_plane(planes[cpos[i]]);
      const uint32_t k2 = i == 8 ? key + 700u : key + 600u + static_cast<uint32_t>(i) * 31u;
      fc = mc_cost<NV>(a, ca + static_cast<int64_t>(i) * n_src * pl, pl, vw, wnorm, cand, x, y, s0, s1, k2);
      if (smooth_w > 1e-6f) fc += smooth_w * smoothness_cost(a, nb, cand, x, y);
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
  float depth_now = depth_from_plane(static_cast<float>(x), static_cast<float>(y), plane_now, c0);

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
  uint32_t sel_out = prev_sel;
  if (((flags >> best) & 1u) && best_cost < cost_now) {
    const Plane bp = to_plane(planes[cpos[best]]);
    const float db = depth_from_plane(static_cast<float>(x), static_cast<float>(y), bp, c0);
    if (db >= a.depth_min && db <= a.depth_max) {
      depth_now = db;
      plane_now = bp;
      cost_now = best_cost;
      sel_out = new_sel;
      copy_costs_to_current(a, cp.cidx, best);
    }
  }

  // Step 4 set-up: the refinement hypotheses (:1669-1743).
  const float depth_perturbation = 0.005f;
  const float normal_perturbation = 0.01f * kPiF;
  const V3 cur_normal = normalize3(V3{plane_now.x, plane_now.y, plane_now.z});
  float depth_rand, depth_pert;
  V3 nrand, npert;
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
  const V3 sn = surface_normal(a, nb, x, y, depth_now);
  const int n_ref = len3(sn) > 0.5f ? 6 : 5;
  // The hypotheses are fixed here, before any of them is scored (:1733-1741).
  const float rd[6] = {depth_rand, depth_now, depth_rand, depth_now, depth_pert, depth_now};
  const V3 rn[6] = {cur_normal, nrand, nrand, npert, cur_normal, sn};
  PM_GLOBAL float4* refp = gpw(a.refp) + cp.cidx;
#pragma unroll
  for (int r = 0; r < 6; ++r)
    if (r < n_ref)
      refp[r * pl] = to_f4(plane_from_depth_normal(static_cast<float>(x), static_cast<float>(y), rd[r], rn[r], c0));
  gpw(a.nref)[cp.cidx] = n_ref;
  PM_GLOBAL float* vwb = gpw(a.vwb) + cp.cidx;
#pragma unroll
  for (int i = 0; i < NV; ++i)
    if (i < n_src) vwb[i * pl] = vw[i];
  gpw(a.wnb)[cp.cidx] = wnorm;
  gpw(a.planes)[idx] = to_f4(plane_now);
  gpw(a.costs)[idx] = cost_now;
  gpw(a.sel)[idx] = sel_out;
  gpw(a.rng)[idx] = make_uint2(s0, s1);
}

template <int HP, typename T, int TX>
__device__ __forceinline__ void refine_cost_body(const PmArgs& a, int b) {
  constexpr int P = (2 * HP + 1) * (2 * HP + 1);
  __shared__ float2 lt[P * kWave];
  int tx, ty;
  if (!colour_tile<TX>(a, b, tx, ty)) return;
  stage_tile_taps<P, TX>(a, tx, ty, lt);
  const int r = threadIdx.y;  // one wave per refinement hypothesis
  ColourPixel cp;
  uint32_t pending = 0u;
  if (colour_lane_pixel<TX>(a, tx, ty, cp) && r < gp(a.nref)[cp.cidx]) {
    const Plane hyp = to_plane(gp(a.refp)[r * a.ref_plane + cp.cidx]);
    float* out = a.ca + static_cast<int64_t>(r) * a.n_src * a.ref_plane + cp.cidx;
    pending = cost_vector<HP, T, true>(a, images_of<T>(a), cp.x, cp.y, cp.cidx, hyp, out, a.ref_plane,
                                       lt + threadIdx.x);
  }
  write_fb_masks(a, ty * colour_tiles_x<TX>(a) + tx, r, pending);
}

template <int NV, int TX>
__device__ __forceinline__ void refine_body(const PmArgs& a, int b) {
  ColourPixel cp;
  if (!colour_pixel<TX>(a, b, cp)) return;
  const int x = cp.x, y = cp.y, idx = cp.idx;
  const int64_t pl = a.ref_plane;
  const PmCam c0 = kcam(a, 0);
  const uint2 st = gp(a.rng)[idx];
  uint32_t s0 = st.x, s1 = st.y;
  const uint32_t key = static_cast<uint32_t>(idx + a.iteration * 7919 + 42);
  Plane plane_now = to_plane(gp(a.planes)[idx]);
  float cost_now = gp(a.costs)[idx];
  float vw[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) vw[i] = i < a.n_src ? gp(a.vwb)[i * pl + cp.cidx] : 0.0f;
  const float wnorm = gp(a.wnb)[cp.cidx];
  const int n_ref = gp(a.nref)[cp.cidx];
  Nbrs nb;
  load_nbrs(a, a.planes, x, y, nb);
  const float smooth_w = a.smooth_w;
  const float* ca = a.ca + cp.cidx;
  // Step 4: score the refinement hypotheses in order (:1744-1771).
  int win = -1;
  for (int r = 0; r < n_ref; ++r) {
    const Plane hyp = to_plane(gp(a.refp)[r * pl + cp.cidx]);
    float tc = mc_cost<NV>(a, ca + static_cast<int64_t>(r) * a.n_src * pl, pl, vw, wnorm, hyp, x, y, s0, s1,
                           key + 800u + static_cast<uint32_t>(r) * 37u);
    if (smooth_w > 1e-6f) tc += smooth_w * smoothness_cost(a, nb, hyp, x, y);
    const float db = depth_from_plane(static_cast<float>(x), static_cast<float>(y), hyp, c0);
    if (!(db < a.depth_min || db > a.depth_max) && tc < cost_now) {
      plane_now = hyp;
      cost_now = tc;
      win = r;
    }
  }
  if (win >= 0) copy_costs_to_current(a, cp.cidx, win);
  gpw(a.planes)[idx] = to_f4(plane_now);
  gpw(a.costs)[idx] = cost_now;
  gpw(a.rng)[idx] = make_uint2(s0, s1);
}

// ---------------------------------------------------------------------------
// acmmp_checkerboard_filter, one colour (opencl_kernels.h:1883-2039)
// ---------------------------------------------------------------------------
template <int TX>
__device__ __forceinline__ void filter_body(const PmArgs& a, int b) {
  const int W = a.width, H = a.height;
  ColourPixel cp;
  if (!colour_pixel<TX>(a, b, cp)) return;
  const int x = cp.x, y = cp.y;
  const int center = cp.idx;
  if (a.costs[center] < 0.001f) return;
  const PmCam c0 = kcam(a, 0);
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
// One thread per pixel, workgroups of kDecodeThreads; outputs a.out.
constexpr int kDecodeThreads = 256;
__device__ __forceinline__ void decode_body(const PmArgs& a, int b) {
  const PmDecodeOut& o = a.out;
  const int W = a.width, H = a.height;
  const int64_t idx = static_cast<int64_t>(b) * kDecodeThreads + threadIdx.x;
  if (idx >= static_cast<int64_t>(W) * H) return;
  const int x = static_cast<int>(idx % W), y = static_cast<int>(idx / W);
  const PmCam c0 = kcam(a, 0);
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
// Kernel entry points.  X_kernel runs one reference (arguments by value,
// workgroup index = blockIdx.x).  X_batch runs the references of a PmBatch
// in one launch, one workgroup per item of any reference: the workgroup
// finds its reference by binary search over the block offsets and loads its
// arguments through the constant address space (scalar loads: the item is
// uniform over the workgroup).  A workgroup does one item only: in a loop
// over items, the stores of one item would precede the next item's loads,
// and uniform loads from global memory that may follow a store cannot be
// scalar loads (this made a grid-stride variant 1.5x slower).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int batch_ref(const PmBatch& bt, int item) {
  const PM_CONST int* base = (const PM_CONST int*)(bt.base);
  int lo = 0, hi = bt.nrefs;  // base[lo] <= item < base[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (base[mid] <= item) lo = mid; else hi = mid;
  }
  return lo;
}

template <typename F>
__device__ __forceinline__ void batch_item(const PmBatch& bt, F&& body) {
  const PM_CONST int* base = (const PM_CONST int*)(bt.base);
  const int item = static_cast<int>(blockIdx.x);
  if (item >= bt.total) return;
  const int r = batch_ref(bt, item);
  const PmArgs a = ((const PM_CONST PmArgs*)(bt.args))[r];
  body(a, item - base[r]);
}

#define PM_BLOCK static_cast<int>(blockIdx.x)

// Occupancy targets (waves per SIMD) of the cost kernels and the fallback;
// 0 = the compiler's choice.  The fallback is compiled for at most
// PM_FALLBACK_WAVES_MAX waves per SIMD: its strided gathers need many loads
// in flight per wave, and more resident waves only compete for L1/L2.
#ifndef PM_COST_WAVES_PER_EU
#define PM_COST_WAVES_PER_EU 0
#endif
#ifndef PM_REFINE_COST_WAVES_PER_EU
#define PM_REFINE_COST_WAVES_PER_EU 0
#endif
#ifndef PM_FALLBACK_WAVES_PER_EU
#define PM_FALLBACK_WAVES_PER_EU 0
#endif
#ifndef PM_FALLBACK_WAVES_MAX
#define PM_FALLBACK_WAVES_MAX 3
#endif
#if PM_COST_WAVES_PER_EU
#define PM_COST_ATTR __attribute__((amdgpu_waves_per_eu(PM_COST_WAVES_PER_EU)))
#else
#define PM_COST_ATTR
#endif
#if PM_REFINE_COST_WAVES_PER_EU
#define PM_REFINE_COST_ATTR __attribute__((amdgpu_waves_per_eu(PM_REFINE_COST_WAVES_PER_EU)))
#else
#define PM_REFINE_COST_ATTR
#endif
#if PM_FALLBACK_WAVES_PER_EU
#define PM_FALLBACK_ATTR __attribute__((amdgpu_waves_per_eu(PM_FALLBACK_WAVES_PER_EU)))
#else
#define PM_FALLBACK_ATTR __attribute__((amdgpu_waves_per_eu(1, PM_FALLBACK_WAVES_MAX)))
#endif

template <int HP, typename T>
__global__ __launch_bounds__(kWave) void refpatch_kernel(PmArgs a) { refpatch_body<HP, T>(a, PM_BLOCK); }
template <int HP, typename T>
__global__ __launch_bounds__(kWave) void refpatch_batch(PmBatch bt) {
  batch_item(bt, [](const PmArgs& a, int b) { refpatch_body<HP, T>(a, b); });
}

template <int HP, typename T>
__global__ __launch_bounds__(kWave) void init_views_kernel(PmArgs a) { init_views_body<HP, T>(a, PM_BLOCK); }
template <int HP, typename T>
__global__ __launch_bounds__(kWave) void init_views_batch(PmBatch bt) {
  batch_item(bt, [](const PmArgs& a, int b) { init_views_body<HP, T>(a, b); });
}
template <int D = 0>
__global__ __launch_bounds__(kWave) void init_select_kernel(PmArgs a) { init_select_body(a, PM_BLOCK); }
template <int D = 0>
__global__ __launch_bounds__(kWave) void init_select_batch(PmBatch bt) {
  batch_item(bt, [](const PmArgs& a, int b) { init_select_body(a, b); });
}

template <int HP, typename T>
__global__ __launch_bounds__(kWave) void plane_sweep_kernel(PmArgs a) { plane_sweep_body<HP, T>(a, PM_BLOCK); }

template <int HP, typename T>
__global__ __launch_bounds__(kWave) void prior_kernel(PmArgs a) { prior_body<HP, T>(a, PM_BLOCK); }
template <int HP, typename T>
__global__ __launch_bounds__(kWave) void prior_batch(PmBatch bt) {
  batch_item(bt, [](const PmArgs& a, int b) { prior_body<HP, T>(a, b); });
}

template <int HP, typename T, int TX>
__global__ __launch_bounds__(kWave * 9) PM_COST_ATTR void candidates_kernel(PmArgs a) {
  candidates_body<HP, T, TX>(a, PM_BLOCK);
}
template <int HP, typename T, int TX>
__global__ __launch_bounds__(kWave * 9) PM_COST_ATTR void candidates_batch(PmBatch bt) {
  batch_item(bt, [](const PmArgs& a, int b) { candidates_body<HP, T, TX>(a, b); });
}

template <int HP, typename T, int TX>
__global__ __launch_bounds__(kWave * 6) PM_REFINE_COST_ATTR void refine_cost_kernel(PmArgs a) {
  refine_cost_body<HP, T, TX>(a, PM_BLOCK);
}
template <int HP, typename T, int TX>
__global__ __launch_bounds__(kWave * 6) PM_REFINE_COST_ATTR void refine_cost_batch(PmBatch bt) {
  batch_item(bt, [](const PmArgs& a, int b) { refine_cost_body<HP, T, TX>(a, b); });
}

template <int HP, typename T, int TX>
__global__ __launch_bounds__(kWave * kFbWaves) PM_FALLBACK_ATTR void fallback_kernel(PmArgs a) {
  fallback_tile_body<HP, T, TX>(a, PM_BLOCK);
}
template <int HP, typename T, int TX>
__global__ __launch_bounds__(kWave * kFbWaves) PM_FALLBACK_ATTR void fallback_batch(PmBatch bt) {
  batch_item(bt, [](const PmArgs& a, int b) { fallback_tile_body<HP, T, TX>(a, b); });
}

// Occupancy target of the select and refine kernels (waves per SIMD) for
// up to NV views.  3 takes select<16> from 173 to 168 VGPRs, so 3 waves fit
// per SIMD instead of 2 (aerial survey, 10 sources: 562 -> 446 ms per step
// of ten 4000x3000 references); at 4 it would spill 136 bytes per lane.
// Up to 12 views 4 waves fit with a spill of at most 44 bytes outside the
// MC draws: select 232 -> 221 ms per aerial step (10 sources, 12-view
// instance), 3.37 -> 3.17 ms per 1600x1200 step with 8 sources.
#ifndef PM_SELECT_WAVES_PER_EU
#define PM_SELECT_WAVES_PER_EU(NV) ((NV) <= 12 ? 4 : 3)
#endif
#define PM_SELECT_ATTR __attribute__((amdgpu_waves_per_eu(PM_SELECT_WAVES_PER_EU(NV))))

template <int NV, int TX>
__global__ __launch_bounds__(kWave) PM_SELECT_ATTR void select_kernel(PmArgs a) { select_body<NV, TX>(a, PM_BLOCK); }
template <int NV, int TX>
__global__ __launch_bounds__(kWave) PM_SELECT_ATTR void select_batch(PmBatch bt) {
  batch_item(bt, [](const PmArgs& a, int b) { select_body<NV, TX>(a, b); });
}

template <int NV, int TX>
__global__ __launch_bounds__(kWave) PM_SELECT_ATTR void refine_kernel(PmArgs a) { refine_body<NV, TX>(a, PM_BLOCK); }
template <int NV, int TX>
__global__ __launch_bounds__(kWave) PM_SELECT_ATTR void refine_batch(PmBatch bt) {
  batch_item(bt, [](const PmArgs& a, int b) { refine_body<NV, TX>(a, b); });
}

template <int TX>
__global__ __launch_bounds__(kWave) void filter_kernel(PmArgs a) { filter_body<TX>(a, PM_BLOCK); }
template <int TX>
__global__ __launch_bounds__(kWave) void filter_batch(PmBatch bt) {
  batch_item(bt, [](const PmArgs& a, int b) { filter_body<TX>(a, b); });
}

__global__ __launch_bounds__(kDecodeThreads) void decode_kernel(PmArgs a) { decode_body(a, PM_BLOCK); }
__global__ __launch_bounds__(kDecodeThreads) void decode_batch(PmBatch bt) {
  batch_item(bt, [](const PmArgs& a, int b) { decode_body(a, b); });
}

#undef PM_BLOCK

}  // namespace
}  // namespace pmhip
