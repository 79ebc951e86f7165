// Device-side building blocks of the PatchMatch path for gfx950.
//
// Numerics contract (shared with the CPU oracle so the two agree bit for
// bit): the library is compiled with -ffp-contract=off, fused multiply-adds
// appear only as explicit fmaf() (bilinear lerps, warp coordinates, NCC
// accumulations), exp/log are pm_expf/pm_logf below, sqrt and '/' are IEEE
// correctly rounded, and bilinear sampling is exact software bilinear with
// texel centres at integer coordinates and clamp-to-edge addressing.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pmhip {

constexpr int kMaxSrc = 16;
constexpr int kMcSamples = 32;                       // opencl_kernels.h:910
constexpr float kPiF = 3.14159274101257324219f;     // M_PI_F

// Camera as packed by the reference host (opencl_utils.h:258-280).
struct PmCam {
  float K[9];
  float R[9];
  float t[3];
  int width;
  int height;
};
static_assert(sizeof(PmCam) == 92, "camera layout");

// Per-view image descriptor.  Images are stored with a one-texel replicated
// border (clamp-to-edge baked in), so texel (x, y) of a w x h view lives at
// base + (y + 1) * pitch + (x + 1), pitch = w + 2.
struct PmView {
  int64_t base;      // element offset of texel (-1, -1)
  int64_t qbase;     // the same in the quad copy (see Sampler)
  int w, h, pitch;
  float cx, cy, cz;  // camera centre -R^T t (sources; opencl_kernels.h:1523-1525)
  float angle_w;     // estimator.cc:380-407
};

struct Plane { float x, y, z, w; };
struct V3 { float x, y, z; };

__device__ __forceinline__ float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float len3(V3 a) { return sqrtf(dot3(a, a)); }
__device__ __forceinline__ V3 normalize3(V3 a) {
  const float l = len3(a);
  if (l == 0.0f) return a;
  return V3{a.x / l, a.y / l, a.z / l};
}
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// e^x: Cody-Waite reduction by ln2, degree-7 Taylor polynomial, 2^k via the
// exponent field; x < -86 -> 0, x > 88 -> +inf.
// Branch-free on the device: the polynomial runs on x clamped to the
// finite range (identical for every x that is not special-cased) and the
// special cases are selected afterwards.
__device__ __forceinline__ float pm_expf(float x) {
  const float xc = fminf(fmaxf(x, -87.0f), 89.0f);
  const float kf = __builtin_rintf(xc * 1.44269502162933349609f);
  float r = fmaf(-kf, 0.693145751953125f, xc);
  r = fmaf(-kf, 1.42860682030941723212e-6f, r);
  float p = 1.98412698412698412698e-4f;
  p = fmaf(p, r, 1.38888888888888888889e-3f);
  p = fmaf(p, r, 8.33333333333333333333e-3f);
  p = fmaf(p, r, 4.16666666666666666667e-2f);
  p = fmaf(p, r, 1.66666666666666666667e-1f);
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  const int k = static_cast<int>(kf);
  float res = p * __uint_as_float(static_cast<uint32_t>(k + 127) << 23);
  res = (x < -86.0f) ? 0.0f : res;
  res = (x > 88.0f) ? __builtin_inff() : res;
  return (x != x) ? x : res;
}

// pm_expf for finite x <= 0 (bilateral weights): the same value as
// pm_expf(x) for every such x -- the upper clamp and the overflow and NaN
// cases cannot occur, so they are left out.
__device__ __forceinline__ float pm_expf_nonpos(float x) {
  const float xc = fmaxf(x, -87.0f);
  const float kf = __builtin_rintf(xc * 1.44269502162933349609f);
  float r = fmaf(-kf, 0.693145751953125f, xc);
  r = fmaf(-kf, 1.42860682030941723212e-6f, r);
  float p = 1.98412698412698412698e-4f;
  p = fmaf(p, r, 1.38888888888888888889e-3f);
  p = fmaf(p, r, 8.33333333333333333333e-3f);
  p = fmaf(p, r, 4.16666666666666666667e-2f);
  p = fmaf(p, r, 1.66666666666666666667e-1f);
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  const int k = static_cast<int>(kf);
  const float res = p * __uint_as_float(static_cast<uint32_t>(k + 127) << 23);
  return (x < -86.0f) ? 0.0f : res;
}

// ln x, x > 0: ln m = 2 atanh((m-1)/(m+1)) with m in [sqrt(1/2), sqrt(2)).
__device__ __forceinline__ float pm_logf(float x) {
  if (!(x > 0.0f)) return x == 0.0f ? -__builtin_inff() : __builtin_nanf("");
  if (x == __builtin_inff()) return x;
  uint32_t u = __float_as_uint(x);
  int e = static_cast<int>((u >> 23) & 0xffu) - 127;
  if (e == -127) {
    x = x * 8388608.0f;
    u = __float_as_uint(x);
    e = static_cast<int>((u >> 23) & 0xffu) - 127 - 23;
  }
  float m = __uint_as_float((u & 0x7fffffu) | 0x3f800000u);
  if (m > 1.41421353816986083984f) { m = m * 0.5f; e += 1; }
  const float f = m - 1.0f;
  const float s = f / (2.0f + f);
  const float s2 = s * s;
  float p = 9.09090909090909090909e-2f;
  p = fmaf(p, s2, 1.11111111111111111111e-1f);
  p = fmaf(p, s2, 1.42857142857142857143e-1f);
  p = fmaf(p, s2, 2.0e-1f);
  p = fmaf(p, s2, 3.33333333333333333333e-1f);
  p = fmaf(p, s2, 1.0f);
  const float lm = (2.0f * s) * p;
  const float fe = static_cast<float>(e);
  return fmaf(fe, 0.693145751953125f, fmaf(fe, 1.42860682030941723212e-6f, lm));
}

// Counter-based per-pixel RNG (opencl_kernels.h:78-97): 4 Philox-style
// rounds, multiplier 0xD256D193, round keys key + i * 0x9E3779B9.
__device__ __forceinline__ float rand_float(uint32_t& s0, uint32_t& s1, uint32_t key) {
#pragma unroll
  for (uint32_t i = 0; i < 4; ++i) {
    // one 32 x 32 -> 64 product gives both halves (a single v_mad_u64_u32)
    const uint64_t prod = static_cast<uint64_t>(s0) * 0xD256D193u;
    const uint32_t lo = static_cast<uint32_t>(prod);
    const uint32_t hi = static_cast<uint32_t>(prod >> 32);
    s0 = s1 ^ (key + i * 0x9E3779B9u) ^ hi;
    s1 = lo;
  }
  return static_cast<float>(s0) / 4294967296.0f;
}

// Geometry (opencl_kernels.h:102-173).
__device__ __forceinline__ V3 project(V3 X, const PmCam& c) {
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

__device__ __forceinline__ V3 backproject(float x, float y, float depth, const PmCam& c) {
  const float fx = c.K[0], fy = c.K[4], cx = c.K[2], cy = c.K[5];
  V3 pc{depth * (x - cx) / fx, depth * (y - cy) / fy, depth};
  V3 d{pc.x - c.t[0], pc.y - c.t[1], pc.z - c.t[2]};
  V3 w;
  w.x = c.R[0] * d.x + c.R[3] * d.y + c.R[6] * d.z;
  w.y = c.R[1] * d.x + c.R[4] * d.y + c.R[7] * d.z;
  w.z = c.R[2] * d.x + c.R[5] * d.y + c.R[8] * d.z;
  return w;
}

__device__ __forceinline__ float depth_from_plane(float x, float y, Plane p, const PmCam& c) {
  const float fx = c.K[0], fy = c.K[4], cx = c.K[2], cy = c.K[5];
  const float denom = p.x * (x - cx) / fx + p.y * (y - cy) / fy + p.z;
  if (fabsf(denom) < 1e-6f) return 0.0f;
  return -p.w / denom;
}

__device__ __forceinline__ Plane plane_from_depth_normal(float x, float y, float depth, V3 n, const PmCam& c) {
  const float fx = c.K[0], fy = c.K[4], cx = c.K[2], cy = c.K[5];
  V3 pt{depth * (x - cx) / fx, depth * (y - cy) / fy, depth};
  const float d = -dot3(n, pt);
  return Plane{n.x, n.y, n.z, d};
}

// Global-address-space view of a pointer: loads become global_load_*
// instead of flat loads (which also tie up the LDS counter).
#if defined(__HIP_DEVICE_COMPILE__)
#define PM_GLOBAL __attribute__((address_space(1)))
#define PM_CONST __attribute__((address_space(4)))
#else
#define PM_GLOBAL
#define PM_CONST
#endif
template <typename T>
__device__ __forceinline__ const PM_GLOBAL T* gp(const T* p) { return (const PM_GLOBAL T*)(p); }
template <typename T>
__device__ __forceinline__ PM_GLOBAL T* gpw(T* p) { return (PM_GLOBAL T*)(p); }

// Texel reads on the padded layout.  Byte offsets inside one view are
// 32-bit, so a load is a 64-bit view base plus a 32-bit per-lane offset.
template <typename T>
__device__ __forceinline__ const PM_GLOBAL char* view_base(const T* img, const PmView& v) {
  return reinterpret_cast<const PM_GLOBAL char*>(gp(img) + v.base);
}

template <typename T>
__device__ __forceinline__ float texel(const T* __restrict__ img, const PmView& v, int ix, int iy) {
  ix = min(max(ix, 0), v.w - 1);
  iy = min(max(iy, 0), v.h - 1);
  const uint32_t o = (__umul24(static_cast<uint32_t>(iy + 1), static_cast<uint32_t>(v.pitch)) +
                      static_cast<uint32_t>(ix + 1)) * static_cast<uint32_t>(sizeof(T));
  T t;
  __builtin_memcpy(&t, view_base(img, v) + o, sizeof(T));
  return static_cast<float>(t);
}

// Bilinear sampling goes through a second copy of every view in which
// texel (x, y) of the padded layout holds its 2 x 2 neighbourhood
// (T[y][x], T[y][x+1], T[y+1][x], T[y+1][x+1]): one aligned 16-byte (f32) or
// 8-byte (f16) load per bilinear sample instead of two unaligned pair loads,
// which halves the work of the texture address unit -- the bottleneck of
// the gather-bound cost kernels.  4x the image memory, a few hundred MB at
// most.  f32 copies keep the horizontal differences instead of the right-hand
// texels, (T[y][x], T[y][x+1] - T[y][x], T[y+1][x], T[y+1][x+1] - T[y+1][x]),
// which are what the lerps use.
template <typename T> struct QuadOf;
template <> struct QuadOf<float> { static constexpr uint32_t bytes = 16; };
template <> struct QuadOf<_Float16> { static constexpr uint32_t bytes = 8; };

// A view prepared for repeated bilinear sampling.
struct Sampler {
  const PM_GLOBAL char* base;  // quad copy of the view (texel (-1, -1) first)
  float pitch_f;               // pitch in texels
  float origin_f;              // index of texel (0, 0): pitch + 1
  float xmax, ymax;            // w - 1, h - 1
};

// A bilinear quad as (t00, d0, t10, d1) with d0 = t01 - t00 and d1 = t11 - t10,
// each difference rounded once to f32.  f32 quads hold this form already.
__device__ __forceinline__ void ld_quad(const PM_GLOBAL char* p, float& t00, float& d0, float& t10, float& d1,
                                        float) {
  const float4 q = *reinterpret_cast<const PM_GLOBAL float4*>(p);
  t00 = q.x; d0 = q.y; t10 = q.z; d1 = q.w;
}
// hi - lo of two packed f16 values as one v_fma_mix_f32 (hi * 1.0 + -lo):
// one rounding of the exact difference, the same f32 value as converting
// both halves and subtracting, in one instruction instead of three.
__device__ __forceinline__ float f16_hi_minus_lo(uint32_t v) {
  float r;
  asm("v_fma_mix_f32 %0, %1, 1.0, -%1 op_sel:[1,0,0] op_sel_hi:[1,0,1]" : "=v"(r) : "v"(v));
  return r;
}
__device__ __forceinline__ void ld_quad(const PM_GLOBAL char* p, float& t00, float& d0, float& t10, float& d1,
                                        _Float16) {
  const uint2 q = *reinterpret_cast<const PM_GLOBAL uint2*>(p);
  t00 = static_cast<float>(__builtin_bit_cast(_Float16, static_cast<unsigned short>(q.x & 0xffffu)));
  d0 = f16_hi_minus_lo(q.x);
  t10 = static_cast<float>(__builtin_bit_cast(_Float16, static_cast<unsigned short>(q.y & 0xffffu)));
  d1 = f16_hi_minus_lo(q.y);
}

// Exact software bilinear at a pixel-centre coordinate with clamp-to-edge:
// clamp to [-1, w-1] x [-1, h-1] (outside that range both taps of an axis
// are the same edge texel, so lerp(a, a, t) == a keeps the clamp-to-edge
// value; one med3 per axis, coordinates reaching here are finite), floor,
// lerp along x on both rows, then along y.  The quad index is formed in float, exact
// because it is an integer below 2^24 (add_view refuses views of more padded texels).
// kClamp = false: the caller has checked that (sx, sy) lies in
// [-1, w-1] x [-1, h-1], where the clamp changes nothing.
template <typename T, bool kClamp = true>
__device__ __forceinline__ float sample(const Sampler& v, float sx, float sy) {
  if constexpr (kClamp) {
    sx = __builtin_amdgcn_fmed3f(sx, -1.0f, v.xmax);
    sy = __builtin_amdgcn_fmed3f(sy, -1.0f, v.ymax);
  }
  const float fx0 = floorf(sx), fy0 = floorf(sy);
  const float ax = sx - fx0, ay = sy - fy0;
  const uint32_t qi = static_cast<uint32_t>(fmaf(fy0, v.pitch_f, fx0 + v.origin_f));
  float t00, d0, t10, d1;
  ld_quad(v.base + qi * QuadOf<T>::bytes, t00, d0, t10, d1, T{});
  const float v0 = fmaf(ax, d0, t00);
  const float v1 = fmaf(ax, d1, t10);
  return fmaf(ay, v1 - v0, v0);
}

// The same bilinear sample and its gradient d/d(sx, sy).  Outside the
// clamp range the padded border repeats the edge texel, so the gradient
// along a clamped axis comes out zero.
template <typename T>
__device__ __forceinline__ float sample_grad(const Sampler& v, float sx, float sy, float& gx, float& gy) {
  sx = __builtin_amdgcn_fmed3f(sx, -1.0f, v.xmax);
  sy = __builtin_amdgcn_fmed3f(sy, -1.0f, v.ymax);
  const float fx0 = floorf(sx), fy0 = floorf(sy);
  const float ax = sx - fx0, ay = sy - fy0;
  const uint32_t qi = static_cast<uint32_t>(fmaf(fy0, v.pitch_f, fx0 + v.origin_f));
  float t00, d0, t10, d1;
  ld_quad(v.base + qi * QuadOf<T>::bytes, t00, d0, t10, d1, T{});
  const float v0 = fmaf(ax, d0, t00);
  const float v1 = fmaf(ax, d1, t10);
  gx = fmaf(ay, d1 - d0, d0);
  gy = v1 - v0;
  return fmaf(ay, v1 - v0, v0);
}

}  // namespace pmhip
