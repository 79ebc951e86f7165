// Differentiable photometric cost of per-pixel plane hypotheses, its
// hand-written adjoint, and a Gauss-Newton refinement driven by it.
//
// Parameters of a pixel: theta = (d, a, b), depth d along the pixel's ray and
// the normal direction u = (a, b, -1) (camera facing, unnormalised).  The
// plane through the pixel's point with normal u gives the depths of the two
// forward neighbours as d_k = d * (u . r0) / (u . r_k), r_k = K^-1 (x_k, y_k, 1),
// and the same local affine warp as the PatchMatch cost (opencl_kernels.h:
// 401-461): a tap (i, j) lands at uv0 + i (uv1 - uv0) + j (uv2 - uv0) in the
// source, uv_k the projection of the three points.  Per selected source view
// the residual is 1 - NCC over the bilateral-weighted stride-1 patch (the
// fused NCC of opencl_kernels.h:471-499, without the textureless fallbacks,
// which are not differentiable).  The cost of a pixel is the mean residual
// over the views that are valid at theta.
//
// Derivatives are forward-mode with three tangents: per tap the bilinear
// gradient (exact for the software bilinear, zero where clamped) times the
// tap position's derivative, accumulated into the derivatives of the three
// NCC sums.  Pixels are independent, so the adjoint of the cost image is the
// per-pixel gradient scaled by the incoming gradient.
#include "pm_kernels.h"

namespace pmhip {

namespace {

__device__ __forceinline__ int colour_index_r(const PmArgs& a, int x, int y) {
  const int hw = (a.width + 1) / 2;
  return ((x + y) & 1) * (a.height * hw) + y * hw + (x >> 1);
}

// p = A (d r) + b for a camera-frame point d r of the reference: A = K_v R_v R0^T,
// b = K_v (t_v - R_v R0^T t0); per source view, 12 floats.
struct VGeom {
  float A[9], b[3];
};

struct Theta {
  float d, a, b;
};

struct ViewTerm {
  bool ok;
  float ncc;
  float dncc[3];  // d NCC / d(d, a, b); only with gradients
};

template <int HP, typename T, bool kGrad>
__device__ __forceinline__ ViewTerm ncc_view(const PmArgs& a, const Sampler& sm, const VGeom& g, int x, int y,
                                             int cidx, const float4& rs, const Theta& th) {
  ViewTerm out{false, 0.0f, {0.0f, 0.0f, 0.0f}};
  const PmCam c0 = ((const PM_CONST PmCam*)(a.cams))[0];
  const float fx = c0.K[0], fy = c0.K[4], cx = c0.K[2], cy = c0.K[5];
  // rays of the pixel and of its +x / +y neighbours
  const float r0x = (static_cast<float>(x) - cx) / fx, r0y = (static_cast<float>(y) - cy) / fy;
  const float r1x = (static_cast<float>(x + 1) - cx) / fx, r2y = (static_cast<float>(y + 1) - cy) / fy;
  const float ur0 = th.a * r0x + th.b * r0y - 1.0f;
  const float ur1 = th.a * r1x + th.b * r0y - 1.0f;
  const float ur2 = th.a * r0x + th.b * r2y - 1.0f;
  if (fabsf(ur0) < 1e-6f || fabsf(ur1) < 1e-6f || fabsf(ur2) < 1e-6f) return out;
  const float g1 = ur0 / ur1, g2 = ur0 / ur2;
  const float d0 = th.d, d1 = th.d * g1, d2 = th.d * g2;
  if (!(d0 > 0.0f && d1 > 0.0f && d2 > 0.0f)) return out;
  // projections p_k = A (d_k r_k) + b
  const float rx[3] = {r0x, r1x, r0x}, ry[3] = {r0y, r0y, r2y}, dk[3] = {d0, d1, d2};
  float u[3], v[3], eu[3], ev[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float qx = g.A[0] * rx[k] + g.A[1] * ry[k] + g.A[2];
    const float qy = g.A[3] * rx[k] + g.A[4] * ry[k] + g.A[5];
    const float qz = g.A[6] * rx[k] + g.A[7] * ry[k] + g.A[8];
    const float px = dk[k] * qx + g.b[0], py = dk[k] * qy + g.b[1], pz = dk[k] * qz + g.b[2];
    if (!(pz > 1e-6f)) return out;
    u[k] = px / pz;
    v[k] = py / pz;
    // d(u, v) / d d_k
    eu[k] = (qx - u[k] * qz) / pz;
    ev[k] = (qy - v[k] * qz) / pz;
  }
  const float jxx = u[1] - u[0], jxy = v[1] - v[0], jyx = u[2] - u[0], jyy = v[2] - v[0];
  const float det = jxx * jyy - jxy * jyx;
  if (det < 0.02f || det > 50.0f) return out;  // the cost's Jacobian gate (opencl_kernels.h:453-456)
  // d d_k / d theta
  float dd[3][3];
  dd[0][0] = 1.0f; dd[0][1] = 0.0f; dd[0][2] = 0.0f;
  dd[1][0] = g1; dd[1][1] = th.d * (r0x * ur1 - ur0 * r1x) / (ur1 * ur1); dd[1][2] = th.d * (r0y * ur1 - ur0 * r0y) / (ur1 * ur1);
  dd[2][0] = g2; dd[2][1] = th.d * (r0x * ur2 - ur0 * r0x) / (ur2 * ur2); dd[2][2] = th.d * (r0y * ur2 - ur0 * r2y) / (ur2 * ur2);
  // d(u_k, v_k) / d theta
  float du[3][3], dv[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int j = 0; j < 3; ++j) { du[k][j] = eu[k] * dd[k][j]; dv[k][j] = ev[k] * dd[k][j]; }

  const float inv_w = 1.0f / rs.y;
  const float mean_r = rs.z * inv_w;
  const float var_r = rs.w * inv_w - mean_r * mean_r;
  if (!(rs.y >= 1e-6f) || var_r < a.var_threshold) return out;

  // centre sample and its gradient (the source values are taken relative to it)
  float gcx = 0.0f, gcy = 0.0f;
  const float scen = sample_grad<T>(sm, u[0], v[0], gcx, gcy);
  float s_s = 0.0f, s_ss = 0.0f, s_rs = 0.0f;
  float ds_s[3] = {0, 0, 0}, ds_ss[3] = {0, 0, 0}, ds_rs[3] = {0, 0, 0};
  // Tap (i, j) sits at uv0 + i (uv1 - uv0) + j (uv2 - uv0): its tangent is
  // du0 + i (du1 - du0) + j (du2 - du0), one FMA per tap from a per-row base;
  // the centre sample's term is the same for every tap.
  float au[3], av[3], bu[3], bv[3], cterm[3];
  if constexpr (kGrad) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      au[j] = du[1][j] - du[0][j];
      av[j] = dv[1][j] - dv[0][j];
      bu[j] = du[2][j] - du[0][j];
      bv[j] = dv[2][j] - dv[0][j];
      cterm[j] = fmaf(gcx, du[0][j], gcy * dv[0][j]);
    }
  }
  constexpr int D = 2 * HP + 1;
  const PM_GLOBAL float2* taps = gp(a.ref_taps);
#pragma unroll 1
  for (int iy = -HP; iy <= HP; ++iy) {
    const float fiy = static_cast<float>(iy);
    float rowu[3], rowv[3];
    if constexpr (kGrad) {
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        rowu[j] = fmaf(fiy, bu[j], du[0][j]);
        rowv[j] = fmaf(fiy, bv[j], dv[0][j]);
      }
    }
#pragma unroll
    for (int ix = -HP; ix <= HP; ++ix) {
      const float fix = static_cast<float>(ix);
      const float sx = u[0] + fix * jxx + fiy * jyx;
      const float sy = v[0] + fix * jxy + fiy * jyy;
      float gx = 0.0f, gy = 0.0f;
      const float sd = sample_grad<T>(sm, sx, sy, gx, gy) - scen;
      const float2 wv = taps[static_cast<int64_t>((iy + HP) * D + (ix + HP)) * a.ref_plane + cidx];
      const float wsd = wv.x * sd;
      s_s += wsd;
      s_ss = fmaf(wsd, sd, s_ss);
      s_rs = fmaf(wv.y, sd, s_rs);
      if constexpr (kGrad) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const float dsx = fmaf(fix, au[j], rowu[j]);
          const float dsy = fmaf(fix, av[j], rowv[j]);
          const float dsd = fmaf(gx, dsx, gy * dsy) - cterm[j];
          ds_s[j] = fmaf(wv.x, dsd, ds_s[j]);
          ds_ss[j] = fmaf(2.0f * wsd, dsd, ds_ss[j]);
          ds_rs[j] = fmaf(wv.y, dsd, ds_rs[j]);
        }
      }
    }
  }
  const float mean_s = s_s * inv_w;
  const float var_s = s_ss * inv_w - mean_s * mean_s;
  if (var_s < a.var_threshold) return out;
  const float inv_den = 1.0f / sqrtf(var_r * var_s);
  const float cov = s_rs * inv_w - mean_r * mean_s;
  out.ok = true;
  out.ncc = cov * inv_den;
  if constexpr (kGrad) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const float dmean = ds_s[j] * inv_w;
      const float dvar = ds_ss[j] * inv_w - 2.0f * mean_s * dmean;
      const float dcov = ds_rs[j] * inv_w - mean_r * dmean;
      out.dncc[j] = dcov * inv_den - 0.5f * out.ncc * dvar / var_s;
    }
  }
  return out;
}

// Mean residual over the valid views of `views` and, with gradients, its
// derivative and the Gauss-Newton sums H = sum J J^T, g = sum J r (J = -dNCC).
// `good`: the views of `views` whose NCC is above 0.1 (those whose
// PatchMatch cost is 1 - NCC, opencl_kernels.h:526-529).  kGate: the sums
// take the good views only (the same sums, in the same order, as a call on
// `good`).
template <int HP, typename T, bool kGrad, bool kGate = false>
__device__ __forceinline__ float pixel_energy(const PmArgs& a, const VGeom* vg, int x, int y, uint32_t views,
                                              const Theta& th, float grad[3], float H[6], float gv[3], int& used,
                                              uint32_t& good) {
  const int cidx = colour_index_r(a, x, y);
  const float4 rs = gp(a.ref_sums)[cidx];
  float e = 0.0f;
  used = 0;
  good = 0u;
  if constexpr (kGrad) {
    for (int j = 0; j < 3; ++j) { grad[j] = 0.0f; gv[j] = 0.0f; }
    for (int j = 0; j < 6; ++j) H[j] = 0.0f;
  }
#pragma unroll 1
  for (int s = 0; s < a.n_src; ++s) {
    if (!((views >> s) & 1u)) continue;
    const PmView pv = gp(a.views)[s + 1];
    const Sampler sm{reinterpret_cast<const PM_GLOBAL char*>(gp(static_cast<const char*>(a.quads))) +
                         pv.qbase * QuadOf<T>::bytes,
                     static_cast<float>(pv.pitch), static_cast<float>(pv.pitch + 1), static_cast<float>(pv.w - 1),
                     static_cast<float>(pv.h - 1)};
    const ViewTerm t = ncc_view<HP, T, kGrad>(a, sm, vg[s], x, y, cidx, rs, th);
    if (!t.ok) continue;
    if (t.ncc > 0.1f) good |= 1u << s;
    else if (kGate) continue;
    ++used;
    const float r = 1.0f - t.ncc;
    e += r;
    if constexpr (kGrad) {
      const float J0 = -t.dncc[0], J1 = -t.dncc[1], J2 = -t.dncc[2];
      grad[0] += J0; grad[1] += J1; grad[2] += J2;
      H[0] += J0 * J0; H[1] += J0 * J1; H[2] += J0 * J2; H[3] += J1 * J1; H[4] += J1 * J2; H[5] += J2 * J2;
      gv[0] += J0 * r; gv[1] += J1 * r; gv[2] += J2 * r;
    }
  }
  if (used == 0) return 0.0f;
  const float inv = 1.0f / static_cast<float>(used);
  if constexpr (kGrad)
    for (int j = 0; j < 3; ++j) grad[j] *= inv;
  return e * inv;
}

// Both kernels are compiled for at most PM_REFINE_WAVES_MAX waves per SIMD:
// with the registers that frees, a thread keeps the gathers of a whole tap
// row and its tangent sums in flight (at 1024x768, N=6: 10 Gauss-Newton
// steps take 7.5 ms at 2 waves, 10.5 ms at the 3 the compiler picks).
#ifndef PM_REFINE_WAVES_MAX
#define PM_REFINE_WAVES_MAX 2
#endif
#define PM_REFINE_ATTR __attribute__((amdgpu_waves_per_eu(1, PM_REFINE_WAVES_MAX)))

// cost[p] and grad[p] = d cost / d(d, a, b) for every reference pixel.
template <int HP, typename T>
__global__ __launch_bounds__(256) PM_REFINE_ATTR void ncc_energy_kernel(PmArgs a, const VGeom* __restrict__ vg,
                                                        const float* __restrict__ depth, const float* __restrict__ ab,
                                                        const uint32_t* __restrict__ views, float* __restrict__ cost,
                                                        float* __restrict__ grad) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<int64_t>(a.width) * a.height) return;
  const int x = static_cast<int>(i % a.width), y = static_cast<int>(i / a.width);
  const Theta th{depth[i], ab[2 * i], ab[2 * i + 1]};
  float gr[3], H[6], gv[3];
  int used;
  uint32_t good;
  const float e = pixel_energy<HP, T, true>(a, vg, x, y, views[i], th, gr, H, gv, used, good);
  cost[i] = e;
  if (grad) {
    grad[3 * i] = gr[0];
    grad[3 * i + 1] = gr[1];
    grad[3 * i + 2] = gr[2];
  }
}

// Solves (H + lambda diag(H) + 1e-9 I) x = -g for symmetric 3x3 H (packed
// upper triangle h00 h01 h02 h11 h12 h22) by Cholesky.
__device__ __forceinline__ bool solve3(const float H[6], const float g[3], float lambda, float x[3]) {
  const float a00 = H[0] * (1.0f + lambda) + 1e-9f, a01 = H[1], a02 = H[2];
  const float a11 = H[3] * (1.0f + lambda) + 1e-9f, a12 = H[4];
  const float a22 = H[5] * (1.0f + lambda) + 1e-9f;
  if (!(a00 > 0.0f)) return false;
  const float l00 = sqrtf(a00);
  const float l10 = a01 / l00, l20 = a02 / l00;
  const float t11 = a11 - l10 * l10;
  if (!(t11 > 0.0f)) return false;
  const float l11 = sqrtf(t11);
  const float l21 = (a12 - l20 * l10) / l11;
  const float t22 = a22 - l20 * l20 - l21 * l21;
  if (!(t22 > 0.0f)) return false;
  const float l22 = sqrtf(t22);
  const float z0 = -g[0] / l00;
  const float z1 = (-g[1] - l10 * z0) / l11;
  const float z2 = (-g[2] - l20 * z0 - l21 * z1) / l22;
  x[2] = z2 / l22;
  x[1] = (z1 - l21 * x[2]) / l11;
  x[0] = (z0 - l10 * x[1] - l20 * x[2]) / l00;
  return true;
}

// One damped Gauss-Newton step per pixel with a view selection; a step is
// kept only where it lowers the pixel's cost.  The views are the pixel's
// selected views whose stride-1 NCC is above 0.1 at the PatchMatch plane:
// for those the PatchMatch cost is 1 - NCC (opencl_kernels.h:526-529), so
// the steps lower the PatchMatch cost itself; views whose cost comes from the
// textureless fallback (not differentiable) are left out, and a step that
// would send a kept view's NCC to 0.1 or below is refused.  The trial point
// is scored with its gradient: when the step is kept, that point, its cost
// and its normal equations are the next step's start (state st, kGnState
// planes of W*H floats: theta, e, used, H, g, views, PatchMatch depth), so a
// step costs one gradient evaluation.  A rejected step would be proposed
// again unchanged (fixed damping), so the pixel is done (used = -1).  The
// first step starts from the PatchMatch planes.  A step whose depth leaves
// kGnDepthRange of the PatchMatch depth ends the pixel's refinement: the
// refinement polishes within the range of PatchMatch's own depth perturbation
// (opencl_kernels.h:1671); a pixel whose photometric optimum lies further away
// is fitting its images' noise, which PatchMatch's smoothness held back.
constexpr int kGnState = 16;
constexpr float kGnDepthRange = 0.005f;

template <int HP, typename T>
__global__ __launch_bounds__(256) PM_REFINE_ATTR void gauss_newton_kernel(PmArgs a, const VGeom* __restrict__ vg,
                                                                          float lambda, float* __restrict__ st,
                                                                          int first) {
  const int64_t n = static_cast<int64_t>(a.width) * a.height;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = static_cast<int>(i % a.width), y = static_cast<int>(i / a.width);
  uint32_t views = a.sel[i];
  if (!views) return;
  float* const s_used = st + 4 * n + i;
  float* const s_views = st + 14 * n + i;
  float* const s_dpm = st + 15 * n + i;
  const PmCam c0 = ((const PM_CONST PmCam*)(a.cams))[0];
  Theta th;
  float e0, H[6], gv[3];
  int used;
  if (first) {
    *s_used = __int_as_float(-1);
    const float4 p = a.planes[i];
    if (!(p.z < -1e-6f)) return;
    const float d = depth_from_plane(static_cast<float>(x), static_cast<float>(y), Plane{p.x, p.y, p.z, p.w}, c0);
    if (!(d > 0.0f)) return;
    th = Theta{d, p.x / -p.z, p.y / -p.z};
    float gr[3];
    uint32_t good;
    // keep the views whose PatchMatch cost is 1 - NCC
    e0 = pixel_energy<HP, T, true, true>(a, vg, x, y, views, th, gr, H, gv, used, good);
    views = good;
    if (!views) return;
    *s_views = __uint_as_float(views);
    *s_dpm = d;
  } else {
    used = __float_as_int(*s_used);
    if (used <= 0) return;
    views = __float_as_uint(*s_views);
    th = Theta{st[i], st[n + i], st[2 * n + i]};
    e0 = st[3 * n + i];
#pragma unroll
    for (int j = 0; j < 6; ++j) H[j] = st[(5 + j) * n + i];
#pragma unroll
    for (int j = 0; j < 3; ++j) gv[j] = st[(11 + j) * n + i];
  }
  float step[3];
  if (used == 0 || !solve3(H, gv, lambda, step)) {
    *s_used = __int_as_float(-1);
    return;
  }
  Theta t1{th.d + step[0], th.a + step[1], th.b + step[2]};
  const float dpm = *s_dpm;
  t1.d = fminf(fmaxf(t1.d, a.depth_min), a.depth_max);
  if (!(fabsf(t1.d - dpm) <= kGnDepthRange * dpm)) {  // not a local polish: the pixel keeps its plane
    *s_used = __int_as_float(-1);
    return;
  }
  float gr1[3], H1[6], gv1[3];
  int used1;
  uint32_t good1;
  const float e1 = pixel_energy<HP, T, true>(a, vg, x, y, views, t1, gr1, H1, gv1, used1, good1);
  if (used1 != used || good1 != views || !(e1 < e0)) {
    *s_used = __int_as_float(-1);
    return;
  }
  const V3 nv = normalize3(V3{t1.a, t1.b, -1.0f});
  const Plane np = plane_from_depth_normal(static_cast<float>(x), static_cast<float>(y), t1.d, nv, c0);
  a.planes[i] = make_float4(np.x, np.y, np.z, np.w);
  st[i] = t1.d;
  st[n + i] = t1.a;
  st[2 * n + i] = t1.b;
  st[3 * n + i] = e1;
  *s_used = __int_as_float(used1);
#pragma unroll
  for (int j = 0; j < 6; ++j) st[(5 + j) * n + i] = H1[j];
#pragma unroll
  for (int j = 0; j < 3; ++j) st[(11 + j) * n + i] = gv1[j];
}

}  // namespace

#define PM_REFINE_DISPATCH(KERNEL, HP, F16, GRID, ...)                                         \
  do {                                                                                           \
    switch ((HP) * 2 + ((F16) ? 1 : 0)) {                                                        \
      case 2: hipLaunchKernelGGL((KERNEL<1, float>), GRID, dim3(256), 0, stream, __VA_ARGS__); break;     \
      case 3: hipLaunchKernelGGL((KERNEL<1, _Float16>), GRID, dim3(256), 0, stream, __VA_ARGS__); break;  \
      case 4: hipLaunchKernelGGL((KERNEL<2, float>), GRID, dim3(256), 0, stream, __VA_ARGS__); break;     \
      case 5: hipLaunchKernelGGL((KERNEL<2, _Float16>), GRID, dim3(256), 0, stream, __VA_ARGS__); break;  \
      case 6: hipLaunchKernelGGL((KERNEL<3, float>), GRID, dim3(256), 0, stream, __VA_ARGS__); break;     \
      case 7: hipLaunchKernelGGL((KERNEL<3, _Float16>), GRID, dim3(256), 0, stream, __VA_ARGS__); break;  \
      default: return hipErrorInvalidValue;                                                      \
    }                                                                                            \
    return hipGetLastError();                                                                    \
  } while (0)

hipError_t launch_ncc_energy(const PmArgs& a, int hp, bool f16, const float* vgeom, const float* depth,
                             const float* ab, const uint32_t* views, float* cost, float* grad, hipStream_t stream) {
  const int64_t n = static_cast<int64_t>(a.width) * a.height;
  const dim3 grid(static_cast<unsigned>((n + 255) / 256));
  const VGeom* vg = reinterpret_cast<const VGeom*>(vgeom);
  PM_REFINE_DISPATCH(ncc_energy_kernel, hp, f16, grid, a, vg, depth, ab, views, cost, grad);
}

int gauss_newton_state_floats() { return kGnState; }

hipError_t launch_gauss_newton(const PmArgs& a, int hp, bool f16, const float* vgeom, float lambda, float* state,
                               bool first, hipStream_t stream) {
  const int64_t n = static_cast<int64_t>(a.width) * a.height;
  const dim3 grid(static_cast<unsigned>((n + 255) / 256));
  const VGeom* vg = reinterpret_cast<const VGeom*>(vgeom);
  const int fst = first ? 1 : 0;
  PM_REFINE_DISPATCH(gauss_newton_kernel, hp, f16, grid, a, vg, lambda, state, fst);
}

}  // namespace pmhip
