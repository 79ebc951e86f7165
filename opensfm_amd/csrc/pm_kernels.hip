// Launchers of the PatchMatch kernels (single reference and batched), and
// the set-up kernels: quad copy of the views, coarse-to-fine transfer,
// confidence, scoring of given planes.  The kernels templated on the image
// storage type are instantiated in pm_cost_f32.hip / pm_cost_f16.hip.
#include <algorithm>
#include <atomic>

#include "pm_body.h"

namespace pmhip {

namespace {

// Quad copy of the padded views (see Sampler): one lane per padded texel.
// f32 views store each quad as (t00, t01 - t00, t10, t11 - t10): the two
// differences are the f32 subtractions the bilinear lerp would otherwise
// do per sample.  f16 views store the four texels.
template <typename T>
__device__ __forceinline__ void put_quad(PM_GLOBAL T* q, T t00, T t01, T t10, T t11) {
  q[0] = t00; q[1] = t01; q[2] = t10; q[3] = t11;
}
template <>
__device__ __forceinline__ void put_quad(PM_GLOBAL float* q, float t00, float t01, float t10, float t11) {
  *reinterpret_cast<PM_GLOBAL float4*>(q) = make_float4(t00, t01 - t00, t10, t11 - t10);
}

template <typename T>
__global__ __launch_bounds__(256) void quad_kernel(PmArgs a, int nviews) {
  for (int vi = 0; vi < nviews; ++vi) {
    const PmView v = gp(a.views)[vi];
    const int64_t n = static_cast<int64_t>(v.pitch) * (v.h + 2);
    const PM_GLOBAL T* src = gp(static_cast<const T*>(a.images)) + v.base;
    PM_GLOBAL T* dst = gpw(static_cast<T*>(a.quads)) + v.qbase * 4;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
      const int x = static_cast<int>(i % v.pitch), y = static_cast<int>(i / v.pitch);
      const int x1 = min(x + 1, v.pitch - 1), y1 = min(y + 1, v.h + 1);
      put_quad(dst + 4 * i, src[static_cast<int64_t>(y) * v.pitch + x], src[static_cast<int64_t>(y) * v.pitch + x1],
               src[static_cast<int64_t>(y1) * v.pitch + x], src[static_cast<int64_t>(y1) * v.pitch + x1]);
    }
  }
}

// Coarse-to-fine transfer between hierarchy levels (acmmp_upsample,
// opencl_kernels.h:1792-1856, with the coarse planes that
// UpsampleDepthNormal builds from the decoded coarse depth and normal,
// estimator.cc:775-791).  The nearest coarse pixel's plane is re-anchored
// at the fine pixel; where it gives no depth the pixel gets a random
// plane.  Costs become 2.0 so that the fine level re-evaluates every plane;
// the random streams are re-seeded.
__global__ __launch_bounds__(256) void upsample_kernel(PmArgs a, const float* __restrict__ cdepth,
                                                       const float* __restrict__ cnormal, int src_w, int src_h) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<int64_t>(a.width) * a.height) return;
  const int x = static_cast<int>(i % a.width), y = static_cast<int>(i / a.width);
  const int idx = static_cast<int>(i);
  uint32_t s0 = static_cast<uint32_t>(idx) * 1099087573u + 2654435769u;
  uint32_t s1 = static_cast<uint32_t>(idx) * 2246822519u + 13u;
  const PmCam c0 = kcam(a, 0);
  const float scale_x = static_cast<float>(src_w) / static_cast<float>(a.width);
  const float scale_y = static_cast<float>(src_h) / static_cast<float>(a.height);
  const int sx = clampi(static_cast<int>(static_cast<float>(x) * scale_x), 0, src_w - 1);
  const int sy = clampi(static_cast<int>(static_cast<float>(y) * scale_y), 0, src_h - 1);
  const int sidx = sy * src_w + sx;
  // coarse plane through the decoded coarse point (estimator.cc:775-791)
  const float d = cdepth[sidx];
  const V3 cn{cnormal[3 * sidx], cnormal[3 * sidx + 1], cnormal[3 * sidx + 2]};
  const float fx = c0.K[0] * static_cast<float>(src_w) / static_cast<float>(a.width);
  const float fy = c0.K[4] * static_cast<float>(src_h) / static_cast<float>(a.height);
  const float cxs = c0.K[2] * static_cast<float>(src_w) / static_cast<float>(a.width);
  const float cys = c0.K[5] * static_cast<float>(src_h) / static_cast<float>(a.height);
  const V3 pt{d * (static_cast<float>(sx) - cxs) / fx, d * (static_cast<float>(sy) - cys) / fy, d};
  const Plane cp{cn.x, cn.y, cn.z, -(cn.x * pt.x + cn.y * pt.y + cn.z * pt.z)};
  const float depth = depth_from_plane(static_cast<float>(x), static_cast<float>(y), cp, c0);
  Plane out;
  if (depth > 0.0f) {
    out = plane_from_depth_normal(static_cast<float>(x), static_cast<float>(y), depth,
                                  normalize3(V3{cp.x, cp.y, cp.z}), c0);
  } else {
    out = random_plane(a, x, y, s0, s1, static_cast<uint32_t>(idx + 7919));
  }
  a.planes[idx] = to_f4(out);
  a.costs[idx] = 2.0f;
  a.rng[idx] = make_uint2(s0, s1);
  a.sel[idx] = 0u;
}

// Confidence after the depth map was post-processed on the host
// (ReadBackResults computes it last, estimator.cc:1211-1222).
__global__ __launch_bounds__(256) void confidence_kernel(int64_t n, const float* __restrict__ depth,
                                                         const float* __restrict__ cost, float* __restrict__ conf) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  conf[i] = depth[i] > 0.0f ? 1.0f / (1.0f + pm_expf((cost[i] - 0.5f) * 5.0f)) : 0.0f;
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
  float* cv = smem + lane;
  float* tmp = smem + a.n_src * kWave + lane;
  full_cost_vector<HP>(a, img, x, y, colour_index(a, x, y), to_plane(pl[i]), cv);
  for (int s = 0; s < a.n_src; ++s) costs[static_cast<int64_t>(i) * a.n_src + s] = cv[s * kWave];
  uint32_t sv = 0u;
  const float ic = initial_cost_and_views(a, cv, tmp, &sv);
  if (init_cost) init_cost[i] = ic;
  if (sel) sel[i] = sv;
}


// ---------------------------------------------------------------------------
// Guard of the Gauss-Newton refinement.  The steps lower the stride-1
// photometric cost of one pixel's views whose NCC is above 0.1, pixel by
// pixel; the PatchMatch objective also has the strided and census fallbacks
// of every selected view (opencl_kernels.h:381-557) and the smoothness term
// against the neighbours' planes (:649-707).  Per pixel whose plane the
// steps moved, the refined plane is kept only if that objective -- the mean
// full cost over the selected views plus smooth_w times the smoothness cost
// against the neighbours' PatchMatch planes -- does not exceed the one of
// the PatchMatch plane `orig`.  (Without the smoothness term, a step that
// fits the images' noise moves a pixel away from its neighbours' surface.)
// ---------------------------------------------------------------------------
// The guard runs on the moved pixels only, compacted into a list first: they
// are a minority of the pixels, and one moved pixel in a wave would
// otherwise hold the whole wave for its full cost vectors.  The list is in pixel order (a count
// per block of 256 pixels, an exclusive prefix of the counts, then each
// block writes its moved pixels from its offset), so the 64 entries a guard
// wave takes are neighbouring pixels whose patches share cache lines.
// list[], the block offsets and *count live in the refinement's state
// buffer, free once its steps are done.
// One-wave workgroups looping over the list, 32 per CU: 8 waves per SIMD hide
// the gathers of the full cost vectors (Gauss-Newton with its guard, ten
// 4000x3000 references: 535 ms at 2048 workgroups, 524 ms at 8192).
constexpr unsigned kGuardWaves = 8192;
constexpr int kGuardBlock = 256;

__device__ __forceinline__ bool plane_moved(const float4* __restrict__ planes, const float4* __restrict__ orig,
                                            int64_t i) {
  const float4 pn = planes[i], po = orig[i];
  return !(pn.x == po.x && pn.y == po.y && pn.z == po.z && pn.w == po.w);
}

__global__ __launch_bounds__(kGuardBlock) void guard_count_kernel(const float4* __restrict__ planes,
                                                                  const float4* __restrict__ orig, int64_t n,
                                                                  int* offs) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kGuardBlock + threadIdx.x;
  const int c = __syncthreads_count(i < n && plane_moved(planes, orig, i));
  if (threadIdx.x == 0) offs[blockIdx.x] = c;
}

// Exclusive prefix of offs[0..nb) in place (one workgroup: a serial run per
// thread over its share, a scan of the 1024 shares in LDS); the total goes
// to *count.
__global__ __launch_bounds__(1024) void guard_scan_kernel(int* offs, int nb, int* count) {
  __shared__ int part[1024];
  const int t = static_cast<int>(threadIdx.x);
  const int per = (nb + 1023) / 1024;
  const int lo = min(t * per, nb), hi = min(lo + per, nb);
  int own = 0;
  for (int k = lo; k < hi; ++k) own += offs[k];
  part[t] = own;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    const int v = t >= d ? part[t - d] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - own;
  for (int k = lo; k < hi; ++k) {
    const int c = offs[k];
    offs[k] = run;
    run += c;
  }
  if (t == 1023) *count = part[1023];
}

__global__ __launch_bounds__(kGuardBlock) void guard_list_kernel(const float4* __restrict__ planes,
                                                                 const float4* __restrict__ orig, int64_t n,
                                                                 const int* __restrict__ offs, int* list) {
  __shared__ int wave_count[kGuardBlock / kWave];
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kGuardBlock + threadIdx.x;
  const bool moved = i < n && plane_moved(planes, orig, i);
  const uint64_t m = __ballot(moved);
  const int lane = static_cast<int>(threadIdx.x % kWave), w = static_cast<int>(threadIdx.x / kWave);
  if (lane == 0) wave_count[w] = __popcll(m);
  __syncthreads();
  int base = offs[blockIdx.x];
  for (int k = 0; k < w; ++k) base += wave_count[k];
  if (moved) list[base + __popcll(m & ((1ull << lane) - 1ull))] = static_cast<int>(i);
}

template <int HP, typename T>
// orig_costs, when not null: the per-view full costs of the PatchMatch
// planes (slot 8 of ca, [view][cidx]), which then need not be scored again.
__global__ __launch_bounds__(kWave) void refine_guard_kernel(PmArgs a, const float4* __restrict__ orig,
                                                             const float* __restrict__ orig_costs,
                                                             const int* __restrict__ list,
                                                             const int* __restrict__ count) {
  extern __shared__ float smem[];
  const int lane = threadIdx.x;
  const int total = __builtin_amdgcn_readfirstlane(*count);
  const T* img = images_of<T>(a);
  float* cv = smem + lane;
#pragma unroll 1
  for (int base = static_cast<int>(blockIdx.x) * kWave; base < total; base += static_cast<int>(gridDim.x) * kWave) {
    if (base + lane >= total) continue;
    const int64_t i = list[base + lane];
    const float4 pn = a.planes[i], po = orig[i];
    const uint32_t sel = a.sel[i];
    const int x = static_cast<int>(i % a.width), y = static_cast<int>(i / a.width);
    Nbrs nb;
    load_nbrs(a, orig, x, y, nb);
    float obj[2];
    const float4 pl[2] = {po, pn};
    const int cidx = colour_index(a, x, y);
#pragma unroll 1
    for (int k = 0; k < 2; ++k) {
      const Plane p = to_plane(pl[k]);
      if (k == 0 && orig_costs) {
        for (int s = 0; s < a.n_src; ++s) cv[s * kWave] = orig_costs[s * a.ref_plane + cidx];
      } else {
        full_cost_vector<HP>(a, img, x, y, cidx, p, cv);
      }
      float sum = 0.0f;
      int n = 0;
      for (int s = 0; s < a.n_src; ++s) {
        if (!((sel >> s) & 1u)) continue;
        sum += cv[s * kWave];
        ++n;
      }
      obj[k] = n > 0 ? sum / static_cast<float>(n) : 0.0f;
      if (a.smooth_w > 1e-6f) obj[k] += a.smooth_w * smoothness_cost(a, nb, p, x, y);
    }
    if (!(obj[1] <= obj[0])) a.planes[i] = po;
  }
}

int64_t xcd_round(int64_t blocks) { return kXcds * ((blocks + kXcds - 1) / kXcds); }

int64_t colour_tiles(int width, int height, int tile_x) {
  const int hw = (width + 1) / 2;
  const int ty = kWave / tile_x;
  return static_cast<int64_t>((hw + tile_x - 1) / tile_x) * ((height + ty - 1) / ty);
}

hipError_t launch_storage(bool f16, int kernel, const PmArgs* a, const PmBatch* bt, unsigned grid, size_t lds, int hp,
                          int tile_x, hipStream_t s) {
  return f16 ? launch_storage_f16(kernel, a, bt, grid, lds, hp, tile_x, s)
             : launch_storage_f32(kernel, a, bt, grid, lds, hp, tile_x, s);
}

unsigned blocks_of(int kernel, const PmArgs& a, int tile_x) {
  return static_cast<unsigned>(kernel_blocks(kernel, a.width, a.height, a.n_src, tile_x));
}

size_t cost_column_lds(int n_src) { return sizeof(float) * 2 * static_cast<size_t>(n_src) * kWave; }

// select / refine / filter / decode: templated on the view count and the tile.
template <int NV, int TX>
void launch_pixel_nv(int kernel, const PmArgs* a, const PmBatch* bt, dim3 g, hipStream_t s) {
  if (kernel == PM_K_SELECT) {
    if (a) hipLaunchKernelGGL((select_kernel<NV, TX>), g, dim3(kWave), 0, s, *a);
    else hipLaunchKernelGGL((select_batch<NV, TX>), g, dim3(kWave), 0, s, *bt);
  } else {
    if (a) hipLaunchKernelGGL((refine_kernel<NV, TX>), g, dim3(kWave), 0, s, *a);
    else hipLaunchKernelGGL((refine_batch<NV, TX>), g, dim3(kWave), 0, s, *bt);
  }
}

template <int TX>
void launch_pixel_tx(int kernel, const PmArgs* a, const PmBatch* bt, dim3 g, int max_src, hipStream_t s) {
  if (kernel == PM_K_FILTER) {
    if (a) hipLaunchKernelGGL((filter_kernel<TX>), g, dim3(kWave), 0, s, *a);
    else hipLaunchKernelGGL((filter_batch<TX>), g, dim3(kWave), 0, s, *bt);
  } else if (max_src <= 8) {
    launch_pixel_nv<8, TX>(kernel, a, bt, g, s);
  } else if (max_src <= 10) {
    launch_pixel_nv<10, TX>(kernel, a, bt, g, s);
  } else if (max_src <= 12) {
    launch_pixel_nv<12, TX>(kernel, a, bt, g, s);
  } else {
    launch_pixel_nv<kMaxSrc, TX>(kernel, a, bt, g, s);
  }
}

hipError_t launch_pixel(int kernel, const PmArgs* a, const PmBatch* bt, unsigned grid, int max_src, int tile_x,
                        hipStream_t s) {
  if (grid == 0) return hipSuccess;
  const dim3 g(grid);
  switch (tile_x) {
    case 16: launch_pixel_tx<16>(kernel, a, bt, g, max_src, s); break;
    case 32: launch_pixel_tx<32>(kernel, a, bt, g, max_src, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
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

bool valid_cost_tile(int tile_x) {
  for (int w : kCostTileWidths)
    if (w == tile_x) return true;
  return false;
}

bool valid_pixel_tile(int tile_x) {
  for (int w : kPixelTileWidths)
    if (w == tile_x) return true;
  return false;
}

int64_t kernel_blocks(int kernel, int width, int height, int n_src, int tile_x) {
  switch (kernel) {
    case PM_K_REFPATCH:
    case PM_K_INIT:
    case PM_K_PRIOR:
    case PM_K_SWEEP:
      return xcd_round(static_cast<int64_t>((width + kFullTileX - 1) / kFullTileX) *
                       ((height + kFullTileY - 1) / kFullTileY));
    case PM_K_INIT_VIEWS:  // one workgroup per (source view, full tile)
      return static_cast<int64_t>(n_src) * xcd_round(static_cast<int64_t>((width + kFullTileX - 1) / kFullTileX) *
                                                     ((height + kFullTileY - 1) / kFullTileY));
    case PM_K_FALLBACK: {
      return xcd_round(colour_tiles(width, height, tile_x));  // one workgroup per colour tile
    }
    case PM_K_DECODE:
      return xcd_round((static_cast<int64_t>(width) * height + kDecodeThreads - 1) / kDecodeThreads);
    default:  // colour-tile grids
      return xcd_round(colour_tiles(width, height, tile_x));
  }
}

int64_t fallback_mask_count(int width, int height, int n_src) {
  int64_t tiles = 0;
  for (int w : kCostTileWidths) tiles = colour_tiles(width, height, w) > tiles ? colour_tiles(width, height, w) : tiles;
  return tiles * 9 * n_src;
}

hipError_t launch_quads(const PmArgs& a, int nviews, bool f16, hipStream_t stream) {
  if (f16) {
    hipLaunchKernelGGL(quad_kernel<_Float16>, dim3(2048), dim3(256), 0, stream, a, nviews);
  } else {
    hipLaunchKernelGGL(quad_kernel<float>, dim3(2048), dim3(256), 0, stream, a, nviews);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Device-resident views (pm_add_view_device): the padded layout of one view
// written from the caller's device image, as upload_level writes it from
// host memory (estimator.cc:313-361): u8 -> value * (1/255), clamp-to-edge
// border, then float16 storage (round to nearest even) or float32 storage
// (optionally rounded through float16 first).  The conversion is the
// integer restatement the host uses, so both paths store the same bits.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t f32_to_f16_rne(float f) {
  const uint32_t x = __float_as_uint(f);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t absx = x & 0x7fffffffu;
  if (absx >= 0x7f800000u) return sign | (absx > 0x7f800000u ? 0x7e00u : 0x7c00u);
  if (absx >= 0x477ff000u) return sign | 0x7c00u;
  if (absx < 0x38800000u) {
    if (absx < 0x33000000u) return sign;
    const uint32_t mant = (absx & 0x7fffffu) | 0x800000u;
    const int shift = 126 - static_cast<int>(absx >> 23);
    const uint32_t q = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u);
    const uint32_t half = 1u << (shift - 1);
    return sign | (q + ((rem > half || (rem == half && (q & 1u))) ? 1u : 0u));
  }
  uint32_t r = (((absx >> 23) - 112u) << 10) | ((absx & 0x7fffffu) >> 13);
  const uint32_t rem = absx & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (r & 1u))) ++r;
  return sign | r;
}

__device__ __forceinline__ float pad_src_value(const unsigned char* src, int format, int64_t stride, int x, int y) {
  const unsigned char* row = src + stride * y;
  if (format == 0) return static_cast<float>(row[x]) * static_cast<float>(1.0 / 255.0);   // PM_U8
  if (format == 1) return reinterpret_cast<const float*>(row)[x];                          // PM_F32
  return static_cast<float>(reinterpret_cast<const _Float16*>(row)[x]);                    // PM_F16
}

__global__ __launch_bounds__(256) void pad_view_kernel(const unsigned char* __restrict__ src, int format,
                                                       int64_t stride, int w, int h, void* __restrict__ dst,
                                                       int64_t base, int pitch, int f16, int quantize_half) {
  const int64_t n = static_cast<int64_t>(pitch) * (h + 2);
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int px = static_cast<int>(i % pitch), py = static_cast<int>(i / pitch);
    const int cx = min(max(px - 1, 0), w - 1), cy = min(max(py - 1, 0), h - 1);
    const float v = pad_src_value(src, format, stride, cx, cy);
    if (f16) {
      static_cast<uint16_t*>(dst)[base + i] = static_cast<uint16_t>(f32_to_f16_rne(v));
    } else {
      float out = v;
      if (quantize_half) out = static_cast<float>(__builtin_bit_cast(_Float16,
                                                  static_cast<uint16_t>(f32_to_f16_rne(v))));
      static_cast<float*>(dst)[base + i] = out;
    }
  }
}

hipError_t launch_pad_view(const void* src, int format, int64_t stride, int w, int h, void* images, int64_t base,
                           int pitch, bool f16, bool quantize_half, hipStream_t stream) {
  const int64_t n = static_cast<int64_t>(pitch) * (h + 2);
  const unsigned grid = static_cast<unsigned>(std::min<int64_t>((n + 255) / 256, 8192));
  hipLaunchKernelGGL(pad_view_kernel, dim3(grid), dim3(256), 0, stream, static_cast<const unsigned char*>(src),
                     format, stride, w, h, images, base, pitch, f16 ? 1 : 0, quantize_half ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_refpatch(const PmArgs& a, int hp, bool f16, hipStream_t stream) {
  return launch_storage(f16, PM_K_REFPATCH, &a, nullptr, blocks_of(PM_K_REFPATCH, a, 0), 0, hp, 0, stream);
}

// Waves per SIMD of the plane sweep and of the init's per-view scoring
// (one-wave workgroups; 0 = no cap).  Their planes gather along whole
// epipolar segments, and fewer waves with more loads in flight each thrash
// L1/L2 less.  The sweep keeps the cap measured on the fused init (at
// 1600x1200, N=8: 7.8 ms uncapped, 6.9 ms at 3, 6.3 ms at 2, f32).  The
// view-major init scoring takes 5.35 / 5.20 / 5.35 / 5.35 ms per step at
// 2 / 3 / 4 / no cap on float32 images and 5.24 / 4.48 / 4.36 / 4.36 ms on
// float16 images.
#ifndef PM_INIT_WAVES
#define PM_INIT_WAVES 2
#endif
#ifndef PM_INIT_VIEWS_WAVES_F32
#define PM_INIT_VIEWS_WAVES_F32 3
#endif
#ifndef PM_INIT_VIEWS_WAVES_F16
#define PM_INIT_VIEWS_WAVES_F16 4
#endif
// LDS bytes per CU of the current device (160 KiB on MI355X), queried once
// per device; the occupancy caps below reserve a share of it per workgroup.
constexpr int kSimdsPerCu = 4;
size_t lds_per_cu() {
  constexpr int kDevices = 64;
  static std::atomic<size_t> cache[kDevices];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kDevices) dev = -1;
  if (dev >= 0) {
    const size_t known = cache[dev].load(std::memory_order_relaxed);
    if (known) return known;
  }
  int bytes = 0;
  if (hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev < 0 ? 0 : dev) !=
          hipSuccess ||
      bytes <= 0) {
    (void)hipGetLastError();
    return 0;  // unknown: no cap
  }
  if (dev >= 0) cache[dev].store(static_cast<size_t>(bytes), std::memory_order_relaxed);
  return static_cast<size_t>(bytes);
}

// Dynamic LDS of a launch of one-wave workgroups: `cols` bytes of cost
// columns, or more to cap the workgroups at `waves` per SIMD (0: no cap).
size_t capped_lds(size_t cols, int waves) {
  const size_t lds = lds_per_cu();
  const size_t cap = waves > 0 && lds > 0 ? lds / (static_cast<size_t>(kSimdsPerCu) * waves + 1) + 1 : 0;
  return cols > cap ? cols : cap;
}
size_t init_lds(int n_src) { return capped_lds(cost_column_lds(n_src), PM_INIT_WAVES); }
size_t init_views_lds(bool f16) { return capped_lds(0, f16 ? PM_INIT_VIEWS_WAVES_F16 : PM_INIT_VIEWS_WAVES_F32); }

hipError_t launch_init(const PmArgs& a, int hp, bool f16, hipStream_t stream) {
  const hipError_t e = launch_storage(f16, PM_K_INIT_VIEWS, &a, nullptr, blocks_of(PM_K_INIT_VIEWS, a, 0),
                                      init_views_lds(f16), hp, 0, stream);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(init_select_kernel<>, dim3(blocks_of(PM_K_INIT, a, 0)), dim3(kWave), cost_column_lds(a.n_src),
                     stream, a);
  return hipGetLastError();
}

hipError_t launch_plane_sweep(const PmArgs& a, int hp, bool f16, int num_planes, hipStream_t stream) {
  PmArgs b = a;
  b.sweep_count = num_planes;
  for (int i = 0; i < num_planes; ++i) {
    b.sweep_index = i;
    const hipError_t e = launch_storage(f16, PM_K_SWEEP, &b, nullptr, blocks_of(PM_K_SWEEP, b, 0),
                                        init_lds(b.n_src), hp, 0, stream);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_prior(const PmArgs& a, int hp, bool f16, hipStream_t stream) {
  return launch_storage(f16, PM_K_PRIOR, &a, nullptr, blocks_of(PM_K_PRIOR, a, 0), cost_column_lds(a.n_src), hp, 0,
                        stream);
}

hipError_t launch_upsample(const PmArgs& a, const float* cdepth, const float* cnormal, int src_w, int src_h,
                           hipStream_t stream) {
  const int64_t n = static_cast<int64_t>(a.width) * a.height;
  hipLaunchKernelGGL(upsample_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, stream, a, cdepth,
                     cnormal, src_w, src_h);
  return hipGetLastError();
}

hipError_t launch_candidates(const PmArgs& a, int hp, bool f16, int cost_x, hipStream_t stream) {
  if (!valid_cost_tile(cost_x)) return hipErrorInvalidValue;
  return launch_storage(f16, PM_K_CANDIDATES, &a, nullptr, blocks_of(PM_K_CANDIDATES, a, cost_x), 0, hp, cost_x,
                        stream);
}

hipError_t launch_refine_cost(const PmArgs& a, int hp, bool f16, int cost_x, hipStream_t stream) {
  if (!valid_cost_tile(cost_x)) return hipErrorInvalidValue;
  return launch_storage(f16, PM_K_REFINE_COST, &a, nullptr, blocks_of(PM_K_REFINE_COST, a, cost_x), 0, hp, cost_x,
                        stream);
}

size_t fallback_cap_lds(int hp, int tile_x, int fallback_waves) {
  if (fallback_waves <= 0 || tile_x <= 0) return 0;
  const size_t lds = lds_per_cu();
  if (lds == 0) return 0;
  const int r = 3 * hp > 2 ? 3 * hp : 2;
  const int taps = (2 * hp + 1) * (2 * hp + 1);
  const size_t static_bytes = sizeof(float) * ((2 * tile_x + 2 * r) * (kWave / tile_x + 2 * r) +
                                               2 * taps * kWave);
  // more than 1/(n+1) of the LDS per workgroup: at most n workgroups per CU,
  // n = waves per SIMD x SIMDs / waves per workgroup
  const size_t wgs = std::max<size_t>(1, static_cast<size_t>(kSimdsPerCu) * fallback_waves / kFbWaves);
  const size_t per_wg = lds / (wgs + 1) + 1;
  return per_wg > static_bytes ? per_wg - static_bytes : 0;
}

hipError_t launch_fallback(const PmArgs& a, int hp, bool f16, int phase, int cost_x, int fallback_waves,
                           hipStream_t stream) {
  if (!valid_cost_tile(cost_x)) return hipErrorInvalidValue;
  PmArgs b = a;
  b.phase = phase;
  return launch_storage(f16, PM_K_FALLBACK, &b, nullptr, blocks_of(PM_K_FALLBACK, a, cost_x),
                        fallback_cap_lds(hp, cost_x, fallback_waves), hp, cost_x, stream);
}

hipError_t launch_select(const PmArgs& a, int pixel_x, hipStream_t stream) {
  return launch_pixel(PM_K_SELECT, &a, nullptr, blocks_of(PM_K_SELECT, a, pixel_x), a.n_src, pixel_x, stream);
}

hipError_t launch_refine(const PmArgs& a, int pixel_x, hipStream_t stream) {
  return launch_pixel(PM_K_REFINE, &a, nullptr, blocks_of(PM_K_REFINE, a, pixel_x), a.n_src, pixel_x, stream);
}

hipError_t launch_filter(const PmArgs& a, int pixel_x, hipStream_t stream) {
  return launch_pixel(PM_K_FILTER, &a, nullptr, blocks_of(PM_K_FILTER, a, pixel_x), a.n_src, pixel_x, stream);
}

hipError_t launch_confidence(int64_t n, const float* depth, const float* cost, float* conf, hipStream_t stream) {
  hipLaunchKernelGGL(confidence_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, stream, n, depth,
                     cost, conf);
  return hipGetLastError();
}

hipError_t launch_decode(const PmArgs& a, const PmDecodeOut& o, hipStream_t stream) {
  PmArgs b = a;
  b.out = o;
  hipLaunchKernelGGL(decode_kernel, dim3(blocks_of(PM_K_DECODE, a, 0)), dim3(kDecodeThreads), 0, stream, b);
  return hipGetLastError();
}

hipError_t launch_batch(int kernel, const PmBatch& bt, unsigned grid, int hp, bool f16, int max_src, int tile_x,
                        int fallback_waves, hipStream_t stream) {
  if (grid == 0 || bt.total == 0) return hipSuccess;
  switch (kernel) {
    case PM_K_REFPATCH:
      return launch_storage(f16, kernel, nullptr, &bt, grid, 0, hp, 0, stream);
    case PM_K_INIT_VIEWS:
      return launch_storage(f16, kernel, nullptr, &bt, grid, init_views_lds(f16), hp, 0, stream);
    case PM_K_INIT:
      hipLaunchKernelGGL(init_select_batch<>, dim3(grid), dim3(kWave), cost_column_lds(max_src), stream, bt);
      return hipGetLastError();
    case PM_K_PRIOR:
      return launch_storage(f16, kernel, nullptr, &bt, grid, cost_column_lds(max_src), hp, 0, stream);
    case PM_K_CANDIDATES:
    case PM_K_REFINE_COST:
    case PM_K_FALLBACK:
      if (!valid_cost_tile(tile_x)) return hipErrorInvalidValue;
      return launch_storage(f16, kernel, nullptr, &bt, grid,
                            kernel == PM_K_FALLBACK ? fallback_cap_lds(hp, tile_x, fallback_waves) : 0, hp, tile_x,
                            stream);
    case PM_K_SELECT:
    case PM_K_REFINE:
    case PM_K_FILTER:
      return launch_pixel(kernel, nullptr, &bt, grid, max_src, tile_x, stream);
    case PM_K_DECODE:
      hipLaunchKernelGGL(decode_batch, dim3(grid), dim3(kDecodeThreads), 0, stream, bt);
      return hipGetLastError();
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_score(const PmArgs& a, int hp, bool f16, int n, const int32_t* xy, const float4* planes,
                        float* costs, float* init_cost, uint32_t* sel, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  const size_t lds = sizeof(float) * 2 * a.n_src * kWave;
  const dim3 grid((n + kWave - 1) / kWave);
  PM_DISPATCH(score_kernel, hp, f16, grid, lds, stream, a, n, xy, planes, costs, init_cost, sel);
}

hipError_t launch_refine_guard(const PmArgs& a, int hp, bool f16, const float4* orig, const float* orig_costs,
                               int* scratch, hipStream_t stream) {
  // scratch: list[n], *count, then the ceil(n / 256) block offsets: at most
  // 2n + 1 ints, within the state buffer's 16 floats per pixel for any n >= 1
  const int64_t n = static_cast<int64_t>(a.width) * a.height;
  const int nb = static_cast<int>((n + kGuardBlock - 1) / kGuardBlock);
  int* count = scratch + n;
  int* offs = scratch + n + 1;
  hipLaunchKernelGGL(guard_count_kernel, dim3(static_cast<unsigned>(nb)), dim3(kGuardBlock), 0, stream, a.planes, orig,
                     n, offs);
  hipLaunchKernelGGL(guard_scan_kernel, dim3(1), dim3(1024), 0, stream, offs, nb, count);
  hipLaunchKernelGGL(guard_list_kernel, dim3(static_cast<unsigned>(nb)), dim3(kGuardBlock), 0, stream, a.planes, orig,
                     n, offs, scratch);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const dim3 grid(kGuardWaves);
  const size_t lds = cost_column_lds(a.n_src);
  PM_DISPATCH(refine_guard_kernel, hp, f16, grid, lds, stream, a, orig, orig_costs, scratch, count);
}

}  // namespace pmhip
