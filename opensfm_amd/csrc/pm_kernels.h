// Host-side launchers of the PatchMatch kernels (pm_kernels.hip, and
// pm_cost_f32.hip / pm_cost_f16.hip for the kernels templated on the image
// storage type).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pm_device.h"

namespace pmhip {

struct PmDecodeOut {
  float* depth;
  float* normal;  // 3 per pixel
  float* cost;
  float* confidence;
};

// Everything a kernel of the path needs about one reference: passed by
// value as the kernel argument of a single-reference launch, or read from
// a device table by a batched launch.  Per-pixel state is row-major over
// the reference view.
struct PmArgs {
  float4* planes;              // (nx, ny, nz, d)
  float* costs;
  uint2* rng;
  uint32_t* sel;
  const float4* prior_planes;  // may be null
  const uint32_t* plane_masks; // may be null
  const float* prev_depths;    // n_src * W * H, may be null
  const float4* snapshot;      // planes before the prior re-init launch
  const void* images;          // padded views, float or _Float16
  void* quads;                 // the same as 2 x 2 quads (bilinear sampling)
  float2* ref_taps;            // precomputed reference taps (w, w*r), strides 1..3
  float4* ref_sums;            // precomputed reference moments, strides 1..3
  int64_t ref_plane;           // entries per tap plane (= 2 * H * ceil(W/2))
  // Camera and view tables in device memory (index 0 = reference).  Read
  // through the constant address space, so wave-uniform indices become
  // scalar loads; per-lane indices become vector loads.
  const PmCam* cams;
  const PmView* views;
  // Colour-pass intermediate state, indexed [slot][cidx] with stride ref_plane.
  float* ca;                   // per-view costs, (9 * n_src) slots
  int* cpos;                   // candidate pixel of each of the 8 slots (-1: none)
  float* vwb;                  // view weights, n_src slots
  float* wnb;                  // sum of the view weights
  float4* refp;                // refinement hypotheses, 6 slots
  int* nref;                   // number of refinement hypotheses (5 or 6)
  uint64_t* fb_masks;          // pending textureless fallbacks, [(tile * 9 + slot) * n_src + view]
  int phase;                   // fallback jobs of the candidate (0) or refinement (1) slots
  // 1: slot 8 of ca already holds, for this colour's pixels, the per-view
  // costs of their current planes (select and refine copy the winning
  // hypothesis's costs there), so the candidates pass does not score it again
  int cur_cached;
  uint32_t prev_depth_mask;
  int width, height, n_src;
  float depth_min, depth_max;
  float spatial_factor, color_factor;
  int top_k;
  float census_weight;
  float var_threshold;
  float smooth_w;
  float geom_weight;
  int anchor_views;
  int color;
  int iteration;
  int sweep_index, sweep_count;  // plane sweep: this launch's plane, planes in all
  PmDecodeOut out;             // decode targets
};

// One launch over many references: workgroups [base[r], base[r+1]) of the
// launch's item space belong to reference r, whose arguments are args[r]
// (both arrays in device memory).  Every base is a multiple of 8 (the XCD
// count), so a workgroup's index within its reference keeps the XCD that
// the tile mapping assumes.
struct PmBatch {
  const PmArgs* args;
  const int* base;   // nrefs + 1 entries
  int nrefs;
  int total;         // base[nrefs]
};

// Kernels of the path, for the batched launcher and the grid sizes.
enum PmKernel {
  PM_K_REFPATCH = 0,
  PM_K_INIT,
  PM_K_PRIOR,
  PM_K_CANDIDATES,
  PM_K_FALLBACK,
  PM_K_SELECT,
  PM_K_REFINE_COST,
  PM_K_REFINE,
  PM_K_FILTER,
  PM_K_DECODE,
  PM_K_SWEEP,
  PM_K_INIT_VIEWS,  // the init's per-view scoring; PM_K_INIT is its per-pixel top-k
  PM_K_COUNT
};

// Colour-tile widths (tiles are width x 64/width same-colour pixels).  The
// cost kernels and the fallback share one width (the fallback masks are per
// tile); the per-pixel kernels (select, refine, filter) take their own.
constexpr int kCostTileWidths[] = {4, 8, 16};
constexpr int kPixelTileWidths[] = {16, 32};
bool valid_cost_tile(int tile_x);
bool valid_pixel_tile(int tile_x);

// Workgroups of one w x h reference with n_src sources for `kernel`
// (tile_x: its colour-tile width; ignored by the full-resolution kernels).
int64_t kernel_blocks(int kernel, int width, int height, int n_src, int tile_x);

// Batched launch of `kernel` over bt with `grid` (= bt.total) workgroups,
// one per item.  max_src: the largest n_src of the batch.
hipError_t launch_batch(int kernel, const PmBatch& bt, unsigned grid, int half_patch, bool f16, int max_src,
                        int tile_x, int fallback_waves, hipStream_t stream);

// All launchers enqueue on `stream` and return the launch status.
hipError_t launch_refpatch(const PmArgs& a, int half_patch, bool f16, hipStream_t stream);
hipError_t launch_quads(const PmArgs& a, int nviews, bool f16, hipStream_t stream);
// Padded layout of one view from a device image (format PM_U8/PM_F32/PM_F16,
// row stride in bytes) at element offset `base` of `images`.
hipError_t launch_pad_view(const void* src, int format, int64_t stride, int w, int h, void* images, int64_t base,
                           int pitch, bool f16, bool quantize_half, hipStream_t stream);
// Length of the fallback mask array for a w x h reference (enough for
// every cost-tile width).
int64_t fallback_mask_count(int width, int height, int n_src);
hipError_t launch_init(const PmArgs& a, int half_patch, bool f16, hipStream_t stream);
hipError_t launch_prior(const PmArgs& a, int half_patch, bool f16, hipStream_t stream);
// Fronto-parallel plane sweep over num_planes log-spaced depths: one launch
// per plane, each keeping the lower cost per pixel (the first writes all).
hipError_t launch_plane_sweep(const PmArgs& a, int half_patch, bool f16, int num_planes, hipStream_t stream);
// Coarse-to-fine transfer from a decoded src_w x src_h level (depth, normal
// 3 per pixel) into the planes/costs/rng/sel of the current level.
hipError_t launch_upsample(const PmArgs& a, const float* cdepth, const float* cnormal, int src_w, int src_h,
                           hipStream_t stream);
// One colour pass (a.color) of iteration a.iteration, in this order:
// candidates, fallback(0), select, refine_cost, fallback(1), refine.
// cost_x / pixel_x: colour-tile widths (kCostTileWidths / kPixelTileWidths).
hipError_t launch_candidates(const PmArgs& a, int half_patch, bool f16, int cost_x, hipStream_t stream);
// fallback_waves > 0: at most that many fallback waves per SIMD (an LDS
// reservation per workgroup caps the workgroups per CU); 0 = no cap.
hipError_t launch_fallback(const PmArgs& a, int half_patch, bool f16, int phase, int cost_x, int fallback_waves,
                           hipStream_t stream);
// Dynamic LDS bytes of a fallback launch capped at fallback_waves per SIMD.
size_t fallback_cap_lds(int half_patch, int tile_x, int fallback_waves);
hipError_t launch_select(const PmArgs& a, int pixel_x, hipStream_t stream);
hipError_t launch_refine_cost(const PmArgs& a, int half_patch, bool f16, int cost_x, hipStream_t stream);
hipError_t launch_refine(const PmArgs& a, int pixel_x, hipStream_t stream);
hipError_t launch_filter(const PmArgs& a, int pixel_x, hipStream_t stream);
hipError_t launch_decode(const PmArgs& a, const PmDecodeOut& o, hipStream_t stream);

// The kernels templated on the image storage type, one translation unit per
// type: a single-reference launch (a != null) or a batched one (bt != null).
hipError_t launch_storage_f32(int kernel, const PmArgs* a, const PmBatch* bt, unsigned grid, size_t lds,
                              int half_patch, int tile_x, hipStream_t stream);
hipError_t launch_storage_f16(int kernel, const PmArgs* a, const PmBatch* bt, unsigned grid, size_t lds,
                              int half_patch, int tile_x, hipStream_t stream);

// Differentiable photometric cost (pm_refine.hip).  vgeom: 12 floats per
// source view (A = K_v R_v R0^T row-major, b = K_v (t_v - R_v R0^T t0)).
// depth: H*W, ab: 2*H*W normal slopes (a, b) of u = (a, b, -1), views:
// H*W view masks; cost: H*W mean residual, grad: 3*H*W (may be null).
hipError_t launch_ncc_energy(const PmArgs& a, int half_patch, bool f16, const float* vgeom, const float* depth,
                             const float* ab, const uint32_t* views, float* cost, float* grad, hipStream_t stream);
// One damped Gauss-Newton step on a.planes over the views of a.sel.  state:
// gauss_newton_state_floats() x W*H floats carried from step to step; first:
// the first step of a refinement (starts from a.planes).
int gauss_newton_state_floats();
// After the Gauss-Newton steps: pixels whose full PatchMatch cost (fallbacks
// included) rose go back to their plane in `orig`.  scratch: W*H + 1 ints
// of device memory (the moved pixels' list and its length).
hipError_t launch_refine_guard(const PmArgs& a, int half_patch, bool f16, const float4* orig,
                               const float* orig_costs, int* scratch, hipStream_t stream);
hipError_t launch_gauss_newton(const PmArgs& a, int half_patch, bool f16, const float* vgeom, float lambda,
                               float* state, bool first, hipStream_t stream);
// conf = logistic(cost) where depth > 0, else 0 (n pixels).
hipError_t launch_confidence(int64_t n, const float* depth, const float* cost, float* conf, hipStream_t stream);
hipError_t launch_score(const PmArgs& a, int half_patch, bool f16, int n, const int32_t* xy,
                        const float4* planes, float* costs, float* init_cost, uint32_t* sel,
                        hipStream_t stream);

}  // namespace pmhip
