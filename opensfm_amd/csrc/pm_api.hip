// C ABI of the PatchMatch estimator (declared in include/opensfm_pm.h).
//
// Host orchestration restating dense::DepthmapEstimator (opensfm/src/lib/
// dense/src/estimator.cc): view staging (:299-510), random init (:515-546),
// prior re-init (:574-614), iterations (:619-680), checkerboard filter
// (:685-719) and decode (:1112-1224).  The reference round-trips every colour
// pass through buffer->image copies (:568-572); here passes update the
// state buffers in place on one HIP stream and nothing leaves HBM until
// pm_download.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/opensfm_pm.h"
#include "pm_kernels.h"

using namespace pmhip;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define PM_HIP(call, what)                                                              \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) return fail(PM_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e_)); \
  } while (0)

// IEEE binary16 round-to-nearest-even from binary32 (bit-exact, handles
// subnormals, overflow to inf and NaN).
uint16_t f32_to_f16_bits(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t absx = x & 0x7fffffffu;
  if (absx >= 0x7f800000u) return static_cast<uint16_t>(sign | (absx > 0x7f800000u ? 0x7e00u : 0x7c00u));
  if (absx >= 0x477ff000u) return static_cast<uint16_t>(sign | 0x7c00u);  // rounds to >= 65520 -> inf
  if (absx < 0x38800000u) {                                                 // result subnormal or zero
    if (absx < 0x33000000u) return static_cast<uint16_t>(sign);            // < 2^-25 -> 0 (ties to even 0)
    const uint32_t mant = (absx & 0x7fffffu) | 0x800000u;
    const int shift = 126 - static_cast<int>(absx >> 23);                   // 14..24
    const uint32_t q = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u);
    const uint32_t half = 1u << (shift - 1);
    uint32_t r = q + ((rem > half || (rem == half && (q & 1u))) ? 1u : 0u);
    return static_cast<uint16_t>(sign | r);
  }
  const uint32_t e = (absx >> 23) - 112u;  // rebias 127 -> 15
  const uint32_t m = absx & 0x7fffffu;
  uint32_t r = (e << 10) | (m >> 13);
  const uint32_t rem = m & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (r & 1u))) ++r;
  return static_cast<uint16_t>(sign | r);
}

float f16_bits_to_f32(uint16_t h) {
  const uint32_t sign = (static_cast<uint32_t>(h) & 0x8000u) << 16;
  const uint32_t e = (h >> 10) & 0x1fu;
  uint32_t m = h & 0x3ffu;
  uint32_t x;
  if (e == 0) {
    if (m == 0) {
      x = sign;
    } else {  // subnormal: normalise
      int k = -1;
      do { ++k; m <<= 1; } while ((m & 0x400u) == 0);
      x = sign | ((112u - k) << 23) | ((m & 0x3ffu) << 13);
    }
  } else if (e == 31) {
    x = sign | 0x7f800000u | (m << 13);
  } else {
    x = sign | ((e + 112u) << 23) | (m << 13);
  }
  float f;
  std::memcpy(&f, &x, 4);
  return f;
}

struct HostView {
  double K[9], R[9], t[3];
  int w, h;
  std::vector<float> pix;    // [0,1] values as given (u8 already /255)
  std::vector<uint8_t> raw;  // the u8 input, kept for the image pyramid
  // A view added from device memory (pm_add_view_device): read from there
  // at upload, pix and raw stay empty.
  const void* dev = nullptr;
  int dev_format = 0;
  int64_t dev_stride = 0;
};

// Device blocks released by idle handles, kept for reuse by size
// (pm_set_buffer_cache), per device.  Runs of many references of one shape
// then stop freeing and re-allocating hundreds of GB between groups.
// Blocks enter the cache only from pm_release_device / pm_destroy, after
// the handle's stream has drained.
struct BlockCache {
  std::mutex mu;
  uint64_t limit = 0, held = 0;
  std::vector<std::pair<size_t, void*>> blocks;
};
constexpr int kMaxCacheDevices = 64;
BlockCache& block_cache(int device) {
  static BlockCache caches[kMaxCacheDevices];
  return caches[device >= 0 && device < kMaxCacheDevices ? device : 0];
}
int current_device() {
  int d = 0;
  return hipGetDevice(&d) == hipSuccess ? d : 0;
}
// Frees cached blocks of the current device until at most keep bytes stay.
void trim_cache(BlockCache& c, uint64_t keep) {
  std::vector<void*> drop;
  {
    std::lock_guard<std::mutex> lock(c.mu);
    while (c.held > keep && !c.blocks.empty()) {
      c.held -= c.blocks.back().first;
      drop.push_back(c.blocks.back().second);
      c.blocks.pop_back();
    }
  }
  for (void* q : drop) (void)hipFree(q);
}
void* device_alloc(size_t bytes) {
  BlockCache& c = block_cache(current_device());
  {
    std::lock_guard<std::mutex> lock(c.mu);
    for (size_t i = 0; i < c.blocks.size(); ++i) {
      if (c.blocks[i].first == bytes) {
        void* q = c.blocks[i].second;
        c.blocks[i] = c.blocks.back();
        c.blocks.pop_back();
        c.held -= bytes;
        return q;
      }
    }
  }
  void* q = nullptr;
  if (hipMalloc(&q, bytes) == hipSuccess) return q;
  (void)hipGetLastError();
  trim_cache(c, 0);  // blocks of other sizes may be what is missing
  if (hipMalloc(&q, bytes) == hipSuccess) return q;
  (void)hipGetLastError();
  return nullptr;
}
void device_park(void* q, size_t bytes) {
  BlockCache& c = block_cache(current_device());
  {
    std::lock_guard<std::mutex> lock(c.mu);
    if (c.held + bytes <= c.limit) {
      c.blocks.emplace_back(bytes, q);
      c.held += bytes;
      return;
    }
  }
  (void)hipFree(q);
}

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  int alloc(size_t count) {
    if (count == n && p) return PM_OK;
    release();
    if (count == 0) return PM_OK;
    p = static_cast<T*>(device_alloc(count * sizeof(T)));
    if (!p) {
      n = 0;
      return fail(PM_E_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
    }
    n = count;
    return PM_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  // Into the block cache (the owner's stream has drained).
  void park() {
    if (p) device_park(p, n * sizeof(T));
    p = nullptr;
    n = 0;
  }
};

}  // namespace

namespace pmhip {
// Error reporting for the other translation units of the C ABI.
int api_fail(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace pmhip

struct pm_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  pm_params params{};
  std::vector<HostView> views;
  float geom_weight = 0.0f;
  struct PrevDepth {
    int w = 0, h = 0;
    std::vector<float> d;  // empty = no previous depth for this source slot
  };
  std::vector<PrevDepth> prev_depths = std::vector<PrevDepth>(PM_MAX_SOURCES);
  int run_levels = 0;        // levels of the run in progress (pm_run_level)
  int next_level = 0;        // next level pm_run_level accepts
  std::vector<float> prior_planes;
  std::vector<uint32_t> prior_masks;
  std::vector<double> sfm_points;  // 3 per point, world frame (SetSfMPoints)

  // device state
  bool uploaded = false;
  bool state_ready = false;
  int W = 0, H = 0, n_src = 0;
  bool f16 = false;
  DevBuf<unsigned char> images;
  DevBuf<unsigned char> quads;  // 2 x 2 quad copy of the padded views
  DevBuf<PmView> gviews;       // view table (index 0 = reference)
  DevBuf<PmCam> gcams;         // cameras
  DevBuf<float> vgeom;         // per source: reference camera frame -> source pixels (12 floats)
  float last_gn_ms = 0.0f;
  DevBuf<float> ca, vwb, wnb;  // colour-pass intermediate state
  DevBuf<float> gn_state;      // Gauss-Newton state when ca is too small (one source)
  DevBuf<float4> gn_orig;      // the PatchMatch planes, for the refinement guard
  DevBuf<int> cpos, nref;
  DevBuf<float4> refp;
  DevBuf<uint64_t> fb_masks;   // pending textureless fallbacks per (tile, slot, view)
  DevBuf<float2> ref_taps;     // reference-side stride-1 taps (w, w*r)
  DevBuf<float4> ref_sums;
  int64_t ref_plane = 0;
  bool ref_ready = false;
  DevBuf<float4> planes, prior, snapshot;
  DevBuf<float> costs, prevd;
  DevBuf<uint2> rng;
  DevBuf<uint32_t> sel, masks;
  uint32_t prev_mask = 0u;
  DevBuf<float> out_depth, out_normal, out_cost, out_conf;
  DevBuf<float> coarse_depth, coarse_normal;  // decoded previous pyramid level
  int coarse_w = 0, coarse_h = 0;
  float last_run_ms = 0.0f, last_sweep_ms = 0.0f;
  int last_sweeps = 0;
  // Per-kernel device time of the last pm_run: HIP events around every
  // launch of each stage, on the handle's stream.
  std::vector<hipEvent_t> kev;
  std::vector<std::pair<int, int>> kev_marks;  // (stage, index of the start event)
  size_t kev_used = 0;
  float stage_ms[PM_STAGES] = {};
  int stage_launches[PM_STAGES] = {};
  bool timing = false;
  // Colour-tile widths of the cost kernels and of the per-pixel kernels
  // (0 = by image storage: see default_cost_tile).
  int cost_tile_x = 0;
  int pixel_tile_x = 0;
  int fallback_waves = -1;  // occupancy cap of the fallback kernel (-1: default, 0: none)
  // Bumped by every change of views, parameters or device buffers, so a
  // batch knows when its argument tables are stale.
  uint64_t gen = 0;
  // Slot 8 of ca holds, for the pixels of colour c, the per-view costs of
  // their current planes since the last colour-c pass (select and refine
  // copy the winner's costs there); cleared by anything else that changes
  // planes or ca, and by a new generation.
  bool cur_valid[2] = {false, false};
  uint64_t cur_gen = 0;

  // Device buffers into the block cache (or freed); the stream has drained.
  // The coarse level's result stays unless keep_coarse is false: the next
  // pyramid level starts from it after the handle released its buffers.
  void park_buffers(bool keep_coarse) {
    images.park(); quads.park(); gviews.park(); gcams.park(); ref_taps.park(); ref_sums.park();
    vgeom.park();
    ca.park(); vwb.park(); wnb.park(); cpos.park(); nref.park(); refp.park();
    fb_masks.park();
    planes.park(); prior.park();
    snapshot.park(); costs.park(); prevd.park(); rng.park(); sel.park(); masks.park();
    out_depth.park(); out_normal.park(); out_cost.park(); out_conf.park();
    gn_state.park();
    gn_orig.park();
    if (!keep_coarse) {
      coarse_depth.park();
      coarse_normal.park();
    }
  }

  ~pm_handle() {
    park_buffers(false);
    for (auto& e : ev) if (e) (void)hipEventDestroy(e);
    for (auto& e : kev) if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

int check_handle(pm_handle* h) {
  if (!h) return fail(PM_E_INVALID, "null pm_handle");
  return PM_OK;
}

// Default colour-tile widths, measured on MI355X at 1600x1200 with 8
// sources: the cost kernels and the fallback run fastest with 4-wide tiles
// (fallback 33.5 ms against 46.2 ms at 16 on float32 images, about equal on
// float16), the per-pixel kernels with 32-wide tiles.
int cost_tile(const pm_handle* h) { return h->cost_tile_x ? h->cost_tile_x : 4; }
int pixel_tile(const pm_handle* h) { return h->pixel_tile_x ? h->pixel_tile_x : 32; }
// Occupancy cap of the fallback (waves per SIMD); by default none: capping
// at 2 or 3 waves was slower in every measured configuration.
int fallback_cap(const pm_handle* h) { return h->fallback_waves > 0 ? h->fallback_waves : 0; }

PmArgs make_args(pm_handle* h) {
  PmArgs a{};
  a.planes = h->planes.p;
  a.costs = h->costs.p;
  a.rng = h->rng.p;
  a.sel = h->sel.p;
  a.prior_planes = h->masks.p ? h->prior.p : nullptr;
  a.plane_masks = h->masks.p;
  a.prev_depths = h->prevd.p;
  a.prev_depth_mask = h->prev_mask;
  a.snapshot = h->snapshot.p;
  a.images = h->images.p;
  a.quads = h->quads.p;
  a.ref_taps = h->ref_taps.p;
  a.ref_sums = h->ref_sums.p;
  a.ref_plane = h->ref_plane;
  a.cams = h->gcams.p;
  a.views = h->gviews.p;
  a.ca = h->ca.p;
  a.cpos = h->cpos.p;
  a.vwb = h->vwb.p;
  a.wnb = h->wnb.p;
  a.refp = h->refp.p;
  a.nref = h->nref.p;
  a.fb_masks = h->fb_masks.p;
  a.phase = 0;
  a.width = h->W;
  a.height = h->H;
  a.n_src = h->n_src;
  const pm_params& p = h->params;
  a.depth_min = p.depth_min;
  a.depth_max = p.depth_max;
  a.spatial_factor = -1.0f / (2.0f * p.sigma_spatial * p.sigma_spatial);
  a.color_factor = -1.0f / (2.0f * p.sigma_color * p.sigma_color);
  a.top_k = p.top_k;
  a.census_weight = p.use_census ? 1.0f : 0.0f;
  a.var_threshold = p.var_threshold;
  a.smooth_w = p.smooth_weight;
  a.geom_weight = h->geom_weight;
  a.anchor_views = p.anchor_views;
  a.color = 0;
  a.iteration = 0;
  return a;
}

// Smooth weight of one iteration: estimator.cc:646-649.
float smooth_weight_at(const pm_params& p, int iter) {
  const float decay = 1.0f - 0.9f * static_cast<float>(iter) /
                                 static_cast<float>(std::max(p.max_iterations - 1, 1));
  return p.smooth_weight * decay;
}

PmCam make_cam(const HostView& v) {
  PmCam c{};
  for (int i = 0; i < 9; ++i) { c.K[i] = static_cast<float>(v.K[i]); c.R[i] = static_cast<float>(v.R[i]); }
  for (int i = 0; i < 3; ++i) c.t[i] = static_cast<float>(v.t[i]);
  c.width = v.w;
  c.height = v.h;
  return c;
}

int validate(pm_handle* h) {
  const pm_params& p = h->params;
  if (h->views.empty()) return fail(PM_E_INVALID, "DepthmapEstimator: no views added");
  if (h->views.size() < 2) return fail(PM_E_INVALID, "DepthmapEstimator: need at least one source view");
  if (static_cast<int>(h->views.size()) - 1 > PM_MAX_SOURCES)
    return fail(PM_E_INVALID, "DepthmapEstimator: at most " + std::to_string(PM_MAX_SOURCES) + " source views");
  if (p.patch_size != 3 && p.patch_size != 5 && p.patch_size != 7)
    return fail(PM_E_INVALID, "patch_size must be 3, 5 or 7");
  if (!(p.depth_min > 0.0f) || !(p.depth_max > p.depth_min))
    return fail(PM_E_INVALID, "depth range must satisfy 0 < min < max");
  if (p.max_iterations < 0) return fail(PM_E_INVALID, "max_iterations must be >= 0");
  if (p.hierarchy_levels < 1 || p.hierarchy_levels > 8)
    return fail(PM_E_INVALID, "hierarchy_levels must be in 1..8");
  // Any image size runs: pyramid levels shrink to at least 1 x 1, as in
  // BuildImagePyramid (estimator.cc:728-735).
  if (!(p.sigma_spatial > 0.0f) || !(p.sigma_color > 0.0f)) return fail(PM_E_INVALID, "sigmas must be > 0");
  if (p.speckle_min_size < 0 || p.gap_max_size < 0) return fail(PM_E_INVALID, "speckle/gap sizes must be >= 0");
  return PM_OK;
}

// Launch `fn` on the handle's stream; while a pm_run is being timed, bracket
// it with a pair of events attributed to `stage`.
template <typename F>
int timed(pm_handle* h, int stage, const char* what, F fn) {
  if (h->timing) {
    while (h->kev.size() < h->kev_used + 2) {
      hipEvent_t e;
      PM_HIP(hipEventCreate(&e), "hipEventCreate");
      h->kev.push_back(e);
    }
    PM_HIP(hipEventRecord(h->kev[h->kev_used], h->stream), "event");
  }
  PM_HIP(fn(), what);
  if (h->timing) {
    PM_HIP(hipEventRecord(h->kev[h->kev_used + 1], h->stream), "event");
    h->kev_marks.emplace_back(stage, static_cast<int>(h->kev_used));
    h->kev_used += 2;
  }
  return PM_OK;
}

// Reference-side patch data for all strides, once per upload (it depends
// only on the reference image and the patch parameters).
int ensure_refpatch(pm_handle* h) {
  if (h->ref_ready) return PM_OK;
  PmArgs a = make_args(h);
  PM_HIP(launch_refpatch(a, h->params.patch_size / 2, h->f16, h->stream), "refpatch kernel launch");
  h->ref_ready = true;
  return PM_OK;
}

}  // namespace

extern "C" {

void pm_default_params(pm_params* o) {
  if (!o) return;
  // DepthmapParams' member initialisers (estimator.h:16-36), field by field.
  o->max_iterations = 3;
  o->patch_size = 5;
  o->sigma_spatial = 3.0f;
  o->sigma_color = 3.0f / 255.0f;
  o->top_k = 4;
  o->use_census = 1;
  o->depth_min = 0.0f;
  o->depth_max = 1.0f;
  o->hierarchy_levels = 1;
  o->smooth_weight = 0.0f;
  o->anchor_views = 0;
  o->checkerboard_filter = 0;
  o->speckle_min_size = 100;
  o->gap_max_size = 7;
  // Not DepthmapParams fields: the kernel's hard-coded variance gate
  // (opencl_kernels.h:254, 516, 519) and this library's storage options.
  o->var_threshold = 1.5e-6f;
  o->image_storage = PM_STORE_F32;
  o->quantize_half = 0;
}

void pm_pipeline_params(pm_params* o) {
  if (!o) return;
  // What the dense pipeline sets through _setup_cluster_params
  // (opensfm/dense/depthmaps.py:154-175) from the config defaults
  // (opensfm/config.py:407-466); the rest as pm_default_params.
  pm_default_params(o);
  o->max_iterations = 4;                // depthmap_max_iterations
  o->patch_size = 5;                    // depthmap_patch_size
  o->sigma_spatial = 1.5f;              // depthmap_sigma_spatial
  o->sigma_color = 25.0f / 255.0f;      // depthmap_sigma_color
  o->top_k = 8;                         // depthmap_num_matching_views
  o->use_census = 1;                    // depthmap_use_census
  o->hierarchy_levels = 3;              // depthmap_hierarchy_levels
  o->smooth_weight = 0.2f;              // depthmap_smooth_weight
  o->anchor_views = 2;                  // depthmap_anchor_views
  o->checkerboard_filter = 0;           // depthmap_checkerboard_filter
  o->speckle_min_size = 0;              // depthmap_speckle_min_size
  o->gap_max_size = 0;                  // depthmap_gap_max_size
}

const char* pm_last_error(void) { return g_last_error.c_str(); }

int pm_device_count(int* count) {
  if (!count) return fail(PM_E_INVALID, "null count");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return PM_OK;
}

int pm_device_name(int device, char* buf, int buflen) {
  if (!buf || buflen <= 0) return fail(PM_E_INVALID, "null buffer");
  hipDeviceProp_t prop;
  PM_HIP(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
  std::snprintf(buf, static_cast<size_t>(buflen), "%s (%s)", prop.name, prop.gcnArchName);
  return PM_OK;
}

int pm_device_memory(int device, uint64_t* total_bytes, uint64_t* free_bytes) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    return fail(PM_E_DEVICE, "no HIP device is available on this system");
  if (device < 0 || device >= n) return fail(PM_E_INVALID, "device index out of range");
  int prev = 0;
  PM_HIP(hipGetDevice(&prev), "hipGetDevice");
  PM_HIP(hipSetDevice(device), "hipSetDevice");
  size_t fr = 0, tot = 0;
  const hipError_t e = hipMemGetInfo(&fr, &tot);
  (void)hipSetDevice(prev);
  if (e != hipSuccess) return fail(PM_E_DEVICE, std::string("hipMemGetInfo: ") + hipGetErrorString(e));
  if (total_bytes) *total_bytes = tot;
  if (free_bytes) *free_bytes = fr;
  return PM_OK;
}

int pm_create(int device, pm_handle** out) {
  if (!out) return fail(PM_E_INVALID, "null output handle");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return fail(PM_E_DEVICE, "DepthmapEstimator: no HIP device is available on this system");
  if (device < 0 || device >= n) return fail(PM_E_INVALID, "device index out of range");
  PM_HIP(hipSetDevice(device), "hipSetDevice");
  pm_handle* h = new pm_handle();
  h->device = device;
  pm_default_params(&h->params);
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return fail(PM_E_DEVICE, "hipStreamCreate failed");
  }
  for (auto& e : h->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      delete h;
      return fail(PM_E_DEVICE, "hipEventCreate failed");
    }
  }
  *out = h;
  return PM_OK;
}

int pm_destroy(pm_handle* h) {
  if (!h) return PM_OK;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  delete h;
  return PM_OK;
}

// Largest view: the bilinear sampler forms a view's quad index in float,
// exact below 2^24, and addresses its rows with 24-bit multiplies, so a view
// has at most 2^24 padded texels (4000 x 4000; the dense pipeline caps
// images at 3200 px, config.py:417) and 65536 pixels per side.
constexpr int64_t kMaxViewTexels = int64_t(1) << 24;
constexpr int kMaxViewSide = 65536;

static int check_view_size(int width, int height, const char* who) {
  if (width <= 0 || height <= 0) return fail(PM_E_INVALID, std::string(who) + ": empty image");
  if (width > kMaxViewSide || height > kMaxViewSide ||
      static_cast<int64_t>(width + 2) * (height + 2) > kMaxViewTexels)
    return fail(PM_E_INVALID, std::string(who) + ": image too large (at most " + std::to_string(kMaxViewSide) +
                                  " pixels per side and " + std::to_string(kMaxViewTexels) + " padded pixels)");
  return PM_OK;
}

int pm_add_view(pm_handle* h, const double* K, const double* R, const double* t, const void* image,
                int format, int width, int height, int64_t row_stride, const uint8_t* /*mask*/) {
  if (int r = check_handle(h)) return r;
  ++h->gen;
  if (!K || !R || !t || !image) return fail(PM_E_INVALID, "add_view: null argument");
  if (int r = check_view_size(width, height, "add_view")) return r;
  int esize = format == PM_U8 ? 1 : format == PM_F32 ? 4 : format == PM_F16 ? 2 : 0;
  if (!esize) return fail(PM_E_INVALID, "add_view: unknown pixel format");
  if (row_stride == 0) row_stride = static_cast<int64_t>(width) * esize;
  if (row_stride < static_cast<int64_t>(width) * esize) return fail(PM_E_INVALID, "add_view: row stride too small");
  HostView v;
  std::memcpy(v.K, K, sizeof(v.K));
  std::memcpy(v.R, R, sizeof(v.R));
  std::memcpy(v.t, t, sizeof(v.t));
  v.w = width;
  v.h = height;
  v.pix.resize(static_cast<size_t>(width) * height);
  const unsigned char* base = static_cast<const unsigned char*>(image);
  const float inv255 = static_cast<float>(1.0 / 255.0);
  for (int y = 0; y < height; ++y) {
    const unsigned char* row = base + row_stride * y;
    float* dst = v.pix.data() + static_cast<size_t>(y) * width;
    for (int x = 0; x < width; ++x) {
      if (format == PM_U8) {
        dst[x] = static_cast<float>(row[x]) * inv255;  // estimator.cc:315
        v.raw.push_back(row[x]);
      } else if (format == PM_F32) {
        std::memcpy(&dst[x], row + 4 * x, 4);
      } else {
        uint16_t b;
        std::memcpy(&b, row + 2 * x, 2);
        dst[x] = f16_bits_to_f32(b);
      }
    }
  }
  h->views.push_back(std::move(v));
  h->uploaded = false;
  h->state_ready = false;
  return PM_OK;
}

int pm_add_view_device(pm_handle* h, const double* K, const double* R, const double* t, const void* image,
                       int format, int width, int height, int64_t row_stride) {
  if (int r = check_handle(h)) return r;
  ++h->gen;
  if (!K || !R || !t || !image) return fail(PM_E_INVALID, "add_view_device: null argument");
  if (int r = check_view_size(width, height, "add_view_device")) return r;
  const int esize = format == PM_U8 ? 1 : format == PM_F32 ? 4 : format == PM_F16 ? 2 : 0;
  if (!esize) return fail(PM_E_INVALID, "add_view_device: unknown pixel format");
  if (row_stride == 0) row_stride = static_cast<int64_t>(width) * esize;
  if (row_stride < static_cast<int64_t>(width) * esize || row_stride % esize != 0)
    return fail(PM_E_INVALID, "add_view_device: bad row stride");
  hipPointerAttribute_t attr{};
  if (hipPointerGetAttributes(&attr, image) != hipSuccess) {
    (void)hipGetLastError();
    return fail(PM_E_INVALID, "add_view_device: the image is not device memory");
  }
  if ((attr.type != hipMemoryTypeDevice && !attr.isManaged) || attr.device != h->device)
    return fail(PM_E_INVALID, "add_view_device: the image is not device memory of the estimator's device");
  HostView v;
  std::memcpy(v.K, K, sizeof(v.K));
  std::memcpy(v.R, R, sizeof(v.R));
  std::memcpy(v.t, t, sizeof(v.t));
  v.w = width;
  v.h = height;
  v.dev = image;
  v.dev_format = format;
  v.dev_stride = row_stride;
  h->views.push_back(std::move(v));
  h->uploaded = false;
  h->state_ready = false;
  return PM_OK;
}

int pm_clear_views(pm_handle* h) {
  if (int r = check_handle(h)) return r;
  ++h->gen;
  h->views.clear();
  h->uploaded = false;
  h->state_ready = false;
  return PM_OK;
}

int pm_num_views(pm_handle* h, int* n) {
  if (int r = check_handle(h)) return r;
  if (!n) return fail(PM_E_INVALID, "null output");
  *n = static_cast<int>(h->views.size());
  return PM_OK;
}

int pm_set_depth_range(pm_handle* h, float dmin, float dmax) {
  if (int r = check_handle(h)) return r;
  ++h->gen;
  h->params.depth_min = dmin;
  h->params.depth_max = dmax;
  h->state_ready = false;
  return PM_OK;
}

int pm_set_params(pm_handle* h, const pm_params* p) {
  if (int r = check_handle(h)) return r;
  ++h->gen;
  if (!p) return fail(PM_E_INVALID, "null params");
  const bool restage = p->image_storage != h->params.image_storage || p->quantize_half != h->params.quantize_half ||
                       p->patch_size != h->params.patch_size;
  if (p->sigma_spatial != h->params.sigma_spatial || p->sigma_color != h->params.sigma_color) h->ref_ready = false;
  h->params = *p;
  if (restage) h->uploaded = false;
  h->state_ready = false;
  return PM_OK;
}

int pm_get_params(pm_handle* h, pm_params* p) {
  if (int r = check_handle(h)) return r;
  if (!p) return fail(PM_E_INVALID, "null params");
  *p = h->params;
  return PM_OK;
}

int pm_set_geom_consistency(pm_handle* h, float weight, const float* const* prev_depths, int n) {
  if (int r = check_handle(h)) return r;
  ++h->gen;
  if (n < 0 || n > PM_MAX_SOURCES) return fail(PM_E_INVALID, "too many previous depth maps");
  if (h->views.empty()) return fail(PM_E_INVALID, "add the reference view before previous depths");
  h->geom_weight = weight;
  if (prev_depths) {
    for (auto& e : h->prev_depths) e = pm_handle::PrevDepth{};
    const int w = h->views[0].w, hh = h->views[0].h;
    for (int i = 0; i < n; ++i)
      if (prev_depths[i]) h->prev_depths[i] = pm_handle::PrevDepth{w, hh, std::vector<float>(prev_depths[i],
                                                                   prev_depths[i] + static_cast<size_t>(w) * hh)};
  }
  h->uploaded = false;
  return PM_OK;
}

int pm_set_previous_depth(pm_handle* h, int slot, const float* depth, int width, int height) {
  if (int r = check_handle(h)) return r;
  ++h->gen;
  if (slot < 0 || slot >= PM_MAX_SOURCES) return fail(PM_E_INVALID, "previous depth: source slot out of range");
  if (!depth) {
    h->prev_depths[slot] = pm_handle::PrevDepth{};
  } else {
    if (width <= 0 || height <= 0) return fail(PM_E_INVALID, "previous depth: empty map");
    h->prev_depths[slot] =
        pm_handle::PrevDepth{width, height, std::vector<float>(depth, depth + static_cast<size_t>(width) * height)};
  }
  h->uploaded = false;
  return PM_OK;
}

int pm_clear_previous_depths(pm_handle* h) {
  if (int r = check_handle(h)) return r;
  ++h->gen;
  for (auto& e : h->prev_depths) e = pm_handle::PrevDepth{};
  h->uploaded = false;
  return PM_OK;
}

int pm_set_prior(pm_handle* h, const float* prior_planes, const uint32_t* masks) {
  if (int r = check_handle(h)) return r;
  ++h->gen;
  if (h->views.empty()) return fail(PM_E_INVALID, "add the reference view before the prior");
  const size_t npix = static_cast<size_t>(h->views[0].w) * h->views[0].h;
  h->sfm_points.clear();
  if (!prior_planes || !masks) {
    h->prior_planes.clear();
    h->prior_masks.clear();
  } else {
    h->prior_planes.assign(prior_planes, prior_planes + 4 * npix);
    h->prior_masks.assign(masks, masks + npix);
  }
  h->uploaded = false;
  return PM_OK;
}

int pm_set_sfm_points(pm_handle* h, const double* points, int num_points) {
  if (int r = check_handle(h)) return r;
  if (num_points < 0 || (num_points > 0 && !points)) return fail(PM_E_INVALID, "set_sfm_points: bad arguments");
  ++h->gen;
  h->prior_planes.clear();
  h->prior_masks.clear();
  h->sfm_points.assign(points, points + 3 * static_cast<size_t>(num_points));
  h->uploaded = false;
  return PM_OK;
}

}  // extern "C"

namespace {

// Pixel-area downscale of a w x h image to nw x nh, the resampling
// cv::resize(INTER_AREA) performs for BuildImagePyramid (estimator.cc:
// 727-756): each output pixel is the mean of the input over its footprint,
// pixels on the footprint's edge weighted by the covered fraction.  Sums in
// double; with round_int the result is rounded to an integer (u8 images).
void resize_area(const float* src, int w, int h, int nw, int nh, bool round_int, float* dst) {
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
          row += wx * static_cast<double>(src[static_cast<size_t>(yy) * w + xx]);
        }
        acc += wy * row;
      }
      double v = acc / ((x1 - x0) * (y1 - y0));
      if (round_int) v = std::nearbyint(v);
      dst[static_cast<size_t>(dy) * nw + dx] = static_cast<float>(v);
    }
  }
}

// Bilinear resize with pixel-centre alignment and edge clamping, the
// cv::resize(INTER_LINEAR) of UploadPreviousDepths (estimator.cc:92-101).
void resize_linear(const float* src, int w, int h, int nw, int nh, float* dst) {
  const double sx = static_cast<double>(w) / nw, sy = static_cast<double>(h) / nh;
  for (int dy = 0; dy < nh; ++dy) {
    double fy = (dy + 0.5) * sy - 0.5;
    int y0 = static_cast<int>(std::floor(fy));
    double ay = fy - y0;
    if (y0 < 0) { y0 = 0; ay = 0.0; }
    if (y0 >= h - 1) { y0 = h - 1; ay = 0.0; }
    const int y1 = std::min(y0 + 1, h - 1);
    for (int dx = 0; dx < nw; ++dx) {
      double fx = (dx + 0.5) * sx - 0.5;
      int x0 = static_cast<int>(std::floor(fx));
      double ax = fx - x0;
      if (x0 < 0) { x0 = 0; ax = 0.0; }
      if (x0 >= w - 1) { x0 = w - 1; ax = 0.0; }
      const int x1 = std::min(x0 + 1, w - 1);
      const double v00 = src[static_cast<size_t>(y0) * w + x0], v01 = src[static_cast<size_t>(y0) * w + x1];
      const double v10 = src[static_cast<size_t>(y1) * w + x0], v11 = src[static_cast<size_t>(y1) * w + x1];
      const double v = (1.0 - ay) * ((1.0 - ax) * v00 + ax * v01) + ay * ((1.0 - ax) * v10 + ax * v11);
      dst[static_cast<size_t>(dy) * nw + dx] = static_cast<float>(v);
    }
  }
}

}  // namespace

namespace pmhip {
void host_resize_linear(const float* src, int w, int h, int nw, int nh, float* dst) {
  resize_linear(src, w, h, nw, nh, dst);
}
}  // namespace pmhip

namespace {

// One pyramid level of a view (BuildImagePyramid, estimator.cc:727-756):
// size int(w * scale), area-resampled pixels, fx, cx scaled by new_w / w
// and fy, cy by new_h / h.  u8 inputs are resampled as u8.
HostView scaled_view(const HostView& v, float scale) {
  HostView o;
  o.w = std::max(static_cast<int>(static_cast<float>(v.w) * scale), 1);
  o.h = std::max(static_cast<int>(static_cast<float>(v.h) * scale), 1);
  std::memcpy(o.K, v.K, sizeof(o.K));
  std::memcpy(o.R, v.R, sizeof(o.R));
  std::memcpy(o.t, v.t, sizeof(o.t));
  const float fsx = static_cast<float>(o.w) / static_cast<float>(v.w);
  const float fsy = static_cast<float>(o.h) / static_cast<float>(v.h);
  o.K[0] *= fsx;
  o.K[2] *= fsx;
  o.K[4] *= fsy;
  o.K[5] *= fsy;
  o.pix.resize(static_cast<size_t>(o.w) * o.h);
  if (!v.raw.empty()) {
    std::vector<float> src(v.raw.begin(), v.raw.end());
    resize_area(src.data(), v.w, v.h, o.w, o.h, true, o.pix.data());
    const float inv255 = static_cast<float>(1.0 / 255.0);
    for (float& f : o.pix) f *= inv255;
  } else {
    resize_area(v.pix.data(), v.w, v.h, o.w, o.h, false, o.pix.data());
  }
  return o;
}

// Speckle removal (RemoveSmallSegments, estimator.cc:1231-1312): 4-connected
// segments of valid depths whose neighbours differ by less than `thr`
// relative depth; segments smaller than min_size lose depth and normal.
void remove_small_segments(std::vector<float>& depth, std::vector<float>& normal, int w, int h, int min_size,
                           float thr) {
  const int npix = w * h;
  std::vector<char> visited(static_cast<size_t>(npix), 0);
  std::vector<int> seg;
  const int dx[4] = {1, -1, 0, 0};
  const int dy[4] = {0, 0, 1, -1};
  for (int idx0 = 0; idx0 < npix; ++idx0) {
    if (visited[idx0] || depth[idx0] <= 0.0f) {
      visited[idx0] = 1;
      continue;
    }
    seg.clear();
    seg.push_back(idx0);
    visited[idx0] = 1;
    for (size_t head = 0; head < seg.size(); ++head) {
      const int cur = seg[head];
      const int cy = cur / w, cx = cur % w;
      const float dc = depth[cur];
      for (int k = 0; k < 4; ++k) {
        const int nx = cx + dx[k], ny = cy + dy[k];
        if (nx < 0 || nx >= w || ny < 0 || ny >= h) continue;
        const int ni = ny * w + nx;
        if (visited[ni]) continue;
        const float dn = depth[ni];
        if (dn <= 0.0f) {
          visited[ni] = 1;
          continue;
        }
        if (std::fabs(dc - dn) / std::max(dc, dn) < thr) {
          visited[ni] = 1;
          seg.push_back(ni);
        }
      }
    }
    if (static_cast<int>(seg.size()) < min_size) {
      for (int i : seg) {
        depth[i] = 0.0f;
        normal[3 * i] = normal[3 * i + 1] = normal[3 * i + 2] = 0.0f;
      }
    }
  }
}

// One gap fill between valid samples a (depth da, normal na) and b, at
// fraction t (GapInterpolation, estimator.cc:1348-1362).
inline void gap_fill(float da, const float* na, float db, const float* nb, float t, float* d, float* n) {
  *d = da + t * (db - da);
  float v[3];
  for (int k = 0; k < 3; ++k) v[k] = (1.0f - t) * na[k] + t * nb[k];
  const float len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (len > 1e-6f)
    for (int k = 0; k < 3; ++k) v[k] /= len;
  for (int k = 0; k < 3; ++k) n[k] = v[k];
}

// Gap interpolation (GapInterpolation, estimator.cc:1319-1418): runs of at
// most max_gap invalid pixels between two valid ones within `thr` relative
// depth are filled linearly, rows first, then columns (only still-empty
// pixels).
void gap_interpolation(std::vector<float>& depth, std::vector<float>& normal, int w, int h, int max_gap, float thr) {
  for (int y = 0; y < h; ++y) {
    int gs = -1;
    for (int x = 0; x < w; ++x) {
      const int i = y * w + x;
      const float d = depth[i];
      if (d <= 0.0f) {
        if (gs < 0) gs = x;
        continue;
      }
      if (gs >= 0) {
        const int lx = gs - 1;
        if (lx >= 0 && x - gs <= max_gap) {
          const float dl = depth[y * w + lx];
          if (std::fabs(dl - d) / std::max(dl, d) < thr) {
            const float nl[3] = {normal[3 * (y * w + lx)], normal[3 * (y * w + lx) + 1], normal[3 * (y * w + lx) + 2]};
            const float nr[3] = {normal[3 * i], normal[3 * i + 1], normal[3 * i + 2]};
            for (int gx = gs; gx < x; ++gx) {
              const float t = static_cast<float>(gx - lx) / static_cast<float>(x - lx);
              gap_fill(dl, nl, d, nr, t, &depth[y * w + gx], &normal[3 * (y * w + gx)]);
            }
          }
        }
      }
      gs = -1;
    }
  }
  for (int x = 0; x < w; ++x) {
    int gs = -1;
    for (int y = 0; y < h; ++y) {
      const int i = y * w + x;
      const float d = depth[i];
      if (d <= 0.0f) {
        if (gs < 0) gs = y;
        continue;
      }
      if (gs >= 0) {
        const int ty = gs - 1;
        if (ty >= 0 && y - gs <= max_gap) {
          const float dt = depth[ty * w + x];
          if (std::fabs(dt - d) / std::max(dt, d) < thr) {
            const float nt[3] = {normal[3 * (ty * w + x)], normal[3 * (ty * w + x) + 1], normal[3 * (ty * w + x) + 2]};
            const float nb[3] = {normal[3 * i], normal[3 * i + 1], normal[3 * i + 2]};
            for (int gy = gs; gy < y; ++gy) {
              const float t = static_cast<float>(gy - ty) / static_cast<float>(y - ty);
              if (depth[gy * w + x] <= 0.0f) gap_fill(dt, nt, d, nb, t, &depth[gy * w + x], &normal[3 * (gy * w + x)]);
            }
          }
        }
      }
      gs = -1;
    }
  }
}

int level_count(const pm_params& p) { return std::max(1, p.hierarchy_levels); }

// Stages the views of pyramid level `level` of `nlevels` in device memory and
// sizes the per-pixel state for it (UploadData, estimator.cc:299-510).
int upload_level(pm_handle* h, int level, int nlevels) {
  PM_HIP(hipSetDevice(h->device), "hipSetDevice");
  ++h->gen;
  // Level views: level l of N is scaled by 1 / 2^(N-1-l) (estimator.cc:182-201).
  std::vector<HostView> scaled;
  const double scale = 1.0 / std::pow(2.0, nlevels - 1 - level);
  bool any_dev = false, any_host = false;
  for (const HostView& v : h->views) (v.dev ? any_dev : any_host) = true;
  if (scale < 0.999) {
    if (any_dev) return fail(PM_E_INVALID, "views added from device memory run at one pyramid level only");
    for (const HostView& v : h->views) scaled.push_back(scaled_view(v, static_cast<float>(scale)));
  }
  const std::vector<HostView>& views = scale < 0.999 ? scaled : h->views;
  const pm_params& p = h->params;
  const int nv = static_cast<int>(views.size());
  h->W = views[0].w;
  h->H = views[0].h;
  h->n_src = nv - 1;
  h->f16 = p.image_storage == PM_STORE_F16;
  const size_t esz = h->f16 ? 2 : 4;

  // Padded images (1-texel replicated border = clamp-to-edge).
  std::vector<PmView> vt(static_cast<size_t>(nv));
  int64_t total = 0, qtotal = 0;
  for (int i = 0; i < nv; ++i) {
    const HostView& v = views[static_cast<size_t>(i)];
    vt[i].qbase = qtotal;
    qtotal += static_cast<int64_t>(v.w + 2) * (v.h + 2);
    vt[i].base = total;
    vt[i].w = v.w;
    vt[i].h = v.h;
    vt[i].pitch = v.w + 2;
    total += static_cast<int64_t>(v.w + 2) * (v.h + 2);
    total = (total + 63) & ~int64_t(63);
  }
  // Host views are padded and converted here; views in device memory by a
  // kernel after the upload (the same conversion, see pad_view_kernel).
  std::vector<unsigned char> host(any_host ? static_cast<size_t>(total) * esz : 0, 0);
  for (int i = 0; i < nv; ++i) {
    const HostView& v = views[static_cast<size_t>(i)];
    if (v.dev) continue;
    for (int y = -1; y <= v.h; ++y) {
      const int cy = std::min(std::max(y, 0), v.h - 1);
      for (int x = -1; x <= v.w; ++x) {
        const int cx = std::min(std::max(x, 0), v.w - 1);
        float val = v.pix[static_cast<size_t>(cy) * v.w + cx];
        const size_t o = static_cast<size_t>(vt[i].base + static_cast<int64_t>(y + 1) * vt[i].pitch + (x + 1));
        if (h->f16 || p.quantize_half) {
          const uint16_t b = f32_to_f16_bits(val);
          if (h->f16) { std::memcpy(&host[o * 2], &b, 2); continue; }
          val = f16_bits_to_f32(b);
        }
        std::memcpy(&host[o * 4], &val, 4);
      }
    }
  }
  // Cameras, source camera centres (float, opencl_kernels.h:1523-1525) and
  // angle weights (float64 then cast, estimator.cc:380-407).
  std::vector<PmCam> cams(static_cast<size_t>(nv));
  for (int i = 0; i < nv; ++i) cams[i] = make_cam(views[static_cast<size_t>(i)]);
  const HostView& v0 = views[0];
  const double c0[3] = {-(v0.R[0] * v0.t[0] + v0.R[3] * v0.t[1] + v0.R[6] * v0.t[2]),
                        -(v0.R[1] * v0.t[0] + v0.R[4] * v0.t[1] + v0.R[7] * v0.t[2]),
                        -(v0.R[2] * v0.t[0] + v0.R[5] * v0.t[1] + v0.R[8] * v0.t[2])};
  for (int i = 1; i < nv; ++i) {
    const PmCam& c = cams[i];
    vt[i].cx = -(c.R[0] * c.t[0] + c.R[3] * c.t[1] + c.R[6] * c.t[2]);
    vt[i].cy = -(c.R[1] * c.t[0] + c.R[4] * c.t[1] + c.R[7] * c.t[2]);
    vt[i].cz = -(c.R[2] * c.t[0] + c.R[5] * c.t[1] + c.R[8] * c.t[2]);
    const HostView& v = views[static_cast<size_t>(i)];
    const double ci[3] = {-(v.R[0] * v.t[0] + v.R[3] * v.t[1] + v.R[6] * v.t[2]),
                          -(v.R[1] * v.t[0] + v.R[4] * v.t[1] + v.R[7] * v.t[2]),
                          -(v.R[2] * v.t[0] + v.R[5] * v.t[1] + v.R[8] * v.t[2])};
    const double b[3] = {ci[0] - c0[0], ci[1] - c0[1], ci[2] - c0[2]};
    const double bl = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    if (bl < 1e-12) {
      vt[i].angle_w = 0.01f;
    } else {
      const double ca = std::fabs(v0.R[6] * (b[0] / bl) + v0.R[7] * (b[1] / bl) + v0.R[8] * (b[2] / bl));
      const double sa = std::sqrt(1.0 - ca * ca);
      vt[i].angle_w = static_cast<float>(std::max(sa, 0.01));
    }
  }
  // Composite maps of the differentiable cost: p = A (camera-frame point) + b
  // with A = K_i R_i R0^T and b = K_i (t_i - R_i R0^T t0), in double.
  std::vector<float> vgeom(static_cast<size_t>(12) * std::max(nv - 1, 1), 0.0f);
  for (int i = 1; i < nv; ++i) {
    const HostView& v = views[static_cast<size_t>(i)];
    double M[9], T[3];  // R_i R0^T and t_i - M t0
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c)
        M[3 * r + c] = v.R[3 * r] * v0.R[3 * c] + v.R[3 * r + 1] * v0.R[3 * c + 1] + v.R[3 * r + 2] * v0.R[3 * c + 2];
    for (int r = 0; r < 3; ++r) T[r] = v.t[r] - (M[3 * r] * v0.t[0] + M[3 * r + 1] * v0.t[1] + M[3 * r + 2] * v0.t[2]);
    float* g = &vgeom[static_cast<size_t>(12) * (i - 1)];
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c)
        g[3 * r + c] = static_cast<float>(v.K[3 * r] * M[c] + v.K[3 * r + 1] * M[3 + c] + v.K[3 * r + 2] * M[6 + c]);
      g[9 + r] = static_cast<float>(v.K[3 * r] * T[0] + v.K[3 * r + 1] * T[1] + v.K[3 * r + 2] * T[2]);
    }
  }
  const size_t npix = static_cast<size_t>(h->W) * h->H;
  if (int r = h->images.alloc(static_cast<size_t>(total) * esz)) return r;
  if (int r = h->quads.alloc(static_cast<size_t>(qtotal) * 4 * esz)) return r;
  if (int r = h->gviews.alloc(vt.size())) return r;
  if (int r = h->gcams.alloc(cams.size())) return r;
  h->ref_plane = 2 * static_cast<int64_t>(h->H) * ((h->W + 1) / 2);
  const int P = (p.patch_size / 2 * 2 + 1) * (p.patch_size / 2 * 2 + 1);
  if (int r = h->ref_taps.alloc(static_cast<size_t>(P) * h->ref_plane)) return r;
  if (int r = h->ref_sums.alloc(static_cast<size_t>(3) * h->ref_plane)) return r;
  const size_t rp = static_cast<size_t>(h->ref_plane);
  if (int r = h->ca.alloc(9 * static_cast<size_t>(h->n_src) * rp)) return r;
  if (int r = h->vwb.alloc(static_cast<size_t>(h->n_src) * rp)) return r;
  if (int r = h->wnb.alloc(rp)) return r;
  if (int r = h->cpos.alloc(8 * rp)) return r;
  if (int r = h->nref.alloc(rp)) return r;
  if (int r = h->refp.alloc(6 * rp)) return r;
  if (int r = h->fb_masks.alloc(static_cast<size_t>(fallback_mask_count(h->W, h->H, h->n_src)))) return r;
  h->ref_ready = false;
  if (int r = h->planes.alloc(npix)) return r;
  if (int r = h->costs.alloc(npix)) return r;
  if (int r = h->rng.alloc(npix)) return r;
  if (int r = h->sel.alloc(npix)) return r;
  if (int r = h->out_depth.alloc(npix)) return r;
  if (int r = h->out_normal.alloc(3 * npix)) return r;
  if (int r = h->out_cost.alloc(npix)) return r;
  if (int r = h->out_conf.alloc(npix)) return r;
  if (any_host)
    PM_HIP(hipMemcpyAsync(h->images.p, host.data(), host.size(), hipMemcpyHostToDevice, h->stream), "upload images");
  for (int i = 0; i < nv; ++i) {
    const HostView& v = views[static_cast<size_t>(i)];
    if (!v.dev) continue;
    PM_HIP(launch_pad_view(v.dev, v.dev_format, v.dev_stride, v.w, v.h, h->images.p, vt[i].base, vt[i].pitch, h->f16,
                           p.quantize_half != 0, h->stream), "pad kernel launch");
  }
  PM_HIP(hipMemcpyAsync(h->gviews.p, vt.data(), vt.size() * sizeof(PmView), hipMemcpyHostToDevice, h->stream),
         "upload view table");
  if (int r = h->vgeom.alloc(vgeom.size())) return r;
  PM_HIP(hipMemcpyAsync(h->vgeom.p, vgeom.data(), vgeom.size() * sizeof(float), hipMemcpyHostToDevice, h->stream),
         "upload view geometry");
  PM_HIP(hipMemcpyAsync(h->gcams.p, cams.data(), cams.size() * sizeof(PmCam), hipMemcpyHostToDevice, h->stream),
         "upload cameras");
  {
    PmArgs qa{};
    qa.images = h->images.p;
    qa.quads = h->quads.p;
    qa.views = h->gviews.p;
    PM_HIP(launch_quads(qa, nv, h->f16, h->stream), "quad kernel launch");
  }
  // Prior (estimator.cc:480-486) and previous depths (:499-508, 66-116).
  if (!h->prior_masks.empty() || !h->sfm_points.empty()) {
    if (int r = h->prior.alloc(npix)) return r;
    if (int r = h->masks.alloc(npix)) return r;
    if (int r = h->snapshot.alloc(npix)) return r;
    std::vector<float> pp(4 * npix);
    std::vector<uint32_t> pm(npix);
    if (!h->sfm_points.empty()) {
      // built for this level from the SfM points (ComputeSfMPlanarPrior per
      // level, estimator.cc:214-218)
      const HostView& v0 = views[0];
      if (int r = pm_sfm_planar_prior(v0.K, v0.R, v0.t, h->W, h->H, p.depth_min, p.depth_max, h->sfm_points.data(),
                                      static_cast<int>(h->sfm_points.size() / 3), pp.data(), pm.data(), nullptr))
        return r;
    } else {
      // Planes are in the reference camera frame, so a coarse level takes the
      // prior of the full-resolution pixel at the centre of its footprint.
      const int fw = h->views[0].w, fh = h->views[0].h;
      for (int y = 0; y < h->H; ++y) {
        const int fy = std::min(static_cast<int>((y + 0.5) * fh / h->H), fh - 1);
        for (int x = 0; x < h->W; ++x) {
          const int fx = std::min(static_cast<int>((x + 0.5) * fw / h->W), fw - 1);
          const size_t si = static_cast<size_t>(fy) * fw + fx, di = static_cast<size_t>(y) * h->W + x;
          std::memcpy(&pp[4 * di], &h->prior_planes[4 * si], 4 * sizeof(float));
          pm[di] = h->prior_masks[si];
        }
      }
    }
    PM_HIP(hipMemcpyAsync(h->prior.p, pp.data(), npix * sizeof(float4), hipMemcpyHostToDevice, h->stream),
           "upload prior");
    PM_HIP(hipMemcpyAsync(h->masks.p, pm.data(), npix * sizeof(uint32_t), hipMemcpyHostToDevice, h->stream),
           "upload masks");
    PM_HIP(hipStreamSynchronize(h->stream), "upload sync");
  } else {
    h->prior.release();
    h->masks.release();
    h->snapshot.release();
  }
  h->prev_mask = 0u;
  bool any_prev = false;
  for (int i = 0; i < h->n_src; ++i) any_prev |= !h->prev_depths[i].d.empty();
  if (any_prev && h->geom_weight > 0.0f) {
    if (int r = h->prevd.alloc(static_cast<size_t>(h->n_src) * npix)) return r;
    std::vector<float> flat(static_cast<size_t>(h->n_src) * npix, 0.0f);
    for (int i = 0; i < h->n_src; ++i) {
      const pm_handle::PrevDepth& e = h->prev_depths[i];
      if (e.d.empty()) continue;
      float* dst = &flat[static_cast<size_t>(i) * npix];
      if (e.w == h->W && e.h == h->H) {
        std::memcpy(dst, e.d.data(), npix * sizeof(float));
      } else {  // UploadPreviousDepths resizes to the level (estimator.cc:92-101)
        resize_linear(e.d.data(), e.w, e.h, h->W, h->H, dst);
      }
      h->prev_mask |= 1u << i;
    }
    PM_HIP(hipMemcpyAsync(h->prevd.p, flat.data(), flat.size() * sizeof(float), hipMemcpyHostToDevice, h->stream),
           "upload previous depths");
    PM_HIP(hipStreamSynchronize(h->stream), "upload sync");
  } else {
    h->prevd.release();
  }
  PM_HIP(hipStreamSynchronize(h->stream), "upload sync");
  h->uploaded = true;
  h->state_ready = false;
  return PM_OK;
}


}  // namespace

extern "C" {

int pm_upload(pm_handle* h) {
  if (int r = check_handle(h)) return r;
  if (int r = validate(h)) return r;
  const int n = level_count(h->params);
  return upload_level(h, n - 1, n);
}

}  // extern "C"

namespace {

// Whether the candidates pass of colour c may take the current planes'
// costs from slot 8 of ca instead of scoring them again.
bool current_cached(const pm_handle* h, int c) { return h->cur_valid[c] && h->cur_gen == h->gen; }
void current_stale(pm_handle* h) { h->cur_valid[0] = h->cur_valid[1] = false; }
void current_filled(pm_handle* h, int c) {
  if (h->cur_gen != h->gen) {
    current_stale(h);
    h->cur_gen = h->gen;
  }
  h->cur_valid[c] = true;
}

// Prior re-seeding (PriorReinit, estimator.cc:574-614), when a prior is set.
int run_prior(pm_handle* h, const PmArgs& a) {
  if (!h->masks.p) return PM_OK;
  PM_HIP(hipMemcpyAsync(h->snapshot.p, h->planes.p, h->planes.n * sizeof(float4), hipMemcpyDeviceToDevice,
                        h->stream), "prior snapshot");
  PM_HIP(launch_prior(a, h->params.patch_size / 2, h->f16, h->stream), "prior kernel launch");
  return PM_OK;
}

// Host post-processing of the decoded level (ReadBackResults,
// estimator.cc:1194-1222): speckle removal, gap filling, then confidence
// again for the changed depth map.
int postprocess(pm_handle* h) {
  const pm_params& p = h->params;
  if (p.speckle_min_size <= 0 && p.gap_max_size <= 0) return PM_OK;
  const size_t npix = static_cast<size_t>(h->W) * h->H;
  std::vector<float> depth(npix), normal(3 * npix);
  PM_HIP(hipMemcpyAsync(depth.data(), h->out_depth.p, npix * 4, hipMemcpyDeviceToHost, h->stream), "download");
  PM_HIP(hipMemcpyAsync(normal.data(), h->out_normal.p, npix * 12, hipMemcpyDeviceToHost, h->stream), "download");
  PM_HIP(hipStreamSynchronize(h->stream), "postprocess sync");
  if (p.speckle_min_size > 0) remove_small_segments(depth, normal, h->W, h->H, p.speckle_min_size, 0.01f);
  if (p.gap_max_size > 0) gap_interpolation(depth, normal, h->W, h->H, p.gap_max_size, 0.025f);
  PM_HIP(hipMemcpyAsync(h->out_depth.p, depth.data(), npix * 4, hipMemcpyHostToDevice, h->stream), "upload");
  PM_HIP(hipMemcpyAsync(h->out_normal.p, normal.data(), npix * 12, hipMemcpyHostToDevice, h->stream), "upload");
  PM_HIP(launch_confidence(static_cast<int64_t>(npix), h->out_depth.p, h->out_cost.p, h->out_conf.p, h->stream),
         "confidence kernel launch");
  PM_HIP(hipStreamSynchronize(h->stream), "postprocess sync");
  return PM_OK;
}

}  // namespace

extern "C" {

int pm_init_state(pm_handle* h) {
  if (int r = check_handle(h)) return r;
  if (!h->uploaded) return fail(PM_E_STATE, "pm_init_state: views are not uploaded (call pm_upload)");
  if (int r = validate(h)) return r;
  PM_HIP(hipSetDevice(h->device), "hipSetDevice");
  PmArgs a = make_args(h);
  const int hp = h->params.patch_size / 2;
  if (int r = ensure_refpatch(h)) return r;
  current_stale(h);
  if (int r = timed(h, PM_STAGE_INIT, "init kernel launch", [&] { return launch_init(a, hp, h->f16, h->stream); }))
    return r;
  if (!h->masks.p) {  // the init leaves the random planes' costs in slot 8; a prior re-seeds planes
    current_filled(h, 0);
    current_filled(h, 1);
  }
  if (int r = run_prior(h, a)) return r;  // estimator.cc:574-614
  h->state_ready = true;
  return PM_OK;
}

int pm_iterate(pm_handle* h, int iteration, int color) {
  if (int r = check_handle(h)) return r;
  if (!h->state_ready) return fail(PM_E_STATE, "pm_iterate: state is not initialised (call pm_init_state)");
  PM_HIP(hipSetDevice(h->device), "hipSetDevice");
  if (int r = ensure_refpatch(h)) return r;
  PmArgs a = make_args(h);
  a.iteration = iteration;
  a.smooth_w = smooth_weight_at(h->params, iteration);
  const int hp = h->params.patch_size / 2;
  const bool f16 = h->f16;
  hipStream_t st = h->stream;
  for (int c = 0; c < 2; ++c) {
    if (color >= 0 && c != color) continue;
    a.color = c;
    a.cur_cached = current_cached(h, c) ? 1 : 0;
    h->cur_valid[c] = false;  // until every launch of this pass went out
    if (int r = timed(h, PM_STAGE_CANDIDATES, "candidates kernel launch",
                      [&] { return launch_candidates(a, hp, f16, cost_tile(h), st); })) return r;
    if (int r = timed(h, PM_STAGE_FALLBACK, "fallback kernel launch",
                      [&] { return launch_fallback(a, hp, f16, 0, cost_tile(h), fallback_cap(h), st); })) return r;
    if (int r = timed(h, PM_STAGE_SELECT, "select kernel launch", [&] { return launch_select(a, pixel_tile(h), st); })) return r;
    if (int r = timed(h, PM_STAGE_REFINE_COST, "refine-cost kernel launch",
                      [&] { return launch_refine_cost(a, hp, f16, cost_tile(h), st); })) return r;
    if (int r = timed(h, PM_STAGE_FALLBACK, "fallback kernel launch",
                      [&] { return launch_fallback(a, hp, f16, 1, cost_tile(h), fallback_cap(h), st); })) return r;
    if (int r = timed(h, PM_STAGE_REFINE, "refine kernel launch", [&] { return launch_refine(a, pixel_tile(h), st); })) return r;
    current_filled(h, c);
  }
  return PM_OK;
}

int pm_filter(pm_handle* h) {
  if (int r = check_handle(h)) return r;
  if (!h->state_ready) return fail(PM_E_STATE, "pm_filter: state is not initialised");
  current_stale(h);
  PmArgs a = make_args(h);
  for (int c = 0; c < 2; ++c) {
    a.color = c;
    PM_HIP(launch_filter(a, pixel_tile(h), h->stream), "filter kernel launch");
  }
  return PM_OK;
}

int pm_decode(pm_handle* h) {
  if (int r = check_handle(h)) return r;
  if (!h->state_ready) return fail(PM_E_STATE, "pm_decode: state is not initialised");
  PmArgs a = make_args(h);
  PmDecodeOut o{h->out_depth.p, h->out_normal.p, h->out_cost.p, h->out_conf.p};
  PM_HIP(launch_decode(a, o, h->stream), "decode kernel launch");
  return PM_OK;
}

int pm_plane_sweep(pm_handle* h, int num_planes) {
  if (int r = check_handle(h)) return r;
  if (num_planes < 1) return fail(PM_E_INVALID, "pm_plane_sweep: num_planes must be at least 1");
  if (!h->uploaded) return fail(PM_E_STATE, "pm_plane_sweep: views are not uploaded (call pm_upload)");
  if (int r = validate(h)) return r;
  PM_HIP(hipSetDevice(h->device), "hipSetDevice");
  if (int r = ensure_refpatch(h)) return r;
  current_stale(h);
  PmArgs a = make_args(h);
  PM_HIP(launch_plane_sweep(a, h->params.patch_size / 2, h->f16, num_planes, h->stream), "plane sweep kernel launch");
  h->state_ready = true;
  if (int r = pm_decode(h)) return r;
  return postprocess(h);
}

}  // extern "C"

namespace {

// One pyramid level of `nlevels` (RunLevel, estimator.cc:172-255): stage
// the level, start from random planes (level 0) or from the decoded
// coarser level, prior re-init, iterations, filter, decode and host
// post-processing; a level below the last is kept for the next one.
int run_level(pm_handle* h, int level, int nlevels) {
  const bool last = level == nlevels - 1;
  if (nlevels > 1 || !h->uploaded)
    if (int r = upload_level(h, level, nlevels)) return r;
  if (level == 0) {
    if (int r = pm_init_state(h)) return r;
  } else {
    if (h->coarse_w <= 0 || !h->coarse_depth.p) return fail(PM_E_STATE, "no coarser level to start from");
    current_stale(h);
    if (int r = ensure_refpatch(h)) return r;
    PmArgs a = make_args(h);
    if (int r = timed(h, PM_STAGE_INIT, "upsample kernel launch", [&] {
          return launch_upsample(a, h->coarse_depth.p, h->coarse_normal.p, h->coarse_w, h->coarse_h, h->stream);
        }))
      return r;
    if (int r = run_prior(h, a)) return r;
    h->state_ready = true;
  }
  if (last && h->timing) PM_HIP(hipEventRecord(h->ev[1], h->stream), "event");
  for (int it = 0; it < h->params.max_iterations; ++it)
    if (int r = pm_iterate(h, it, -1)) return r;
  if (last && h->timing) PM_HIP(hipEventRecord(h->ev[2], h->stream), "event");
  if (h->params.checkerboard_filter)
    if (int r = pm_filter(h)) return r;
  if (int r = pm_decode(h)) return r;
  if (int r = postprocess(h)) return r;
  if (!last) {  // keep the decoded level for the next one (estimator.cc:238-246)
    const size_t npix = static_cast<size_t>(h->W) * h->H;
    ++h->gen;  // a batch's captured upsample launches may point at the buffers replaced here
    if (int r = h->coarse_depth.alloc(npix)) return r;
    if (int r = h->coarse_normal.alloc(3 * npix)) return r;
    PM_HIP(hipMemcpyAsync(h->coarse_depth.p, h->out_depth.p, npix * sizeof(float), hipMemcpyDeviceToDevice,
                          h->stream), "keep coarse depth");
    PM_HIP(hipMemcpyAsync(h->coarse_normal.p, h->out_normal.p, 3 * npix * sizeof(float), hipMemcpyDeviceToDevice,
                          h->stream), "keep coarse normal");
    h->coarse_w = h->W;
    h->coarse_h = h->H;
  }
  return PM_OK;
}

}  // namespace

extern "C" {

int pm_run(pm_handle* h) {
  if (int r = check_handle(h)) return r;
  if (int r = validate(h)) return r;
  const int nlevels = level_count(h->params);
  PM_HIP(hipSetDevice(h->device), "hipSetDevice");
  h->timing = true;
  h->kev_used = 0;
  h->kev_marks.clear();
  PM_HIP(hipEventRecord(h->ev[0], h->stream), "event");
  // Coarse to fine (Run, estimator.cc:134-139).
  for (int level = 0; level < nlevels; ++level) {
    if (int r = run_level(h, level, nlevels)) {
      h->timing = false;
      return r;
    }
  }
  h->timing = false;
  h->next_level = nlevels;
  PM_HIP(hipEventRecord(h->ev[3], h->stream), "event");
  PM_HIP(hipStreamSynchronize(h->stream), "pm_run");
  PM_HIP(hipEventElapsedTime(&h->last_run_ms, h->ev[0], h->ev[3]), "event time");
  PM_HIP(hipEventElapsedTime(&h->last_sweep_ms, h->ev[1], h->ev[2]), "event time");
  h->last_sweeps = 2 * h->params.max_iterations;
  for (int k = 0; k < PM_STAGES; ++k) { h->stage_ms[k] = 0.0f; h->stage_launches[k] = 0; }
  for (const auto& m : h->kev_marks) {
    float ms = 0.0f;
    PM_HIP(hipEventElapsedTime(&ms, h->kev[m.second], h->kev[m.second + 1]), "event time");
    h->stage_ms[m.first] += ms;
    h->stage_launches[m.first] += 1;
  }
  return PM_OK;
}

int pm_run_level(pm_handle* h, int level) {
  if (int r = check_handle(h)) return r;
  if (int r = validate(h)) return r;
  const int nlevels = level_count(h->params);
  if (level < 0 || level >= nlevels) return fail(PM_E_INVALID, "pm_run_level: level out of range");
  if (level == 0) h->next_level = 0;
  if (level != h->next_level) return fail(PM_E_STATE, "pm_run_level: levels run in order, starting from 0");
  PM_HIP(hipSetDevice(h->device), "hipSetDevice");
  if (int r = run_level(h, level, nlevels)) return r;
  PM_HIP(hipStreamSynchronize(h->stream), "pm_run_level");
  h->next_level = level + 1;
  return PM_OK;
}

int pm_level_size(pm_handle* h, int* width, int* height) {
  if (int r = check_handle(h)) return r;
  if (!width || !height) return fail(PM_E_INVALID, "null output");
  if (!h->uploaded) return fail(PM_E_STATE, "pm_level_size: no level is staged");
  *width = h->W;
  *height = h->H;
  return PM_OK;
}

int pm_release_device(pm_handle* h) {
  if (int r = check_handle(h)) return r;
  ++h->gen;
  PM_HIP(hipSetDevice(h->device), "hipSetDevice");
  PM_HIP(hipStreamSynchronize(h->stream), "pm_release_device");
  h->park_buffers(true);
  h->uploaded = false;
  h->state_ready = false;
  h->ref_ready = false;
  return PM_OK;
}

int pm_ncc_cost_grad(pm_handle* h, const float* depth, const float* ab, const uint32_t* views, float* cost,
                     float* grad, int device_pointers) {
  if (int r = check_handle(h)) return r;
  if (!h->uploaded) return fail(PM_E_STATE, "pm_ncc_cost_grad: views are not uploaded (call pm_upload)");
  if (!depth || !ab || !views || !cost) return fail(PM_E_INVALID, "pm_ncc_cost_grad: null argument");
  PM_HIP(hipSetDevice(h->device), "hipSetDevice");
  if (int r = ensure_refpatch(h)) return r;
  const PmArgs a = make_args(h);
  const int hp = h->params.patch_size / 2;
  const size_t npix = static_cast<size_t>(h->W) * h->H;
  if (device_pointers) {
    PM_HIP(launch_ncc_energy(a, hp, h->f16, h->vgeom.p, depth, ab, views, cost, grad, h->stream),
           "energy kernel launch");
    PM_HIP(hipStreamSynchronize(h->stream), "pm_ncc_cost_grad");
    return PM_OK;
  }
  DevBuf<float> dd, dab, dc, dg;
  DevBuf<uint32_t> dv;
  if (int r = dd.alloc(npix)) return r;
  if (int r = dab.alloc(2 * npix)) return r;
  if (int r = dv.alloc(npix)) return r;
  if (int r = dc.alloc(npix)) return r;
  if (grad)
    if (int r = dg.alloc(3 * npix)) return r;
  PM_HIP(hipMemcpyAsync(dd.p, depth, npix * 4, hipMemcpyHostToDevice, h->stream), "upload");
  PM_HIP(hipMemcpyAsync(dab.p, ab, 2 * npix * 4, hipMemcpyHostToDevice, h->stream), "upload");
  PM_HIP(hipMemcpyAsync(dv.p, views, npix * 4, hipMemcpyHostToDevice, h->stream), "upload");
  PM_HIP(launch_ncc_energy(a, hp, h->f16, h->vgeom.p, dd.p, dab.p, dv.p, dc.p, grad ? dg.p : nullptr, h->stream),
         "energy kernel launch");
  PM_HIP(hipMemcpyAsync(cost, dc.p, npix * 4, hipMemcpyDeviceToHost, h->stream), "download");
  if (grad) PM_HIP(hipMemcpyAsync(grad, dg.p, 3 * npix * 4, hipMemcpyDeviceToHost, h->stream), "download");
  PM_HIP(hipStreamSynchronize(h->stream), "pm_ncc_cost_grad");
  return PM_OK;
}

int pm_gauss_newton(pm_handle* h, int steps, float damping) {
  if (int r = check_handle(h)) return r;
  if (!h->state_ready) return fail(PM_E_STATE, "pm_gauss_newton: no state to refine (call pm_run)");
  if (steps < 0 || !(damping >= 0.0f)) return fail(PM_E_INVALID, "pm_gauss_newton: bad steps or damping");
  PM_HIP(hipSetDevice(h->device), "hipSetDevice");
  if (int r = ensure_refpatch(h)) return r;
  // The PatchMatch planes' per-view costs, for the guard: slot 8 of ca when
  // it is current for both colours (no filter or state change since the
  // last passes)
  const bool pm_costs = current_cached(h, 0) && current_cached(h, 1);
  current_stale(h);  // its state takes ca, and it moves planes
  const PmArgs a = make_args(h);
  const int hp = h->params.patch_size / 2;
  PM_HIP(hipEventRecord(h->ev[0], h->stream), "event");
  // The per-pixel state between steps lives in the colour passes' candidate
  // cost buffer, idle after a run (the next run rewrites it), when it is big
  // enough (two or more sources); else in a buffer of its own.
  const size_t need = static_cast<size_t>(gauss_newton_state_floats()) * h->W * h->H;
  float* state = h->ca.p;
  if (h->ca.n < need) {
    if (int r = h->gn_state.alloc(need)) return r;
    state = h->gn_state.p;
  }
  const size_t npix = static_cast<size_t>(h->W) * h->H;
  if (steps > 0) {
    if (int r = h->gn_orig.alloc(npix)) return r;
    PM_HIP(hipMemcpyAsync(h->gn_orig.p, h->planes.p, npix * sizeof(float4), hipMemcpyDeviceToDevice, h->stream),
           "keep the PatchMatch planes");
  }
  for (int i = 0; i < steps; ++i)
    PM_HIP(launch_gauss_newton(a, hp, h->f16, h->vgeom.p, damping, state, i == 0, h->stream),
           "Gauss-Newton kernel launch");
  // the guard's list of moved pixels, its count and its block offsets take
  // the state buffer, free after the last step (2n + 1 ints at most; it holds
  // gauss_newton_state_floats() = 16 floats per pixel)
  // slot 8 lies past the state (16 floats per pixel) whenever the state
  // lives in ca, which takes at least two sources
  const size_t slot8 = static_cast<size_t>(8) * h->n_src * static_cast<size_t>(h->ref_plane);
  const float* orig_costs = pm_costs && (state != h->ca.p || need <= slot8) ? h->ca.p + slot8 : nullptr;
  if (steps > 0)
    PM_HIP(launch_refine_guard(a, hp, h->f16, h->gn_orig.p, orig_costs, reinterpret_cast<int*>(state), h->stream),
           "refinement guard launch");
  PM_HIP(hipEventRecord(h->ev[1], h->stream), "event");
  if (int r = pm_decode(h)) return r;
  if (int r = postprocess(h)) return r;
  PM_HIP(hipStreamSynchronize(h->stream), "pm_gauss_newton");
  PM_HIP(hipEventElapsedTime(&h->last_gn_ms, h->ev[0], h->ev[1]), "event time");
  return PM_OK;
}

int pm_last_refine_timing(pm_handle* h, float* ms) {
  if (int r = check_handle(h)) return r;
  if (ms) *ms = h->last_gn_ms;
  return PM_OK;
}

int pm_last_stage_timing(pm_handle* h, int stage, float* ms, int* launches) {
  if (int r = check_handle(h)) return r;
  if (stage < 0 || stage >= PM_STAGES) return fail(PM_E_INVALID, "unknown stage");
  if (ms) *ms = h->stage_ms[stage];
  if (launches) *launches = h->stage_launches[stage];
  return PM_OK;
}

int pm_last_timing(pm_handle* h, float* run_ms, float* sweep_ms, int* sweeps) {
  if (int r = check_handle(h)) return r;
  if (run_ms) *run_ms = h->last_run_ms;
  if (sweep_ms) *sweep_ms = h->last_sweep_ms;
  if (sweeps) *sweeps = h->last_sweeps;
  return PM_OK;
}

int pm_download(pm_handle* h, float* depth, float* plane, float* normal, float* cost, uint32_t* views,
                float* confidence) {
  if (int r = check_handle(h)) return r;
  if (!h->state_ready) return fail(PM_E_STATE, "pm_download: nothing computed yet");
  PM_HIP(hipSetDevice(h->device), "hipSetDevice");
  const size_t npix = static_cast<size_t>(h->W) * h->H;
  auto cp = [&](void* dst, const void* src, size_t bytes) -> int {
    if (!dst) return PM_OK;
    PM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream), "download");
    return PM_OK;
  };
  if (int r = cp(depth, h->out_depth.p, npix * 4)) return r;
  if (int r = cp(plane, h->planes.p, npix * 16)) return r;
  if (int r = cp(normal, h->out_normal.p, npix * 12)) return r;
  if (int r = cp(cost, h->out_cost.p, npix * 4)) return r;
  if (int r = cp(views, h->sel.p, npix * 4)) return r;
  if (int r = cp(confidence, h->out_conf.p, npix * 4)) return r;
  PM_HIP(hipStreamSynchronize(h->stream), "download sync");
  return PM_OK;
}

int pm_compute(pm_handle* h, float* depth, float* plane, float* normal, float* cost, uint32_t* views,
               float* confidence) {
  if (int r = check_handle(h)) return r;
  if (int r = validate(h)) return r;
  // with several levels pm_run stages every level itself
  if (level_count(h->params) == 1)
    if (int r = pm_upload(h)) return r;
  if (int r = pm_run(h)) return r;
  return pm_download(h, depth, plane, normal, cost, views, confidence);
}

int pm_read_state(pm_handle* h, float* planes, float* costs, uint32_t* rng, uint32_t* sel) {
  if (int r = check_handle(h)) return r;
  if (!h->state_ready) return fail(PM_E_STATE, "pm_read_state: state is not initialised");
  PM_HIP(hipSetDevice(h->device), "hipSetDevice");
  const size_t npix = static_cast<size_t>(h->W) * h->H;
  if (planes) PM_HIP(hipMemcpyAsync(planes, h->planes.p, npix * 16, hipMemcpyDeviceToHost, h->stream), "read");
  if (costs) PM_HIP(hipMemcpyAsync(costs, h->costs.p, npix * 4, hipMemcpyDeviceToHost, h->stream), "read");
  if (rng) PM_HIP(hipMemcpyAsync(rng, h->rng.p, npix * 8, hipMemcpyDeviceToHost, h->stream), "read");
  if (sel) PM_HIP(hipMemcpyAsync(sel, h->sel.p, npix * 4, hipMemcpyDeviceToHost, h->stream), "read");
  PM_HIP(hipStreamSynchronize(h->stream), "read sync");
  return PM_OK;
}

int pm_write_state(pm_handle* h, const float* planes, const float* costs, const uint32_t* rng,
                   const uint32_t* sel) {
  if (int r = check_handle(h)) return r;
  if (!h->uploaded) return fail(PM_E_STATE, "pm_write_state: views are not uploaded (call pm_upload)");
  if (!planes || !costs || !rng || !sel) return fail(PM_E_INVALID, "pm_write_state needs all four buffers");
  PM_HIP(hipSetDevice(h->device), "hipSetDevice");
  const size_t npix = static_cast<size_t>(h->W) * h->H;
  PM_HIP(hipMemcpyAsync(h->planes.p, planes, npix * 16, hipMemcpyHostToDevice, h->stream), "write");
  PM_HIP(hipMemcpyAsync(h->costs.p, costs, npix * 4, hipMemcpyHostToDevice, h->stream), "write");
  PM_HIP(hipMemcpyAsync(h->rng.p, rng, npix * 8, hipMemcpyHostToDevice, h->stream), "write");
  PM_HIP(hipMemcpyAsync(h->sel.p, sel, npix * 4, hipMemcpyHostToDevice, h->stream), "write");
  PM_HIP(hipStreamSynchronize(h->stream), "write sync");
  h->state_ready = true;
  current_stale(h);
  return PM_OK;
}

int pm_score_planes(pm_handle* h, int n, const int32_t* xy, const float* planes, float* costs, float* init_cost,
                    uint32_t* sel) {
  if (int r = check_handle(h)) return r;
  if (!h->uploaded) return fail(PM_E_STATE, "pm_score_planes: views are not uploaded (call pm_upload)");
  if (n < 0 || (n > 0 && (!xy || !planes || !costs))) return fail(PM_E_INVALID, "pm_score_planes: bad arguments");
  if (n == 0) return PM_OK;
  for (int i = 0; i < n; ++i)
    if (xy[2 * i] < 0 || xy[2 * i] >= h->W || xy[2 * i + 1] < 0 || xy[2 * i + 1] >= h->H)
      return fail(PM_E_INVALID, "pm_score_planes: pixel outside the reference view");
  PM_HIP(hipSetDevice(h->device), "hipSetDevice");
  DevBuf<int32_t> dxy;
  DevBuf<float4> dpl;
  DevBuf<float> dcost, dinit;
  DevBuf<uint32_t> dsel;
  if (int r = dxy.alloc(2 * static_cast<size_t>(n))) return r;
  if (int r = dpl.alloc(static_cast<size_t>(n))) return r;
  if (int r = dcost.alloc(static_cast<size_t>(n) * h->n_src)) return r;
  if (int r = dinit.alloc(static_cast<size_t>(n))) return r;
  if (int r = dsel.alloc(static_cast<size_t>(n))) return r;
  int rc = PM_OK;
  if (int r = ensure_refpatch(h)) return r;
  do {
    PmArgs a = make_args(h);
    if (hipMemcpyAsync(dxy.p, xy, 8 * static_cast<size_t>(n), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipMemcpyAsync(dpl.p, planes, 16 * static_cast<size_t>(n), hipMemcpyHostToDevice, h->stream) != hipSuccess) {
      rc = fail(PM_E_DEVICE, "pm_score_planes: upload failed");
      break;
    }
    hipError_t e = launch_score(a, h->params.patch_size / 2, h->f16, n, dxy.p, dpl.p, dcost.p, dinit.p, dsel.p,
                                h->stream);
    if (e != hipSuccess) { rc = fail(PM_E_DEVICE, std::string("score kernel: ") + hipGetErrorString(e)); break; }
    if (hipMemcpyAsync(costs, dcost.p, 4 * static_cast<size_t>(n) * h->n_src, hipMemcpyDeviceToHost, h->stream) !=
        hipSuccess) { rc = fail(PM_E_DEVICE, "pm_score_planes: download failed"); break; }
    if (init_cost) (void)hipMemcpyAsync(init_cost, dinit.p, 4 * static_cast<size_t>(n), hipMemcpyDeviceToHost, h->stream);
    if (sel) (void)hipMemcpyAsync(sel, dsel.p, 4 * static_cast<size_t>(n), hipMemcpyDeviceToHost, h->stream);
    e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) rc = fail(PM_E_DEVICE, std::string("pm_score_planes: ") + hipGetErrorString(e));
  } while (0);
  dxy.release(); dpl.release(); dcost.release(); dinit.release(); dsel.release();
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free scoring.  pm_score_planes stages its (pixel, plane) pairs
// and results in device buffers it allocates per call; callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops)
// use pm_score_planes_noalloc instead, which stages them in this file-scope
// array.  Each device holds its own copy at its own address, taken once per
// device with hipGetSymbolAddress (the host-side name is not a device
// address).  Every call on a device shares it: calls that overlap on one
// device race.  Nothing in it needs clearing: the score kernel writes every
// output of the pairs it is given.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t kNoAllocPlanes = 0;                                                    // float4 per pair
constexpr size_t kNoAllocXy = kNoAllocPlanes + 16 * static_cast<size_t>(PM_NOALLOC_MAX_PAIRS);  // 2 int32
constexpr size_t kNoAllocCosts = kNoAllocXy + 8 * static_cast<size_t>(PM_NOALLOC_MAX_PAIRS);    // n_src floats
constexpr size_t kNoAllocInit = kNoAllocCosts + 4 * static_cast<size_t>(PM_NOALLOC_MAX_PAIRS) * PM_MAX_SOURCES;
constexpr size_t kNoAllocSel = kNoAllocInit + 4 * static_cast<size_t>(PM_NOALLOC_MAX_PAIRS);
constexpr size_t kNoAllocBytes = kNoAllocSel + 4 * static_cast<size_t>(PM_NOALLOC_MAX_PAIRS);  // 6 MiB
static_assert(kNoAllocBytes == 96u * PM_NOALLOC_MAX_PAIRS, "scratch layout");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Device address of `scratch` on each device, taken on the first
// allocation-free call there (with that device current).
std::mutex g_scratch_mutex;
std::vector<unsigned char*> g_scratch_ptr;

int scratch_of(int device, unsigned char** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (device < 0) return fail(PM_E_INVALID, "scratch: bad device");
  if (g_scratch_ptr.size() <= static_cast<size_t>(device)) g_scratch_ptr.resize(static_cast<size_t>(device) + 1, nullptr);
  unsigned char*& ptr = g_scratch_ptr[static_cast<size_t>(device)];
  if (!ptr) {
    void* p = nullptr;
    PM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) return fail(PM_E_DEVICE, "hipGetSymbolAddress(scratch) returned no address");
    ptr = static_cast<unsigned char*>(p);
  }
  *out = ptr;
  return PM_OK;
}

}  // namespace

extern "C" {

int pm_score_planes_noalloc(pm_handle* h, int n, const int32_t* xy, const float* planes, float* costs,
                            float* init_cost, uint32_t* sel) {
  if (int r = check_handle(h)) return r;
  if (!h->uploaded) return fail(PM_E_STATE, "pm_score_planes_noalloc: views are not uploaded (call pm_upload)");
  if (n < 0 || (n > 0 && (!xy || !planes || !costs)))
    return fail(PM_E_INVALID, "pm_score_planes_noalloc: bad arguments");
  if (n > PM_NOALLOC_MAX_PAIRS)
    return fail(PM_E_INVALID, "pm_score_planes_noalloc: at most " + std::to_string(PM_NOALLOC_MAX_PAIRS) +
                                  " (pixel, plane) pairs per call (asked for " + std::to_string(n) + ")");
  if (n == 0) return PM_OK;
  for (int i = 0; i < n; ++i)
    if (xy[2 * i] < 0 || xy[2 * i] >= h->W || xy[2 * i + 1] < 0 || xy[2 * i + 1] >= h->H)
      return fail(PM_E_INVALID, "pm_score_planes_noalloc: pixel outside the reference view");
  PM_HIP(hipSetDevice(h->device), "hipSetDevice");
  unsigned char* base = nullptr;
  if (int r = scratch_of(h->device, &base)) return r;
  if (int r = ensure_refpatch(h)) return r;  // a launch into the handle's buffers, no allocation
  auto* dpl = reinterpret_cast<float4*>(base + kNoAllocPlanes);
  auto* dxy = reinterpret_cast<int32_t*>(base + kNoAllocXy);
  auto* dcost = reinterpret_cast<float*>(base + kNoAllocCosts);
  auto* dinit = reinterpret_cast<float*>(base + kNoAllocInit);
  auto* dsel = reinterpret_cast<uint32_t*>(base + kNoAllocSel);
  const size_t un = static_cast<size_t>(n);
  PM_HIP(hipMemcpyAsync(dxy, xy, 8 * un, hipMemcpyHostToDevice, h->stream), "pm_score_planes_noalloc: upload");
  PM_HIP(hipMemcpyAsync(dpl, planes, 16 * un, hipMemcpyHostToDevice, h->stream), "pm_score_planes_noalloc: upload");
  PmArgs a = make_args(h);
  PM_HIP(launch_score(a, h->params.patch_size / 2, h->f16, n, dxy, dpl, dcost, dinit, dsel, h->stream),
         "pm_score_planes_noalloc: score kernel");
  PM_HIP(hipMemcpyAsync(costs, dcost, 4 * un * h->n_src, hipMemcpyDeviceToHost, h->stream),
         "pm_score_planes_noalloc: download");
  if (init_cost)
    PM_HIP(hipMemcpyAsync(init_cost, dinit, 4 * un, hipMemcpyDeviceToHost, h->stream),
           "pm_score_planes_noalloc: download");
  if (sel)
    PM_HIP(hipMemcpyAsync(sel, dsel, 4 * un, hipMemcpyDeviceToHost, h->stream), "pm_score_planes_noalloc: download");
  PM_HIP(hipStreamSynchronize(h->stream), "pm_score_planes_noalloc");
  return PM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Colour-tile widths of a single estimator.
// ---------------------------------------------------------------------------
extern "C" {

int pm_set_tiles(pm_handle* h, int cost_tile_x, int pixel_tile_x) {
  if (int r = check_handle(h)) return r;
  if (cost_tile_x != 0 && !valid_cost_tile(cost_tile_x))
    return fail(PM_E_INVALID, "cost tile width must be 0 (default), 4, 8 or 16");
  if (pixel_tile_x != 0 && !valid_pixel_tile(pixel_tile_x))
    return fail(PM_E_INVALID, "pixel tile width must be 0 (default), 16 or 32");
  ++h->gen;
  h->cost_tile_x = cost_tile_x;
  h->pixel_tile_x = pixel_tile_x;
  return PM_OK;
}

int pm_set_fallback_occupancy(pm_handle* h, int waves_per_simd) {
  if (int r = check_handle(h)) return r;
  if (waves_per_simd < -1 || waves_per_simd > 8) return fail(PM_E_INVALID, "fallback occupancy: -1, 0 or 1..8");
  ++h->gen;
  h->fallback_waves = waves_per_simd;
  return PM_OK;
}

int pm_get_tiles(pm_handle* h, int* cost_tile_x, int* pixel_tile_x) {
  if (int r = check_handle(h)) return r;
  if (cost_tile_x) *cost_tile_x = cost_tile(h);
  if (pixel_tile_x) *pixel_tile_x = pixel_tile(h);
  return PM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched runs: the references of a pm_batch go through one launch per
// kernel of the schedule (init, 6 kernels per colour pass, decode), every
// launch covering the workgroups of all references, and the whole schedule
// is captured once into a HIP graph.  DepthmapClusterEstimator::Run
// (cluster_estimator.cc:82-156) runs the references one after the other;
// the results here are bit-identical to running each reference alone
// (same kernels, same per-reference arguments and workgroup mapping).
// ---------------------------------------------------------------------------
struct pm_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  std::vector<pm_handle*> refs;
  bool use_graph = true;
  int level = 0;  // pyramid level every reference runs (pm_batch_set_level)
  // Requested tile widths: 0 = autotune.
  int want_cost_x = 0, want_pixel_x = 0;
  // Configuration in use.
  int cost_x = 4, pixel_x = 32;
  int fallback_waves = 0;  // the first reference's fallback occupancy cap
  float tune_ms = 0.0f, last_ms = 0.0f;
  std::vector<float> tune_log;  // per trial: cost_x, pixel_x, 6 stage times (ms)
  // Prepared state.
  bool prepared = false;
  std::vector<uint64_t> gens;
  int hp = 0, max_src = 0, iters = 0, cus = 0, nslices = 0;
  bool f16 = false;
  DevBuf<PmArgs> args;  // [slice][reference]
  DevBuf<int> bases;    // block offsets per (grid kind, tile width), nrefs + 1 each
  std::vector<int> host_bases;
  std::vector<std::pair<int, int>> base_keys;  // (grid kind, tile width) of each base array
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // Per-stage device time of the last eager run (no graph): events around
  // every launch, as for a single handle.
  bool timing = false;
  std::vector<hipEvent_t> kev;
  std::vector<std::pair<int, int>> kev_marks;  // (stage, index of the start event)
  size_t kev_used = 0;
  float stage_ms[PM_STAGES] = {};
  int stage_launches[PM_STAGES] = {};

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~pm_batch() {
    drop_graph();
    args.release();
    bases.release();
    for (auto& e : ev) if (e) (void)hipEventDestroy(e);
    for (auto& e : kev) if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// Stage of a kernel for the per-stage timing (-1: not attributed).
int stage_of(int kernel) {
  switch (kernel) {
    case PM_K_INIT_VIEWS: case PM_K_INIT: return PM_STAGE_INIT;
    case PM_K_CANDIDATES: return PM_STAGE_CANDIDATES;
    case PM_K_FALLBACK: return PM_STAGE_FALLBACK;
    case PM_K_SELECT: return PM_STAGE_SELECT;
    case PM_K_REFINE_COST: return PM_STAGE_REFINE_COST;
    case PM_K_REFINE: return PM_STAGE_REFINE;
    default: return -1;
  }
}

// The grid kind of a kernel: kernels of one kind share a workgroup layout.
int grid_kind(int kernel) {
  switch (kernel) {
    case PM_K_REFPATCH: case PM_K_INIT: case PM_K_PRIOR: return PM_K_INIT;
    case PM_K_CANDIDATES: case PM_K_REFINE_COST: return PM_K_CANDIDATES;
    case PM_K_SELECT: case PM_K_REFINE: case PM_K_FILTER: return PM_K_SELECT;
    default: return kernel;  // PM_K_FALLBACK, PM_K_DECODE
  }
}

// Argument slice of (iteration, colour, fallback phase); slice 0 holds the
// init / decode arguments.
int batch_slice(int it, int color, int phase) { return 1 + ((it * 2 + color) * 2 + phase); }

// Launch view of the first `nrefs` references for `kernel` with tile width tile_x.
PmBatch batch_view(const pm_batch* b, int kernel, int tile_x, int slice, int nrefs) {
  const int kind = grid_kind(kernel);
  const int tile = (kind == PM_K_INIT || kind == PM_K_INIT_VIEWS || kind == PM_K_DECODE) ? 0 : tile_x;
  const int n = static_cast<int>(b->refs.size());
  for (size_t i = 0; i < b->base_keys.size(); ++i) {
    if (b->base_keys[i].first == kind && b->base_keys[i].second == tile) {
      const int off = static_cast<int>(i) * (n + 1);
      return PmBatch{b->args.p + static_cast<size_t>(slice) * n, b->bases.p + off, nrefs,
                     b->host_bases[static_cast<size_t>(off + nrefs)]};
    }
  }
  return PmBatch{nullptr, nullptr, 0, 0};
}

int batch_launch(pm_batch* b, int kernel, int slice, int nrefs, int tile_x) {
  const PmBatch v = batch_view(b, kernel, tile_x, slice, nrefs);
  if (!v.args) return fail(PM_E_INVALID, "batch: no launch layout for this tile width");
  const unsigned grid = static_cast<unsigned>(v.total);
  const int stage = b->timing ? stage_of(kernel) : -1;
  if (stage >= 0) {
    while (b->kev.size() < b->kev_used + 2) {
      hipEvent_t e;
      PM_HIP(hipEventCreate(&e), "hipEventCreate");
      b->kev.push_back(e);
    }
    PM_HIP(hipEventRecord(b->kev[b->kev_used], b->stream), "event");
  }
  PM_HIP(launch_batch(kernel, v, grid, b->hp, b->f16, b->max_src, tile_x, b->fallback_waves, b->stream),
         "batched kernel launch");
  if (stage >= 0) {
    PM_HIP(hipEventRecord(b->kev[b->kev_used + 1], b->stream), "event");
    b->kev_marks.emplace_back(stage, static_cast<int>(b->kev_used));
    b->kev_used += 2;
  }
  return PM_OK;
}

// Random init of the first nrefs references: per-view scoring, then top-k.
int batch_init(pm_batch* b, int nrefs) {
  if (int r = batch_launch(b, PM_K_INIT_VIEWS, 0, nrefs, 0)) return r;
  return batch_launch(b, PM_K_INIT, 0, nrefs, 0);
}

// One colour pass of iteration `it` over the first nrefs references;
// stage_ev (optional): 7 events bracketing the 6 kernels.
int batch_pass(pm_batch* b, int it, int c, int nrefs, int cx, int px, hipEvent_t* stage_ev) {
  const int s0 = batch_slice(it, c, 0), s1 = batch_slice(it, c, 1);
  const int kernels[6] = {PM_K_CANDIDATES, PM_K_FALLBACK, PM_K_SELECT, PM_K_REFINE_COST, PM_K_FALLBACK, PM_K_REFINE};
  const int slices[6] = {s0, s0, s0, s1, s1, s1};
  for (int k = 0; k < 6; ++k) {
    if (stage_ev) PM_HIP(hipEventRecord(stage_ev[k], b->stream), "event");
    const int tx = (kernels[k] == PM_K_SELECT || kernels[k] == PM_K_REFINE) ? px : cx;
    if (int r = batch_launch(b, kernels[k], slices[k], nrefs, tx)) return r;
  }
  if (stage_ev) PM_HIP(hipEventRecord(stage_ev[6], b->stream), "event");
  return PM_OK;
}

// Prior re-init of the references that have a prior (PriorReinit,
// estimator.cc:574-614), one launch each.
int batch_priors(pm_batch* b) {
  for (pm_handle* h : b->refs) {
    if (!h->masks.p) continue;
    PM_HIP(hipMemcpyAsync(h->snapshot.p, h->planes.p, h->planes.n * sizeof(float4), hipMemcpyDeviceToDevice,
                          b->stream), "prior snapshot");
    PM_HIP(launch_prior(make_args(h), b->hp, b->f16, b->stream), "prior kernel launch");
  }
  return PM_OK;
}

// The whole schedule of one run (Run, estimator.cc:134-255, one level).
int batch_schedule(pm_batch* b) {
  const int n = static_cast<int>(b->refs.size());
  if (b->level == 0) {
    if (int r = batch_init(b, n)) return r;
  } else {
    // finer levels start from each reference's coarser result
    // (UpsampleDepthNormal, estimator.cc:761-821): one small launch each
    for (pm_handle* h : b->refs) {
      const bool mark = b->timing;
      if (mark) {
        while (b->kev.size() < b->kev_used + 2) {
          hipEvent_t e;
          PM_HIP(hipEventCreate(&e), "hipEventCreate");
          b->kev.push_back(e);
        }
        PM_HIP(hipEventRecord(b->kev[b->kev_used], b->stream), "event");
      }
      PM_HIP(launch_upsample(make_args(h), h->coarse_depth.p, h->coarse_normal.p, h->coarse_w, h->coarse_h,
                             b->stream), "upsample kernel launch");
      if (mark) {
        PM_HIP(hipEventRecord(b->kev[b->kev_used + 1], b->stream), "event");
        b->kev_marks.emplace_back(PM_STAGE_INIT, static_cast<int>(b->kev_used));
        b->kev_used += 2;
      }
    }
  }
  if (int r = batch_priors(b)) return r;
  for (int it = 0; it < b->iters; ++it)
    for (int c = 0; c < 2; ++c)
      if (int r = batch_pass(b, it, c, n, b->cost_x, b->pixel_x, nullptr)) return r;
  if (b->refs[0]->params.checkerboard_filter)
    for (int c = 0; c < 2; ++c)
      if (int r = batch_launch(b, PM_K_FILTER, batch_slice(0, c, 0), n, b->pixel_x)) return r;
  return batch_launch(b, PM_K_DECODE, 0, n, 0);
}

struct TuneResult {
  int cost_x = 4, pixel_x = 32;
};
std::vector<std::pair<std::vector<int>, TuneResult>> g_tune_cache;  // key: device, half patch, f16
std::mutex g_tune_mutex;

// Autotuner: on a prefix of the batch (about 8 Mpx of references), time
// one colour pass per candidate configuration, each from a fresh random
// init, with events around every kernel; keep the cost-tile width with the
// least cost-kernel time (candidates, refine_cost and both fallbacks) and
// the pixel-tile width with the least select + refine time.  All
// configurations give bit-identical results, so only time is compared.
int batch_tune(pm_batch* b) {
  const int n = static_cast<int>(b->refs.size());
  const std::vector<int> key = {b->device, b->hp, b->f16 ? 1 : 0};
  TuneResult best;
  bool cached = false;
  {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    for (const auto& e : g_tune_cache)
      if (e.first == key) { best = e.second; cached = true; }
  }
  const bool need = (b->want_cost_x == 0 || b->want_pixel_x == 0) && !cached;
  b->tune_ms = 0.0f;
  if (need) {
    int k = 0;
    int64_t px = 0;
    while (k < n && px < 8000000) { px += static_cast<int64_t>(b->refs[k]->W) * b->refs[k]->H; ++k; }
    hipEvent_t ev[9];
    for (auto& e : ev) PM_HIP(hipEventCreate(&e), "hipEventCreate");
    auto trial = [&](int cx, int pxw, float* st) -> int {
      if (int r = batch_init(b, k)) return r;
      if (int r = batch_pass(b, 0, 0, k, cx, pxw, ev)) return r;
      PM_HIP(hipEventSynchronize(ev[6]), "tune sync");
      for (int j = 0; j < 6; ++j) PM_HIP(hipEventElapsedTime(&st[j], ev[j], ev[j + 1]), "event time");
      b->tune_log.insert(b->tune_log.end(), {static_cast<float>(cx), static_cast<float>(pxw), st[0], st[1], st[2],
                                             st[3], st[4], st[5]});
      return PM_OK;
    };
    PM_HIP(hipEventRecord(ev[7], b->stream), "event");
    b->tune_log.clear();
    float st[6];
    int rc = trial(kCostTileWidths[0], kPixelTileWidths[1], st);  // warm-up
    float best_cost = 1e30f, best_pix = 1e30f;
    for (int cx : kCostTileWidths) {
      if (rc) break;
      rc = trial(cx, kPixelTileWidths[1], st);
      const float c = st[0] + st[1] + st[3] + st[4];
      if (c < best_cost) { best_cost = c; best.cost_x = cx; }
    }
    for (int pw : kPixelTileWidths) {
      if (rc) break;
      rc = trial(best.cost_x, pw, st);
      const float p = st[2] + st[5];
      if (p < best_pix) { best_pix = p; best.pixel_x = pw; }
    }
    if (!rc) {
      PM_HIP(hipEventRecord(ev[8], b->stream), "event");
      PM_HIP(hipEventSynchronize(ev[8]), "tune sync");
      PM_HIP(hipEventElapsedTime(&b->tune_ms, ev[7], ev[8]), "event time");
      std::lock_guard<std::mutex> lock(g_tune_mutex);
      g_tune_cache.emplace_back(key, best);
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (rc) return rc;
  }
  b->cost_x = b->want_cost_x ? b->want_cost_x : best.cost_x;
  b->pixel_x = b->want_pixel_x ? b->want_pixel_x : best.pixel_x;
  return PM_OK;
}

int batch_validate(pm_batch* b) {
  if (b->refs.empty()) return fail(PM_E_INVALID, "batch: no references added");
  const pm_params& p0 = b->refs[0]->params;
  if (b->level >= level_count(p0)) return fail(PM_E_INVALID, "batch: level out of range for the references' pyramid");
  for (pm_handle* h : b->refs) {
    if (h->device != b->device) return fail(PM_E_INVALID, "batch: every reference must be on the batch's device");
    if (int r = validate(h)) return r;
    // A finer level starts from the level just below it, once: the coarse
    // buffers then hold this level's result, not the one it started from.
    if (b->level > 0 && h->next_level != b->level)
      return fail(PM_E_STATE, "batch: levels run in order, each once (level " + std::to_string(b->level) +
                                  " needs level " + std::to_string(b->level - 1) + " just before it)");
    const pm_params& p = h->params;
    if (level_count(p) != level_count(p0))
      return fail(PM_E_INVALID, "batch: every reference must have the same number of pyramid levels");
    if (p.patch_size != p0.patch_size || p.image_storage != p0.image_storage ||
        p.max_iterations != p0.max_iterations || p.checkerboard_filter != p0.checkerboard_filter)
      return fail(PM_E_INVALID,
                  "batch: patch size, image storage, iterations and checkerboard filter must be the same for all "
                  "references");
  }
  return PM_OK;
}

// Stage every reference, build the argument slices and block offsets, tune,
// capture the schedule.
int batch_prepare(pm_batch* b) {
  b->prepared = false;
  b->drop_graph();
  if (int r = batch_validate(b)) return r;
  PM_HIP(hipSetDevice(b->device), "hipSetDevice");
  const int n = static_cast<int>(b->refs.size());
  {
    // Stage the references on host threads (image conversion and padding
    // are host work; each handle uploads on its own stream).
    std::vector<int> rcs(static_cast<size_t>(n), PM_OK);
    std::vector<std::string> msgs(static_cast<size_t>(n));
    std::atomic<int> next{0};
    auto work = [&]() {
      if (hipSetDevice(b->device) != hipSuccess) return;
      for (int i = next++; i < n; i = next++) {
        pm_handle* h = b->refs[static_cast<size_t>(i)];
        int r = PM_OK;
        const int nl = level_count(h->params);
        if (nl > 1) {
          r = upload_level(h, b->level, nl);
          if (r == PM_OK && b->level > 0 && (h->coarse_w <= 0 || !h->coarse_depth.p))
            r = fail(PM_E_STATE, "batch: no coarser level to start from (run the levels in order)");
        } else if (!h->uploaded || h->W != h->views[0].w || h->H != h->views[0].h) {
          r = upload_level(h, 0, 1);
        }
        if (r == PM_OK) r = ensure_refpatch(h);
        if (r == PM_OK && hipStreamSynchronize(h->stream) != hipSuccess) r = fail(PM_E_DEVICE, "batch staging");
        rcs[static_cast<size_t>(i)] = r;
        if (r != PM_OK) msgs[static_cast<size_t>(i)] = g_last_error;
      }
    };
    const int nt = std::max(1, std::min({n, 16, static_cast<int>(std::thread::hardware_concurrency())}));
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    for (int i = 0; i < n; ++i)
      if (rcs[static_cast<size_t>(i)] != PM_OK) return fail(rcs[static_cast<size_t>(i)], msgs[static_cast<size_t>(i)]);
  }
  const pm_params& p0 = b->refs[0]->params;
  b->hp = p0.patch_size / 2;
  b->f16 = p0.image_storage == PM_STORE_F16;
  b->iters = p0.max_iterations;
  b->fallback_waves = fallback_cap(b->refs[0]);
  b->max_src = 0;
  for (pm_handle* h : b->refs) b->max_src = std::max(b->max_src, h->n_src);
  hipDeviceProp_t prop;
  PM_HIP(hipGetDeviceProperties(&prop, b->device), "hipGetDeviceProperties");
  b->cus = prop.multiProcessorCount;
  // Argument slices.
  const int iter_slices = std::max(b->iters, 1);
  b->nslices = 1 + iter_slices * 4;
  std::vector<PmArgs> tab(static_cast<size_t>(b->nslices) * n);
  for (int r = 0; r < n; ++r) {
    pm_handle* h = b->refs[r];
    PmArgs a = make_args(h);
    a.out = PmDecodeOut{h->out_depth.p, h->out_normal.p, h->out_cost.p, h->out_conf.p};
    tab[static_cast<size_t>(r)] = a;
    for (int it = 0; it < iter_slices; ++it)
      for (int c = 0; c < 2; ++c)
        for (int ph = 0; ph < 2; ++ph) {
          PmArgs s = a;
          s.iteration = it;
          s.smooth_w = smooth_weight_at(h->params, it);
          s.color = c;
          s.phase = ph;
          // a batch run starts from new planes: at level 0 the init leaves
          // the random planes' costs in slot 8, unless a prior re-seeds
          // planes; after an upsample iteration 0 scores the current planes
          s.cur_cached = (it > 0 || (b->level == 0 && !h->masks.p)) ? 1 : 0;
          tab[static_cast<size_t>(batch_slice(it, c, ph)) * n + r] = s;
        }
  }
  // Block offsets per grid kind and tile width.
  b->base_keys.clear();
  b->base_keys.emplace_back(PM_K_INIT, 0);
  b->base_keys.emplace_back(PM_K_INIT_VIEWS, 0);
  b->base_keys.emplace_back(PM_K_DECODE, 0);
  for (int w : kCostTileWidths) {
    b->base_keys.emplace_back(PM_K_CANDIDATES, w);
    b->base_keys.emplace_back(PM_K_FALLBACK, w);
  }
  for (int w : kPixelTileWidths) b->base_keys.emplace_back(PM_K_SELECT, w);
  b->host_bases.assign(b->base_keys.size() * (n + 1), 0);
  for (size_t k = 0; k < b->base_keys.size(); ++k) {
    int64_t acc = 0;
    int* row = &b->host_bases[k * (n + 1)];
    for (int r = 0; r < n; ++r) {
      row[r] = static_cast<int>(acc);
      const pm_handle* h = b->refs[r];
      acc += kernel_blocks(b->base_keys[k].first, h->W, h->H, h->n_src, b->base_keys[k].second);
      if (acc > 0x7fffffff) return fail(PM_E_INVALID, "batch: too many workgroups for one launch");
    }
    row[n] = static_cast<int>(acc);
  }
  if (int r = b->args.alloc(tab.size())) return r;
  if (int r = b->bases.alloc(b->host_bases.size())) return r;
  PM_HIP(hipMemcpyAsync(b->args.p, tab.data(), tab.size() * sizeof(PmArgs), hipMemcpyHostToDevice, b->stream),
         "upload batch arguments");
  PM_HIP(hipMemcpyAsync(b->bases.p, b->host_bases.data(), b->host_bases.size() * sizeof(int),
                        hipMemcpyHostToDevice, b->stream), "upload batch offsets");
  PM_HIP(hipStreamSynchronize(b->stream), "batch upload");
  if (int r = batch_tune(b)) return r;
  if (b->use_graph) {
    PM_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "begin graph capture");
    const int rc = batch_schedule(b);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    PM_HIP(e, "end graph capture");
    b->graph = g;
    PM_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "graph instantiate");
  }
  b->gens.clear();
  for (pm_handle* h : b->refs) b->gens.push_back(h->gen);
  b->prepared = true;
  return PM_OK;
}

bool batch_stale(const pm_batch* b) {
  if (!b->prepared || b->gens.size() != b->refs.size()) return true;
  for (size_t i = 0; i < b->refs.size(); ++i)
    if (!b->refs[i]->uploaded || b->refs[i]->gen != b->gens[i]) return true;
  return false;
}

int check_batch(pm_batch* b) { return b ? PM_OK : fail(PM_E_INVALID, "null pm_batch"); }

}  // namespace

extern "C" {

int pm_batch_create(int device, pm_batch** out) {
  if (!out) return fail(PM_E_INVALID, "null output batch");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(PM_E_DEVICE, "no HIP device is available");
  if (device < 0 || device >= n) return fail(PM_E_INVALID, "device index out of range");
  PM_HIP(hipSetDevice(device), "hipSetDevice");
  pm_batch* b = new pm_batch();
  b->device = device;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return fail(PM_E_DEVICE, "hipStreamCreate failed");
  }
  for (auto& e : b->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      delete b;
      return fail(PM_E_DEVICE, "hipEventCreate failed");
    }
  }
  *out = b;
  return PM_OK;
}

int pm_batch_destroy(pm_batch* b) {
  if (!b) return PM_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  delete b;
  return PM_OK;
}

int pm_batch_add(pm_batch* b, pm_handle* h) {
  if (int r = check_batch(b)) return r;
  if (int r = check_handle(h)) return r;
  for (pm_handle* o : b->refs)
    if (o == h) return fail(PM_E_INVALID, "batch: reference added twice");
  b->refs.push_back(h);
  b->prepared = false;
  return PM_OK;
}

int pm_batch_clear(pm_batch* b) {
  if (int r = check_batch(b)) return r;
  b->refs.clear();
  b->prepared = false;
  b->drop_graph();
  return PM_OK;
}

int pm_batch_size(pm_batch* b, int* n) {
  if (int r = check_batch(b)) return r;
  if (!n) return fail(PM_E_INVALID, "null output");
  *n = static_cast<int>(b->refs.size());
  return PM_OK;
}

int pm_batch_set_config(pm_batch* b, int cost_tile_x, int pixel_tile_x, int use_graph) {
  if (int r = check_batch(b)) return r;
  if (cost_tile_x != 0 && !valid_cost_tile(cost_tile_x)) return fail(PM_E_INVALID, "cost tile width: 0, 4, 8 or 16");
  if (pixel_tile_x != 0 && !valid_pixel_tile(pixel_tile_x)) return fail(PM_E_INVALID, "pixel tile width: 0, 16 or 32");
  b->want_cost_x = cost_tile_x;
  b->want_pixel_x = pixel_tile_x;
  b->use_graph = use_graph != 0;
  b->prepared = false;
  return PM_OK;
}

int pm_batch_set_level(pm_batch* b, int level) {
  if (int r = check_batch(b)) return r;
  if (level < 0) return fail(PM_E_INVALID, "batch: negative level");
  b->level = level;
  b->prepared = false;
  return PM_OK;
}

int pm_batch_get_config(pm_batch* b, int* cost_tile_x, int* pixel_tile_x, float* tune_ms) {
  if (int r = check_batch(b)) return r;
  if (cost_tile_x) *cost_tile_x = b->cost_x;
  if (pixel_tile_x) *pixel_tile_x = b->pixel_x;
  if (tune_ms) *tune_ms = b->tune_ms;
  return PM_OK;
}

int pm_batch_tune_log(pm_batch* b, float* out, int capacity, int* count) {
  if (int r = check_batch(b)) return r;
  const int n = static_cast<int>(b->tune_log.size());
  if (count) *count = n;
  if (out) std::memcpy(out, b->tune_log.data(), sizeof(float) * static_cast<size_t>(std::min(n, capacity)));
  return PM_OK;
}

void pm_batch_reset_tuning(void) {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  g_tune_cache.clear();
}

int pm_batch_prepare(pm_batch* b) {
  if (int r = check_batch(b)) return r;
  return batch_prepare(b);
}

int pm_batch_run(pm_batch* b) {
  if (int r = check_batch(b)) return r;
  if (b->level > 0)  // before a graph replay too: its upsample launches read the coarse buffers
    for (pm_handle* h : b->refs)
      if (h->next_level != b->level)
        return fail(PM_E_STATE, "batch: levels run in order, each once (level " + std::to_string(b->level) +
                                    " needs level " + std::to_string(b->level - 1) + " just before it)");
  if (batch_stale(b))
    if (int r = batch_prepare(b)) return r;
  PM_HIP(hipSetDevice(b->device), "hipSetDevice");
  PM_HIP(hipEventRecord(b->ev[0], b->stream), "event");
  for (int k = 0; k < PM_STAGES; ++k) { b->stage_ms[k] = 0.0f; b->stage_launches[k] = 0; }
  if (b->use_graph) {
    PM_HIP(hipGraphLaunch(b->exec, b->stream), "graph launch");
  } else {
    b->timing = true;
    b->kev_used = 0;
    b->kev_marks.clear();
    const int rc = batch_schedule(b);
    b->timing = false;
    if (rc) return rc;
  }
  PM_HIP(hipEventRecord(b->ev[1], b->stream), "event");
  PM_HIP(hipStreamSynchronize(b->stream), "batch run");
  PM_HIP(hipEventElapsedTime(&b->last_ms, b->ev[0], b->ev[1]), "event time");
  for (const auto& m : b->kev_marks) {
    float ms = 0.0f;
    PM_HIP(hipEventElapsedTime(&ms, b->kev[m.second], b->kev[m.second + 1]), "event time");
    b->stage_ms[m.first] += ms;
    b->stage_launches[m.first] += 1;
  }
  b->kev_marks.clear();
  for (pm_handle* h : b->refs) {
    h->state_ready = true;
    // the last passes left slot 8 current for both colours, unless the
    // filter changed planes after them
    current_stale(h);
    if (b->iters > 0 && !h->params.checkerboard_filter) {
      current_filled(h, 0);
      current_filled(h, 1);
    }
    if (int r = postprocess(h)) return r;
    h->next_level = b->level + 1;
    if (b->level < level_count(h->params) - 1) {  // keep the decoded level for the next one
      const size_t npix = static_cast<size_t>(h->W) * h->H;
      ++h->gen;  // captured upsample launches may point at the buffers replaced here
      if (int r = h->coarse_depth.alloc(npix)) return r;
      if (int r = h->coarse_normal.alloc(3 * npix)) return r;
      PM_HIP(hipMemcpyAsync(h->coarse_depth.p, h->out_depth.p, npix * sizeof(float), hipMemcpyDeviceToDevice,
                            h->stream), "keep coarse depth");
      PM_HIP(hipMemcpyAsync(h->coarse_normal.p, h->out_normal.p, 3 * npix * sizeof(float),
                            hipMemcpyDeviceToDevice, h->stream), "keep coarse normal");
      h->coarse_w = h->W;
      h->coarse_h = h->H;
      PM_HIP(hipStreamSynchronize(h->stream), "keep coarse level");
    }
  }
  return PM_OK;
}

int pm_batch_timing(pm_batch* b, float* run_ms) {
  if (int r = check_batch(b)) return r;
  if (run_ms) *run_ms = b->last_ms;
  return PM_OK;
}

int pm_batch_stage_timing(pm_batch* b, int stage, float* ms, int* launches) {
  if (int r = check_batch(b)) return r;
  if (stage < 0 || stage >= PM_STAGES) return fail(PM_E_INVALID, "unknown stage");
  if (ms) *ms = b->stage_ms[stage];
  if (launches) *launches = b->stage_launches[stage];
  return PM_OK;
}

}  // extern "C"

extern "C" {

int pm_set_buffer_cache(int device, uint64_t max_bytes) {
  int n = 0;
  PM_HIP(hipGetDeviceCount(&n), "hipGetDeviceCount");
  if (device < 0 || device >= n || device >= kMaxCacheDevices) return fail(PM_E_INVALID, "no such device");
  int prev = 0;
  PM_HIP(hipGetDevice(&prev), "hipGetDevice");
  PM_HIP(hipSetDevice(device), "hipSetDevice");
  BlockCache& c = block_cache(device);
  {
    std::lock_guard<std::mutex> lock(c.mu);
    c.limit = max_bytes;
  }
  trim_cache(c, max_bytes);
  PM_HIP(hipSetDevice(prev), "hipSetDevice");
  return PM_OK;
}

int pm_buffer_cache_bytes(int device, uint64_t* held) {
  if (!held) return fail(PM_E_INVALID, "null output");
  if (device < 0 || device >= kMaxCacheDevices) return fail(PM_E_INVALID, "no such device");
  BlockCache& c = block_cache(device);
  std::lock_guard<std::mutex> lock(c.mu);
  *held = c.held;
  return PM_OK;
}

int pm_device_bytes(pm_handle* h, uint64_t* bytes) {
  if (int r = check_handle(h)) return r;
  if (!bytes) return fail(PM_E_INVALID, "null output");
  if (int r = validate(h)) return r;
  const pm_params& p = h->params;
  const uint64_t esz = p.image_storage == PM_STORE_F16 ? 2 : 4;
  uint64_t total = 0;
  for (const HostView& v : h->views) total += static_cast<uint64_t>(v.w + 2) * (v.h + 2) * esz * 5 + 64 * esz;
  const uint64_t w = h->views[0].w, hh = h->views[0].h, npix = w * hh;
  const uint64_t rp = 2 * hh * ((w + 1) / 2);
  const uint64_t ns = h->views.size() - 1;
  const uint64_t P = static_cast<uint64_t>(p.patch_size / 2 * 2 + 1) * (p.patch_size / 2 * 2 + 1);
  total += rp * (P * 8 + 3 * 16 + 9 * ns * 4 + ns * 4 + 4 + 8 * 4 + 4 + 6 * 16);
  total += static_cast<uint64_t>(fallback_mask_count(static_cast<int>(w), static_cast<int>(hh), static_cast<int>(ns))) * 8;
  total += npix * (16 + 4 + 8 + 4 + 4 + 12 + 4 + 4);
  if (!h->prior_masks.empty() || !h->sfm_points.empty()) total += npix * (16 + 4 + 16);
  for (const auto& e : h->prev_depths)
    if (!e.d.empty()) { total += ns * npix * 4; break; }
  *bytes = total;
  return PM_OK;
}

}  // extern "C"
