// pybind11 module `pydense`: the DepthmapEstimator surface of the reference
// type stub (opensfm/src/lib/dense/pydense.pyi:42-59) plus the parameter
// setters of DepthmapClusterEstimator (opensfm/src/lib/dense/python/
// pybind.cc:64-104), on top of the C ABI in include/opensfm_pm.h.
// Errors raise RuntimeError with the library's message; the GIL is released
// around device work (as depthmap_bind.h:61-64 does).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <initializer_list>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "lz4_frame.h"
#include "opensfm_pm.h"

namespace py = pybind11;

namespace {

std::vector<py::ssize_t> shape(std::initializer_list<py::ssize_t> s) { return std::vector<py::ssize_t>(s); }

void check(int rc) {
  if (rc != PM_OK) throw std::runtime_error(pm_last_error());
}

std::vector<double> as_f64(py::handle obj, size_t n, const char* what) {
  auto a = py::array_t<double, py::array::c_style | py::array::forcecast>::ensure(obj);
  if (!a || static_cast<size_t>(a.size()) != n)
    throw std::invalid_argument(std::string(what) + ": expected " + std::to_string(n) + " values");
  return std::vector<double>(a.data(), a.data() + n);
}

// Adds one view to a handle: u8, f16 or (anything else, cast to) f32 (H, W).
void add_view_to(pm_handle* h, py::handle K, py::handle R, py::handle t, py::array image, py::object mask) {
  auto k = as_f64(K, 9, "K");
  auto r = as_f64(R, 9, "R");
  auto tt = as_f64(t, 3, "t");
  if (image.ndim() != 2) throw std::invalid_argument("image must be a 2-D (H, W) array");
  int fmt;
  py::array img;
  if (image.dtype().is(py::dtype::of<uint8_t>())) {
    fmt = PM_U8;
    img = py::array_t<uint8_t, py::array::c_style>::ensure(image);
  } else if (image.dtype().kind() == 'f' && image.dtype().itemsize() == 2) {
    fmt = PM_F16;
    img = py::array::ensure(image, py::array::c_style);
  } else {
    fmt = PM_F32;
    img = py::array_t<float, py::array::c_style | py::array::forcecast>::ensure(image);
  }
  std::vector<uint8_t> m;
  if (!mask.is_none()) {
    auto ma = py::array_t<uint8_t, py::array::c_style | py::array::forcecast>::ensure(mask);
    m.assign(ma.data(), ma.data() + ma.size());
  }
  const void* data = img.data();
  const int w = static_cast<int>(img.shape(1)), hgt = static_cast<int>(img.shape(0));
  py::gil_scoped_release nogil;  // img keeps the buffer alive
  check(pm_add_view(h, k.data(), r.data(), tt.data(), data, fmt, w, hgt, 0, m.empty() ? nullptr : m.data()));
}

// (N, 3) float64 world points -> pm_set_sfm_points.
void set_points(pm_handle* h, py::handle points) {
  auto a = py::array_t<double, py::array::c_style | py::array::forcecast>::ensure(points);
  if (!a || (a.size() > 0 && (a.ndim() != 2 || a.shape(1) != 3)))
    throw std::invalid_argument("SfM points must be an (N, 3) array");
  check(pm_set_sfm_points(h, a.size() ? a.data() : nullptr, static_cast<int>(a.size() / 3)));
}

// pm_params <-> {field: value}, the field names of pm_params.
py::dict params_dict(const pm_params& p) {
  py::dict d;
  d["max_iterations"] = p.max_iterations;
  d["patch_size"] = p.patch_size;
  d["sigma_spatial"] = p.sigma_spatial;
  d["sigma_color"] = p.sigma_color;
  d["top_k"] = p.top_k;
  d["use_census"] = static_cast<bool>(p.use_census);
  d["depth_min"] = p.depth_min;
  d["depth_max"] = p.depth_max;
  d["hierarchy_levels"] = p.hierarchy_levels;
  d["smooth_weight"] = p.smooth_weight;
  d["anchor_views"] = p.anchor_views;
  d["checkerboard_filter"] = static_cast<bool>(p.checkerboard_filter);
  d["var_threshold"] = p.var_threshold;
  d["image_storage"] = p.image_storage == PM_STORE_F16 ? "f16" : "f32";
  d["quantize_half"] = static_cast<bool>(p.quantize_half);
  d["speckle_min_size"] = p.speckle_min_size;
  d["gap_max_size"] = p.gap_max_size;
  return d;
}

int storage_code(const std::string& s) {
  if (s == "f32") return PM_STORE_F32;
  if (s == "f16") return PM_STORE_F16;
  throw std::invalid_argument("image storage must be 'f32' or 'f16'");
}

void params_from_dict(pm_params& p, py::dict d) {
  for (auto item : d) {
    const std::string k = py::str(item.first);
    py::handle v = item.second;
    if (k == "max_iterations") p.max_iterations = v.cast<int>();
    else if (k == "patch_size") p.patch_size = v.cast<int>();
    else if (k == "sigma_spatial") p.sigma_spatial = v.cast<float>();
    else if (k == "sigma_color") p.sigma_color = v.cast<float>();
    else if (k == "top_k") p.top_k = v.cast<int>();
    else if (k == "use_census") p.use_census = v.cast<bool>() ? 1 : 0;
    else if (k == "depth_min") p.depth_min = v.cast<float>();
    else if (k == "depth_max") p.depth_max = v.cast<float>();
    else if (k == "hierarchy_levels") p.hierarchy_levels = v.cast<int>();
    else if (k == "smooth_weight") p.smooth_weight = v.cast<float>();
    else if (k == "anchor_views") p.anchor_views = v.cast<int>();
    else if (k == "checkerboard_filter") p.checkerboard_filter = v.cast<bool>() ? 1 : 0;
    else if (k == "var_threshold") p.var_threshold = v.cast<float>();
    else if (k == "image_storage") p.image_storage = storage_code(v.cast<std::string>());
    else if (k == "quantize_half") p.quantize_half = v.cast<bool>() ? 1 : 0;
    else if (k == "speckle_min_size") p.speckle_min_size = v.cast<int>();
    else if (k == "gap_max_size") p.gap_max_size = v.cast<int>();
    else throw std::invalid_argument("unknown parameter '" + k + "'");
  }
}

using F32Array = py::array_t<float, py::array::c_style | py::array::forcecast>;
using U32Array = py::array_t<uint32_t, py::array::c_style | py::array::forcecast>;

// A previous depth map: None, or a non-empty (H, W) array (returned as
// float32, C order); anything else raises ValueError.
F32Array depth_map_array(py::handle item) {
  if (item.is_none()) return F32Array();
  auto a = F32Array::ensure(item);
  if (!a || a.ndim() != 2 || a.shape(0) <= 0 || a.shape(1) <= 0)
    throw std::invalid_argument("previous depth map must be a non-empty (H, W) array");
  return a;
}

// Planar prior of a w x h reference: planes (h, w, 4) and masks (h, w);
// anything else raises ValueError.
std::pair<F32Array, U32Array> prior_arrays(py::handle planes, py::handle masks, py::ssize_t w, py::ssize_t h) {
  auto p = F32Array::ensure(planes);
  auto m = U32Array::ensure(masks);
  if (!p || p.ndim() != 3 || p.shape(0) != h || p.shape(1) != w || p.shape(2) != 4)
    throw std::invalid_argument("prior planes must be an (H, W, 4) array of the reference view's size (" +
                                std::to_string(h) + ", " + std::to_string(w) + ", 4)");
  if (!m || m.ndim() != 2 || m.shape(0) != h || m.shape(1) != w)
    throw std::invalid_argument("prior masks must be an (H, W) array of the reference view's size (" +
                                std::to_string(h) + ", " + std::to_string(w) + ")");
  return {p, m};
}

class Estimator {
 public:
  explicit Estimator(int device) { check(pm_create(device, &h_)); }
  ~Estimator() { pm_destroy(h_); }
  Estimator(const Estimator&) = delete;
  Estimator& operator=(const Estimator&) = delete;

  void add_view(py::handle K, py::handle R, py::handle t, py::array image, py::object mask) {
    add_view_to(h_, K, R, t, image, mask);
  }

  void add_view_device(py::handle K, py::handle R, py::handle t, uintptr_t data, const std::string& dtype, int width,
                       int height, int64_t row_stride) {
    auto k = as_f64(K, 9, "K");
    auto r = as_f64(R, 9, "R");
    auto tt = as_f64(t, 3, "t");
    const int fmt = dtype == "u8" ? PM_U8 : dtype == "f32" ? PM_F32 : dtype == "f16" ? PM_F16 : -1;
    if (fmt < 0) throw std::invalid_argument("dtype must be 'u8', 'f16' or 'f32'");
    check(pm_add_view_device(h_, k.data(), r.data(), tt.data(), reinterpret_cast<const void*>(data), fmt, width,
                             height, row_stride));
    if (ref_w_ < 0) {
      ref_w_ = width;
      ref_h_ = height;
    }
  }

  // num_planes: the depths of compute_brute_force (PatchMatch does not use it).
  void set_depth_range(float dmin, float dmax, int num_planes) {
    check(pm_set_depth_range(h_, dmin, dmax));
    num_planes_ = num_planes;
  }

  template <typename F>
  void update(F f) {
    pm_params p;
    check(pm_get_params(h_, &p));
    f(p);
    check(pm_set_params(h_, &p));
  }

  py::dict params() {
    pm_params p;
    check(pm_get_params(h_, &p));
    return params_dict(p);
  }

  // Any subset of the params() fields, by name.
  void set_params(py::dict d) {
    update([&](pm_params& p) { params_from_dict(p, d); });
  }

  py::list compute_patch_match() { return as_list(compute()); }

  // PatchMatch already samples each pixel's views (the Monte-Carlo view
  // draws of acmmp_patchmatch, opencl_kernels.h:910-969), so the sampled
  // variant of the stub (pydense.pyi:55) is the same computation.
  py::list compute_patch_match_sample() { return as_list(compute()); }

  // Fronto-parallel plane sweep over the set_depth_range num_planes depths
  // (pydense.pyi:53).
  py::list compute_brute_force() {
    {
      py::gil_scoped_release nogil;
      check(pm_upload(h_));
      check(pm_plane_sweep(h_, num_planes_));
    }
    return as_list(download());
  }

  static py::list as_list(py::dict r) {
    py::list out;
    out.append(r["depth"]);
    out.append(r["normal"]);
    out.append(r["cost"]);
    out.append(r["views"]);
    return out;
  }

  py::dict compute() {
    pm_params p;
    check(pm_get_params(h_, &p));
    {
      py::gil_scoped_release nogil;
      if (p.hierarchy_levels <= 1) check(pm_upload(h_));  // several levels: pm_run stages each one
      check(pm_run(h_));
    }
    return download();
  }

  void upload() {
    py::gil_scoped_release nogil;
    check(pm_upload(h_));
  }
  void run() {
    py::gil_scoped_release nogil;
    check(pm_run(h_));
  }
  void init_state() {
    py::gil_scoped_release nogil;
    check(pm_init_state(h_));
  }
  void iterate(int iteration, int color) {
    py::gil_scoped_release nogil;
    check(pm_iterate(h_, iteration, color));
  }
  void filter() { check(pm_filter(h_)); }
  void decode() { check(pm_decode(h_)); }

  // Every decoded map, or only the ones named in ``maps``.
  py::dict download(py::object maps = py::none()) {
    static const char* kNames[] = {"depth", "plane", "normal", "cost", "views", "confidence"};
    bool want[6] = {true, true, true, true, true, true};
    if (!maps.is_none()) {
      for (bool& b : want) b = false;
      for (py::handle m : maps) {
        const std::string name = py::str(m);
        int k = 0;
        while (k < 6 && name != kNames[k]) ++k;
        if (k == 6) throw std::invalid_argument("download: unknown map '" + name + "'");
        want[k] = true;
      }
    }
    const auto [w, h] = staged_size();
    py::array_t<float> depth, plane, normal, cost, conf;
    py::array_t<uint32_t> views;
    if (want[0]) depth = py::array_t<float>(shape({h, w}));
    if (want[1]) plane = py::array_t<float>(shape({h, w, 4}));
    if (want[2]) normal = py::array_t<float>(shape({h, w, 3}));
    if (want[3]) cost = py::array_t<float>(shape({h, w}));
    if (want[4]) views = py::array_t<uint32_t>(shape({h, w}));
    if (want[5]) conf = py::array_t<float>(shape({h, w}));
    {
      py::gil_scoped_release nogil;
      check(pm_download(h_, want[0] ? depth.mutable_data() : nullptr, want[1] ? plane.mutable_data() : nullptr,
                        want[2] ? normal.mutable_data() : nullptr, want[3] ? cost.mutable_data() : nullptr,
                        want[4] ? views.mutable_data() : nullptr, want[5] ? conf.mutable_data() : nullptr));
    }
    py::dict d;
    if (want[0]) d["depth"] = depth;
    if (want[1]) d["plane"] = plane;
    if (want[2]) d["normal"] = normal;
    if (want[3]) d["cost"] = cost;
    if (want[4]) d["views"] = views;
    if (want[5]) d["confidence"] = conf;
    return d;
  }
  void release_device() { check(pm_release_device(h_)); }

  py::tuple read_state() {
    const auto [w, h] = staged_size();
    py::array_t<float> planes(shape({h, w, 4})), costs(shape({h, w}));
    py::array_t<uint32_t> rng(shape({h, w, 2})), sel(shape({h, w}));
    check(pm_read_state(h_, planes.mutable_data(), costs.mutable_data(), rng.mutable_data(), sel.mutable_data()));
    return py::make_tuple(planes, costs, rng, sel);
  }

  void write_state(py::array_t<float, py::array::c_style | py::array::forcecast> planes,
                   py::array_t<float, py::array::c_style | py::array::forcecast> costs,
                   py::array_t<uint32_t, py::array::c_style | py::array::forcecast> rng,
                   py::array_t<uint32_t, py::array::c_style | py::array::forcecast> sel) {
    const auto [w, h] = staged_size();
    const py::ssize_t n = static_cast<py::ssize_t>(w) * h;
    if (planes.size() != 4 * n || costs.size() != n || rng.size() != 2 * n || sel.size() != n)
      throw std::invalid_argument("write_state: arrays do not match the reference view size");
    check(pm_write_state(h_, planes.data(), costs.data(), rng.data(), sel.data()));
  }

  py::tuple score_planes(py::array_t<int32_t, py::array::c_style | py::array::forcecast> xy,
                         py::array_t<float, py::array::c_style | py::array::forcecast> planes) {
    if (xy.ndim() != 2 || xy.shape(1) != 2) throw std::invalid_argument("xy must be (n, 2)");
    const py::ssize_t n = xy.shape(0);
    if (planes.size() != 4 * n) throw std::invalid_argument("planes must be (n, 4)");
    int nv = 0;
    check(pm_num_views(h_, &nv));
    py::array_t<float> costs(shape({n, static_cast<py::ssize_t>(std::max(nv - 1, 0))}));
    py::array_t<float> init(shape({n}));
    py::array_t<uint32_t> sel(shape({n}));
    {
      py::gil_scoped_release nogil;
      check(pm_score_planes(h_, static_cast<int>(n), xy.data(), planes.data(), costs.mutable_data(),
                            init.mutable_data(), sel.mutable_data()));
    }
    return py::make_tuple(costs, init, sel);
  }

  // Photometric cost and d cost / d(depth, a, b) of per-pixel planes
  // (pm_ncc_cost_grad) on host arrays.
  py::tuple ncc_cost_grad(py::array_t<float, py::array::c_style | py::array::forcecast> depth,
                          py::array_t<float, py::array::c_style | py::array::forcecast> ab,
                          py::array_t<uint32_t, py::array::c_style | py::array::forcecast> views) {
    const auto [w, h] = staged_size();
    const py::ssize_t n = w * h;
    if (depth.size() != n || ab.size() != 2 * n || views.size() != n)
      throw std::invalid_argument("ncc_cost_grad: arrays do not match the reference view size");
    py::array_t<float> cost(shape({h, w})), grad(shape({h, w, 3}));
    {
      py::gil_scoped_release nogil;
      check(pm_ncc_cost_grad(h_, depth.data(), ab.data(), views.data(), cost.mutable_data(), grad.mutable_data(), 0));
    }
    return py::make_tuple(cost, grad);
  }

  // The same on device buffers given by address (torch tensors' data_ptr()).
  void ncc_cost_grad_device(uintptr_t depth, uintptr_t ab, uintptr_t views, uintptr_t cost, uintptr_t grad) {
    py::gil_scoped_release nogil;
    check(pm_ncc_cost_grad(h_, reinterpret_cast<const float*>(depth), reinterpret_cast<const float*>(ab),
                           reinterpret_cast<const uint32_t*>(views), reinterpret_cast<float*>(cost),
                           reinterpret_cast<float*>(grad), 1));
  }

  void gauss_newton(int steps, float damping) {
    py::gil_scoped_release nogil;
    check(pm_gauss_newton(h_, steps, damping));
  }

  float last_refine_ms() {
    float ms = 0.0f;
    check(pm_last_refine_timing(h_, &ms));
    return ms;
  }

  py::dict stage_timing() {
    static const char* names[PM_STAGES] = {"init", "candidates", "fallback", "select", "refine_cost", "refine"};
    py::dict d;
    for (int k = 0; k < PM_STAGES; ++k) {
      float ms = 0;
      int n = 0;
      check(pm_last_stage_timing(h_, k, &ms, &n));
      d[names[k]] = py::make_tuple(ms, n);
    }
    return d;
  }

  py::dict last_timing() {
    float run_ms = 0, sweep_ms = 0;
    int sweeps = 0;
    check(pm_last_timing(h_, &run_ms, &sweep_ms, &sweeps));
    py::dict d;
    d["run_ms"] = run_ms;
    d["sweep_ms"] = sweep_ms;
    d["sweep_launches"] = sweeps;
    return d;
  }

  // Each previous depth map is an (H, W) array of its own size; the library
  // resizes it to every pyramid level (SetPreviousDepths takes maps of any
  // size, estimator.cc:59-116).
  void set_geom_consistency(float weight, py::list prev) {
    if (py::len(prev) > PM_MAX_SOURCES)
      throw std::invalid_argument("set_geom_consistency: at most " + std::to_string(PM_MAX_SOURCES) +
                                  " previous depth maps");
    std::vector<F32Array> maps;
    std::vector<bool> given;
    for (auto item : prev) {  // validate every map before any change
      given.push_back(!item.is_none());
      maps.push_back(depth_map_array(item));
    }
    check(pm_set_geom_consistency(h_, weight, nullptr, 0));
    check(pm_clear_previous_depths(h_));
    for (size_t i = 0; i < maps.size(); ++i) {
      if (!given[i]) continue;
      check(pm_set_previous_depth(h_, static_cast<int>(i), maps[i].data(), static_cast<int>(maps[i].shape(1)),
                                  static_cast<int>(maps[i].shape(0))));
    }
  }

  // planes (H, W, 4) and masks (H, W) at the reference view's size.
  void set_prior(py::object planes, py::object masks) {
    if (planes.is_none() || masks.is_none()) {
      check(pm_set_prior(h_, nullptr, nullptr));
      return;
    }
    const auto [w, h] = ref_size();
    auto pm = prior_arrays(planes, masks, w, h);
    check(pm_set_prior(h_, pm.first.data(), pm.second.data()));
  }

 private:
  std::pair<py::ssize_t, py::ssize_t> ref_size() {
    if (ref_w_ < 0) throw std::runtime_error("DepthmapEstimator: no views added");
    return {ref_w_, ref_h_};
  }
  // Size of the pyramid level staged on the device (the decoded maps and the
  // per-pixel state have it; it is the reference size after a full run).
  std::pair<py::ssize_t, py::ssize_t> staged_size() {
    (void)ref_size();
    int w = 0, h = 0;
    check(pm_level_size(h_, &w, &h));
    return {w, h};
  }

 public:
  void note_size(py::array image) {
    if (ref_w_ < 0) {
      ref_w_ = image.shape(1);
      ref_h_ = image.shape(0);
    }
  }
  void clear() {
    check(pm_clear_views(h_));
    ref_w_ = ref_h_ = -1;
  }

  pm_handle* handle() const { return h_; }
  void set_tiles(int cost_x, int pixel_x) { check(pm_set_tiles(h_, cost_x, pixel_x)); }
  py::tuple tiles() {
    int c = 0, p = 0;
    check(pm_get_tiles(h_, &c, &p));
    return py::make_tuple(c, p);
  }
  uint64_t device_bytes() {
    uint64_t b = 0;
    check(pm_device_bytes(h_, &b));
    return b;
  }

 private:
  pm_handle* h_ = nullptr;
  py::ssize_t ref_w_ = -1, ref_h_ = -1;
  int num_planes_ = 0;
};

// Many reference estimators run through one set of launches (pm_batch_*).
// Results land in each estimator (download() / compute-style dicts).
class Batch {
 public:
  explicit Batch(int device) { check(pm_batch_create(device, &b_)); }
  ~Batch() { pm_batch_destroy(b_); }
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;

  void add(Estimator& e) { check(pm_batch_add(b_, e.handle())); }
  void clear() { check(pm_batch_clear(b_)); }
  int size() {
    int n = 0;
    check(pm_batch_size(b_, &n));
    return n;
  }
  void set_config(int cost_x, int pixel_x, bool use_graph) {
    check(pm_batch_set_config(b_, cost_x, pixel_x, use_graph ? 1 : 0));
  }
  void set_level(int level) { check(pm_batch_set_level(b_, level)); }
  py::dict config() {
    int c = 0, p = 0;
    float ms = 0.0f;
    check(pm_batch_get_config(b_, &c, &p, &ms));
    py::dict d;
    d["cost_tile_x"] = c;
    d["pixel_tile_x"] = p;
    d["tune_ms"] = ms;
    return d;
  }
  py::array_t<float> tune_log() {
    int n = 0;
    check(pm_batch_tune_log(b_, nullptr, 0, &n));
    py::array_t<float> out(shape({static_cast<py::ssize_t>(n / 8), 8}));
    check(pm_batch_tune_log(b_, out.mutable_data(), n, &n));
    return out;
  }
  void prepare() {
    py::gil_scoped_release nogil;
    check(pm_batch_prepare(b_));
  }
  float run() {
    {
      py::gil_scoped_release nogil;
      check(pm_batch_run(b_));
    }
    float ms = 0.0f;
    check(pm_batch_timing(b_, &ms));
    return ms;
  }
  py::dict stage_timing() {
    static const char* names[PM_STAGES] = {"init", "candidates", "fallback", "select", "refine_cost", "refine"};
    py::dict d;
    for (int k = 0; k < PM_STAGES; ++k) {
      float ms = 0;
      int n = 0;
      check(pm_batch_stage_timing(b_, k, &ms, &n));
      d[names[k]] = py::make_tuple(ms, n);
    }
    return d;
  }

 private:
  pm_batch* b_ = nullptr;
};

// GPUDepthmapCleaner (depthmap_bind.h:120-162, pybind.cc:125-143).
class Cleaner {
 public:
  Cleaner() { check(pm_cleaner_create(0, &c_)); }
  ~Cleaner() { pm_cleaner_destroy(c_); }
  Cleaner(const Cleaner&) = delete;
  Cleaner& operator=(const Cleaner&) = delete;

  void set_device(int d) {
    int n = 0;
    check(pm_cleaner_num_views(c_, &n));
    if (n > 0) throw std::runtime_error("GPUDepthmapCleaner.set_device: call before add_view");
    pm_cleaner* c = nullptr;
    check(pm_cleaner_create(d, &c));
    pm_cleaner_destroy(c_);
    c_ = c;
    push();
  }
  void set_same_depth_threshold(float v) { same_ = v; push(); }
  void set_min_consistent_views(int v) { min_views_ = v; push(); }
  void set_carving_threshold(float v) { carve_ = v; push(); }
  void set_max_carved_views(int v) { max_carved_ = v; push(); }
  void set_grazing_cos_threshold(float v) { grazing_ = v; push(); }

  int add_view(py::handle K, py::handle R, py::handle t, py::handle depth, py::object normal) {
    auto k = as_f64(K, 9, "K");
    auto r = as_f64(R, 9, "R");
    auto tt = as_f64(t, 3, "t");
    auto d = py::array_t<float, py::array::c_style | py::array::forcecast>::ensure(depth);
    if (!d || d.ndim() != 2) throw std::invalid_argument("depth must be a 2-D (H, W) array");
    const py::ssize_t h = d.shape(0), w = d.shape(1);
    std::vector<float> nrm;
    if (!normal.is_none()) {
      // (3, H*W) as PixelData3f (depthmap_bind.h:135-139), or (H, W, 3)
      py::array a = py::array::ensure(normal);
      if (a.ndim() == 2 && a.shape(0) == 3 && a.shape(1) == h * w) {
        auto f = py::array_t<float, py::array::f_style | py::array::forcecast>::ensure(a);
        nrm.assign(f.data(), f.data() + 3 * h * w);  // column-major (3, N) = interleaved
      } else if (a.ndim() == 3 && a.shape(0) == h && a.shape(1) == w && a.shape(2) == 3) {
        auto cc = py::array_t<float, py::array::c_style | py::array::forcecast>::ensure(a);
        nrm.assign(cc.data(), cc.data() + 3 * h * w);
      } else {
        throw std::invalid_argument("normal must be (3, H*W) or (H, W, 3)");
      }
    }
    int idx = -1;
    check(pm_cleaner_add_view(c_, k.data(), r.data(), tt.data(), d.data(), static_cast<int>(w), static_cast<int>(h),
                              nrm.empty() ? nullptr : nrm.data(), &idx));
    return idx;
  }

  py::array_t<float> clean(int ref_idx, py::handle neighbor_ids) {
    auto ids = py::array_t<int, py::array::c_style | py::array::forcecast>::ensure(neighbor_ids);
    if (!ids) throw std::invalid_argument("neighbor_ids must be an integer array");
    int n = 0;
    check(pm_cleaner_num_views(c_, &n));
    if (ref_idx < 0 || ref_idx >= n)
      throw std::out_of_range("Clean: ref_idx " + std::to_string(ref_idx) + " out of range [0, " + std::to_string(n) +
                              ")");
    int w = 0, h = 0;
    check(pm_cleaner_view_size(c_, ref_idx, &w, &h));
    py::array_t<float> out(shape({h, w}));
    std::vector<int> v(ids.data(), ids.data() + ids.size());
    float* o = out.mutable_data();
    {
      py::gil_scoped_release nogil;
      check(pm_cleaner_clean(c_, ref_idx, v.data(), static_cast<int>(v.size()), o));
    }
    return out;
  }
  float last_ms() {
    float ms = 0.0f;
    check(pm_cleaner_last_timing(c_, &ms));
    return ms;
  }
  void clear() { check(pm_cleaner_clear(c_)); }

 private:
  void push() { check(pm_cleaner_set_params(c_, same_, min_views_, carve_, max_carved_, grazing_)); }
  pm_cleaner* c_ = nullptr;
  float same_ = 0.01f, carve_ = 0.05f, grazing_ = 0.2f;
  int min_views_ = 3, max_carved_ = 2;
};

// Per-device memory reservations of reserve_device_memory.
uint64_t& reserved_bytes(int device) {
  static std::vector<uint64_t> r(64, 0);
  if (device < 0 || device >= static_cast<int>(r.size())) throw std::invalid_argument("device index out of range");
  return r[static_cast<size_t>(device)];
}

// DepthmapClusterEstimator (cluster_estimator.cc:42-156, bound at
// pybind.cc:64-123): several reference views, each with its own sources,
// run level by level; between levels a reference receives the previous
// level's depth maps of the references its geometric links name.  Every
// reference has its own handle; its device memory is released after each
// level, as the reference releases its GPU buffers.
class ClusterEstimator {
 public:
  ClusterEstimator() { pm_default_params(&params_); }
  ~ClusterEstimator() { clear(); }
  ClusterEstimator(const ClusterEstimator&) = delete;
  ClusterEstimator& operator=(const ClusterEstimator&) = delete;

  pm_params& params() { return params_; }
  void set_geom_consistency_weight(float w) { geom_weight_ = w; }
  void set_device(int d) { device_ = d; }

  int begin_ref_view(py::handle K, py::handle R, py::handle t, py::array image, float dmin, float dmax) {
    pm_handle* h = nullptr;
    check(pm_create(device_, &h));
    refs_.push_back(Ref{h, dmin, dmax, {}});
    add_view_to(h, K, R, t, image, py::none());
    return static_cast<int>(refs_.size()) - 1;
  }

  void add_source_view(py::handle K, py::handle R, py::handle t, py::array image) {
    if (refs_.empty()) throw std::runtime_error("DepthmapClusterEstimator: no ref view started");
    add_view_to(refs_.back().h, K, R, t, image, py::none());
  }

  // SetSfMPoints of the current reference (cluster_estimator.cc; called
  // per reference by depthmaps.py:235-242).
  void set_sfm_points(py::array points) {
    if (refs_.empty()) throw std::runtime_error("DepthmapClusterEstimator: no ref view started");
    set_points(refs_.back().h, points);
  }

  void add_geom_link(int ref_idx, int source_view_idx, int from_ref_idx) {
    if (ref_idx < 0 || ref_idx >= static_cast<int>(refs_.size()))
      throw std::runtime_error("DepthmapClusterEstimator::AddGeomLink: invalid ref_idx");
    refs_[ref_idx].links.emplace_back(source_view_idx, from_ref_idx);
  }

  // Run (cluster_estimator.cc:82-156): [depth, normal, cost, confidence] per reference.
  py::list run() {
    const int K = static_cast<int>(refs_.size());
    if (K == 0) throw std::runtime_error("DepthmapClusterEstimator: no reference views");
    for (auto& r : refs_) {
      pm_params p = params_;
      p.depth_min = r.dmin;
      p.depth_max = r.dmax;
      check(pm_set_params(r.h, &p));
      check(pm_set_geom_consistency(r.h, geom_weight_, nullptr, 0));
    }
    const int levels = std::max(1, params_.hierarchy_levels);
    struct Map {
      int w = 0, h = 0;
      std::vector<float> d;
    };
    std::vector<Map> prev(static_cast<size_t>(K));
    std::vector<std::vector<float>> out(static_cast<size_t>(K));
    std::vector<std::pair<int, int>> size(static_cast<size_t>(K));
    {
      // Every level runs all references as one batch (one launch per kernel,
      // one HIP graph), in groups that fit the free device memory, instead
      // of one reference after the other; finer levels start from each
      // reference's coarser result and see the linked references' previous
      // level through the geometric-consistency term.
      py::gil_scoped_release nogil;
      uint64_t total_mem = 0, free_mem = 0;
      check(pm_device_memory(device_, &total_mem, &free_mem));
      const uint64_t budget = free_mem / 10 * 9;
      for (int level = 0; level < levels; ++level) {
        const bool last = level == levels - 1;
        std::vector<Map> cur(static_cast<size_t>(K));
        for (int i = 0; i < K; ++i) {
          Ref& r = refs_[i];
          check(pm_clear_previous_depths(r.h));
          if (level > 0 && geom_weight_ > 0.0f) {
            for (const auto& link : r.links) {
              const int from = link.second;
              if (from >= 0 && from < K && !prev[from].d.empty())
                check(pm_set_previous_depth(r.h, link.first, prev[from].d.data(), prev[from].w, prev[from].h));
            }
          }
        }
        int i = 0;
        while (i < K) {
          pm_batch* b = nullptr;
          check(pm_batch_create(device_, &b));
          std::vector<int> group;
          uint64_t used = 0;
          try {
            for (; i < K; ++i) {
              uint64_t need = 0;  // at the finest level: an upper bound for the coarser ones
              check(pm_device_bytes(refs_[i].h, &need));
              if (!group.empty() && used + need > budget) break;
              check(pm_batch_add(b, refs_[i].h));
              group.push_back(i);
              used += need;
            }
            check(pm_batch_set_level(b, level));
            check(pm_batch_run(b));
            for (int j : group) {
              int w = 0, h = 0;
              check(pm_level_size(refs_[j].h, &w, &h));
              const size_t n = static_cast<size_t>(w) * h;
              if (last) {
                std::vector<float>& o = out[j];
                o.resize(6 * n);  // depth, normal (3), cost, confidence
                check(pm_download(refs_[j].h, o.data(), nullptr, o.data() + n, o.data() + 4 * n, nullptr,
                                  o.data() + 5 * n));
                size[j] = {w, h};
              } else {
                cur[j].w = w;
                cur[j].h = h;
                cur[j].d.resize(n);
                check(pm_download(refs_[j].h, cur[j].d.data(), nullptr, nullptr, nullptr, nullptr, nullptr));
              }
              check(pm_release_device(refs_[j].h));
            }
          } catch (...) {
            pm_batch_destroy(b);
            throw;
          }
          pm_batch_destroy(b);
        }
        prev = std::move(cur);
      }
    }
    py::list results;
    for (int i = 0; i < K; ++i) {
      const py::ssize_t w = size[i].first, h = size[i].second, n = w * h;
      const float* o = out[i].data();
      py::array_t<float> depth(shape({h, w})), normal(shape({h, w, 3})), cost(shape({h, w})), conf(shape({h, w}));
      std::memcpy(depth.mutable_data(), o, n * sizeof(float));
      std::memcpy(normal.mutable_data(), o + n, 3 * n * sizeof(float));
      std::memcpy(cost.mutable_data(), o + 4 * n, n * sizeof(float));
      std::memcpy(conf.mutable_data(), o + 5 * n, n * sizeof(float));
      py::list entry;
      entry.append(depth);
      entry.append(normal);
      entry.append(cost);
      entry.append(conf);
      results.append(entry);
    }
    return results;
  }

  void clear() {
    for (auto& r : refs_) pm_destroy(r.h);
    refs_.clear();
  }

 private:
  struct Ref {
    pm_handle* h;
    float dmin, dmax;
    std::vector<std::pair<int, int>> links;  // (source view slot, reference it takes depths from)
  };
  pm_params params_{};
  float geom_weight_ = 0.0f;
  int device_ = 0;
  std::vector<Ref> refs_;
};

}  // namespace

PYBIND11_MODULE(pydense, m) {
  m.doc() = "MI355X PatchMatch multi-view depth estimator (pydense.DepthmapEstimator surface)";
  // The argument checks of set_prior / set_geom_consistency, callable without
  // a device (they raise ValueError as those methods do).
  m.def("_check_prior_arrays", [](py::handle planes, py::handle masks, py::ssize_t width, py::ssize_t height) {
    prior_arrays(planes, masks, width, height);
  }, py::arg("planes"), py::arg("masks"), py::arg("width"), py::arg("height"));
  m.def("_check_depth_map", [](py::handle item) { depth_map_array(item); }, py::arg("depth"));
  m.def("default_params", []() {
    pm_params p;
    pm_default_params(&p);
    return params_dict(p);
  }, "DepthmapParams' defaults (estimator.h:16-36): what a new estimator starts from");
  m.def("pipeline_params", []() {
    pm_params p;
    pm_pipeline_params(&p);
    return params_dict(p);
  }, "The dense pipeline's parameters (depthmaps.py:154-175 from config.py:407-466)");
  py::class_<Estimator>(m, "DepthmapEstimator")
      .def(py::init<int>(), py::arg("device") = 0)
      .def("add_view",
           [](Estimator& e, py::handle K, py::handle R, py::handle t, py::array image, py::object mask) {
             e.add_view(K, R, t, image, mask);
             e.note_size(image);
           },
           py::arg("K"), py::arg("R"), py::arg("t"), py::arg("image"), py::arg("mask") = py::none())
      .def("add_view_device", &Estimator::add_view_device, py::arg("K"), py::arg("R"), py::arg("t"),
           py::arg("data_ptr"), py::arg("dtype"), py::arg("width"), py::arg("height"), py::arg("row_stride") = 0,
           "add_view for an image in device memory (e.g. a torch tensor's data_ptr(); dtype 'u8', 'f16' or "
           "'f32'); it is read at the next upload and must stay valid until then")
      .def("clear_views", [](Estimator& e) { e.clear(); })
      .def("set_depth_range", &Estimator::set_depth_range, py::arg("min_depth"), py::arg("max_depth"),
           py::arg("num_planes") = 0)
      .def("set_patchmatch_iterations", [](Estimator& e, int n) { e.update([n](pm_params& p) { p.max_iterations = n; }); })
      .def("set_patch_size", [](Estimator& e, int n) { e.update([n](pm_params& p) { p.patch_size = n; }); })
      .def("set_min_patch_sd",
           // patch SD in 0..255 grey levels -> weighted variance gate in [0,1]^2
           [](Estimator& e, float sd) {
             e.update([sd](pm_params& p) { p.var_threshold = (sd / 255.0f) * (sd / 255.0f); });
           })
      .def("set_sigma_spatial", [](Estimator& e, float v) { e.update([v](pm_params& p) { p.sigma_spatial = v; }); })
      .def("set_sigma_color", [](Estimator& e, float v) { e.update([v](pm_params& p) { p.sigma_color = v; }); })
      .def("set_top_k", [](Estimator& e, int v) { e.update([v](pm_params& p) { p.top_k = v; }); })
      .def("set_use_census", [](Estimator& e, bool v) { e.update([v](pm_params& p) { p.use_census = v; }); })
      .def("set_smooth_weight", [](Estimator& e, float v) { e.update([v](pm_params& p) { p.smooth_weight = v; }); })
      .def("set_anchor_views", [](Estimator& e, int v) { e.update([v](pm_params& p) { p.anchor_views = v; }); })
      .def("set_checkerboard_filter",
           [](Estimator& e, bool v) { e.update([v](pm_params& p) { p.checkerboard_filter = v; }); })
      .def("set_hierarchy_levels", [](Estimator& e, int v) { e.update([v](pm_params& p) { p.hierarchy_levels = v; }); })
      .def("set_var_threshold", [](Estimator& e, float v) { e.update([v](pm_params& p) { p.var_threshold = v; }); })
      .def("set_image_storage",
           [](Estimator& e, const std::string& s) {
             const int code = storage_code(s);
             e.update([code](pm_params& p) { p.image_storage = code; });
           })
      .def("set_quantize_half", [](Estimator& e, bool v) { e.update([v](pm_params& p) { p.quantize_half = v; }); })
      .def("set_speckle_min_size", [](Estimator& e, int v) { e.update([v](pm_params& p) { p.speckle_min_size = v; }); })
      .def("set_gap_max_size", [](Estimator& e, int v) { e.update([v](pm_params& p) { p.gap_max_size = v; }); })
      .def("set_geom_consistency", &Estimator::set_geom_consistency, py::arg("weight"), py::arg("prev_depths"))
      .def("set_prior", &Estimator::set_prior, py::arg("prior_planes"), py::arg("masks"))
      .def("set_sfm_points", [](Estimator& e, py::handle pts) { set_points(e.handle(), pts); }, py::arg("points"),
           "SfM points (N, 3) of the reference: planar prior per level (ComputeSfMPlanarPrior)")
      .def("params", &Estimator::params)
      .def("set_params", &Estimator::set_params, py::arg("params"),
           "Set any subset of the params() fields, e.g. set_params(pydense.pipeline_params())")
      .def("compute_patch_match", &Estimator::compute_patch_match,
           "Run PatchMatch; returns [depth (H,W), plane normal (H,W,3), score (H,W), selected-view bitmask (H,W)]")
      .def("compute_patch_match_sample", &Estimator::compute_patch_match_sample,
           "compute_patch_match (PatchMatch samples each pixel's views already)")
      .def("compute_brute_force", &Estimator::compute_brute_force,
           "Fronto-parallel plane sweep over set_depth_range's num_planes log-spaced depths; returns the same "
           "list as compute_patch_match")
      .def("compute", &Estimator::compute, "Run PatchMatch; returns a dict of every decoded map")
      .def("upload", &Estimator::upload)
      .def("run", &Estimator::run)
      .def("download", &Estimator::download, py::arg("maps") = py::none(),
           "Decoded maps of the last run: all, or those named in maps (depth, plane, normal, cost, views, "
           "confidence)")
      .def("release_device", &Estimator::release_device,
           "Free the device memory of this estimator (views stay on the host; the next run uploads again)")
      .def("init_state", &Estimator::init_state)
      .def("iterate", &Estimator::iterate, py::arg("iteration"), py::arg("color") = -1)
      .def("filter", &Estimator::filter)
      .def("decode", &Estimator::decode)
      .def("read_state", &Estimator::read_state)
      .def("write_state", &Estimator::write_state)
      .def("score_planes", &Estimator::score_planes, py::arg("xy"), py::arg("planes"))
      .def("last_timing", &Estimator::last_timing)
      .def("ncc_cost_grad", &Estimator::ncc_cost_grad, py::arg("depth"), py::arg("ab"), py::arg("views"),
           "Photometric cost (H,W) and its gradient (H,W,3) w.r.t. (depth, a, b) of the normal (a, b, -1)")
      .def("ncc_cost_grad_device", &Estimator::ncc_cost_grad_device, py::arg("depth"), py::arg("ab"),
           py::arg("views"), py::arg("cost"), py::arg("grad"))
      .def("gauss_newton", &Estimator::gauss_newton, py::arg("steps") = 10, py::arg("damping") = 1e-3f,
           "Gauss-Newton refinement of the PatchMatch result; read it with download()")
      .def("last_refine_ms", &Estimator::last_refine_ms)
      .def("stage_timing", &Estimator::stage_timing,
           "Per-stage device time of the last run: {stage: (milliseconds, launches)}")
      .def("set_tiles", &Estimator::set_tiles, py::arg("cost_tile_x"), py::arg("pixel_tile_x"),
           "Colour-tile widths of the cost kernels (4, 8, 16) and the per-pixel kernels (16, 32)")
      .def("tiles", &Estimator::tiles)
      .def("device_bytes", &Estimator::device_bytes, "Device memory a run of this estimator allocates")
      .def_static("is_available", []() {
        int n = 0;
        return pm_device_count(&n) == PM_OK && n > 0;
      })
      .def_static("set_buffer_cache", [](int i, uint64_t max_bytes) { check(pm_set_buffer_cache(i, max_bytes)); },
                  py::arg("device"), py::arg("max_bytes"),
                  "Keep up to max_bytes of released device buffers for reuse by later estimators (0: off)")
      .def_static("buffer_cache_bytes", [](int i) {
        uint64_t held = 0;
        check(pm_buffer_cache_bytes(i, &held));
        return held;
      }, py::arg("device") = 0)
      .def_static("device_free_memory", [](int i) {
        uint64_t total = 0, free_b = 0;
        check(pm_device_memory(i, &total, &free_b));
        return free_b;
      }, py::arg("device") = 0, "Free memory of a device in bytes")
      .def_static("num_devices", []() {
        int n = 0;
        pm_device_count(&n);
        return n;
      })
      .def_static("device_name", [](int i) {
        char buf[256];
        check(pm_device_name(i, buf, sizeof(buf)));
        return std::string(buf);
      });
  m.attr("MAX_SOURCES") = PM_MAX_SOURCES;

  using C = ClusterEstimator;
  auto set = [](auto field) {
    return [field](C& c, decltype(std::declval<pm_params&>().*field) v) { c.params().*field = v; };
  };
  m.def(
      "sfm_planar_prior",
      [](py::handle K, py::handle R, py::handle t, int w, int h, float dmin, float dmax, py::handle points) {
        auto k = as_f64(K, 9, "K");
        auto r = as_f64(R, 9, "R");
        auto tt = as_f64(t, 3, "t");
        auto a = py::array_t<double, py::array::c_style | py::array::forcecast>::ensure(points);
        if (!a || (a.size() > 0 && (a.ndim() != 2 || a.shape(1) != 3)))
          throw std::invalid_argument("points must be an (N, 3) array");
        py::array_t<float> planes(shape({h, w, 4}));
        py::array_t<uint32_t> masks(shape({h, w}));
        int ntri = 0;
        check(pm_sfm_planar_prior(k.data(), r.data(), tt.data(), w, h, dmin, dmax, a.size() ? a.data() : nullptr,
                                  static_cast<int>(a.size() / 3), planes.mutable_data(), masks.mutable_data(), &ntri));
        return py::make_tuple(planes, masks, ntri);
      },
      py::arg("K"), py::arg("R"), py::arg("t"), py::arg("width"), py::arg("height"), py::arg("depth_min"),
      py::arg("depth_max"), py::arg("points"),
      "ComputeSfMPlanarPrior of one image: (prior planes (H, W, 4), masks (H, W), triangles kept)");
  m.def(
      "delaunay",
      [](py::handle xy) {
        auto a = py::array_t<double, py::array::c_style | py::array::forcecast>::ensure(xy);
        if (!a || (a.size() > 0 && (a.ndim() != 2 || a.shape(1) != 2)))
          throw std::invalid_argument("points must be an (N, 2) array");
        const int n = static_cast<int>(a.size() / 2);
        int count = 0;
        check(pm_delaunay(a.data(), n, nullptr, 0, &count));
        py::array_t<int> tris(shape({count, 3}));
        check(pm_delaunay(a.data(), n, tris.mutable_data(), count, &count));
        return tris;
      },
      py::arg("xy"), "Delaunay triangles of (N, 2) points, canonical order");
  m.def(
      "select_neighbors",
      [](py::handle origins, py::handle coords, const std::vector<std::vector<int>>& vis,
         const std::vector<int>& track_counts, int num_neighbors, int min_point_best, int min_point_all,
         double theta_min_deg, double theta_max_deg) {
        auto o = py::array_t<double, py::array::c_style | py::array::forcecast>::ensure(origins);
        auto c = py::array_t<double, py::array::c_style | py::array::forcecast>::ensure(coords);
        if (!o || !c) throw std::invalid_argument("origins / coords must be float arrays");
        const int N = static_cast<int>(o.size() / 3), P = static_cast<int>(c.size() / 3);
        if (static_cast<int>(vis.size()) != P || static_cast<int>(track_counts.size()) != P)
          throw std::invalid_argument("one visibility list and one track count per super-point");
        std::vector<int> off(1, 0), shots;
        for (const auto& v : vis) {
          shots.insert(shots.end(), v.begin(), v.end());
          off.push_back(static_cast<int>(shots.size()));
        }
        std::vector<int> best(static_cast<size_t>(N) * (1 + num_neighbors)), bc(N), all(static_cast<size_t>(N) * N),
            ac(N);
        {
          py::gil_scoped_release nogil;
          check(pm_select_neighbors(N, o.data(), P, c.data(), off.data(), shots.data(), track_counts.data(),
                                    num_neighbors, min_point_best, min_point_all, theta_min_deg, theta_max_deg,
                                    best.data(), bc.data(), all.data(), ac.data()));
        }
        py::list b, a;
        for (int i = 0; i < N; ++i) {
          py::list bi, ai;
          for (int k = 0; k < bc[i]; ++k) bi.append(best[static_cast<size_t>(i) * (1 + num_neighbors) + k]);
          for (int k = 0; k < ac[i]; ++k) ai.append(all[static_cast<size_t>(i) * N + k]);
          b.append(bi);
          a.append(ai);
        }
        return py::make_tuple(b, a);
      },
      py::arg("origins"), py::arg("coords"), py::arg("visibility"), py::arg("track_counts"),
      py::arg("num_neighbors"), py::arg("min_point_best"), py::arg("min_point_all"), py::arg("theta_min_deg"),
      py::arg("theta_max_deg"),
      "SelectNeighbors: (best, all) neighbour index lists per shot, self first");

  // LZ4 frames of the depth-map files (dataset.py:1416-1452).
  m.def(
      "lz4_frame_compress",
      [](py::buffer data, int block_max_id, bool content_checksum) {
        py::buffer_info bi = data.request();
        const auto* p = static_cast<const uint8_t*>(bi.ptr);
        const size_t n = static_cast<size_t>(bi.size * bi.itemsize);
        std::vector<uint8_t> out;
        {
          py::gil_scoped_release nogil;
          out = lz4f::compress_frame(p, n, block_max_id, content_checksum);
        }
        return py::bytes(reinterpret_cast<const char*>(out.data()), out.size());
      },
      py::arg("data"), py::arg("block_max_id") = 7, py::arg("content_checksum") = false,
      "One LZ4 frame of data (independent blocks of 64 KB << 2*(block_max_id - 4), content size stored)");
  m.def(
      "lz4_frame_decompress",
      [](py::buffer data) {
        py::buffer_info bi = data.request();
        const auto* p = static_cast<const uint8_t*>(bi.ptr);
        const size_t n = static_cast<size_t>(bi.size * bi.itemsize);
        std::vector<uint8_t> out;
        std::string err;
        bool ok;
        {
          py::gil_scoped_release nogil;
          ok = lz4f::decompress_frames(p, n, out, err);
        }
        if (!ok) throw std::runtime_error(err);
        return py::bytes(reinterpret_cast<const char*>(out.data()), out.size());
      },
      py::arg("data"), "Content of one or more concatenated LZ4 frames");
  m.def(
      "xxh32",
      [](py::buffer data, uint32_t seed) {
        py::buffer_info bi = data.request();
        return lz4f::xxh32(static_cast<const uint8_t*>(bi.ptr), static_cast<size_t>(bi.size * bi.itemsize), seed);
      },
      py::arg("data"), py::arg("seed") = 0);

  py::class_<Cleaner>(m, "GPUDepthmapCleaner")
      .def(py::init<>())
      .def("set_same_depth_threshold", &Cleaner::set_same_depth_threshold)
      .def("set_min_consistent_views", &Cleaner::set_min_consistent_views)
      .def("set_device", &Cleaner::set_device)
      .def("set_carving_threshold", &Cleaner::set_carving_threshold)
      .def("set_max_carved_views", &Cleaner::set_max_carved_views)
      .def("set_grazing_cos_threshold", &Cleaner::set_grazing_cos_threshold)
      .def("add_view", [](Cleaner& c, py::handle K, py::handle R, py::handle t, py::handle depth) {
        return c.add_view(K, R, t, depth, py::none());
      })
      .def("add_view_with_normal", [](Cleaner& c, py::handle K, py::handle R, py::handle t, py::handle depth,
                                      py::object normal) { return c.add_view(K, R, t, depth, normal); })
      .def("clean", &Cleaner::clean, py::arg("ref_idx"), py::arg("neighbor_ids"))
      .def("clear", &Cleaner::clear)
      .def("last_ms", &Cleaner::last_ms, "Device time of the last clean's kernel")
      .def_static("is_available", []() {
        int n = 0;
        return pm_device_count(&n) == PM_OK && n > 0;
      });

  py::class_<Batch>(m, "DepthmapBatch",
                    "Many DepthmapEstimator references in one launch per kernel, captured in a HIP graph")
      .def(py::init<int>(), py::arg("device") = 0)
      .def("add", &Batch::add, py::arg("estimator"), py::keep_alive<1, 2>())
      .def("clear", &Batch::clear)
      .def("size", &Batch::size)
      .def("set_level", &Batch::set_level, py::arg("level"),
           "Pyramid level to run (0 = coarsest); finer levels start from each reference's previous result")
      .def("set_config", &Batch::set_config, py::arg("cost_tile_x") = 0, py::arg("pixel_tile_x") = 0,
           py::arg("use_graph") = true, "tile widths: 0 = chosen by the autotuner")
      .def("config", &Batch::config)
      .def("tune_log", &Batch::tune_log)
      .def("prepare", &Batch::prepare)
      .def("run", &Batch::run, "Run every reference; returns the device time in ms")
      .def("stage_timing", &Batch::stage_timing,
           "Per-stage device time of the last run without a graph: {stage: (milliseconds, launches)}")
      .def_static("reset_tuning", []() { pm_batch_reset_tuning(); });

  py::class_<C>(m, "DepthmapClusterEstimator")
      .def(py::init<>())
      .def("set_max_iterations", set(&pm_params::max_iterations))
      .def("set_patch_size", set(&pm_params::patch_size))
      // images are used at the size they are given
      .def("set_max_image_size", [](C&, int) {})
      .def("set_hierarchy_levels", set(&pm_params::hierarchy_levels))
      .def("set_sigma_spatial", set(&pm_params::sigma_spatial))
      .def("set_sigma_color", set(&pm_params::sigma_color))
      .def("set_top_k", set(&pm_params::top_k))
      .def("set_use_census", [](C& c, bool v) { c.params().use_census = v; })
      .def("set_smooth_weight", set(&pm_params::smooth_weight))
      .def("set_anchor_views", set(&pm_params::anchor_views))
      .def("set_checkerboard_filter", [](C& c, bool v) { c.params().checkerboard_filter = v; })
      .def("set_speckle_min_size", set(&pm_params::speckle_min_size))
      .def("set_gap_max_size", set(&pm_params::gap_max_size))
      .def("set_geom_consistency_weight", &C::set_geom_consistency_weight)
      .def("params", [](C& c) { return params_dict(c.params()); })
      .def("set_params", [](C& c, py::dict d) { params_from_dict(c.params(), d); }, py::arg("params"))
      .def("set_device", &C::set_device)
      .def("begin_ref_view", &C::begin_ref_view, py::arg("K"), py::arg("R"), py::arg("t"), py::arg("image"),
           py::arg("depth_min"), py::arg("depth_max"))
      .def("add_source_view", &C::add_source_view, py::arg("K"), py::arg("R"), py::arg("t"), py::arg("image"))
      .def("set_sfm_points", &C::set_sfm_points, py::arg("points"))
      .def("add_geom_link", &C::add_geom_link, py::arg("ref_idx"), py::arg("source_view_idx"),
           py::arg("from_ref_idx"))
      .def("run", &C::run, "Run every reference view; returns [depth, normal, cost, confidence] per reference")
      .def("clear", &C::clear)
      .def_static("is_available", []() {
        int n = 0;
        return pm_device_count(&n) == PM_OK && n > 0;
      })
      .def_static("num_devices", []() {
        int n = 0;
        pm_device_count(&n);
        return n;
      })
      .def_static("device_name", [](int i) {
        char buf[256];
        check(pm_device_name(i, buf, sizeof(buf)));
        return std::string(buf);
      })
      .def_static("device_is_gpu", [](int i) {
        int n = 0;
        return pm_device_count(&n) == PM_OK && i >= 0 && i < n;
      })
      .def_static("device_memory_bytes", [](int i) {
        uint64_t total = 0;
        check(pm_device_memory(i, &total, nullptr));
        return total;
      })
      .def_static("gpu_memory_bytes", []() {
        int n = 0;
        uint64_t best = 0;
        if (pm_device_count(&n) != PM_OK) return best;
        for (int i = 0; i < n; ++i) {
          uint64_t t = 0;
          if (pm_device_memory(i, &t, nullptr) == PM_OK) best = std::max(best, t);
        }
        return best;
      })
      // Reservations are bookkeeping for callers that share a device
      // between clusters (available = total - reserved).
      .def_static("device_available_memory", [](int i) {
        uint64_t total = 0;
        check(pm_device_memory(i, &total, nullptr));
        const uint64_t r = reserved_bytes(i);
        return total > r ? total - r : 0;
      })
      .def_static("reserve_device_memory", [](int i, uint64_t bytes) { reserved_bytes(i) += bytes; })
      .def_static("release_device_memory", [](int i, uint64_t bytes) {
        uint64_t& r = reserved_bytes(i);
        r = r > bytes ? r - bytes : 0;
      });
}
