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
#include <cstring>
#include <initializer_list>
#include <stdexcept>
#include <string>
#include <vector>

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

class Estimator {
 public:
  explicit Estimator(int device) { check(pm_create(device, &h_)); }
  ~Estimator() { pm_destroy(h_); }
  Estimator(const Estimator&) = delete;
  Estimator& operator=(const Estimator&) = delete;

  void add_view(py::handle K, py::handle R, py::handle t, py::array image, py::object mask) {
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
    check(pm_add_view(h_, k.data(), r.data(), tt.data(), img.data(), fmt, static_cast<int>(img.shape(1)),
                      static_cast<int>(img.shape(0)), 0, m.empty() ? nullptr : m.data()));
  }

  void set_depth_range(float dmin, float dmax, int /*num_planes*/) { check(pm_set_depth_range(h_, dmin, dmax)); }

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
    return d;
  }

  py::list compute_patch_match() {
    py::dict r = compute();
    py::list out;
    out.append(r["depth"]);
    out.append(r["normal"]);
    out.append(r["cost"]);
    out.append(r["views"]);
    return out;
  }

  py::dict compute() {
    {
      py::gil_scoped_release nogil;
      check(pm_upload(h_));
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

  py::dict download() {
    const auto [w, h] = ref_size();
    py::array_t<float> depth(shape({h, w})), plane(shape({h, w, 4})), normal(shape({h, w, 3})),
        cost(shape({h, w})), conf(shape({h, w}));
    py::array_t<uint32_t> views(shape({h, w}));
    {
      py::gil_scoped_release nogil;
      check(pm_download(h_, depth.mutable_data(), plane.mutable_data(), normal.mutable_data(),
                        cost.mutable_data(), views.mutable_data(), conf.mutable_data()));
    }
    py::dict d;
    d["depth"] = depth;
    d["plane"] = plane;
    d["normal"] = normal;
    d["cost"] = cost;
    d["views"] = views;
    d["confidence"] = conf;
    return d;
  }

  py::tuple read_state() {
    const auto [w, h] = ref_size();
    py::array_t<float> planes(shape({h, w, 4})), costs(shape({h, w}));
    py::array_t<uint32_t> rng(shape({h, w, 2})), sel(shape({h, w}));
    check(pm_read_state(h_, planes.mutable_data(), costs.mutable_data(), rng.mutable_data(), sel.mutable_data()));
    return py::make_tuple(planes, costs, rng, sel);
  }

  void write_state(py::array_t<float, py::array::c_style | py::array::forcecast> planes,
                   py::array_t<float, py::array::c_style | py::array::forcecast> costs,
                   py::array_t<uint32_t, py::array::c_style | py::array::forcecast> rng,
                   py::array_t<uint32_t, py::array::c_style | py::array::forcecast> sel) {
    const auto [w, h] = ref_size();
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

  void set_geom_consistency(float weight, py::list prev) {
    std::vector<py::array_t<float, py::array::c_style | py::array::forcecast>> keep;
    std::vector<const float*> ptrs;
    for (auto item : prev) {
      if (item.is_none()) {
        ptrs.push_back(nullptr);
      } else {
        keep.push_back(py::array_t<float, py::array::c_style | py::array::forcecast>::ensure(item));
        ptrs.push_back(keep.back().data());
      }
    }
    check(pm_set_geom_consistency(h_, weight, ptrs.data(), static_cast<int>(ptrs.size())));
  }

  void set_prior(py::object planes, py::object masks) {
    if (planes.is_none() || masks.is_none()) {
      check(pm_set_prior(h_, nullptr, nullptr));
      return;
    }
    auto p = py::array_t<float, py::array::c_style | py::array::forcecast>::ensure(planes);
    auto m = py::array_t<uint32_t, py::array::c_style | py::array::forcecast>::ensure(masks);
    check(pm_set_prior(h_, p.data(), m.data()));
  }

 private:
  std::pair<py::ssize_t, py::ssize_t> ref_size() {
    if (ref_w_ < 0) throw std::runtime_error("DepthmapEstimator: no views added");
    return {ref_w_, ref_h_};
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

 private:
  pm_handle* h_ = nullptr;
  py::ssize_t ref_w_ = -1, ref_h_ = -1;
};

}  // namespace

PYBIND11_MODULE(pydense, m) {
  m.doc() = "MI355X PatchMatch multi-view depth estimator (pydense.DepthmapEstimator surface)";
  py::class_<Estimator>(m, "DepthmapEstimator")
      .def(py::init<int>(), py::arg("device") = 0)
      .def("add_view",
           [](Estimator& e, py::handle K, py::handle R, py::handle t, py::array image, py::object mask) {
             e.add_view(K, R, t, image, mask);
             e.note_size(image);
           },
           py::arg("K"), py::arg("R"), py::arg("t"), py::arg("image"), py::arg("mask") = py::none())
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
             if (s != "f32" && s != "f16") throw std::invalid_argument("image storage must be 'f32' or 'f16'");
             e.update([s](pm_params& p) { p.image_storage = s == "f16" ? PM_STORE_F16 : PM_STORE_F32; });
           })
      .def("set_quantize_half", [](Estimator& e, bool v) { e.update([v](pm_params& p) { p.quantize_half = v; }); })
      .def("set_geom_consistency", &Estimator::set_geom_consistency, py::arg("weight"), py::arg("prev_depths"))
      .def("set_prior", &Estimator::set_prior, py::arg("prior_planes"), py::arg("masks"))
      .def("params", &Estimator::params)
      .def("compute_patch_match", &Estimator::compute_patch_match,
           "Run PatchMatch; returns [depth (H,W), plane normal (H,W,3), score (H,W), selected-view bitmask (H,W)]")
      .def("compute", &Estimator::compute, "Run PatchMatch; returns a dict of every decoded map")
      .def("upload", &Estimator::upload)
      .def("run", &Estimator::run)
      .def("download", &Estimator::download)
      .def("init_state", &Estimator::init_state)
      .def("iterate", &Estimator::iterate, py::arg("iteration"), py::arg("color") = -1)
      .def("filter", &Estimator::filter)
      .def("decode", &Estimator::decode)
      .def("read_state", &Estimator::read_state)
      .def("write_state", &Estimator::write_state)
      .def("score_planes", &Estimator::score_planes, py::arg("xy"), py::arg("planes"))
      .def("last_timing", &Estimator::last_timing)
      .def("stage_timing", &Estimator::stage_timing,
           "Per-stage device time of the last run: {stage: (milliseconds, launches)}")
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
      });
  m.attr("MAX_SOURCES") = PM_MAX_SOURCES;
}
