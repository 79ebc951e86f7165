// Host code around the PatchMatch path: the SfM planar prior
// (DepthmapEstimator::ComputeSfMPlanarPrior, estimator.cc:838-1107) and the
// neighbour selection that picks each reference's sources
// (sfm::dense_helpers::SelectNeighbors, dense_helpers.cc:331-568).
//
// Differences from the reference, both at third-party boundaries:
//   * the triangulation is this file's Bowyer-Watson Delaunay (cv::Subdiv2D
//     in the reference): for points in general position the triangle set is
//     the same, but the order in which triangles are rasterised (the later
//     one owns the pixels of a shared edge) and the order of a triangle's
//     vertices in the parametric rasteriser are canonical here (vertex
//     indices ascending) instead of OpenCV's internal edge order;
//   * the plane through a triangle is the cross product of its edges in
//     double precision (Eigen's float JacobiSVD null vector in the
//     reference): the same plane up to float rounding;
//   * a triangle two of whose vertices share a pixel (so the pixel -> point
//     map gives one SfM point for both, estimator.cc:914-917) is skipped;
//     the reference fits an arbitrary null vector of the rank-2 system;
//   * candidate neighbours of equal score are ordered by index (the
//     reference's std::sort over an unordered_map leaves ties unspecified).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/opensfm_pm.h"

namespace pmhip {
int api_fail(int code, const std::string& msg);
}  // namespace pmhip

using namespace pmhip;

namespace {

struct P2 {
  double x, y;
};

// > 0 when d lies inside the circumcircle of the counter-clockwise a, b, c.
double incircle(const P2& a, const P2& b, const P2& c, const P2& d) {
  const double adx = a.x - d.x, ady = a.y - d.y;
  const double bdx = b.x - d.x, bdy = b.y - d.y;
  const double cdx = c.x - d.x, cdy = c.y - d.y;
  const double ad = adx * adx + ady * ady, bd = bdx * bdx + bdy * bdy, cd = cdx * cdx + cdy * cdy;
  return adx * (bdy * cd - bd * cdy) - ady * (bdx * cd - bd * cdx) + ad * (bdx * cdy - bdy * cdx);
}

double orient(const P2& a, const P2& b, const P2& c) { return (b.x - a.x) * (c.y - a.y) - (b.y - a.y) * (c.x - a.x); }

// Bowyer-Watson Delaunay triangulation of the points; exact duplicates are
// inserted once (the first copy).  Output in canonical order: the vertex
// indices of a triangle ascending, the triangles sorted.
std::vector<std::array<int, 3>> delaunay(const std::vector<P2>& pts) {
  const int n = static_cast<int>(pts.size());
  std::vector<std::array<int, 3>> out;
  if (n < 3) return out;
  double minx = pts[0].x, maxx = pts[0].x, miny = pts[0].y, maxy = pts[0].y;
  for (const P2& p : pts) {
    minx = std::min(minx, p.x); maxx = std::max(maxx, p.x);
    miny = std::min(miny, p.y); maxy = std::max(maxy, p.y);
  }
  const double span = std::max({maxx - minx, maxy - miny, 1.0});
  const double mx = 0.5 * (minx + maxx), my = 0.5 * (miny + maxy);
  std::vector<P2> v(pts);
  const double big = 1.0e5 * span;  // super-triangle, far enough that no hull circumcircle reaches it
  v.push_back({mx - big, my - 0.75 * big});
  v.push_back({mx + big, my - 0.75 * big});
  v.push_back({mx, my + big});
  struct Tri {
    int a, b, c;
  };
  std::vector<Tri> tris = {{n, n + 1, n + 2}};
  std::unordered_map<uint64_t, int> seen;  // exact duplicates
  std::vector<std::pair<int, int>> edges;
  std::vector<Tri> keep;
  for (int i = 0; i < n; ++i) {
    uint64_t kx, ky;
    std::memcpy(&kx, &pts[i].x, 8);
    std::memcpy(&ky, &pts[i].y, 8);
    const uint64_t key = kx * 0x9E3779B97F4A7C15ull ^ (ky + (kx << 6) + (kx >> 2));
    auto it = seen.find(key);
    if (it != seen.end() && pts[it->second].x == pts[i].x && pts[it->second].y == pts[i].y) continue;
    seen.emplace(key, i);
    edges.clear();
    keep.clear();
    for (const Tri& t : tris) {
      if (incircle(v[t.a], v[t.b], v[t.c], v[i]) > 0.0) {
        edges.emplace_back(t.a, t.b);
        edges.emplace_back(t.b, t.c);
        edges.emplace_back(t.c, t.a);
      } else {
        keep.push_back(t);
      }
    }
    // boundary of the cavity: edges that appear once (an inner edge appears
    // as (u, v) and (v, u))
    std::unordered_map<uint64_t, int> count;
    for (const auto& e : edges) {
      const uint64_t k = (static_cast<uint64_t>(std::min(e.first, e.second)) << 32) |
                         static_cast<uint32_t>(std::max(e.first, e.second));
      count[k]++;
    }
    for (const auto& e : edges) {
      const uint64_t k = (static_cast<uint64_t>(std::min(e.first, e.second)) << 32) |
                         static_cast<uint32_t>(std::max(e.first, e.second));
      if (count[k] == 1) keep.push_back(Tri{e.first, e.second, i});
    }
    tris.swap(keep);
  }
  for (const Tri& t : tris) {
    if (t.a >= n || t.b >= n || t.c >= n) continue;
    std::array<int, 3> r = {t.a, t.b, t.c};
    if (orient(v[r[0]], v[r[1]], v[r[2]]) == 0.0) continue;  // degenerate
    std::sort(r.begin(), r.end());
    out.push_back(r);
  }
  std::sort(out.begin(), out.end());
  return out;
}

}  // namespace

extern "C" {

int pm_delaunay(const double* xy, int n, int* triangles, int capacity, int* count) {
  if (n < 0 || (n > 0 && !xy) || !count) return api_fail(PM_E_INVALID, "pm_delaunay: bad arguments");
  std::vector<P2> pts(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) pts[i] = {xy[2 * i], xy[2 * i + 1]};
  const auto tris = delaunay(pts);
  *count = static_cast<int>(tris.size());
  if (triangles)
    for (int i = 0; i < std::min(capacity, *count); ++i)
      for (int k = 0; k < 3; ++k) triangles[3 * i + k] = tris[static_cast<size_t>(i)][k];
  return PM_OK;
}

// ComputeSfMPlanarPrior (estimator.cc:838-1107) for one w x h level with
// intrinsics K (of that level), R, t of the reference.
int pm_sfm_planar_prior(const double* K, const double* R, const double* t, int w, int h, float depth_min,
                        float depth_max, const double* points, int num_points, float* prior_planes, uint32_t* masks,
                        int* num_triangles) {
  if (!K || !R || !t || w <= 0 || h <= 0 || num_points < 0 || (num_points > 0 && !points) || !prior_planes || !masks)
    return api_fail(PM_E_INVALID, "pm_sfm_planar_prior: bad arguments");
  const size_t npix = static_cast<size_t>(w) * h;
  for (size_t i = 0; i < npix; ++i) {  // unmasked: (0, 0, -1, 0), mask 0 (:1061-1062)
    prior_planes[4 * i] = 0.0f;
    prior_planes[4 * i + 1] = 0.0f;
    prior_planes[4 * i + 2] = -1.0f;
    prior_planes[4 * i + 3] = 0.0f;
    masks[i] = 0u;
  }
  if (num_triangles) *num_triangles = 0;
  // float intrinsics of the CLCamera (:839-840)
  const float fx = static_cast<float>(K[0]), fy = static_cast<float>(K[4]);
  const float cx = static_cast<float>(K[2]), cy = static_cast<float>(K[5]);
  // 1. projections (:856-888)
  struct Proj {
    float u, v, depth, cam[3];
  };
  std::vector<Proj> proj;
  proj.reserve(static_cast<size_t>(num_points));
  for (int i = 0; i < num_points; ++i) {
    const double* X = points + 3 * i;
    const double xc = R[0] * X[0] + R[1] * X[1] + R[2] * X[2] + t[0];
    const double yc = R[3] * X[0] + R[4] * X[1] + R[5] * X[2] + t[1];
    const double zc = R[6] * X[0] + R[7] * X[1] + R[8] * X[2] + t[2];
    if (zc <= 0.0) continue;
    const float depth = static_cast<float>(zc);
    if (depth < depth_min || depth > depth_max) continue;
    const float u = static_cast<float>(fx * (xc / zc) + cx);
    const float v = static_cast<float>(fy * (yc / zc) + cy);
    if (u < 0.0f || u >= static_cast<float>(w) || v < 0.0f || v >= static_cast<float>(h)) continue;
    proj.push_back(Proj{u, v, depth, {static_cast<float>(xc), static_cast<float>(yc), static_cast<float>(zc)}});
  }
  if (proj.size() < 4) return PM_OK;  // (:898-907)
  // 2. triangulation; vertex -> projection through the truncated pixel (:910-921, 929-940)
  std::vector<P2> pts(proj.size());
  std::unordered_map<int, int> px_to_proj;
  for (size_t i = 0; i < proj.size(); ++i) {
    pts[i] = {proj[i].u, proj[i].v};
    px_to_proj[static_cast<int>(proj[i].v) * w + static_cast<int>(proj[i].u)] = static_cast<int>(i);
  }
  const auto tris = delaunay(pts);
  // 3. quality gates and planes (:927-1001)
  struct TriData {
    int x[3], y[3];
    float plane[4];
  };
  std::vector<TriData> td;
  for (const auto& tri : tris) {
    TriData d{};
    bool ok = true;
    int pi[3];
    for (int k = 0; k < 3; ++k) {
      d.x[k] = static_cast<int>(pts[tri[k]].x);
      d.y[k] = static_cast<int>(pts[tri[k]].y);
      if (d.x[k] < 0 || d.x[k] >= w || d.y[k] < 0 || d.y[k] >= h) ok = false;
      auto it = ok ? px_to_proj.find(d.y[k] * w + d.x[k]) : px_to_proj.end();
      if (it == px_to_proj.end()) ok = false; else pi[k] = it->second;
    }
    if (!ok) continue;
    const float d1 = proj[pi[0]].depth, d2 = proj[pi[1]].depth, d3 = proj[pi[2]].depth;
    const float dmn = std::min({d1, d2, d3}), dmx = std::max({d1, d2, d3});
    if (dmn > 0.0f && dmx / dmn > 1.5f) continue;
    const float ax = static_cast<float>(d.x[1] - d.x[0]), ay = static_cast<float>(d.y[1] - d.y[0]);
    const float bx = static_cast<float>(d.x[2] - d.x[0]), by = static_cast<float>(d.y[2] - d.y[0]);
    if (0.5f * std::abs(ax * by - ay * bx) > 20000.0f) continue;
    const float* X1 = proj[pi[0]].cam;
    const float* X2 = proj[pi[1]].cam;
    const float* X3 = proj[pi[2]].cam;
    const double e1[3] = {double(X2[0]) - X1[0], double(X2[1]) - X1[1], double(X2[2]) - X1[2]};
    const double e2[3] = {double(X3[0]) - X1[0], double(X3[1]) - X1[1], double(X3[2]) - X1[2]};
    double nn[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    double dd = -(nn[0] * X1[0] + nn[1] * X1[1] + nn[2] * X1[2]);
    const double l4 = std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2] + dd * dd);
    if (!(l4 > 0.0)) continue;
    const double ln = std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
    if (ln / l4 < 1e-8) continue;  // the unit null vector's normal part (:985-989)
    double s = (dd < 0.0 ? -1.0 : 1.0) / ln;  // unit normal, d >= 0 (:990-993)
    for (int k = 0; k < 3; ++k) d.plane[k] = static_cast<float>(nn[k] * s);
    d.plane[3] = static_cast<float>(dd * s);
    td.push_back(d);
  }
  if (num_triangles) *num_triangles = static_cast<int>(td.size());
  if (td.empty()) return PM_OK;
  // 5. parametric rasterisation; a later triangle owns shared pixels (:1029-1053)
  std::vector<int> owner(npix, 0);
  for (size_t ti = 0; ti < td.size(); ++ti) {
    const TriData& d = td[ti];
    const float L01 = std::hypot(static_cast<float>(d.x[0] - d.x[1]), static_cast<float>(d.y[0] - d.y[1]));
    const float L02 = std::hypot(static_cast<float>(d.x[0] - d.x[2]), static_cast<float>(d.y[0] - d.y[2]));
    const float L12 = std::hypot(static_cast<float>(d.x[1] - d.x[2]), static_cast<float>(d.y[1] - d.y[2]));
    const float max_edge = std::max(L01, std::max(L02, L12));
    if (max_edge < 1.0f) continue;
    const float step = 1.0f / max_edge;
    const float x1 = static_cast<float>(d.x[0]), x2 = static_cast<float>(d.x[1]), x3 = static_cast<float>(d.x[2]);
    const float y1 = static_cast<float>(d.y[0]), y2 = static_cast<float>(d.y[1]), y3 = static_cast<float>(d.y[2]);
    for (float p = 0.0f; p < 1.0f; p += step) {
      for (float q = 0.0f; q < 1.0f - p; q += step) {
        const float r = 1.0f - p - q;
        const int px = static_cast<int>(p * x1 + q * x2 + r * x3);
        const int py = static_cast<int>(p * y1 + q * y2 + r * y3);
        if (px >= 0 && px < w && py >= 0 && py < h) owner[static_cast<size_t>(py) * w + px] = static_cast<int>(ti) + 1;
      }
    }
  }
  // 6. planes of the pixels whose depth is in range (:1064-1090)
  for (int y = 0; y < h; ++y) {
    for (int x = 0; x < w; ++x) {
      const size_t idx = static_cast<size_t>(y) * w + x;
      if (owner[idx] <= 0) continue;
      const float* pl = td[static_cast<size_t>(owner[idx] - 1)].plane;
      const float rx = (static_cast<float>(x) - cx) / fx, ry = (static_cast<float>(y) - cy) / fy;
      const float denom = pl[0] * rx + pl[1] * ry + pl[2] * 1.0f;
      const float dep = (std::abs(denom) > 1e-6f) ? -pl[3] / denom : 0.0f;
      if (dep >= depth_min && dep <= depth_max) {
        std::memcpy(&prior_planes[4 * idx], pl, 4 * sizeof(float));
        masks[idx] = 1u;
      }
    }
  }
  return PM_OK;
}

// SelectNeighbors (dense_helpers.cc:331-568).  Super-point p has centroid
// coords[3p..3p+2], track_counts[p] tracks and is seen by the shots
// vis_shots[vis_offsets[p] .. vis_offsets[p+1]) (indices into origins).
// best: num_shots x (1 + num_neighbors), self first then the top-scoring
// shots, -1 padded; all_out: num_shots x num_shots, best then the other
// shots sharing > min_point_all tracks, -1 padded; counts per shot.
int pm_select_neighbors(int num_shots, const double* origins, int num_points, const double* coords,
                        const int* vis_offsets, const int* vis_shots, const int* track_counts, int num_neighbors,
                        int min_point_best, int min_point_all, double theta_min_deg, double theta_max_deg, int* best,
                        int* best_counts, int* all_out, int* all_counts) {
  if (num_shots < 0 || num_points < 0 || num_neighbors < 0 || !best || !best_counts || !all_out || !all_counts ||
      (num_shots > 0 && !origins) || (num_points > 0 && (!coords || !vis_offsets || !vis_shots || !track_counts)))
    return api_fail(PM_E_INVALID, "pm_select_neighbors: bad arguments");
  const int N = num_shots;
  const double kPi = 3.14159265358979323846;
  const double theta_min = theta_min_deg * kPi / 180.0, theta_max = theta_max_deg * kPi / 180.0;
  const double cos_lo = std::cos(theta_max), cos_hi = std::cos(theta_min);
  struct PairScore {
    long long score_best = 0, count = 0;
    double sum_cos = 0.0;
  };
  std::vector<std::unordered_map<int, PairScore>> per(static_cast<size_t>(N));
  std::vector<int> cams;
  std::vector<std::array<double, 4>> cv;
  for (int p = 0; p < num_points; ++p) {  // pair scores (:390-445)
    cams.clear();
    for (int k = vis_offsets[p]; k < vis_offsets[p + 1]; ++k)
      if (vis_shots[k] >= 0 && vis_shots[k] < N) cams.push_back(vis_shots[k]);
    const int nc = static_cast<int>(cams.size());
    if (nc < 2) continue;
    const double* P = coords + 3 * p;
    const int nt = track_counts[p];
    cv.assign(static_cast<size_t>(nc), {0, 0, 0, 0});
    for (int c = 0; c < nc; ++c) {
      const double* o = origins + 3 * cams[c];
      cv[c] = {o[0] - P[0], o[1] - P[1], o[2] - P[2], 0.0};
      cv[c][3] = std::sqrt(cv[c][0] * cv[c][0] + cv[c][1] * cv[c][1] + cv[c][2] * cv[c][2]);
    }
    for (int ci = 0; ci < nc; ++ci) {
      if (cv[ci][3] < 1e-12) continue;
      for (int cj = ci + 1; cj < nc; ++cj) {
        if (cv[cj][3] < 1e-12) continue;
        const double dot = cv[ci][0] * cv[cj][0] + cv[ci][1] * cv[cj][1] + cv[ci][2] * cv[cj][2];
        const double c = std::min(1.0, std::max(-1.0, dot / (cv[ci][3] * cv[cj][3])));
        const bool good = c > cos_lo && c < cos_hi;
        for (int side = 0; side < 2; ++side) {
          PairScore& ps = per[side ? cams[cj] : cams[ci]][side ? cams[ci] : cams[cj]];
          ps.count += nt;
          ps.sum_cos += c * nt;
          ps.score_best += good ? nt : 0;
        }
      }
    }
  }
  for (int i = 0; i < N; ++i) {  // per-shot selection (:470-550)
    struct Cand {
      int idx;
      long long score;
    };
    std::vector<Cand> ns_best, ns_all;
    for (const auto& kv : per[i]) {
      const PairScore& ps = kv.second;
      if (ps.count == 0) continue;
      if (ps.count > min_point_all) ns_all.push_back({kv.first, ps.count});
      if (ps.count <= min_point_best) continue;
      const double avg = std::acos(std::min(1.0, std::max(-1.0, ps.sum_cos / static_cast<double>(ps.count))));
      if (avg < theta_min || avg > theta_max) continue;
      if (ps.score_best > min_point_best) ns_best.push_back({kv.first, ps.score_best});
    }
    std::sort(ns_best.begin(), ns_best.end(),
              [](const Cand& a, const Cand& b) { return a.score != b.score ? a.score > b.score : a.idx < b.idx; });
    std::sort(ns_all.begin(), ns_all.end(), [](const Cand& a, const Cand& b) { return a.idx < b.idx; });
    int* b = best + static_cast<size_t>(i) * (1 + num_neighbors);
    for (int k = 0; k <= num_neighbors; ++k) b[k] = -1;
    int nb = 0;
    b[nb++] = i;
    for (int j = 0; j < std::min(num_neighbors, static_cast<int>(ns_best.size())); ++j) b[nb++] = ns_best[j].idx;
    best_counts[i] = nb;
    std::vector<char> in_best(static_cast<size_t>(N), 0);  // best_set (:531-537)
    in_best[i] = 1;
    int bs = 1;
    for (const Cand& c : ns_best)
      if (bs <= num_neighbors) { in_best[c.idx] = 1; ++bs; }
    int* a = all_out + static_cast<size_t>(i) * N;
    for (int k = 0; k < N; ++k) a[k] = -1;
    int na = 0;
    for (int k = 0; k < nb; ++k) a[na++] = b[k];
    for (const Cand& c : ns_all)
      if (!in_best[c.idx] && na < N) a[na++] = c.idx;
    all_counts[i] = na;
  }
  return PM_OK;
}

}  // extern "C"
