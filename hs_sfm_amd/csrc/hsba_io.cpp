// hsba_io.cpp — problems from files: the hs_sfm text formats and the
// "Bundle Adjustment in the Large" (BAL) format, as host-side scenes that
// feed hsba_problem_create / hsba_lm_solve, and writing results back.
//
// hs_sfm formats (include/hs_sfm/sfm_file_io/*_loader.hpp): whitespace
// separated, a count first.  The camera_shared problem is assembled the way
// BundleAdjustmentOptimizor::operator() does it
// (sfm_pipeline/bundle_adjustment_optimizor.hpp:59-203):
//   * a track enters the bundle when track_point_map holds it; its views are
//     kept when the image has an extrinsic and the view is not a blunder; a
//     track left with no view is dropped (:71-101);
//   * feature_maps[k] = (extrinsic id, bundle track index), keys in track
//     order (:104-117);
//   * image_camera_map[extrinsic e] = image_intrinsic_map[image of e], every
//     extrinsic must belong to an image (:120-134);
//   * x = [points of the bundle tracks | per extrinsic (angle-axis, -R C) |
//     per camera [k1 k2 k3 d1 d2 f skew px py pr]] (:146-187), y = the keys
//     (:188-203).
// Writing back follows :228-257: points by bundle_point_map, position
// C = -R^T t, intrinsics from the camera block.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "hsba_host.h"

namespace {

constexpr int kNI = HSBA_NUM_INTRINSICS;

// ---- rotations (angle-axis <-> matrix, row-major 3x3) --------------------
void aa_to_matrix(const double r[3], double R[9]) {
  const double t2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  if (t2 < 1e-30) {
    // first order: R = I + [r]x
    R[0] = 1; R[1] = -r[2]; R[2] = r[1];
    R[3] = r[2]; R[4] = 1; R[5] = -r[0];
    R[6] = -r[1]; R[7] = r[0]; R[8] = 1;
    return;
  }
  const double t = std::sqrt(t2), c = std::cos(t), s = std::sin(t), v = 1.0 - c;
  const double k[3] = {r[0] / t, r[1] / t, r[2] / t};
  R[0] = c + k[0] * k[0] * v;        R[1] = k[0] * k[1] * v - k[2] * s; R[2] = k[0] * k[2] * v + k[1] * s;
  R[3] = k[1] * k[0] * v + k[2] * s; R[4] = c + k[1] * k[1] * v;        R[5] = k[1] * k[2] * v - k[0] * s;
  R[6] = k[2] * k[0] * v - k[1] * s; R[7] = k[2] * k[1] * v + k[0] * s; R[8] = c + k[2] * k[2] * v;
}

// Logarithm of a rotation matrix; near theta = pi the axis comes from the
// symmetric part (R + R^T)/2 = cos(t) I + (1 - cos(t)) k k^T.
void matrix_to_aa(const double R[9], double r[3]) {
  const double tr = R[0] + R[4] + R[8];
  const double w[3] = {R[7] - R[5], R[2] - R[6], R[3] - R[1]};  // 2 sin(t) k
  const double s2 = 0.5 * std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);  // sin(t)
  const double c = std::max(-1.0, std::min(1.0, 0.5 * (tr - 1.0)));
  const double t = std::atan2(s2, c);
  if (s2 > 1e-7) {
    const double f = t / (2.0 * s2);
    r[0] = f * w[0]; r[1] = f * w[1]; r[2] = f * w[2];
    return;
  }
  if (c > 0) {  // t ~ 0: r ~ w / 2
    r[0] = 0.5 * w[0]; r[1] = 0.5 * w[1]; r[2] = 0.5 * w[2];
    return;
  }
  // t ~ pi: k k^T = (S - cos t I) / (1 - cos t), take the largest diagonal
  const double v = 1.0 - c;
  double kk[3] = {std::sqrt(std::max(0.0, (R[0] - c) / v)), std::sqrt(std::max(0.0, (R[4] - c) / v)),
                  std::sqrt(std::max(0.0, (R[8] - c) / v))};
  int m = 0;
  if (kk[1] > kk[m]) m = 1;
  if (kk[2] > kk[m]) m = 2;
  const double S[3][3] = {{R[0], 0.5 * (R[1] + R[3]), 0.5 * (R[2] + R[6])},
                          {0.5 * (R[1] + R[3]), R[4], 0.5 * (R[5] + R[7])},
                          {0.5 * (R[2] + R[6]), 0.5 * (R[5] + R[7]), R[8]}};
  double k[3];
  for (int j = 0; j < 3; ++j) k[j] = (j == m) ? kk[m] : S[m][j] / (v * kk[m]);
  // sign from the antisymmetric part (sin t k), when it is resolvable
  if (w[0] * k[0] + w[1] * k[1] + w[2] * k[2] < 0) { k[0] = -k[0]; k[1] = -k[1]; k[2] = -k[2]; }
  const double n = std::sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
  for (int j = 0; j < 3; ++j) r[j] = t * k[j] / n;
}

template <typename X>
bool read_count(std::ifstream &f, X &n) {
  long long v = -1;
  if (!(f >> v) || v < 0) return false;
  // every counted entry takes at least one byte of the file, so a count
  // beyond the file's size is malformed (and would only exhaust memory)
  const std::streampos here = f.tellg();
  f.seekg(0, std::ios::end);
  const long long size = (long long)f.tellg();
  f.seekg(here);
  if (!f || size < 0 || v > size) return false;
  n = (X)v;
  return true;
}

// ObjectIndexMap text form: n, then per object 0 (none) or id + 1
bool load_index_map(const char *path, std::vector<int64_t> &m) {
  std::ifstream f(path);
  if (!f.is_open()) return false;
  int64_t n;
  if (!read_count(f, n)) return false;
  m.assign(n, -1);
  for (int64_t i = 0; i < n; ++i) {
    long long id;
    if (!(f >> id) || id < 0) return false;
    m[i] = id == 0 ? -1 : id - 1;
  }
  return true;
}

bool write_text(const char *path, const std::string &s) {
  FILE *f = std::fopen(path, "w");
  if (!f) return false;
  const bool ok = std::fwrite(s.data(), 1, s.size(), f) == s.size();
  return std::fclose(f) == 0 && ok;
}

void put(std::string &out, double v, const char *sep) {
  char b[40];
  std::snprintf(b, sizeof(b), "%.17g%s", v, sep);
  out += b;
}

uint32_t enabled_bits(uint32_t comp) {
  uint32_t en = 0;
  if (comp & HSBA_COMPUTE_RADIAL_DISTORTION) en |= 0x7u;
  if (comp & HSBA_COMPUTE_DECENTERING_DISTORTION) en |= 0x18u;
  if (comp & HSBA_COMPUTE_INTRINSIC_PARAMS) en |= 0x3E0u;
  return en;
}

}  // namespace

struct hsba_scene {
  int kind = 0;  // 0 hs_sfm files, 1 BAL
  hsba_problem_desc desc;
  std::vector<int32_t> fm, icm;
  std::vector<double> intr;  // [C][10]: held values
  std::vector<double> x0, y;
  std::vector<int32_t> free_idx;  // free intrinsic indices, camera-block order
  // input entities, for writing back
  std::vector<double> points_in;      // [P_file][3]
  std::vector<double> extr_in;        // [E][6] angle-axis, centre (hs_sfm) / [C][9] camera (BAL)
  std::vector<double> intr_in;        // [C][10]
  std::vector<int64_t> bundle_point;  // bundle point -> file point
  std::vector<int32_t> bal_obs_cam, bal_obs_pt;
  std::vector<double> bal_obs;        // as read
};

extern "C" {

int hsba_scene_load_sfm(const hsba_sfm_files *files, uint32_t comp, uint32_t free_mask, hsba_scene **out) {
  return hsba::guarded([&]() -> int {
    if (!files || !out || !files->tracks || !files->intrinsics || !files->extrinsics || !files->points ||
        !files->image_intrinsic_map || !files->image_extrinsic_map || !files->track_point_map ||
        (files->num_keysets > 0 && !files->keysets) || files->num_keysets < 0)
      return HSBA_ERR_INVALID;
    *out = nullptr;
    std::unique_ptr<hsba_scene> sc(new hsba_scene());
    sc->kind = 0;
    // ---- files
    struct View { int64_t image, key; int blunder; };
    std::vector<std::vector<View>> tracks;
    {
      std::ifstream f(files->tracks);
      int64_t n;
      if (!f.is_open() || !read_count(f, n)) return HSBA_ERR_INVALID;
      tracks.resize(n);
      for (int64_t t = 0; t < n; ++t) {
        int64_t nv;
        if (!read_count(f, nv)) return HSBA_ERR_INVALID;
        tracks[t].resize(nv);
        for (int64_t j = 0; j < nv; ++j) {
          long long im, key;
          int flag;
          if (!(f >> im >> key >> flag) || im < 0 || key < 0) return HSBA_ERR_INVALID;
          tracks[t][j] = {im, key, flag};
        }
        // the reference keeps one flag per (track, image): the last one read
        // wins (tracks_loader.hpp:48-66)
        for (int64_t j = 0; j < nv; ++j)
          for (int64_t j2 = j + 1; j2 < nv; ++j2)
            if (tracks[t][j2].image == tracks[t][j].image) tracks[t][j].blunder = tracks[t][j2].blunder;
      }
    }
    std::vector<std::vector<double>> keysets(files->num_keysets);
    for (int64_t i = 0; i < files->num_keysets; ++i) {
      if (!files->keysets[i]) continue;  // an image without keys
      std::ifstream f(files->keysets[i]);
      int64_t n;
      if (!f.is_open() || !read_count(f, n)) return HSBA_ERR_INVALID;
      keysets[i].resize(2 * n);
      for (int64_t k = 0; k < 2 * n; ++k)
        if (!(f >> keysets[i][k])) return HSBA_ERR_INVALID;
    }
    int64_t C = 0, E = 0, Pf = 0;
    {
      std::ifstream f(files->intrinsics);
      if (!f.is_open() || !read_count(f, C)) return HSBA_ERR_INVALID;
      sc->intr_in.assign(C * kNI, 0.0);
      for (int64_t c = 0; c < C; ++c) {
        double *q = &sc->intr_in[c * kNI];
        if (!(f >> q[HSBA_FOCAL] >> q[HSBA_SKEW] >> q[HSBA_PRINCIPAL_X] >> q[HSBA_PRINCIPAL_Y] >>
              q[HSBA_PIXEL_RATIO]))
          return HSBA_ERR_INVALID;
      }
    }
    {
      std::ifstream f(files->extrinsics);
      if (!f.is_open() || !read_count(f, E)) return HSBA_ERR_INVALID;
      sc->extr_in.resize(E * 6);
      for (int64_t k = 0; k < 6 * E; ++k)
        if (!(f >> sc->extr_in[k])) return HSBA_ERR_INVALID;
    }
    {
      std::ifstream f(files->points);
      if (!f.is_open() || !read_count(f, Pf)) return HSBA_ERR_INVALID;
      sc->points_in.resize(Pf * 3);
      for (int64_t k = 0; k < 3 * Pf; ++k)
        if (!(f >> sc->points_in[k])) return HSBA_ERR_INVALID;
    }
    std::vector<int64_t> img_intr, img_extr, track_pt;
    if (!load_index_map(files->image_intrinsic_map, img_intr) || !load_index_map(files->image_extrinsic_map, img_extr) ||
        !load_index_map(files->track_point_map, track_pt))
      return HSBA_ERR_INVALID;
    // ---- bundle tracks (:71-101)
    std::vector<int32_t> &fm = sc->fm;
    std::vector<double> &y = sc->y;
    int64_t P = 0;
    for (size_t t = 0; t < tracks.size(); ++t) {
      if (t >= track_pt.size() || track_pt[t] < 0) continue;
      const int64_t before = (int64_t)fm.size();
      for (const View &v : tracks[t]) {
        if (v.image >= (int64_t)img_extr.size() || img_extr[v.image] < 0 || v.blunder == 1) continue;
        if (img_extr[v.image] >= E) return HSBA_ERR_INVALID;
        if (v.image >= (int64_t)keysets.size() || 2 * v.key + 1 >= (int64_t)keysets[v.image].size())
          return HSBA_ERR_INVALID;
        fm.push_back((int32_t)img_extr[v.image]);
        fm.push_back((int32_t)P);
        y.push_back(keysets[v.image][2 * v.key]);
        y.push_back(keysets[v.image][2 * v.key + 1]);
      }
      if ((int64_t)fm.size() == before) continue;
      if (track_pt[t] >= Pf) return HSBA_ERR_INVALID;
      sc->bundle_point.push_back(track_pt[t]);
      ++P;
    }
    // ---- image -> camera map over extrinsics (:120-134)
    std::vector<int64_t> extr_img(E, -1);
    for (size_t i = 0; i < img_extr.size(); ++i)
      if (img_extr[i] >= 0 && img_extr[i] < E) extr_img[img_extr[i]] = (int64_t)i;
    sc->icm.resize(E);
    for (int64_t e = 0; e < E; ++e) {
      const int64_t im = extr_img[e];
      if (im < 0 || im >= (int64_t)img_intr.size() || img_intr[im] < 0 || img_intr[im] >= C) return HSBA_ERR_INVALID;
      sc->icm[e] = (int32_t)img_intr[im];
    }
    // ---- x (:146-187)
    const uint32_t en = enabled_bits(comp & 7u);
    const uint32_t fb = free_mask ? (free_mask & 0x3FFu) : en;
    if (fb & ~en) return HSBA_ERR_INVALID;
    for (int j = 0; j < kNI; ++j)
      if (fb & (1u << j)) sc->free_idx.push_back(j);
    const int64_t d = (int64_t)sc->free_idx.size();
    std::vector<double> &x = sc->x0;
    x.reserve(3 * P + 6 * E + d * C);
    for (int64_t i = 0; i < P; ++i)
      for (int j = 0; j < 3; ++j) x.push_back(sc->points_in[3 * sc->bundle_point[i] + j]);
    for (int64_t e = 0; e < E; ++e) {
      const double *ex = &sc->extr_in[6 * e];
      double R[9];
      aa_to_matrix(ex, R);
      for (int j = 0; j < 3; ++j) x.push_back(ex[j]);
      for (int a = 0; a < 3; ++a) x.push_back(-(R[3 * a] * ex[3] + R[3 * a + 1] * ex[4] + R[3 * a + 2] * ex[5]));
    }
    for (int64_t c = 0; c < C; ++c)
      for (int64_t q = 0; q < d; ++q) x.push_back(sc->intr_in[c * kNI + sc->free_idx[q]]);
    sc->intr = sc->intr_in;
    // ---- description
    hsba_problem_desc &ds = sc->desc;
    std::memset(&ds, 0, sizeof(ds));
    ds.num_points = P;
    ds.num_images = E;
    ds.num_cameras = C;
    ds.num_keys = (int64_t)fm.size() / 2;
    ds.feature_maps = fm.data();
    ds.image_camera_map = sc->icm.data();
    ds.intrinsic_computations_mask = comp & 7u;
    ds.intrinsic_free_mask = free_mask & 0x3FFu;
    ds.intrinsic_defaults = sc->intr.data();
    *out = sc.release();
    return HSBA_OK;
  });
}

int hsba_scene_load_bal(const char *path, hsba_scene **out) {
  return hsba::guarded([&]() -> int {
    if (!path || !out) return HSBA_ERR_INVALID;
    *out = nullptr;
    std::ifstream f(path);
    int64_t C, P, N;
    if (!f.is_open() || !read_count(f, C) || !read_count(f, P) || !read_count(f, N)) return HSBA_ERR_INVALID;
    std::unique_ptr<hsba_scene> sc(new hsba_scene());
    sc->kind = 1;
    sc->bal_obs_cam.resize(N);
    sc->bal_obs_pt.resize(N);
    sc->bal_obs.resize(2 * N);
    sc->fm.resize(2 * N);
    sc->y.resize(2 * N);
    for (int64_t k = 0; k < N; ++k) {
      long long c, p;
      double u, v;
      if (!(f >> c >> p >> u >> v) || c < 0 || c >= C || p < 0 || p >= P) return HSBA_ERR_INVALID;
      sc->bal_obs_cam[k] = (int32_t)c;
      sc->bal_obs_pt[k] = (int32_t)p;
      sc->bal_obs[2 * k] = u;
      sc->bal_obs[2 * k + 1] = v;
      sc->fm[2 * k] = (int32_t)c;
      sc->fm[2 * k + 1] = (int32_t)p;
      // camera frame flipped by diag(1, -1, -1): the normalised v changes sign
      sc->y[2 * k] = u;
      sc->y[2 * k + 1] = -v;
    }
    sc->extr_in.resize(9 * C);
    for (int64_t k = 0; k < 9 * C; ++k)
      if (!(f >> sc->extr_in[k])) return HSBA_ERR_INVALID;
    sc->points_in.resize(3 * P);
    for (int64_t k = 0; k < 3 * P; ++k)
      if (!(f >> sc->points_in[k])) return HSBA_ERR_INVALID;
    // camera_shared: one image and one camera group per BAL camera; free
    // intrinsics {k1, k2, f} in camera-block order, px = py = skew = k3 = d = 0, pr = 1
    sc->free_idx = {HSBA_K1, HSBA_K2, HSBA_FOCAL};
    sc->icm.resize(C);
    sc->intr.assign(C * kNI, 0.0);
    std::vector<double> &x = sc->x0;
    x.reserve(3 * P + 9 * C);
    for (int64_t k = 0; k < 3 * P; ++k) x.push_back(sc->points_in[k]);
    for (int64_t c = 0; c < C; ++c) {
      const double *cam = &sc->extr_in[9 * c];
      double R[9], Rf[9], r[3];
      aa_to_matrix(cam, R);
      for (int j = 0; j < 3; ++j) {
        Rf[j] = R[j];
        Rf[3 + j] = -R[3 + j];
        Rf[6 + j] = -R[6 + j];
      }
      matrix_to_aa(Rf, r);
      x.push_back(r[0]); x.push_back(r[1]); x.push_back(r[2]);
      x.push_back(cam[3]); x.push_back(-cam[4]); x.push_back(-cam[5]);
      sc->icm[c] = (int32_t)c;
      double *q = &sc->intr[c * kNI];
      q[HSBA_FOCAL] = cam[6];
      q[HSBA_K1] = cam[7];
      q[HSBA_K2] = cam[8];
      q[HSBA_PIXEL_RATIO] = 1.0;
    }
    for (int64_t c = 0; c < C; ++c) {
      const double *cam = &sc->extr_in[9 * c];
      x.push_back(cam[7]); x.push_back(cam[8]); x.push_back(cam[6]);  // k1 k2 f
    }
    hsba_problem_desc &ds = sc->desc;
    std::memset(&ds, 0, sizeof(ds));
    ds.num_points = P;
    ds.num_images = C;
    ds.num_cameras = C;
    ds.num_keys = N;
    ds.feature_maps = sc->fm.data();
    ds.image_camera_map = sc->icm.data();
    ds.intrinsic_computations_mask = HSBA_COMPUTE_RADIAL_DISTORTION | HSBA_COMPUTE_INTRINSIC_PARAMS;
    ds.intrinsic_free_mask = (1u << HSBA_K1) | (1u << HSBA_K2) | (1u << HSBA_FOCAL);
    ds.intrinsic_defaults = sc->intr.data();
    *out = sc.release();
    return HSBA_OK;
  });
}

int hsba_scene_get(const hsba_scene *sc, hsba_scene_view *view) {
  return hsba::guarded([&]() -> int {
    if (!sc || !view) return HSBA_ERR_INVALID;
    view->desc = sc->desc;
    view->x0 = sc->x0.data();
    view->y_obs = sc->y.data();
    view->x_size = (int64_t)sc->x0.size();
    view->y_size = (int64_t)sc->y.size();
    return HSBA_OK;
  });
}

int hsba_scene_store_sfm(hsba_scene *sc, const double *x, const char *intrinsics, const char *extrinsics,
                         const char *points) {
  return hsba::guarded([&]() -> int {
    if (!sc || !x || sc->kind != 0) return HSBA_ERR_INVALID;
    const int64_t P = sc->desc.num_points, E = sc->desc.num_images, C = sc->desc.num_cameras;
    const int64_t d = (int64_t)sc->free_idx.size();
    if (points) {
      std::vector<double> pts = sc->points_in;
      for (int64_t i = 0; i < P; ++i)
        for (int j = 0; j < 3; ++j) pts[3 * sc->bundle_point[i] + j] = x[3 * i + j];
      std::string s = std::to_string(pts.size() / 3) + "\n";
      for (size_t i = 0; i < pts.size(); i += 3) {
        put(s, pts[i], " "); put(s, pts[i + 1], " "); put(s, pts[i + 2], "\n");
      }
      if (!write_text(points, s)) return HSBA_ERR_GENERIC;
    }
    if (extrinsics) {
      std::string s = std::to_string(E) + "\n";
      for (int64_t e = 0; e < E; ++e) {
        const double *im = x + 3 * P + 6 * e;
        double R[9];
        aa_to_matrix(im, R);
        double c[3];
        for (int a = 0; a < 3; ++a) c[a] = -(R[a] * im[3] + R[3 + a] * im[4] + R[6 + a] * im[5]);  // -R^T t
        put(s, im[0], " "); put(s, im[1], " "); put(s, im[2], " ");
        put(s, c[0], " "); put(s, c[1], " "); put(s, c[2], "\n");
      }
      if (!write_text(extrinsics, s)) return HSBA_ERR_GENERIC;
    }
    if (intrinsics) {
      std::string s = std::to_string(C) + "\n";
      for (int64_t c = 0; c < C; ++c) {
        double q[kNI];
        for (int j = 0; j < kNI; ++j) q[j] = sc->intr[c * kNI + j];
        for (int64_t k = 0; k < d; ++k) q[sc->free_idx[k]] = x[3 * P + 6 * E + d * c + k];
        // intrinsic_params_set_saver.hpp writes f, skew, px, py, pr, one per line
        put(s, q[HSBA_FOCAL], "\n"); put(s, q[HSBA_SKEW], "\n"); put(s, q[HSBA_PRINCIPAL_X], "\n");
        put(s, q[HSBA_PRINCIPAL_Y], "\n"); put(s, q[HSBA_PIXEL_RATIO], "\n");
      }
      if (!write_text(intrinsics, s)) return HSBA_ERR_GENERIC;
    }
    return HSBA_OK;
  });
}

int hsba_scene_store_bal(hsba_scene *sc, const double *x, const char *path) {
  return hsba::guarded([&]() -> int {
    if (!sc || !x || !path || sc->kind != 1) return HSBA_ERR_INVALID;
    const int64_t P = sc->desc.num_points, C = sc->desc.num_cameras, N = sc->desc.num_keys;
    std::string s = std::to_string(C) + " " + std::to_string(P) + " " + std::to_string(N) + "\n";
    for (int64_t k = 0; k < N; ++k) {
      s += std::to_string(sc->bal_obs_cam[k]) + " " + std::to_string(sc->bal_obs_pt[k]) + " ";
      put(s, sc->bal_obs[2 * k], " ");
      put(s, sc->bal_obs[2 * k + 1], "\n");
    }
    for (int64_t c = 0; c < C; ++c) {
      const double *im = x + 3 * P + 6 * c;
      const double *q = x + 3 * P + 6 * C + 3 * c;  // k1 k2 f
      double R[9], Rb[9], r[3];
      aa_to_matrix(im, R);
      for (int j = 0; j < 3; ++j) {
        Rb[j] = R[j];
        Rb[3 + j] = -R[3 + j];
        Rb[6 + j] = -R[6 + j];
      }
      matrix_to_aa(Rb, r);
      const double cam[9] = {r[0], r[1], r[2], im[3], -im[4], -im[5], q[2], q[0], q[1]};
      for (int j = 0; j < 9; ++j) put(s, cam[j], "\n");
    }
    for (int64_t k = 0; k < 3 * P; ++k) put(s, x[k], "\n");
    return write_text(path, s) ? HSBA_OK : HSBA_ERR_GENERIC;
  });
}

int hsba_scene_destroy(hsba_scene *sc) {
  return hsba::guarded([&]() -> int {
    delete sc;
    return HSBA_OK;
  });
}

}  // extern "C"
