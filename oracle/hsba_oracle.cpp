// hsba_oracle.cpp — FP64 CPU restatement of hs_sfm's camera_shared bundle
// adjustment path.  TEST INFRASTRUCTURE ONLY: the parity checker for the HIP
// library and the CPU baseline of the benchmark.  Nothing in the product path
// links or calls this file.
//
// Every function cites the reference file:line it restates (paths relative to
// the hs_sfm repository root; BA/ = include/hs_sfm/bundle_adjustment/).
//
// Deviations from the reference, by design:
//   * theta == 0 in the Rodrigues derivative: the reference divides by theta
//     (BA/camera_shared_analytical_jacobian_matrix_calculator.hpp:354-380) and
//     returns NaN; this restatement uses the Taylor limit below theta = 1e-4.
//   * SimplicialLDLT (BA/camera_shared_normal_equation_solver.hpp:453) is
//     replaced by a dense LDL^T without pivoting; it fails on a zero pivot
//     exactly like SimplicialLDLT's info() != Success.
//   * The camera-rhs loop's uninitialised point_id
//     (BA/camera_shared_normal_equation_solver.hpp:429) is read as point_id = 0.
//   * The LM loop of the absent hs_optimizor module is restated from
//     Madsen, Nielsen & Tingleff (2004), "Methods for non-linear least squares
//     problems", Algorithm 3.16, with the reference's cost F = r^T W r (no 1/2).
//
// Parity status per function (checks live in tests/test_oracle_*.py):
//   project / evaluate   pinned: closed form, (C,R) identity, theta = 0 branch
//   jacobian             pinned: complex step + central FD, reference's fixed input
//   build_normal         pinned: dense J^T W J / J^T W r brute force, 7 FixMasks
//   reduced system/solve pinned: literal reference loops, (N + mu I) delta = g
//   lm                   pinned by invariants only (the hs_optimizor trajectory
//                        itself is parity unpinned: the module is not in the tree)

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <set>
#include <vector>

#include "../include/hsba.h"

namespace {

constexpr int kFull = HSBA_NUM_INTRINSICS;

struct Constraint {
  int kind;        // 0 point, 1 image, 2 camera
  int64_t entity;  // point / image / camera id
  int param;       // index inside the entity's x block
  int64_t x_index; // absolute x index
};

struct OProblem {
  int64_t P = 0, I = 0, C = 0, N = 0;
  std::vector<int32_t> key_img, key_pt, icm;
  uint32_t comp_mask = 0, fix = 0;
  uint32_t enabled = 0, free_bits = 0;
  int d = 0;
  int free_idx[kFull] = {0};
  std::vector<double> cam_base;  // [C][10]
  std::vector<double> fix_points, fix_images, fix_cameras;
  std::vector<Constraint> cons;
  int64_t n_point_cons = 0, n_image_cons = 0, n_camera_cons = 0;
  // Full 3x3 weight blocks of the point constraints ([records][3][3], each
  // record constraining all three coordinates), as the GCP blocks of
  // BAGCPConstrainedYCovarianceInverse (BA/ba_gcp_constrained_y_covariance_inverse.hpp:14-23).
  // Empty: diagonal per-row weights.
  std::vector<double> pblocks;
  // Robust loss on each key's s = r^T Lambda r (0 none, 1 Huber with scale
  // loss_a): cost rho(s), normal equations reweighted by rho'(s).  The
  // reference's in-house path has none; its production caller uses
  // ceres::HuberLoss(1.0) (BA/camera_shared_ceres_optimizor.hpp:253).
  int loss = 0;
  double loss_a = 1.0;
  int64_t x_pt = 0, x_img = 0, x_cam = 0, x_size = 0, y_size = 0;
  bool pts_free = true, imgs_free = true, cams_free = true;

  // last build (builder outputs, camera_shared_normal_matrix.hpp layout)
  std::vector<double> V, Ui, Uc, Wpi, Wpc, Wic, g;
  std::vector<int64_t> pt_begin, pt_keys;  // keys grouped by point, key order
  std::vector<int64_t> img_nkeys, cam_nkeys;
  bool built = false;
};

uint32_t enabled_bits(uint32_t comp) {
  uint32_t b = 0;
  if (comp & HSBA_COMPUTE_RADIAL_DISTORTION) b |= 0x7u;
  if (comp & HSBA_COMPUTE_DECENTERING_DISTORTION) b |= 0x18u;
  if (comp & HSBA_COMPUTE_INTRINSIC_PARAMS) b |= 0x3E0u;
  return b;
}

// Values of a parameter when its group is disabled: they make every stage an
// identity, exactly as the reference skips it
// (BA/camera_shared_vector_function.hpp:645-683).
double identity_value(int j) {
  return (j == HSBA_FOCAL || j == HSBA_PIXEL_RATIO) ? 1.0 : 0.0;
}

int setup(OProblem &o, const hsba_problem_desc *d) {
  if (!d || d->num_points < 0 || d->num_images < 0 || d->num_cameras < 0 ||
      d->num_keys < 0)
    return HSBA_ERR_INVALID;
  o.P = d->num_points; o.I = d->num_images; o.C = d->num_cameras; o.N = d->num_keys;
  o.comp_mask = d->intrinsic_computations_mask & 7u;
  o.fix = d->fix_mask & 7u;
  if (o.fix == 7u) return HSBA_ERR_INVALID;
  o.pts_free = !(o.fix & HSBA_FIX_POINTS);
  o.imgs_free = !(o.fix & HSBA_FIX_IMAGES);
  o.cams_free = !(o.fix & HSBA_FIX_CAMERAS);
  o.enabled = enabled_bits(o.comp_mask);
  o.free_bits = d->intrinsic_free_mask ? d->intrinsic_free_mask : o.enabled;
  if (o.free_bits & ~o.enabled) return HSBA_ERR_INVALID;
  o.d = 0;
  for (int j = 0; j < kFull; ++j)
    if (o.free_bits & (1u << j)) o.free_idx[o.d++] = j;
  o.key_img.resize(o.N); o.key_pt.resize(o.N); o.icm.resize(o.I);
  for (int64_t k = 0; k < o.N; ++k) {
    o.key_img[k] = d->feature_maps[2 * k];
    o.key_pt[k] = d->feature_maps[2 * k + 1];
    if (o.key_img[k] < 0 || o.key_img[k] >= o.I || o.key_pt[k] < 0 || o.key_pt[k] >= o.P)
      return HSBA_ERR_INVALID;
  }
  for (int64_t i = 0; i < o.I; ++i) {
    o.icm[i] = d->image_camera_map[i];
    if (o.icm[i] < 0 || o.icm[i] >= o.C) return HSBA_ERR_INVALID;
  }
  o.cam_base.assign(o.C * kFull, 0.0);
  for (int64_t c = 0; c < o.C; ++c)
    for (int j = 0; j < kFull; ++j) {
      double v = identity_value(j);
      if ((o.enabled >> j) & 1u)
        v = d->intrinsic_defaults ? d->intrinsic_defaults[c * kFull + j] : identity_value(j);
      o.cam_base[c * kFull + j] = v;
    }
  if (!o.pts_free) o.fix_points.assign(d->fix_points, d->fix_points + 3 * o.P);
  if (!o.imgs_free) o.fix_images.assign(d->fix_images, d->fix_images + 6 * o.I);
  if (!o.cams_free) o.fix_cameras.assign(d->fix_cameras, d->fix_cameras + o.d * o.C);
  // x layout, BA/camera_shared_vector_function.hpp:349-354, 427-430
  o.x_pt = 0;
  o.x_img = o.pts_free ? 3 * o.P : 0;
  o.x_cam = o.x_img + (o.imgs_free ? 6 * o.I : 0);
  o.x_size = o.x_cam + (o.cams_free ? o.d * o.C : 0);
  // constraint rows in y order, BA/camera_shared_vector_function.hpp:485-629
  // and BA/camera_shared_jacobian_matrix.hpp:391-511
  o.cons.clear();
  if (o.pts_free)
    for (int64_t q = 0; q < d->num_point_constraints; ++q) {
      const hsba_constraint &c = d->point_constraints[q];
      if (c.id < 0 || c.id >= o.P) return HSBA_ERR_INVALID;
      for (int b = 0; b < 3; ++b)
        if (c.mask & (1u << b)) o.cons.push_back({0, c.id, b, o.x_pt + 3 * c.id + b});
    }
  o.n_point_cons = (int64_t)o.cons.size();
  if (o.imgs_free)
    for (int64_t q = 0; q < d->num_image_constraints; ++q) {
      const hsba_constraint &c = d->image_constraints[q];
      if (c.id < 0 || c.id >= o.I) return HSBA_ERR_INVALID;
      if (c.mask & 1u)
        for (int b = 0; b < 3; ++b) o.cons.push_back({1, c.id, b, o.x_img + 6 * c.id + b});
      for (int b = 1; b < 4; ++b)
        if (c.mask & (1u << b)) o.cons.push_back({1, c.id, 2 + b, o.x_img + 6 * c.id + 2 + b});
    }
  o.n_image_cons = (int64_t)o.cons.size() - o.n_point_cons;
  if (o.cams_free)
    for (int64_t q = 0; q < d->num_camera_constraints; ++q) {
      const hsba_constraint &c = d->camera_constraints[q];
      if (c.id < 0 || c.id >= o.C) return HSBA_ERR_INVALID;
      // a constraint on a disabled group is an error, :549-557
      if ((c.mask & 0x3FFu) & ~o.enabled) return HSBA_ERR_INVALID;
      for (int j = 0; j < kFull; ++j) {
        if (!(c.mask & (1u << j))) continue;
        int local = -1;
        for (int q2 = 0; q2 < o.d; ++q2)
          if (o.free_idx[q2] == j) local = q2;
        if (local < 0) return HSBA_ERR_INVALID;  // held parameter
        o.cons.push_back({2, c.id, local, o.x_cam + o.d * c.id + local});
      }
    }
  o.n_camera_cons = (int64_t)o.cons.size() - o.n_point_cons - o.n_image_cons;
  o.y_size = 2 * o.N + (int64_t)o.cons.size();
  // keys grouped by point (stable, key order inside a point)
  o.pt_begin.assign(o.P + 1, 0);
  for (int64_t k = 0; k < o.N; ++k) o.pt_begin[o.key_pt[k] + 1]++;
  for (int64_t p = 0; p < o.P; ++p) o.pt_begin[p + 1] += o.pt_begin[p];
  o.pt_keys.resize(o.N);
  {
    std::vector<int64_t> fill(o.pt_begin.begin(), o.pt_begin.end() - 1);
    for (int64_t k = 0; k < o.N; ++k) o.pt_keys[fill[o.key_pt[k]]++] = k;
  }
  o.img_nkeys.assign(o.I, 0); o.cam_nkeys.assign(o.C, 0);
  for (int64_t k = 0; k < o.N; ++k) { o.img_nkeys[o.key_img[k]]++; o.cam_nkeys[o.icm[o.key_img[k]]]++; }
  return HSBA_OK;
}

template <typename T>
void get_point(const OProblem &o, int64_t p, const T *x, T out[3]) {
  for (int j = 0; j < 3; ++j)
    out[j] = o.pts_free ? x[o.x_pt + 3 * p + j] : T(o.fix_points[3 * p + j]);
}
template <typename T>
void get_image(const OProblem &o, int64_t i, const T *x, T out[6]) {
  for (int j = 0; j < 6; ++j)
    out[j] = o.imgs_free ? x[o.x_img + 6 * i + j] : T(o.fix_images[6 * i + j]);
}
template <typename T>
void get_camera(const OProblem &o, int64_t c, const T *x, T full[kFull]) {
  for (int j = 0; j < kFull; ++j) full[j] = T(o.cam_base[c * kFull + j]);
  for (int q = 0; q < o.d; ++q)
    full[o.free_idx[q]] = o.cams_free ? x[o.x_cam + o.d * c + q] : T(o.fix_cameras[o.d * c + q]);
}

// WorldPointToNormalizedKey, BA/camera_shared_vector_function.hpp:688-710.
// Templated so the complex-step check can run it on std::complex<double>.
template <typename T>
T tnorm(const T r[3]) { return std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]); }
double re(double v) { return v; }
double re(const std::complex<double> &v) { return v.real(); }

template <typename T>
void normalized_key(const T X[3], const T r[3], const T t[3], T n[2]) {
  T theta = tnorm(r);
  T y[3];
  if (re(theta) == 0.0) {
    for (int j = 0; j < 3; ++j) y[j] = X[j] + t[j];
  } else {
    T k[3] = {r[0] / theta, r[1] / theta, r[2] / theta};
    T ct = std::cos(theta), st = std::sin(theta);
    T kx[3] = {k[1] * X[2] - k[2] * X[1], k[2] * X[0] - k[0] * X[2], k[0] * X[1] - k[1] * X[0]};
    T kd = k[0] * X[0] + k[1] * X[1] + k[2] * X[2];
    for (int j = 0; j < 3; ++j) y[j] = ct * X[j] + st * kx[j] + (T(1) - ct) * kd * k[j] + t[j];
  }
  n[0] = y[0] / y[2];
  n[1] = y[1] / y[2];
}

// WorldPointToImageKey, BA/camera_shared_vector_function.hpp:634-686 with
// RadialDistortor (sfm_utility/radial_distortor.hpp:14-27), DecenteringDistortor
// (sfm_utility/decentering_distortor.hpp:14-24) and NormalizedKeyToImageKey (:731-744).
template <typename T>
void image_key(const T X[3], const T img[6], const T cam[kFull], T uv[2]) {
  T n[2];
  normalized_key(X, img, img + 3, n);
  T r2 = n[0] * n[0] + n[1] * n[1], r4 = r2 * r2, r6 = r2 * r4;
  T coeff = cam[HSBA_K1] * r2 + cam[HSBA_K2] * r4 + cam[HSBA_K3] * r6;
  T rd[2] = {coeff * n[0], coeff * n[1]};
  T d1 = cam[HSBA_D1], d2 = cam[HSBA_D2];
  T dd[2] = {T(2) * d1 * n[0] * n[1] + d2 * (r2 + T(2) * n[0] * n[0]),
             T(2) * d2 * n[0] * n[1] + d1 * (r2 + T(2) * n[1] * n[1])};
  T m[2] = {n[0] + (rd[0] + dd[0]), n[1] + (rd[1] + dd[1])};
  uv[0] = cam[HSBA_FOCAL] * m[0] + cam[HSBA_SKEW] * m[1] + cam[HSBA_PRINCIPAL_X];
  uv[1] = cam[HSBA_FOCAL] * cam[HSBA_PIXEL_RATIO] * m[1] + cam[HSBA_PRINCIPAL_Y];
}

template <typename T>
void project_key(const OProblem &o, int64_t k, const T *x, T uv[2]) {
  T X[3], img[6], cam[kFull];
  get_point(o, o.key_pt[k], x, X);
  get_image(o, o.key_img[k], x, img);
  get_camera(o, o.icm[o.key_img[k]], x, cam);
  image_key(X, img, cam, uv);
}

// CameraSharedVectorFunction::operator(), BA/camera_shared_vector_function.hpp:412-632
void evaluate(const OProblem &o, const double *x, double *y) {
  for (int64_t k = 0; k < o.N; ++k) project_key(o, k, x, y + 2 * k);
  for (size_t q = 0; q < o.cons.size(); ++q) y[2 * o.N + q] = x[o.cons[q].x_index];
}

// WorldPointToNormalizedPointDerivation,
// BA/camera_shared_analytical_jacobian_matrix_calculator.hpp:334-429.
// pe: d n / d (r, t) as [2][6]; pp: d n / d X as [2][3].
void normalized_derivative(const double X[3], const double r[3], const double t[3],
                           double pe[2][6], double pp[2][3]) {
  double theta2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  double theta = std::sqrt(theta2);
  double dot = r[0] * X[0] + r[1] * X[1] + r[2] * X[2];
  double ct = std::cos(theta), st = std::sin(theta);
  double s, c, ds, dc, minus_st_over_theta;
  if (theta < 1e-4) {  // Taylor limit (the reference has no guard)
    s = 1.0 - theta2 / 6.0;
    c = 0.5 - theta2 / 24.0;
    ds = -1.0 / 3.0 + theta2 / 30.0;    // (ct*theta - st) / theta^3
    dc = -1.0 / 12.0 + theta2 / 180.0;  // (st*theta - 2(1-ct)) / theta^4
    minus_st_over_theta = -s;
  } else {
    s = st / theta;
    c = (1.0 - ct) / theta2;
    ds = (ct * theta - st) / theta2 / theta;
    dc = (st * theta - 2.0 * (1.0 - ct)) / (theta2 * theta2);
    minus_st_over_theta = -st / theta;
  }
  double v[3] = {r[1] * X[2] - r[2] * X[1], r[2] * X[0] - r[0] * X[2], r[0] * X[1] - r[1] * X[0]};
  double y[3];
  for (int j = 0; j < 3; ++j) y[j] = ct * X[j] + s * v[j] + c * dot * r[j] + t[j];
  // dy/dr = X (-st/theta r)^T + v (ds r)^T + s dv/dr + u (dc r)^T + c du/dr
  double dvdr[3][3] = {{0, X[2], -X[1]}, {-X[2], 0, X[0]}, {X[1], -X[0], 0}};
  double dydr[3][3], dydX[3][3];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      double dudr = (a == b ? dot : 0.0) + r[a] * X[b];
      dydr[a][b] = X[a] * minus_st_over_theta * r[b] + v[a] * ds * r[b] + s * dvdr[a][b] +
                   dot * r[a] * dc * r[b] + c * dudr;
    }
  double dvdX[3][3] = {{0, -r[2], r[1]}, {r[2], 0, -r[0]}, {-r[1], r[0], 0}};
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      dydX[a][b] = (a == b ? ct : 0.0) + s * dvdX[a][b] + c * r[a] * r[b];
  double iz2 = 1.0 / (y[2] * y[2]);
  for (int m = 0; m < 3; ++m) {
    for (int a = 0; a < 2; ++a) {
      pe[a][m] = (dydr[a][m] * y[2] - y[a] * dydr[2][m]) * iz2;
      double dyt_a = (a == m) ? 1.0 : 0.0, dyt_2 = (m == 2) ? 1.0 : 0.0;
      pe[a][3 + m] = (dyt_a * y[2] - y[a] * dyt_2) * iz2;
      pp[a][m] = (dydX[a][m] * y[2] - y[a] * dydX[2][m]) * iz2;
    }
  }
}

// Per-key analytic Jacobian, BA/camera_shared_analytical_jacobian_matrix_calculator.hpp:94-321
// (radial :431-470, decentering :472-503, intrinsic layers :505-557).
// Jp [2][3], Ji [2][6], Jc [2][d] (free camera parameters only).
void key_jacobian(const OProblem &o, int64_t k, const double *x, double Jp[6], double Ji[12],
                  double *Jc) {
  double X[3], img[6], cam[kFull];
  get_point(o, o.key_pt[k], x, X);
  get_image(o, o.key_img[k], x, img);
  get_camera(o, o.icm[o.key_img[k]], x, cam);
  double n[2];
  normalized_key(X, img, img + 3, n);
  double pe[2][6], pp[2][3];
  normalized_derivative(X, img, img + 3, pe, pp);
  double r2 = n[0] * n[0] + n[1] * n[1], r4 = r2 * r2, r6 = r2 * r4;
  double k1 = cam[HSBA_K1], k2 = cam[HSBA_K2], k3 = cam[HSBA_K3];
  double d1 = cam[HSBA_D1], d2 = cam[HSBA_D2];
  double coeff = k1 * r2 + k2 * r4 + k3 * r6;
  // radius derivatives, computed from the undistorted pn (both distortions use it)
  double dpp[2][3], dpe[2][6];
  for (int m = 0; m < 3; ++m) {
    double pr2 = 2 * n[0] * pp[0][m] + 2 * n[1] * pp[1][m];
    double cd = k1 * pr2 + k2 * 2 * r2 * pr2 + k3 * 3 * r4 * pr2;
    double rad0 = cd * n[0] + coeff * pp[0][m], rad1 = cd * n[1] + coeff * pp[1][m];
    double cross = pp[0][m] * n[1] + pp[1][m] * n[0];
    double dec0 = 2 * d1 * cross + d2 * (pr2 + 4 * n[0] * pp[0][m]);
    double dec1 = 2 * d2 * cross + d1 * (pr2 + 4 * n[1] * pp[1][m]);
    dpp[0][m] = pp[0][m] + rad0 + dec0;
    dpp[1][m] = pp[1][m] + rad1 + dec1;
  }
  for (int m = 0; m < 6; ++m) {
    double pr2 = 2 * n[0] * pe[0][m] + 2 * n[1] * pe[1][m];
    double cd = k1 * pr2 + k2 * 2 * r2 * pr2 + k3 * 3 * r4 * pr2;
    double rad0 = cd * n[0] + coeff * pe[0][m], rad1 = cd * n[1] + coeff * pe[1][m];
    double cross = pe[0][m] * n[1] + pe[1][m] * n[0];
    double dec0 = 2 * d1 * cross + d2 * (pr2 + 4 * n[0] * pe[0][m]);
    double dec1 = 2 * d2 * cross + d1 * (pr2 + 4 * n[1] * pe[1][m]);
    dpe[0][m] = pe[0][m] + rad0 + dec0;
    dpe[1][m] = pe[1][m] + rad1 + dec1;
  }
  double f = cam[HSBA_FOCAL], sk = cam[HSBA_SKEW], pr = cam[HSBA_PIXEL_RATIO];
  for (int m = 0; m < 3; ++m) {
    Jp[m] = f * dpp[0][m] + sk * dpp[1][m];
    Jp[3 + m] = f * pr * dpp[1][m];
  }
  for (int m = 0; m < 6; ++m) {
    Ji[m] = f * dpe[0][m] + sk * dpe[1][m];
    Ji[6 + m] = f * pr * dpe[1][m];
  }
  if (!Jc) return;
  // distorted normalized key
  double dd0 = 2 * d1 * n[0] * n[1] + d2 * (r2 + 2 * n[0] * n[0]);
  double dd1 = 2 * d2 * n[0] * n[1] + d1 * (r2 + 2 * n[1] * n[1]);
  double m0 = n[0] + (coeff * n[0] + dd0), m1 = n[1] + (coeff * n[1] + dd1);
  // full 2x10 camera derivative, columns [k1 k2 k3 d1 d2 f skew px py pr]
  double nd[2][kFull] = {{r2 * n[0], r4 * n[0], r6 * n[0], 2 * n[0] * n[1], r2 + 2 * n[0] * n[0]},
                         {r2 * n[1], r4 * n[1], r6 * n[1], r2 + 2 * n[1] * n[1], 2 * n[0] * n[1]}};
  double full[2][kFull];
  for (int j = 0; j < 5; ++j) {
    full[0][j] = f * nd[0][j] + sk * nd[1][j];
    full[1][j] = f * pr * nd[1][j];
  }
  full[0][5] = m0; full[0][6] = m1; full[0][7] = 1; full[0][8] = 0; full[0][9] = 0;
  full[1][5] = pr * m1; full[1][6] = 0; full[1][7] = 0; full[1][8] = 1; full[1][9] = f * m1;
  for (int q = 0; q < o.d; ++q) {
    Jc[q] = full[0][o.free_idx[q]];
    Jc[o.d + q] = full[1][o.free_idx[q]];
  }
}

// Forward finite differences,
// BA/camera_shared_forward_finite_difference_jacobian_matrix_calculator.hpp:50-230:
// step = max(|delta * x_j|, min_delta).  central != 0 uses symmetric steps.
void key_jacobian_fd(const OProblem &o, int64_t k, const double *x, double delta,
                     double min_delta, int central, double Jp[6], double Ji[12], double *Jc) {
  double X[3], img[6], cam[kFull], uv0[2];
  int64_t p = o.key_pt[k], i = o.key_img[k], c = o.icm[i];
  get_point(o, p, x, X); get_image(o, i, x, img); get_camera(o, c, x, cam);
  image_key(X, img, cam, uv0);
  auto fd = [&](double &v, double *out, int stride) {
    double h = std::max(std::fabs(delta * v), min_delta), keep = v, up[2], dn[2];
    v = keep + h; image_key(X, img, cam, up);
    if (central) {
      v = keep - h; image_key(X, img, cam, dn);
      out[0] = (up[0] - dn[0]) / (2 * h); out[stride] = (up[1] - dn[1]) / (2 * h);
    } else {
      out[0] = (up[0] - uv0[0]) / h; out[stride] = (up[1] - uv0[1]) / h;
    }
    v = keep;
  };
  for (int m = 0; m < 3; ++m) fd(X[m], Jp + m, 3);
  for (int m = 0; m < 6; ++m) fd(img[m], Ji + m, 6);
  if (Jc)
    for (int q = 0; q < o.d; ++q) fd(cam[o.free_idx[q]], Jc + q, o.d);
}

// Complex-step derivative (Squire & Trapp 1998): exact to rounding, no
// subtractive cancellation.  Independent of the analytic chain rule.
void key_jacobian_cs(const OProblem &o, int64_t k, const double *x, double Jp[6], double Ji[12],
                     double *Jc) {
  typedef std::complex<double> Z;
  const double h = 1e-30;
  double Xd[3], imgd[6], camd[kFull];
  int64_t p = o.key_pt[k], i = o.key_img[k], c = o.icm[i];
  get_point(o, p, x, Xd); get_image(o, i, x, imgd); get_camera(o, c, x, camd);
  Z X[3], img[6], cam[kFull], uv[2];
  for (int j = 0; j < 3; ++j) X[j] = Xd[j];
  for (int j = 0; j < 6; ++j) img[j] = imgd[j];
  for (int j = 0; j < kFull; ++j) cam[j] = camd[j];
  auto cs = [&](Z &v, double *out, int stride) {
    Z keep = v; v = Z(keep.real(), h);
    image_key(X, img, cam, uv);
    out[0] = uv[0].imag() / h; out[stride] = uv[1].imag() / h;
    v = keep;
  };
  for (int m = 0; m < 3; ++m) cs(X[m], Jp + m, 3);
  for (int m = 0; m < 6; ++m) cs(img[m], Ji + m, 6);
  if (Jc)
    for (int q = 0; q < o.d; ++q) cs(cam[o.free_idx[q]], Jc + q, o.d);
}

// Huber: rho(s) = s (s <= a^2), 2 a sqrt(s) - a^2 (s > a^2); rho'(s) = min(1, a / sqrt(s)).
double rho(const OProblem &o, double s) {
  return (o.loss == 1 && s > o.loss_a * o.loss_a) ? 2.0 * o.loss_a * std::sqrt(s) - o.loss_a * o.loss_a : s;
}
double rho1(const OProblem &o, double s) {
  return (o.loss == 1 && s > o.loss_a * o.loss_a) ? o.loss_a / std::sqrt(s) : 1.0;
}

double key_w(const double *kb, int64_t k, int a, int b) {
  if (!kb) return a == b ? 1.0 : 0.0;
  return kb[4 * k + 2 * a + b];
}
double cons_w(const double *cw, size_t q) { return cw ? cw[q] : 1.0; }

// CameraSharedNormalEquationBuilder, BA/camera_shared_normal_equation_builder.hpp:111-462.
// Stores V_p, U_i, U_c, the per-key coupling blocks W_pi = Jp^T L Ji, W_pc = Jp^T L Jc,
// the per-image W_ic = sum Ji^T L Jc (one camera per image), and g = J^T L r.
int build_normal(OProblem &o, const double *x, const double *y_obs, const double *kb,
                 const double *cw) {
  const int d = o.d;
  o.V.assign(o.P * 9, 0.0); o.Ui.assign(o.I * 36, 0.0); o.Uc.assign(o.C * d * d, 0.0);
  o.Wpi.assign(o.N * 18, 0.0); o.Wpc.assign(o.N * 3 * d, 0.0); o.Wic.assign(o.I * 6 * d, 0.0);
  o.g.assign(o.x_size, 0.0);
  std::vector<double> Jc(2 * d + 1), LJc(2 * d + 1);
  for (int64_t k = 0; k < o.N; ++k) {
    double Jp[6], Ji[12], uv[2];
    key_jacobian(o, k, x, Jp, Ji, Jc.data());
    project_key(o, k, x, uv);
    double r[2] = {uv[0] - y_obs[2 * k], uv[1] - y_obs[2 * k + 1]};
    double L[2][2] = {{key_w(kb, k, 0, 0), key_w(kb, k, 0, 1)}, {key_w(kb, k, 1, 0), key_w(kb, k, 1, 1)}};
    if (o.loss) {
      // iteratively reweighted: Lambda -> rho'(s) Lambda
      const double s = r[0] * (L[0][0] * r[0] + L[0][1] * r[1]) + r[1] * (L[1][0] * r[0] + L[1][1] * r[1]);
      const double w = rho1(o, s);
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) L[a][b] *= w;
    }
    double Lr[2] = {L[0][0] * r[0] + L[0][1] * r[1], L[1][0] * r[0] + L[1][1] * r[1]};
    // L * Ji, L * Jc
    double LJi[12], LJp[6];
    for (int m = 0; m < 6; ++m) {
      LJi[m] = L[0][0] * Ji[m] + L[0][1] * Ji[6 + m];
      LJi[6 + m] = L[1][0] * Ji[m] + L[1][1] * Ji[6 + m];
    }
    for (int m = 0; m < 3; ++m) {
      LJp[m] = L[0][0] * Jp[m] + L[0][1] * Jp[3 + m];
      LJp[3 + m] = L[1][0] * Jp[m] + L[1][1] * Jp[3 + m];
    }
    for (int m = 0; m < d; ++m) {
      LJc[m] = L[0][0] * Jc[m] + L[0][1] * Jc[d + m];
      LJc[d + m] = L[1][0] * Jc[m] + L[1][1] * Jc[d + m];
    }
    int64_t p = o.key_pt[k], i = o.key_img[k], c = o.icm[i];
    if (o.pts_free) {
      for (int a = 0; a < 3; ++a) {
        for (int b = 0; b < 3; ++b) o.V[9 * p + 3 * a + b] += Jp[a] * LJp[b] + Jp[3 + a] * LJp[3 + b];
        o.g[o.x_pt + 3 * p + a] += Jp[a] * Lr[0] + Jp[3 + a] * Lr[1];
      }
    }
    if (o.imgs_free) {
      for (int a = 0; a < 6; ++a) {
        for (int b = 0; b < 6; ++b) o.Ui[36 * i + 6 * a + b] += Ji[a] * LJi[b] + Ji[6 + a] * LJi[6 + b];
        o.g[o.x_img + 6 * i + a] += Ji[a] * Lr[0] + Ji[6 + a] * Lr[1];
      }
    }
    if (o.cams_free) {
      for (int a = 0; a < d; ++a) {
        for (int b = 0; b < d; ++b)
          o.Uc[(c * d + a) * d + b] += Jc[a] * LJc[b] + Jc[d + a] * LJc[d + b];
        o.g[o.x_cam + d * c + a] += Jc[a] * Lr[0] + Jc[d + a] * Lr[1];
      }
    }
    if (o.pts_free && o.imgs_free)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 6; ++b) o.Wpi[18 * k + 6 * a + b] = Jp[a] * LJi[b] + Jp[3 + a] * LJi[6 + b];
    if (o.pts_free && o.cams_free)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < d; ++b) o.Wpc[(3 * k + a) * d + b] = Jp[a] * LJc[b] + Jp[3 + a] * LJc[d + b];
    if (o.imgs_free && o.cams_free)
      for (int a = 0; a < 6; ++a)
        for (int b = 0; b < d; ++b) o.Wic[(6 * i + a) * d + b] += Ji[a] * LJc[b] + Ji[6 + a] * LJc[d + b];
  }
  // point constraints with 3x3 blocks: N_pp += B, g_p += B r
  // (BA/ba_gcp_constrained_normal_equation_builder.hpp:59-72)
  const size_t q_first = o.pblocks.empty() ? 0 : (size_t)o.n_point_cons;
  for (size_t q0 = 0; q0 < q_first; q0 += 3) {
    const double *B = &o.pblocks[3 * q0];
    const int64_t p = o.cons[q0].entity;
    double r[3];
    for (int b = 0; b < 3; ++b) r[b] = x[o.cons[q0 + b].x_index] - y_obs[2 * o.N + q0 + b];
    for (int a = 0; a < 3; ++a) {
      double s = 0.0;
      for (int b = 0; b < 3; ++b) {
        o.V[9 * p + 3 * a + b] += B[3 * a + b];
        s += B[3 * a + b] * r[b];
      }
      o.g[o.x_pt + 3 * p + a] += s;
    }
  }
  // constraint rows: weights on the diagonal, w * r in the gradient (:309-349, :418-459)
  for (size_t q = q_first; q < o.cons.size(); ++q) {
    const Constraint &c = o.cons[q];
    double w = cons_w(cw, q);
    double r = x[c.x_index] - y_obs[2 * o.N + q];
    if (c.kind == 0) o.V[9 * c.entity + 4 * c.param] += w;
    else if (c.kind == 1) o.Ui[36 * c.entity + 7 * c.param] += w;
    else o.Uc[(c.entity * d + c.param) * d + c.param] += w;
    o.g[c.x_index] += w * r;
  }
  o.built = true;
  return HSBA_OK;
}

// Mahalanobis distance, BA/camera_shared_mahalanobis_distance_calculator.hpp:28-61.
double mahalanobis(const OProblem &o, const double *x, const double *y_obs, const double *kb,
                   const double *cw) {
  double F = 0.0;
  for (int64_t k = 0; k < o.N; ++k) {
    double uv[2];
    project_key(o, k, x, uv);
    double r0 = uv[0] - y_obs[2 * k], r1 = uv[1] - y_obs[2 * k + 1];
    F += rho(o, r0 * (key_w(kb, k, 0, 0) * r0 + key_w(kb, k, 0, 1) * r1) +
                r1 * (key_w(kb, k, 1, 0) * r0 + key_w(kb, k, 1, 1) * r1));
  }
  // 3x3 point blocks: r^T B r (BA/ba_gcp_constrained_mahalanobis_distance_calculator.hpp:50-58)
  const size_t q_first = o.pblocks.empty() ? 0 : (size_t)o.n_point_cons;
  for (size_t q0 = 0; q0 < q_first; q0 += 3) {
    const double *B = &o.pblocks[3 * q0];
    double r[3];
    for (int b = 0; b < 3; ++b) r[b] = x[o.cons[q0 + b].x_index] - y_obs[2 * o.N + q0 + b];
    for (int a = 0; a < 3; ++a) F += r[a] * (B[3 * a] * r[0] + B[3 * a + 1] * r[1] + B[3 * a + 2] * r[2]);
  }
  for (size_t q = q_first; q < o.cons.size(); ++q) {
    double r = x[o.cons[q].x_index] - y_obs[2 * o.N + q];
    F += r * cons_w(cw, q) * r;
  }
  return F;
}

// 3x3 inverse by cofactors (Eigen's fixed-size inverse, used at
// BA/camera_shared_normal_equation_solver.hpp:148).
void inv3(const double A[9], double B[9]) {
  double c00 = A[4] * A[8] - A[5] * A[7], c01 = A[5] * A[6] - A[3] * A[8], c02 = A[3] * A[7] - A[4] * A[6];
  double det = A[0] * c00 + A[1] * c01 + A[2] * c02;
  double id = 1.0 / det;
  B[0] = c00 * id; B[3] = c01 * id; B[6] = c02 * id;
  B[1] = (A[2] * A[7] - A[1] * A[8]) * id;
  B[4] = (A[0] * A[8] - A[2] * A[6]) * id;
  B[7] = (A[1] * A[6] - A[0] * A[7]) * id;
  B[2] = (A[1] * A[5] - A[2] * A[4]) * id;
  B[5] = (A[2] * A[3] - A[0] * A[5]) * id;
  B[8] = (A[0] * A[4] - A[1] * A[3]) * id;
}

int64_t reduced_size(const OProblem &o) {
  return (o.imgs_free ? 6 * o.I : 0) + (o.cams_free ? o.d * o.C : 0);
}

// Reduced system by point-wise accumulation; mathematically the loops of
// ComputeReducedSystem (BA/camera_shared_normal_equation_solver.hpp:116-446)
// regrouped by point.  S is dense row-major [n][n].
int reduced_system_pointwise(const OProblem &o, double mu, std::vector<double> &S,
                             std::vector<double> &rhs) {
  const int d = o.d;
  const int64_t n = reduced_size(o), cb = o.imgs_free ? 6 * o.I : 0;
  S.assign(n * n, 0.0); rhs.assign(n, 0.0);
  // "exist": a diagonal block is emitted only if some point couples it (:217-233, :288-304)
  if (o.imgs_free)
    for (int64_t i = 0; i < o.I; ++i) if (o.img_nkeys[i] == 0) return HSBA_ERR_GENERIC;
  if (o.cams_free)
    for (int64_t c = 0; c < o.C; ++c) if (o.cam_nkeys[c] == 0) return HSBA_ERR_GENERIC;
  if (o.imgs_free)
    for (int64_t i = 0; i < o.I; ++i) {
      for (int a = 0; a < 6; ++a) {
        for (int b = 0; b < 6; ++b) S[(6 * i + a) * n + 6 * i + b] = o.Ui[36 * i + 6 * a + b];
        S[(6 * i + a) * n + 6 * i + a] += mu;
        rhs[6 * i + a] = o.g[o.x_img + 6 * i + a];
      }
    }
  if (o.cams_free)
    for (int64_t c = 0; c < o.C; ++c)
      for (int a = 0; a < d; ++a) {
        for (int b = 0; b < d; ++b) S[(cb + d * c + a) * n + cb + d * c + b] = o.Uc[(c * d + a) * d + b];
        S[(cb + d * c + a) * n + cb + d * c + a] += mu;
        rhs[cb + d * c + a] = o.g[o.x_cam + d * c + a];
      }
  if (o.imgs_free && o.cams_free)
    for (int64_t i = 0; i < o.I; ++i) {
      int64_t c = o.icm[i];
      for (int a = 0; a < 6; ++a)
        for (int b = 0; b < d; ++b) {
          double w = o.Wic[(6 * i + a) * d + b];
          S[(6 * i + a) * n + cb + d * c + b] += w;
          S[(cb + d * c + b) * n + 6 * i + a] += w;
        }
    }
  if (!o.pts_free) return HSBA_OK;
  // per point: columns touched = images (6 each) then distinct cameras (d each)
  std::vector<int64_t> cols;
  std::vector<double> Wp, Y;
  for (int64_t p = 0; p < o.P; ++p) {
    double Vs[9], Vi[9];
    for (int j = 0; j < 9; ++j) Vs[j] = o.V[9 * p + j];
    for (int j = 0; j < 3; ++j) Vs[4 * j] += mu;
    inv3(Vs, Vi);
    cols.clear();
    // W_p as a 3 x m matrix, column-major in `cols` order
    Wp.clear();
    std::map<int64_t, size_t> cam_col;  // camera -> first column of its block
    for (int64_t q = o.pt_begin[p]; q < o.pt_begin[p + 1]; ++q) {
      int64_t k = o.pt_keys[q];
      int64_t i = o.key_img[k];
      if (o.imgs_free)
        for (int b = 0; b < 6; ++b) {
          cols.push_back(6 * i + b);
          for (int a = 0; a < 3; ++a) Wp.push_back(o.Wpi[18 * k + 6 * a + b]);
        }
      if (o.cams_free) {
        int64_t c = o.icm[i];
        auto it = cam_col.find(c);
        if (it == cam_col.end()) {
          cam_col[c] = cols.size();
          for (int b = 0; b < d; ++b) {
            cols.push_back(cb + d * c + b);
            for (int a = 0; a < 3; ++a) Wp.push_back(o.Wpc[(3 * k + a) * d + b]);
          }
        } else {  // several images of one camera group: W_pc blocks sum
          for (int b = 0; b < d; ++b)
            for (int a = 0; a < 3; ++a) Wp[3 * (it->second + b) + a] += o.Wpc[(3 * k + a) * d + b];
        }
      }
    }
    const size_t m = cols.size();
    Y.assign(3 * m, 0.0);  // Y = Vinv * W_p
    for (size_t j = 0; j < m; ++j)
      for (int a = 0; a < 3; ++a)
        Y[3 * j + a] = Vi[3 * a] * Wp[3 * j] + Vi[3 * a + 1] * Wp[3 * j + 1] + Vi[3 * a + 2] * Wp[3 * j + 2];
    const double *gp = &o.g[o.x_pt + 3 * p];
    for (size_t u = 0; u < m; ++u) {
      double *Srow = &S[cols[u] * n];
      const double *wu = &Wp[3 * u];
      for (size_t v = 0; v < m; ++v) {
        const double *yv = &Y[3 * v];
        Srow[cols[v]] -= wu[0] * yv[0] + wu[1] * yv[1] + wu[2] * yv[2];
      }
      const double *yu = &Y[3 * u];
      rhs[cols[u]] -= yu[0] * gp[0] + yu[1] * gp[1] + yu[2] * gp[2];
    }
  }
  return HSBA_OK;
}

// The reference's literal loops (BA/camera_shared_normal_equation_solver.hpp:141-443)
// over (weighted block, block) pairs x all points with sparse look-ups.  Only for
// small problems: cost O((I^2 + C^2 + IC) P).
int reduced_system_literal(const OProblem &o, double mu, std::vector<double> &S,
                           std::vector<double> &rhs) {
  const int d = o.d;
  const int64_t n = reduced_size(o), cb = o.imgs_free ? 6 * o.I : 0;
  if (!o.pts_free || (!o.imgs_free && !o.cams_free)) return HSBA_ERR_GENERIC;
  // W_pi and W_pc maps keyed by (point, image) / (point, camera); duplicates sum
  std::map<std::pair<int64_t, int64_t>, std::vector<double>> wpi, wpc;
  for (int64_t k = 0; k < o.N; ++k) {
    int64_t p = o.key_pt[k], i = o.key_img[k], c = o.icm[i];
    if (o.imgs_free) {
      auto &b = wpi[{p, i}]; b.resize(18, 0.0);
      for (int j = 0; j < 18; ++j) b[j] += o.Wpi[18 * k + j];
    }
    if (o.cams_free) {
      auto &b = wpc[{p, c}]; b.resize(3 * d, 0.0);
      for (int j = 0; j < 3 * d; ++j) b[j] += o.Wpc[3 * d * k + j];
    }
  }
  // weighted blocks Y^T = W^T Vinv per (image, point) and (camera, point), :141-196
  std::map<std::pair<int64_t, int64_t>, std::vector<double>> yip, ycp;
  for (int64_t p = 0; p < o.P; ++p) {
    double Vs[9], Vi[9];
    for (int j = 0; j < 9; ++j) Vs[j] = o.V[9 * p + j];
    for (int j = 0; j < 3; ++j) Vs[4 * j] += mu;
    inv3(Vs, Vi);
    if (o.imgs_free)
      for (int64_t i = 0; i < o.I; ++i) {
        auto it = wpi.find({p, i});
        if (it == wpi.end()) continue;
        std::vector<double> y(18, 0.0);  // [6][3]
        for (int a = 0; a < 6; ++a)
          for (int b = 0; b < 3; ++b)
            for (int c2 = 0; c2 < 3; ++c2) y[3 * a + b] += it->second[6 * c2 + a] * Vi[3 * c2 + b];
        yip[{i, p}] = y;
      }
    if (o.cams_free)
      for (int64_t c = 0; c < o.C; ++c) {
        auto it = wpc.find({p, c});
        if (it == wpc.end()) continue;
        std::vector<double> y(3 * d, 0.0);  // [d][3]
        for (int a = 0; a < d; ++a)
          for (int b = 0; b < 3; ++b)
            for (int c2 = 0; c2 < 3; ++c2) y[3 * a + b] += it->second[d * c2 + a] * Vi[3 * c2 + b];
        ycp[{c, p}] = y;
      }
  }
  S.assign(n * n, 0.0); rhs.assign(n, 0.0);
  std::vector<char> written(n * n, 0);
  if (o.imgs_free)  // :200-268
    for (int64_t wi = 0; wi < o.I; ++wi)
      for (int64_t i = wi; i < o.I; ++i) {
        double blk[36] = {0};
        if (wi == i) {
          for (int j = 0; j < 36; ++j) blk[j] = o.Ui[36 * i + j];
          for (int a = 0; a < 6; ++a) blk[7 * a] += mu;
        }
        bool exist = false;
        for (int64_t p = 0; p < o.P; ++p) {
          auto y = yip.find({wi, p});
          auto w = wpi.find({p, i});
          if (y == yip.end() || w == wpi.end()) continue;
          for (int a = 0; a < 6; ++a)
            for (int b = 0; b < 6; ++b)
              for (int c2 = 0; c2 < 3; ++c2) blk[6 * a + b] -= y->second[3 * a + c2] * w->second[6 * c2 + b];
          exist = true;
        }
        if (!exist) continue;
        for (int a = 0; a < 6; ++a)
          for (int b = 0; b < 6; ++b) {
            S[(6 * wi + a) * n + 6 * i + b] += blk[6 * a + b];
            if (wi != i) S[(6 * i + b) * n + 6 * wi + a] += blk[6 * a + b];
          }
      }
  if (o.cams_free)  // :270-341
    for (int64_t wc = 0; wc < o.C; ++wc)
      for (int64_t c = wc; c < o.C; ++c) {
        std::vector<double> blk(d * d, 0.0);
        if (wc == c) {
          for (int j = 0; j < d * d; ++j) blk[j] = o.Uc[c * d * d + j];
          for (int a = 0; a < d; ++a) blk[(d + 1) * a] += mu;
        }
        bool exist = false;
        for (int64_t p = 0; p < o.P; ++p) {
          auto y = ycp.find({wc, p});
          auto w = wpc.find({p, c});
          if (y == ycp.end() || w == wpc.end()) continue;
          for (int a = 0; a < d; ++a)
            for (int b = 0; b < d; ++b)
              for (int c2 = 0; c2 < 3; ++c2) blk[d * a + b] -= y->second[3 * a + c2] * w->second[d * c2 + b];
          exist = true;
        }
        if (!exist) continue;
        for (int a = 0; a < d; ++a)
          for (int b = 0; b < d; ++b) {
            S[(cb + d * wc + a) * n + cb + d * c + b] += blk[d * a + b];
            if (wc != c) S[(cb + d * c + b) * n + cb + d * wc + a] += blk[d * a + b];
          }
      }
  if (o.imgs_free && o.cams_free)  // :343-392
    for (int64_t i = 0; i < o.I; ++i)
      for (int64_t c = 0; c < o.C; ++c) {
        std::vector<double> blk(6 * d, 0.0);
        if (o.icm[i] == c)
          for (int j = 0; j < 6 * d; ++j) blk[j] = o.Wic[6 * d * i + j];
        bool exist = false;
        for (int64_t p = 0; p < o.P; ++p) {
          auto y = yip.find({i, p});
          auto w = wpc.find({p, c});
          if (y == yip.end() || w == wpc.end()) continue;
          for (int a = 0; a < 6; ++a)
            for (int b = 0; b < d; ++b)
              for (int c2 = 0; c2 < 3; ++c2) blk[d * a + b] -= y->second[3 * a + c2] * w->second[d * c2 + b];
          exist = true;
        }
        if (!exist) continue;
        for (int a = 0; a < 6; ++a)
          for (int b = 0; b < d; ++b) {
            S[(6 * i + a) * n + cb + d * c + b] += blk[d * a + b];
            S[(cb + d * c + b) * n + 6 * i + a] += blk[d * a + b];
          }
      }
  // rhs, :398-443
  if (o.imgs_free)
    for (int64_t i = 0; i < o.I; ++i)
      for (int a = 0; a < 6; ++a) {
        double v = o.g[o.x_img + 6 * i + a];
        for (int64_t p = 0; p < o.P; ++p) {
          auto y = yip.find({i, p});
          if (y == yip.end()) continue;
          for (int c2 = 0; c2 < 3; ++c2) v -= y->second[3 * a + c2] * o.g[o.x_pt + 3 * p + c2];
        }
        rhs[6 * i + a] = v;
      }
  if (o.cams_free)
    for (int64_t c = 0; c < o.C; ++c)
      for (int a = 0; a < d; ++a) {
        double v = o.g[o.x_cam + d * c + a];
        for (int64_t p = 0; p < o.P; ++p) {
          auto y = ycp.find({c, p});
          if (y == ycp.end()) continue;
          for (int c2 = 0; c2 < 3; ++c2) v -= y->second[3 * a + c2] * o.g[o.x_pt + 3 * p + c2];
        }
        rhs[cb + d * c + a] = v;
      }
  // a diagonal block never written makes the system singular
  for (int64_t j = 0; j < n; ++j) if (S[j * n + j] == 0.0) return HSBA_ERR_GENERIC;
  return HSBA_OK;
}

// Dense LDL^T without pivoting in place of Eigen::SimplicialLDLT
// (BA/camera_shared_normal_equation_solver.hpp:448-466); fails on a zero pivot.
int ldlt_solve(std::vector<double> &A, int64_t n, std::vector<double> &b) {
  // right-looking on the lower triangle: after step k, A[i][k] = l_ik and
  // A[k][k] = d_k; col holds the undivided column l_jk * d_k.
  std::vector<double> col(n);
  for (int64_t k = 0; k < n; ++k) {
    double dk = A[k * n + k];
    if (dk == 0.0 || !std::isfinite(dk)) return HSBA_ERR_GENERIC;
    for (int64_t j = k + 1; j < n; ++j) col[j] = A[j * n + k];
    for (int64_t i = k + 1; i < n; ++i) {
      double *rowi = &A[i * n];
      double lik = col[i] / dk;
      if (lik != 0.0)
        for (int64_t j = k + 1; j <= i; ++j) rowi[j] -= lik * col[j];
      rowi[k] = lik;
    }
  }
  // forward L z = b, diagonal D, backward L^T x = y
  for (int64_t i = 0; i < n; ++i) {
    double v = b[i];
    const double *rowi = &A[i * n];
    for (int64_t j = 0; j < i; ++j) v -= rowi[j] * b[j];
    b[i] = v;
  }
  for (int64_t i = 0; i < n; ++i) b[i] /= A[i * n + i];
  for (int64_t i = n - 1; i >= 0; --i) {
    double v = b[i];
    for (int64_t j = i + 1; j < n; ++j) v -= A[j * n + i] * b[j];
    b[i] = v;
  }
  for (int64_t i = 0; i < n; ++i) if (!std::isfinite(b[i])) return HSBA_ERR_GENERIC;
  return HSBA_OK;
}

// CameraSharedNormalEquationSolver::operator() and its special cases
// (BA/camera_shared_normal_equation_solver.hpp:75-114, 468-706).
int solve(OProblem &o, double mu, double *delta, int literal) {
  if (!o.built) return HSBA_ERR_GENERIC;
  const int d = o.d;
  const int64_t cb = o.imgs_free ? 6 * o.I : 0;
  std::fill(delta, delta + o.x_size, 0.0);
  if (o.pts_free && !o.imgs_free && !o.cams_free) {  // SolvePointsOnly :540-569
    for (int64_t p = 0; p < o.P; ++p) {
      double Vs[9], Vi[9];
      for (int j = 0; j < 9; ++j) Vs[j] = o.V[9 * p + j];
      for (int j = 0; j < 3; ++j) Vs[4 * j] += mu;
      inv3(Vs, Vi);
      for (int a = 0; a < 3; ++a)
        delta[3 * p + a] = Vi[3 * a] * o.g[3 * p] + Vi[3 * a + 1] * o.g[3 * p + 1] + Vi[3 * a + 2] * o.g[3 * p + 2];
    }
    return HSBA_OK;
  }
  std::vector<double> S, rhs;
  int rc = (literal && o.pts_free) ? reduced_system_literal(o, mu, S, rhs)
                                   : reduced_system_pointwise(o, mu, S, rhs);
  if (rc) return rc;
  const int64_t n = reduced_size(o);
  if (ldlt_solve(S, n, rhs)) return HSBA_ERR_GENERIC;
  const int64_t r0 = o.pts_free ? 3 * o.P : 0;
  for (int64_t j = 0; j < n; ++j) delta[r0 + j] = rhs[j];
  if (!o.pts_free) return HSBA_OK;
  // BackSubstitution :468-538 (point-wise: every W block of the point)
  for (int64_t p = 0; p < o.P; ++p) {
    double Vs[9], Vi[9], t[3];
    for (int j = 0; j < 9; ++j) Vs[j] = o.V[9 * p + j];
    for (int j = 0; j < 3; ++j) Vs[4 * j] += mu;
    inv3(Vs, Vi);
    for (int a = 0; a < 3; ++a) t[a] = o.g[3 * p + a];
    for (int64_t q = o.pt_begin[p]; q < o.pt_begin[p + 1]; ++q) {
      int64_t k = o.pt_keys[q], i = o.key_img[k], c = o.icm[i];
      for (int a = 0; a < 3; ++a) {
        if (o.imgs_free)
          for (int b = 0; b < 6; ++b) t[a] -= o.Wpi[18 * k + 6 * a + b] * delta[r0 + 6 * i + b];
        if (o.cams_free)
          for (int b = 0; b < d; ++b) t[a] -= o.Wpc[(3 * k + a) * d + b] * delta[r0 + cb + d * c + b];
      }
    }
    for (int a = 0; a < 3; ++a) delta[3 * p + a] = Vi[3 * a] * t[0] + Vi[3 * a + 1] * t[1] + Vi[3 * a + 2] * t[2];
  }
  return HSBA_OK;
}

double max_diag(const OProblem &o) {
  double m = -1e308;
  if (o.pts_free) for (int64_t p = 0; p < o.P; ++p) for (int a = 0; a < 3; ++a) m = std::max(m, o.V[9 * p + 4 * a]);
  if (o.imgs_free) for (int64_t i = 0; i < o.I; ++i) for (int a = 0; a < 6; ++a) m = std::max(m, o.Ui[36 * i + 7 * a]);
  if (o.cams_free) for (int64_t c = 0; c < o.C; ++c) for (int a = 0; a < o.d; ++a) m = std::max(m, o.Uc[(c * o.d + a) * o.d + a]);
  return m;
}
double inf_norm(const OProblem &o) {
  double m = -1e308;
  for (double v : o.g) m = std::max(m, std::fabs(v));
  return m;
}

// ---- sparsity of N and of the reduced system --------------------------------
// Coupling sets in the reference's block order
// (BA/camera_shared_normal_equation_builder.hpp:137-205): every key adds a
// (point, image) triplet whose block id is the triplet count + 1 (:141-153);
// (point, camera) and (image, camera) triplets go through a std::set ordered
// by (row, col) (comparator :66-69), inserted with id size() + 1, so a pair
// seen again keeps the id of its first insertion (:157-204).  Blocks are
// stored at id - 1 (BA/camera_shared_normal_matrix.hpp:315-362).
struct OTriplet {
  int64_t row, col, value;
  bool operator<(const OTriplet &b) const { return row < b.row || (!(b.row < row) && col < b.col); }
};

// (row, col) of every block, indexed by block id - 1.
void coupling_set(const OProblem &o, int which, std::vector<int64_t> &rc) {
  rc.clear();
  if (which == 0) {
    if (!o.pts_free || !o.imgs_free) return;
    for (int64_t k = 0; k < o.N; ++k) { rc.push_back(o.key_pt[k]); rc.push_back(o.key_img[k]); }
    return;
  }
  if (!o.cams_free || o.d == 0 || (which == 1 ? !o.pts_free : !o.imgs_free)) return;
  std::set<OTriplet> set;
  for (int64_t k = 0; k < o.N; ++k) {
    const int64_t row = which == 1 ? o.key_pt[k] : o.key_img[k];
    set.insert(OTriplet{row, (int64_t)o.icm[o.key_img[k]], (int64_t)set.size() + 1});
  }
  rc.assign(2 * set.size(), 0);
  for (const OTriplet &t : set) { rc[2 * (t.value - 1)] = t.row; rc[2 * (t.value - 1) + 1] = t.col; }
}

// Image pairs (a <= b) that share a point and, per pair, the (key on a, key
// on b) of every point seeing both, points in id order.  These are the image
// blocks S_ab that the reference's Schur loops emit (the `exist` rule of
// BA/camera_shared_normal_equation_solver.hpp:200-268: a block exists when
// some point couples both images).
void pair_structure(const OProblem &o, std::vector<int64_t> &pairs, std::vector<int64_t> &ptr,
                    std::vector<int64_t> &inc) {
  std::map<std::pair<int64_t, int64_t>, std::vector<int64_t>> m;
  for (int64_t p = 0; p < o.P; ++p)
    for (int64_t qa = o.pt_begin[p]; qa < o.pt_begin[p + 1]; ++qa)
      for (int64_t qb = o.pt_begin[p]; qb < o.pt_begin[p + 1]; ++qb) {
        const int64_t ka = o.pt_keys[qa], kb = o.pt_keys[qb];
        if (o.key_img[ka] > o.key_img[kb]) continue;
        std::vector<int64_t> &v = m[{o.key_img[ka], o.key_img[kb]}];
        v.push_back(ka);
        v.push_back(kb);
      }
  pairs.clear(); ptr.assign(1, 0); inc.clear();
  for (const auto &e : m) {
    pairs.push_back(e.first.first);
    pairs.push_back(e.first.second);
    inc.insert(inc.end(), e.second.begin(), e.second.end());
    ptr.push_back((int64_t)inc.size() / 2);
  }
}

}  // namespace

extern "C" {

struct oracle_lm_trace {
  double mu, cost, cost_new, rho, step_norm;
  int32_t accepted, reserved_;
};

void *oracle_create(const hsba_problem_desc *desc, int *rc) {
  OProblem *o = new OProblem();
  int r = setup(*o, desc);
  if (rc) *rc = r;
  if (r) { delete o; return nullptr; }
  return o;
}
void oracle_destroy(void *h) { delete static_cast<OProblem *>(h); }

int oracle_sizes(void *h, int64_t *x_size, int64_t *y_size, int64_t *d, int64_t *n_reduced) {
  OProblem &o = *static_cast<OProblem *>(h);
  if (x_size) *x_size = o.x_size;
  if (y_size) *y_size = o.y_size;
  if (d) *d = o.d;
  if (n_reduced) *n_reduced = reduced_size(o);
  return 0;
}

// Full 3x3 weight blocks for the point constraints ([records][3][3]); every
// point constraint must constrain all three coordinates.  NULL clears them.
int oracle_set_point_blocks(void *h, const double *blocks) {
  OProblem &o = *static_cast<OProblem *>(h);
  if (!blocks) { o.pblocks.clear(); return 0; }
  if (o.n_point_cons % 3) return HSBA_ERR_INVALID;
  for (int64_t q = 0; q < o.n_point_cons; ++q)
    if (o.cons[q].param != q % 3) return HSBA_ERR_INVALID;
  o.pblocks.assign(blocks, blocks + 3 * o.n_point_cons);
  return 0;
}

// Robust loss on the key residuals: 0 none, 1 Huber with scale a.
int oracle_set_loss(void *h, int loss, double a) {
  OProblem &o = *static_cast<OProblem *>(h);
  if (loss != 0 && loss != 1) return HSBA_ERR_INVALID;
  o.loss = loss;
  o.loss_a = a > 0.0 ? a : 1.0;
  return 0;
}

int oracle_evaluate(void *h, const double *x, double *y) {
  evaluate(*static_cast<OProblem *>(h), x, y);
  return 0;
}

// mode: 0 analytic, 1 forward FD, 2 central FD, 3 complex step
int oracle_jacobian(void *h, const double *x, int mode, double delta, double min_delta,
                    double *Jp, double *Ji, double *Jc) {
  OProblem &o = *static_cast<OProblem *>(h);
  for (int64_t k = 0; k < o.N; ++k) {
    double *jc = Jc ? Jc + 2 * o.d * k : nullptr;
    if (mode == 0) key_jacobian(o, k, x, Jp + 6 * k, Ji + 12 * k, jc);
    else if (mode == 3) key_jacobian_cs(o, k, x, Jp + 6 * k, Ji + 12 * k, jc);
    else key_jacobian_fd(o, k, x, delta, min_delta, mode == 2, Jp + 6 * k, Ji + 12 * k, jc);
  }
  return 0;
}

int oracle_build_normal(void *h, const double *x, const double *y_obs, const double *key_blocks,
                        const double *cons_weights, double *V, double *Ui, double *Uc,
                        double *Wpi, double *Wpc, double *Wic, double *g) {
  OProblem &o = *static_cast<OProblem *>(h);
  int rc = build_normal(o, x, y_obs, key_blocks, cons_weights);
  if (rc) return rc;
  if (V) std::copy(o.V.begin(), o.V.end(), V);
  if (Ui) std::copy(o.Ui.begin(), o.Ui.end(), Ui);
  if (Uc) std::copy(o.Uc.begin(), o.Uc.end(), Uc);
  if (Wpi) std::copy(o.Wpi.begin(), o.Wpi.end(), Wpi);
  if (Wpc) std::copy(o.Wpc.begin(), o.Wpc.end(), Wpc);
  if (Wic) std::copy(o.Wic.begin(), o.Wic.end(), Wic);
  if (g) std::copy(o.g.begin(), o.g.end(), g);
  return 0;
}

// Coupling sets and blocks of the last build (block-id order); counts first
// (any array pointer may be null).  which: 0 point-image, 1 point-camera,
// 2 image-camera.  Blocks: [n][3][6], [n][3][d], [n][6][d].
int oracle_coupling(void *h, int which, int64_t *count, int32_t *rc_out, double *blocks) {
  OProblem &o = *static_cast<OProblem *>(h);
  std::vector<int64_t> rc;
  coupling_set(o, which, rc);
  const int64_t n = (int64_t)rc.size() / 2;
  *count = n;
  if (rc_out)
    for (int64_t j = 0; j < 2 * n; ++j) rc_out[j] = (int32_t)rc[j];
  if (!blocks || n == 0) return 0;
  if (!o.built) return HSBA_ERR_GENERIC;
  const int d = o.d;
  if (which == 0) {
    std::copy(o.Wpi.begin(), o.Wpi.end(), blocks);
    return 0;
  }
  // AddPointCameraBlock / AddImageCameraBlock (BA/camera_shared_normal_matrix.hpp:335-366):
  // every key's block is added at its pair's id
  std::map<std::pair<int64_t, int64_t>, int64_t> id;
  for (int64_t q = 0; q < n; ++q) id[{rc[2 * q], rc[2 * q + 1]}] = q;
  const int rows = which == 1 ? 3 : 6;
  std::fill(blocks, blocks + n * rows * d, 0.0);
  if (which == 1) {
    for (int64_t k = 0; k < o.N; ++k) {
      double *b = blocks + id[{o.key_pt[k], (int64_t)o.icm[o.key_img[k]]}] * 3 * d;
      for (int j = 0; j < 3 * d; ++j) b[j] += o.Wpc[3 * d * k + j];
    }
  } else {
    // the builder's per-image sums (one camera per image)
    for (int64_t q = 0; q < n; ++q)
      std::copy(o.Wic.begin() + 6 * d * rc[2 * q], o.Wic.begin() + 6 * d * (rc[2 * q] + 1), blocks + q * 6 * d);
  }
  return 0;
}

// Pair structure of the reduced system: counts first, then the arrays
// (pairs [n][2], ptr [n+1], incidences [m][2] as caller key indices).
int oracle_pairs(void *h, int64_t *num_pairs, int64_t *num_inc, int32_t *pairs, int64_t *ptr, int32_t *inc) {
  OProblem &o = *static_cast<OProblem *>(h);
  std::vector<int64_t> pr, pp, ic;
  if (o.pts_free && (o.imgs_free || (o.cams_free && o.d > 0))) pair_structure(o, pr, pp, ic);
  else pp.assign(1, 0);
  *num_pairs = (int64_t)pr.size() / 2;
  *num_inc = (int64_t)ic.size() / 2;
  if (pairs) for (size_t j = 0; j < pr.size(); ++j) pairs[j] = (int32_t)pr[j];
  if (ptr) std::copy(pp.begin(), pp.end(), ptr);
  if (inc) for (size_t j = 0; j < ic.size(); ++j) inc[j] = (int32_t)ic[j];
  return 0;
}

int oracle_reduced_system(void *h, double mu, int literal, double *S, double *rhs) {
  OProblem &o = *static_cast<OProblem *>(h);
  if (!o.built) return HSBA_ERR_GENERIC;
  std::vector<double> s, r;
  int rc = literal ? reduced_system_literal(o, mu, s, r) : reduced_system_pointwise(o, mu, s, r);
  if (S && !s.empty()) std::copy(s.begin(), s.end(), S);
  if (rhs && !r.empty()) std::copy(r.begin(), r.end(), rhs);
  return rc;
}

int oracle_solve(void *h, double mu, int literal, double *delta) {
  return solve(*static_cast<OProblem *>(h), mu, delta, literal);
}

double oracle_max_diagonal(void *h) { return max_diag(*static_cast<OProblem *>(h)); }
double oracle_gradient_inf_norm(void *h) { return inf_norm(*static_cast<OProblem *>(h)); }
double oracle_gradient_dot(void *h, const double *delta) {
  OProblem &o = *static_cast<OProblem *>(h);
  double s = 0.0;
  for (int64_t j = 0; j < o.x_size; ++j) s += o.g[j] * delta[j];
  return s;
}
double oracle_mahalanobis(void *h, const double *x, const double *y_obs, const double *key_blocks,
                          const double *cons_weights) {
  return mahalanobis(*static_cast<OProblem *>(h), x, y_obs, key_blocks, cons_weights);
}

// Levenberg–Marquardt, Madsen–Nielsen–Tingleff Alg. 3.16 in the reference's
// conventions (r = f(x) - y, (N + mu I) delta = g, x_new = x - delta,
// F = r^T W r).  `literal` selects the reference's literal Schur loops.
int oracle_lm(void *h, const double *x0, const double *y_obs, const double *key_blocks,
              const double *cons_weights, const hsba_lm_options *opt, int literal,
              double *x_out, hsba_lm_stats *stats, oracle_lm_trace *trace) {
  OProblem &o = *static_cast<OProblem *>(h);
  std::vector<double> x(x0, x0 + o.x_size), xn(o.x_size), delta(o.x_size);
  hsba_lm_stats st;
  std::memset(&st, 0, sizeof(st));
  double F = mahalanobis(o, x.data(), y_obs, key_blocks, cons_weights);
  st.initial_cost = F;
  build_normal(o, x.data(), y_obs, key_blocks, cons_weights);
  double mu = opt->tau * max_diag(o), nu = 2.0;
  bool found = inf_norm(o) <= opt->eps1;
  st.stop_reason = found ? HSBA_STOP_GRADIENT : HSBA_STOP_MAX_ITERATIONS;
  int k = 0;
  while (!found && k < opt->max_iterations) {
    ++k;
    if (solve(o, mu, delta.data(), literal)) {
      // a damped system that does not factor is a rejected step (the
      // hs_optimizor rule on solver failure is not available; this is the
      // device's rule)
      oracle_lm_trace tf = {mu, F, 0.0, 0.0, 0.0, 0, 0};
      if (trace) trace[k - 1] = tf;
      mu *= nu; nu *= 2.0;
      if (!(mu < 1.0e300)) { st.status = HSBA_ERR_NOT_SPD; st.stop_reason = HSBA_STOP_FAILED; break; }
      continue;
    }
    double dn = 0.0, xn2 = 0.0;
    for (int64_t j = 0; j < o.x_size; ++j) { dn += delta[j] * delta[j]; xn2 += x[j] * x[j]; }
    dn = std::sqrt(dn);
    oracle_lm_trace tr = {mu, F, 0.0, 0.0, dn, 0, 0};
    if (dn <= opt->eps2 * (std::sqrt(xn2) + opt->eps2)) {
      found = true; st.stop_reason = HSBA_STOP_STEP;
      if (trace) trace[k - 1] = tr;
      break;
    }
    for (int64_t j = 0; j < o.x_size; ++j) xn[j] = x[j] - delta[j];
    double Fn = mahalanobis(o, xn.data(), y_obs, key_blocks, cons_weights);
    double den = 0.0;
    for (int64_t j = 0; j < o.x_size; ++j) den += delta[j] * (mu * delta[j] + o.g[j]);
    double rho = (F - Fn) / den;
    tr.cost_new = Fn; tr.rho = rho;
    if (rho > 0.0 && std::isfinite(Fn)) {
      x.swap(xn); F = Fn;
      build_normal(o, x.data(), y_obs, key_blocks, cons_weights);
      found = inf_norm(o) <= opt->eps1;
      if (found) st.stop_reason = HSBA_STOP_GRADIENT;
      double t = 2.0 * rho - 1.0;
      mu *= std::max(1.0 / 3.0, 1.0 - t * t * t);
      nu = 2.0;
      st.accepted++;
      tr.accepted = 1;
    } else {
      mu *= nu; nu *= 2.0;
    }
    if (trace) trace[k - 1] = tr;
  }
  st.iterations = k;
  st.final_cost = F;
  st.final_mu = mu;
  std::copy(x.begin(), x.end(), x_out);
  if (stats) *stats = st;
  return st.status;
}

// Dense brute force for tiny problems: J (y_size x x_size, row-major) from the
// analytic per-key blocks plus identity constraint rows, the dense inverse
// covariance, N = J^T W J and g = J^T W r.  Used to pin build_normal.
int oracle_dense_normal(void *h, const double *x, const double *y_obs, const double *key_blocks,
                        const double *cons_weights, double *J, double *N, double *g) {
  OProblem &o = *static_cast<OProblem *>(h);
  const int64_t nx = o.x_size, ny = o.y_size;
  const int d = o.d;
  std::vector<double> Jd(ny * nx, 0.0), r(ny), y(ny), Wr(ny);
  std::vector<double> Jc(2 * d + 1);
  for (int64_t k = 0; k < o.N; ++k) {
    double Jp[6], Ji[12];
    key_jacobian(o, k, x, Jp, Ji, Jc.data());
    int64_t p = o.key_pt[k], i = o.key_img[k], c = o.icm[i];
    for (int a = 0; a < 2; ++a) {
      double *row = &Jd[(2 * k + a) * nx];
      if (o.pts_free) for (int b = 0; b < 3; ++b) row[o.x_pt + 3 * p + b] = Jp[3 * a + b];
      if (o.imgs_free) for (int b = 0; b < 6; ++b) row[o.x_img + 6 * i + b] = Ji[6 * a + b];
      if (o.cams_free) for (int b = 0; b < d; ++b) row[o.x_cam + d * c + b] += Jc[d * a + b];
    }
  }
  for (size_t q = 0; q < o.cons.size(); ++q) Jd[(2 * o.N + q) * nx + o.cons[q].x_index] = 1.0;
  evaluate(o, x, y.data());
  for (int64_t j = 0; j < ny; ++j) r[j] = y[j] - y_obs[j];
  // per-key robust weight rho'(s), s = r^T Lambda r (1 without a loss)
  std::vector<double> kw(o.N, 1.0);
  for (int64_t k = 0; k < o.N; ++k) {
    double s = 0.0;
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) s += r[2 * k + a] * key_w(key_blocks, k, a, b) * r[2 * k + b];
    kw[k] = rho1(o, s);
  }
  for (int64_t k = 0; k < o.N; ++k)
    for (int a = 0; a < 2; ++a)
      Wr[2 * k + a] = kw[k] * (key_w(key_blocks, k, a, 0) * r[2 * k] + key_w(key_blocks, k, a, 1) * r[2 * k + 1]);
  // the inverse covariance of the constraint rows: diagonal weights, or 3x3
  // blocks over each point constraint's three rows
  auto cons_weight = [&](size_t q, size_t j) -> double {
    if (!o.pblocks.empty() && q < (size_t)o.n_point_cons) {
      const size_t q0 = q - q % 3;
      if (j < q0 || j >= q0 + 3) return 0.0;
      return o.pblocks[3 * q0 + 3 * (q - q0) + (j - q0)];
    }
    return j == q ? cons_w(cons_weights, q) : 0.0;
  };
  for (size_t q = 0; q < o.cons.size(); ++q) {
    double s = 0.0;
    for (size_t j = 0; j < o.cons.size(); ++j) s += cons_weight(q, j) * r[2 * o.N + j];
    Wr[2 * o.N + q] = s;
  }
  if (J) std::copy(Jd.begin(), Jd.end(), J);
  if (g)
    for (int64_t c = 0; c < nx; ++c) {
      double s = 0.0;
      for (int64_t j = 0; j < ny; ++j) s += Jd[j * nx + c] * Wr[j];
      g[c] = s;
    }
  if (N) {
    // W J then J^T (W J)
    std::vector<double> WJ(ny * nx, 0.0);
    for (int64_t k = 0; k < o.N; ++k)
      for (int a = 0; a < 2; ++a)
        for (int64_t c = 0; c < nx; ++c)
          WJ[(2 * k + a) * nx + c] = kw[k] * (key_w(key_blocks, k, a, 0) * Jd[(2 * k) * nx + c] +
                                              key_w(key_blocks, k, a, 1) * Jd[(2 * k + 1) * nx + c]);
    for (size_t q = 0; q < o.cons.size(); ++q)
      for (int64_t c = 0; c < nx; ++c) {
        double s = 0.0;
        for (size_t j = 0; j < o.cons.size(); ++j) s += cons_weight(q, j) * Jd[(2 * o.N + j) * nx + c];
        WJ[(2 * o.N + q) * nx + c] = s;
      }
    for (int64_t a = 0; a < nx; ++a)
      for (int64_t b = 0; b < nx; ++b) {
        double s = 0.0;
        for (int64_t j = 0; j < ny; ++j) s += Jd[j * nx + a] * WJ[j * nx + b];
        N[a * nx + b] = s;
      }
  }
  return 0;
}

}  // extern "C"
