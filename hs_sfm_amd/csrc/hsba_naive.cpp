// hsba_naive.cpp — the ba_naive bundle-adjustment family over the same device
// engine.
//
// ba_naive (include/hs_sfm/bundle_adjustment/ba_naive_*.hpp) is the
// camera_shared algorithm restricted to 6-DoF cameras in normalised image
// coordinates:
//   * feature = (R(r) p + t).xy / (R(r) p + t).z, Rodrigues as in
//     BANaiveVectorFunction::PointProjectToFeature (ba_naive_vector_function.hpp:170-196),
//     which is the camera_shared projection with one image per camera and
//     identity intrinsics (f = pr = 1, no skew, principal point or distortion);
//   * x = [cameras 6C | points 3P] (ba_naive_vector_function.hpp:117-135)
//     instead of camera_shared's [points 3P | images 6I];
//   * y = features / f, 2 per feature in feature order.
// So each entry point here re-indexes x (and the gradient / step) and calls
// the camera_shared entry point on a problem with num_images = num_cameras,
// image_camera_map = identity and no intrinsic group.  The analytic Jacobian
// of the reference fails (-1) at an exactly zero rotation
// (ba_naive_analytical_jacobian_matrix_calculator.hpp:104-108); the entry
// points that evaluate it here report the same condition as
// HSBA_ERR_INVALID.
//
// ba_gcp_constrained (ba_gcp_constrained_*.hpp) is ba_naive whose last G
// points are ground control points with observed coordinates: y gains 3G
// rows equal to those points' x entries, weighted by 3x3 blocks.  These are
// camera_shared point constraints on all three coordinates with full 3x3
// weight blocks (hsba_ycov_inverse::point_constraint_blocks), in GCP order.
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "hsba_host.h"

struct hsba_naive_problem {
  hsba_problem *p = nullptr;
  int64_t C = 0, P = 0, F = 0, G = 0;
  std::vector<double> xs;  // camera_shared-layout scratch
  ~hsba_naive_problem() {
    if (p) hsba_problem_destroy(p);
  }
};

struct hsba_gcp_problem : hsba_naive_problem {};

namespace {

// [cameras 6C | points 3P] -> [points 3P | images 6C]
void to_shared(const hsba_naive_problem &q, const double *x, double *xs) {
  std::memcpy(xs, x + 6 * q.C, sizeof(double) * 3 * q.P);
  std::memcpy(xs + 3 * q.P, x, sizeof(double) * 6 * q.C);
}

void to_naive(const hsba_naive_problem &q, const double *xs, double *x) {
  std::memcpy(x, xs + 3 * q.P, sizeof(double) * 6 * q.C);
  std::memcpy(x + 6 * q.C, xs, sizeof(double) * 3 * q.P);
}

bool zero_rotation(const hsba_naive_problem &q, const double *x) {
  for (int64_t c = 0; c < q.C; ++c)
    if (x[6 * c] == 0.0 && x[6 * c + 1] == 0.0 && x[6 * c + 2] == 0.0) return true;
  return false;
}

// The camera_shared problem behind a naive (G = 0) or GCP-constrained one.
int create_shared(hsba_naive_problem &q, const int32_t *feature_maps, const hsba_options *options) {
  std::vector<int32_t> icm(q.C);
  for (int64_t c = 0; c < q.C; ++c) icm[c] = (int32_t)c;
  std::vector<hsba_constraint> gcps(q.G);
  for (int64_t i = 0; i < q.G; ++i) {
    gcps[i].id = q.P - q.G + i;
    gcps[i].mask = 7u;  // X, Y and Z observed
    gcps[i].reserved_ = 0;
  }
  hsba_problem_desc d;
  std::memset(&d, 0, sizeof(d));
  d.num_points = q.P;
  d.num_images = q.C;
  d.num_cameras = q.C;
  d.num_keys = q.F;
  d.feature_maps = feature_maps;       // (camera, point) == (image, point)
  d.image_camera_map = icm.data();
  d.intrinsic_computations_mask = 0;   // identity intrinsics: normalised coordinates
  d.fix_mask = 0;
  d.num_point_constraints = q.G;
  d.point_constraints = q.G ? gcps.data() : nullptr;
  const int rc = hsba_problem_create(&d, options, &q.p);
  if (rc) return rc;
  q.xs.resize(3 * q.P + 6 * q.C);
  return HSBA_OK;
}

}  // namespace

extern "C" {

int hsba_naive_problem_create(const hsba_naive_desc *desc, const hsba_options *options,
                              hsba_naive_problem **problem) {
  return hsba::guarded([&]() -> int {
    if (!desc || !problem || !desc->feature_maps || desc->num_cameras <= 0 || desc->num_points < 0 ||
        desc->num_features < 0)
      return HSBA_ERR_INVALID;
    *problem = nullptr;
    std::unique_ptr<hsba_naive_problem> q(new hsba_naive_problem());
    q->C = desc->num_cameras;
    q->P = desc->num_points;
    q->F = desc->num_features;
    const int rc = create_shared(*q, desc->feature_maps, options);
    if (rc) return rc;
    *problem = q.release();
    return HSBA_OK;
  });
}

int hsba_naive_problem_destroy(hsba_naive_problem *problem) {
  return hsba::guarded([&]() -> int {
    delete problem;
    return HSBA_OK;
  });
}

int hsba_naive_evaluate(hsba_naive_problem *q, const double *x, double *y) {
  return hsba::guarded([&]() -> int {
    if (!q || !x || !y) return HSBA_ERR_INVALID;
    to_shared(*q, x, q->xs.data());
    return hsba_evaluate(q->p, q->xs.data(), y);
  });
}

int hsba_naive_jacobian(hsba_naive_problem *q, const double *x, double *camera_blocks, double *point_blocks) {
  return hsba::guarded([&]() -> int {
    if (!q || !x) return HSBA_ERR_INVALID;
    if (zero_rotation(*q, x)) return HSBA_ERR_INVALID;
    to_shared(*q, x, q->xs.data());
    return hsba_jacobian(q->p, q->xs.data(), point_blocks, camera_blocks, nullptr);
  });
}

int hsba_naive_build_normal(hsba_naive_problem *q, const double *x, const double *y_obs,
                            const double *feature_cov_inverse, double *camera_blocks, double *point_blocks,
                            double *gradient) {
  return hsba::guarded([&]() -> int {
    if (!q || !x || !y_obs) return HSBA_ERR_INVALID;
    if (zero_rotation(*q, x)) return HSBA_ERR_INVALID;
    to_shared(*q, x, q->xs.data());
    hsba_ycov_inverse w = {feature_cov_inverse, nullptr, nullptr};
    std::vector<double> gs(gradient ? q->xs.size() : 0);
    const int rc = hsba_build_normal(q->p, q->xs.data(), y_obs, &w, point_blocks, camera_blocks, nullptr,
                                     gradient ? gs.data() : nullptr);
    if (!rc && gradient) to_naive(*q, gs.data(), gradient);
    return rc;
  });
}

int hsba_naive_solve(hsba_naive_problem *q, double mu, double *delta_x) {
  return hsba::guarded([&]() -> int {
    if (!q || !delta_x) return HSBA_ERR_INVALID;
    std::vector<double> ds(q->xs.size());
    const int rc = hsba_solve(q->p, mu, ds.data());
    if (!rc) to_naive(*q, ds.data(), delta_x);
    return rc;
  });
}

int hsba_naive_lm_solve(hsba_naive_problem *q, const double *x0, const double *y_obs,
                        const double *feature_cov_inverse, const hsba_lm_options *lm, double *x_out,
                        hsba_lm_stats *stats) {
  return hsba::guarded([&]() -> int {
    if (!q || !x0 || !y_obs || !lm || !x_out) return HSBA_ERR_INVALID;
    if (zero_rotation(*q, x0)) return HSBA_ERR_INVALID;
    to_shared(*q, x0, q->xs.data());
    hsba_ycov_inverse w = {feature_cov_inverse, nullptr, nullptr};
    std::vector<double> xo(q->xs.size());
    const int rc = hsba_lm_solve(q->p, q->xs.data(), y_obs, &w, lm, xo.data(), stats);
    to_naive(*q, xo.data(), x_out);
    return rc;
  });
}

// ---- ba_gcp_constrained -------------------------------------------------

int hsba_gcp_problem_create(const hsba_gcp_desc *desc, const hsba_options *options,
                            hsba_gcp_problem **problem) {
  return hsba::guarded([&]() -> int {
    if (!desc || !problem || !desc->feature_maps || desc->num_cameras <= 0 || desc->num_points < 0 ||
        desc->num_features < 0 || desc->num_gcps < 0 || desc->num_gcps > desc->num_points)
      return HSBA_ERR_INVALID;
    *problem = nullptr;
    std::unique_ptr<hsba_gcp_problem> q(new hsba_gcp_problem());
    q->C = desc->num_cameras;
    q->P = desc->num_points;
    q->F = desc->num_features;
    q->G = desc->num_gcps;
    const int rc = create_shared(*q, desc->feature_maps, options);
    if (rc) return rc;
    *problem = q.release();
    return HSBA_OK;
  });
}

int hsba_gcp_problem_destroy(hsba_gcp_problem *problem) {
  return hsba::guarded([&]() -> int {
    delete problem;
    return HSBA_OK;
  });
}

int hsba_gcp_evaluate(hsba_gcp_problem *q, const double *x, double *y) {
  return hsba::guarded([&]() -> int {
    // camera_shared y = [keys 2F | point constraint rows], i.e. the GCP
    // coordinates in GCP order
    return hsba_naive_evaluate(q, x, y);
  });
}

int hsba_gcp_jacobian(hsba_gcp_problem *q, const double *x, double *camera_blocks, double *point_blocks) {
  return hsba::guarded([&]() -> int {
    return hsba_naive_jacobian(q, x, camera_blocks, point_blocks);
  });
}

int hsba_gcp_build_normal(hsba_gcp_problem *q, const double *x, const double *y_obs,
                          const double *feature_cov_inverse, const double *gcp_blocks, double *camera_blocks,
                          double *point_blocks, double *gradient) {
  return hsba::guarded([&]() -> int {
    if (!q || !x || !y_obs) return HSBA_ERR_INVALID;
    if (zero_rotation(*q, x)) return HSBA_ERR_INVALID;
    to_shared(*q, x, q->xs.data());
    hsba_ycov_inverse w = {feature_cov_inverse, nullptr, gcp_blocks};
    std::vector<double> gs(gradient ? q->xs.size() : 0);
    const int rc = hsba_build_normal(q->p, q->xs.data(), y_obs, &w, point_blocks, camera_blocks, nullptr,
                                     gradient ? gs.data() : nullptr);
    if (!rc && gradient) to_naive(*q, gs.data(), gradient);
    return rc;
  });
}

int hsba_gcp_solve(hsba_gcp_problem *q, double mu, double *delta_x) { return hsba_naive_solve(q, mu, delta_x); }

int hsba_gcp_mahalanobis(hsba_gcp_problem *q, const double *x, const double *y_obs,
                         const double *feature_cov_inverse, const double *gcp_blocks, double *value) {
  return hsba::guarded([&]() -> int {
    if (!q || !x || !y_obs || !value) return HSBA_ERR_INVALID;
    to_shared(*q, x, q->xs.data());
    hsba_ycov_inverse w = {feature_cov_inverse, nullptr, gcp_blocks};
    return hsba_mahalanobis(q->p, q->xs.data(), y_obs, &w, value);
  });
}

int hsba_gcp_lm_solve(hsba_gcp_problem *q, const double *x0, const double *y_obs,
                      const double *feature_cov_inverse, const double *gcp_blocks, const hsba_lm_options *lm,
                      double *x_out, hsba_lm_stats *stats) {
  return hsba::guarded([&]() -> int {
    if (!q || !x0 || !y_obs || !lm || !x_out) return HSBA_ERR_INVALID;
    if (zero_rotation(*q, x0)) return HSBA_ERR_INVALID;
    to_shared(*q, x0, q->xs.data());
    hsba_ycov_inverse w = {feature_cov_inverse, nullptr, gcp_blocks};
    std::vector<double> xo(q->xs.size());
    const int rc = hsba_lm_solve(q->p, q->xs.data(), y_obs, &w, lm, xo.data(), stats);
    to_naive(*q, xo.data(), x_out);
    return rc;
  });
}

}  // extern "C"
