/*
 * hsba.h — C ABI of the MI355X Levenberg–Marquardt bundle adjuster.
 *
 * Drop-in boundary for the camera_shared bundle-adjustment path of hs_sfm
 * (include/hs_sfm/bundle_adjustment/camera_shared_*.hpp).  Every entry point
 * takes plain pointers and sizes in the reference's own parameter layouts:
 *
 *   x = [ points 3P | images 6I | cameras d*C ]   (a block is absent when fixed)
 *       camera_shared_vector_function.hpp:349-354, 364-410
 *   y = [ keys 2N | point constraints | image constraints | camera constraints ]
 *       camera_shared_vector_function.hpp:356-362, 483-629
 *   feature_maps[k] = (image_id, point_id)      camera_shared_vector_function.hpp:36, 433-435
 *   image_camera_map[image] = camera group id   camera_shared_vector_function.hpp:39, 435
 *
 * Residuals follow the reference sign r = f(x) - y
 * (unit_test/bundle_adjustment/test_camera_shared_normal_equation_builder.hpp:463),
 * the step solves (N + mu I) delta = g with g = J^T Sigma^-1 r and the update
 * is x_new = x - delta.
 *
 * Return codes: 0 = ok, negative = failure.  The reference uses
 * `typedef int Err` with 0 / -1 (camera_shared_normal_equation_solver.hpp:455-464);
 * callers that want exactly that map every nonzero code to -1.
 *
 * Device state lives behind the opaque handles; no device pointer crosses
 * this boundary.  A handle is used by one host thread at a time.
 */
#ifndef HSBA_H_
#define HSBA_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ------------------------------------------------------ */
#define HSBA_OK                 0
#define HSBA_ERR_GENERIC       -1  /* reference's only failure value       */
#define HSBA_ERR_INVALID       -2  /* bad sizes / ids / masks              */
#define HSBA_ERR_NOT_SPD       -3  /* factorization met a non-positive pivot */
#define HSBA_ERR_NAN           -4  /* non-finite value in the iteration    */
#define HSBA_ERR_OOM           -5  /* device allocation failed             */
#define HSBA_ERR_DEVICE        -6  /* HIP runtime error                    */
#define HSBA_ERR_ADMISSION     -7  /* job larger than the server's HBM budget */
#define HSBA_ERR_UNSUPPORTED   -8  /* option combination not implemented   */
#define HSBA_ERR_NOT_READY     -9  /* server job still running             */
#define HSBA_ERR_TOO_LARGE    -10  /* beyond the fixed scratch of an allocation-free entry point, or
                                       * beyond the size limit of the dense covariance route */

/* ---- masks (bit i == reference enum value i) ---------------------------- */
/* IntrinsicComputationsMask, camera_shared_common_types.hpp:14-22 */
#define HSBA_COMPUTE_RADIAL_DISTORTION       (1u << 0)
#define HSBA_COMPUTE_DECENTERING_DISTORTION  (1u << 1)
#define HSBA_COMPUTE_INTRINSIC_PARAMS        (1u << 2)
/* FixMask, camera_shared_common_types.hpp:97-104 */
#define HSBA_FIX_POINTS   (1u << 0)
#define HSBA_FIX_IMAGES   (1u << 1)
#define HSBA_FIX_CAMERAS  (1u << 2)

/* Full per-camera intrinsic vector, in the reference's camera-block column
 * order [k1 k2 k3 | d1 d2 | f skew px py pr]
 * (camera_shared_analytical_jacobian_matrix_calculator.hpp:229-271). */
enum {
  HSBA_K1 = 0, HSBA_K2, HSBA_K3, HSBA_D1, HSBA_D2,
  HSBA_FOCAL, HSBA_SKEW, HSBA_PRINCIPAL_X, HSBA_PRINCIPAL_Y, HSBA_PIXEL_RATIO,
  HSBA_NUM_INTRINSICS
};

/* Constraint records (camera_shared_common_types.hpp:24-95).
 *   point : mask bits 0..2 = POINT_CONSTRAIN_{X,Y,Z}
 *   image : mask bits 0..3 = IMAGE_CONSTRAIN_{ROTATION,POSITION_X,_Y,_Z}
 *   camera: mask bits 0..2 = radial {K1,K2,K3}, bits 3..4 = decentering
 *           {D1,D2}, bits 5..9 = intrinsic {F,SKEW,PX,PY,PR}
 */
typedef struct {
  int64_t  id;
  uint32_t mask;
  uint32_t reserved_;
} hsba_constraint;

/* Problem description: the state of CameraSharedVectorFunction
 * (camera_shared_vector_function.hpp:20-410). */
typedef struct {
  int64_t num_points;
  int64_t num_images;
  int64_t num_cameras;
  int64_t num_keys;
  const int32_t *feature_maps;      /* [num_keys][2] = (image_id, point_id) */
  const int32_t *image_camera_map;  /* [num_images] */
  uint32_t intrinsic_computations_mask;
  /* Extension with no reference counterpart: which parameters of the enabled
   * groups are free, as bits over the HSBA_K1..HSBA_PIXEL_RATIO order.  0
   * means "all parameters of the enabled groups", which is exactly the
   * reference layout.  Held parameters take their value from
   * intrinsic_defaults.  Used for the 9-parameter (r, t, f, k1, k2) camera. */
  uint32_t intrinsic_free_mask;
  const double *intrinsic_defaults; /* [num_cameras][10] or NULL (zeros, f=pr=1) */
  uint32_t fix_mask;
  uint32_t reserved_;
  const double *fix_points;         /* [P][3] when HSBA_FIX_POINTS  */
  const double *fix_images;         /* [I][6] when HSBA_FIX_IMAGES  */
  const double *fix_cameras;        /* [C][d] when HSBA_FIX_CAMERAS, d = free-parameter count */
  int64_t num_point_constraints;
  const hsba_constraint *point_constraints;
  int64_t num_image_constraints;
  const hsba_constraint *image_constraints;
  int64_t num_camera_constraints;
  const hsba_constraint *camera_constraints;
} hsba_problem_desc;

/* Inverse observation covariance: CameraSharedYCovarianceInverse
 * (camera_shared_y_covariance_inverse.hpp:18-149). */
typedef struct {
  const double *key_blocks;         /* [num_keys][2][2] row-major, NULL = identity */
  const double *constraint_weights; /* [num constraint rows], NULL = all 1 */
  /* Full inverse covariance of each point constraint, [num_point_constraints][3][3]
   * row-major and symmetric, or NULL.  When given, every point constraint must
   * constrain all three coordinates; its three rows then take this block and
   * their constraint_weights entries are ignored.  This is the GCP weighting of
   * BAGCPConstrainedYCovarianceInverse::gcp_blocks
   * (ba_gcp_constrained_y_covariance_inverse.hpp:14-23); camera_shared itself
   * only has diagonal constraint weights. */
  const double *point_constraint_blocks;
} hsba_ycov_inverse;

#define HSBA_PRECISION_FP64   0     /* every kernel in FP64 ("GPU-plain")   */
#define HSBA_PRECISION_MIXED  1     /* FP32 J / Hessian / Schur, FP64 factor, cost, gain */

#define HSBA_SOLVER_AUTO   0
#define HSBA_SOLVER_DENSE  1        /* dense FP64 Cholesky of the reduced system */
#define HSBA_SOLVER_PCG    2        /* block-Jacobi PCG on the implicit Schur complement */

/* Loss on each key's squared whitened residual s = r^T Sigma^-1 r.  The
 * in-house LM path has none (TRIVIAL: rho(s) = s); the production caller
 * (sfm_pipeline/bundle_adjustment_optimizor.hpp) runs Ceres with
 * ceres::HuberLoss(1.0) per key (camera_shared_ceres_optimizor.hpp:230-481).
 * HUBER with scale a: rho(s) = s for s <= a^2, 2 a sqrt(s) - a^2 beyond, so the
 * cost is sum rho(s) (no 1/2, like the rest of this ABI).  Its second
 * derivative is never positive, so the robustified Gauss-Newton model
 * (Triggs et al. 2000, "Bundle adjustment — a modern synthesis", sec. 4.3)
 * reduces to reweighting: each key's residual and Jacobian rows are scaled by
 * sqrt(rho'(s)).  Constraint rows are never robustified. */
#define HSBA_LOSS_TRIVIAL  0
#define HSBA_LOSS_HUBER    1

typedef struct {
  int32_t precision;                /* HSBA_PRECISION_* */
  int32_t solver;                   /* HSBA_SOLVER_*    */
  double  pcg_rtol;                 /* relative residual tolerance (default 1e-6) */
  int32_t pcg_max_iterations;       /* default 500 */
  int32_t point_group_lanes;        /* lanes per point in point kernels: 0 = by each point's track
                                       length in the linearization (2 / 4 / 8), 8 elsewhere */
  int32_t loss;                     /* HSBA_LOSS_* (0 = trivial)             */
  int32_t pair_kernel;              /* HSBA_PAIRS_* (0 = chosen by the autotuner) */
  double  loss_scale;               /* a of the Huber loss (0 = 1.0)         */
} hsba_options;

/* Camera-pair Schur kernel (no reference counterpart).  TUNED: the library's
 * autotuner times both kernels on the first problem of each window class
 * (precision, camera block width, incidences per pair) and keeps the faster
 * one for every later window of that class (hsba_tuning_table). */
#define HSBA_PAIRS_TUNED  0
#define HSBA_PAIRS_VALU   1         /* one wavefront per pair, register tiles */
#define HSBA_PAIRS_MFMA   2         /* one workgroup per pair, matrix cores (camera block <= 15) */
#define HSBA_PAIRS_MFMA_WAVE 3      /* one wavefront per pair, matrix cores (camera block <= 15) */

/* Optimizer settings: CameraSharedLevenbergMarquardtOptimizor ctor
 * (camera_shared_levenberg_marquardt_optimizor.hpp:36-41). */
typedef struct {
  int32_t max_iterations;
  int32_t reserved_;
  double  tau;                      /* mu0 = tau * max diag(N)  */
  double  eps1;                     /* stop if ||g||_inf <= eps1 */
  double  eps2;                     /* stop if ||delta|| <= eps2 (||x|| + eps2) */
} hsba_lm_options;

#define HSBA_STOP_MAX_ITERATIONS  0
#define HSBA_STOP_GRADIENT        1
#define HSBA_STOP_STEP            2
#define HSBA_STOP_FAILED          3

typedef struct {
  int32_t status;                   /* HSBA_OK or a negative code            */
  int32_t stop_reason;              /* HSBA_STOP_*                           */
  int32_t iterations;               /* LM iterations run (accepted + rejected) */
  int32_t accepted;
  double  initial_cost;             /* r^T Sigma^-1 r at x0 (no 1/2, mahalanobis_distance_calculator.hpp:28-61) */
  double  final_cost;
  double  final_mu;
  int64_t pcg_iterations;           /* summed over LM iterations (PCG solver) */
} hsba_lm_stats;

typedef struct hsba_problem hsba_problem;

/* Library / device information. */
const char *hsba_version(void);
int hsba_device_count(int *count);

/* Builds the device-resident problem: observation orderings, point and image
 * CSR, camera-pair incidence lists, reduced-system layout.  Mirrors the
 * sparsity pass of CameraSharedNormalEquationBuilder::ComputeNormalMatrix
 * (camera_shared_normal_equation_builder.hpp:137-205). */
int hsba_problem_create(const hsba_problem_desc *desc,
                        const hsba_options *options,   /* NULL = defaults */
                        hsba_problem **problem);
int hsba_problem_destroy(hsba_problem *problem);

/* Sizes: GetXSize / GetYSize (camera_shared_vector_function.hpp:349-362). */
int hsba_problem_sizes(const hsba_problem *problem,
                       int64_t *x_size, int64_t *y_size,
                       int64_t *camera_params, int64_t *reduced_size);

/* CameraSharedVectorFunction::operator() (camera_shared_vector_function.hpp:412-632). */
int hsba_evaluate(hsba_problem *problem, const double *x, double *y);

/* CameraSharedAnalyticalJacobianMatrixCalculator::operator()
 * (camera_shared_analytical_jacobian_matrix_calculator.hpp:66-332).
 * Blocks in key order, row-major: point [N][2][3], image [N][2][6],
 * camera [N][2][d].  A pointer may be NULL to skip that block. */
int hsba_jacobian(hsba_problem *problem, const double *x,
                  double *point_blocks, double *image_blocks,
                  double *camera_blocks);

/* CameraSharedNormalEquationBuilder::operator()
 * (camera_shared_normal_equation_builder.hpp:92-462) at x with observations
 * y_obs.  Keeps N and g resident for hsba_solve and also copies out:
 *   point_blocks  [P][3][3]   image_blocks [I][6][6]   camera_blocks [C][d][d]
 *   gradient      [x_size]    (reference x layout)
 * Any output pointer may be NULL. */
int hsba_build_normal(hsba_problem *problem, const double *x,
                      const double *y_obs, const hsba_ycov_inverse *ycov,
                      double *point_blocks, double *image_blocks,
                      double *camera_blocks, double *gradient);

/* Sparsity of the normal equations and of the reduced camera system, as built
 * at hsba_problem_create (host memory only; every index is exact).
 *
 * Coupling sets of N in the reference's block order
 * (CameraSharedNormalEquationBuilder::ComputeNormalMatrix,
 * camera_shared_normal_equation_builder.hpp:137-205; block storage
 * camera_shared_normal_matrix.hpp:315-443):
 *   point_image  [num_point_image][2]  (point, image) of every key, in key
 *                order (block id = key + 1, :141-153); present when points
 *                and images are free
 *   point_camera [num_point_camera][2] distinct (point, camera), in the order
 *                in which they first appear in key order (the std::set
 *                insertion with value size()+1, :157-179)
 *   image_camera [num_image_camera][2] distinct (image, camera), likewise
 *                (:180-204)
 * Device layout of the keys and of the Schur complement (no reference
 * counterpart; the reference loops over all image pairs and points):
 *   key_order    [num_keys]      caller key index of each point-major slot
 *                                (a point's keys contiguous, images ascending)
 *   point_ptr    [num_points+1]  slots of point p: point_ptr[p] .. point_ptr[p+1]
 *   pairs        [num_pairs][2]  image pairs (a <= b) sharing a point, in
 *                                (a, b) order: the nonzero image blocks S_ab
 *                                of camera_shared_normal_equation_solver.hpp:198-268
 *   pair_ptr     [num_pairs+1]   incidences of pair q: pair_ptr[q] .. pair_ptr[q+1]
 *   incidences   [num_incidences][2]  (caller key on image a, caller key on
 *                                image b) of every point seeing both, in point order
 * Any array may be NULL; sizes is always filled.  Under the PCG solver only
 * the diagonal pairs (a, a) exist. */
typedef struct {
  int64_t num_keys;
  int64_t num_points;
  int64_t num_point_image;
  int64_t num_point_camera;
  int64_t num_image_camera;
  int64_t num_pairs;
  int64_t num_incidences;
} hsba_structure_sizes;
int hsba_problem_structure(const hsba_problem *problem, hsba_structure_sizes *sizes,
                           int32_t *key_order, int32_t *point_ptr, int32_t *point_image,
                           int32_t *point_camera, int32_t *image_camera, int32_t *pairs,
                           int64_t *pair_ptr, int32_t *incidences);

/* Coupling blocks of N at the last hsba_build_normal, in the order of the
 * sets of hsba_problem_structure (camera_shared_normal_equation_builder.hpp:249-307):
 *   point_image  [num_point_image][3][6]   W_pi = J_p^T Sigma^-1 J_i of each key
 *   point_camera [num_point_camera][3][d]  W_pc summed over the point's keys on
 *                                          the camera's images
 *   image_camera [num_image_camera][6][d]  W_ic summed over the image's keys
 * Any pointer may be NULL. */
int hsba_coupling_blocks(hsba_problem *problem, double *point_image, double *point_camera,
                         double *image_camera);

/* CameraSharedNormalEquationSolver::operator() on the normal equations of the
 * last hsba_build_normal, damped by mu (camera_shared_normal_equation_solver.hpp:75-114). */
int hsba_solve(hsba_problem *problem, double mu, double *delta_x);

/* Reduced camera system of the last build, damped by mu, dense row-major
 * [reduced_size][reduced_size] and rhs [reduced_size]
 * (ComputeReducedSystem, camera_shared_normal_equation_solver.hpp:116-446). */
int hsba_reduced_system(hsba_problem *problem, double mu,
                        double *reduced_matrix, double *reduced_rhs);

/* Posterior covariance at x: CameraSharedAnalyticalXCovarianceCalculator
 * (camera_shared_analytical_x_covariance_calculator.hpp:45-90) forms the dense
 * N = J^T Sigma^-1 J and its pseudo-inverse by SVD.  Here the diagonal blocks
 * of N^-1 come from the Schur complement on the device (S = U - W^T V^-1 W is
 * factored and inverted, each point block is L_p^-T (I + Z_p S^-1 Z_p^T) L_p^-1),
 * which equals the reference's pseudo-inverse when N is nonsingular, i.e. when
 * the gauge is fixed (constraints, GCPs or fixed blocks).  A singular N (free
 * gauge) fails with HSBA_ERR_NOT_SPD.  Outputs (any may be NULL): point blocks
 * [P][3][3], image blocks [I][6][6], camera blocks [C][d][d], and the dense
 * inverse of the reduced system [reduced_size][reduced_size] (image and camera
 * block of N^-1).  y_obs may be NULL (N depends on y only through a robust
 * loss).  Dense-solver problems only (HSBA_ERR_UNSUPPORTED under PCG). */
int hsba_covariance(hsba_problem *problem, const double *x, const double *y_obs,
                    const hsba_ycov_inverse *ycov, double *point_blocks, double *image_blocks,
                    double *camera_blocks, double *reduced);

/* hsba_covariance without any device allocation, for callers that cannot
 * call hipMalloc (code under HIP graph capture, tight serving loops,
 * harnesses that forbid it).  No counterpart in the reference.  The same
 * kernels and results as hsba_covariance; the device buffers of the inverse
 * (L^-T and S^-1 of the reduced system, the point blocks) come from one fixed
 * scratch region per GPU (8.5 MiB, held until the process ends).  Limits: a
 * reduced system of at most 512 rows after padding to whole 64-row tiles and
 * at most 65536 points; beyond them it returns HSBA_ERR_TOO_LARGE and does
 * nothing.  Every call on one GPU shares that GPU's region: calls that overlap
 * on one GPU (several host threads, or problems on different streams) race and
 * must be serialised by the caller. */
int hsba_covariance_static(hsba_problem *problem, const double *x, const double *y_obs,
                           const hsba_ycov_inverse *ycov, double *point_blocks, double *image_blocks,
                           double *camera_blocks, double *reduced);

/* Full dense posterior covariance at x: the output of
 * CameraSharedAnalyticalXCovarianceCalculator::operator()
 * (camera_shared_analytical_x_covariance_calculator.hpp:45-90), the
 * pseudo-inverse of the dense normal matrix N = J^T Sigma^-1 J in which
 * singular values at or below 1e-6 are dropped (:73-87), written row-major
 * into x_covariance[x_size][x_size] in the x layout (points | images |
 * cameras; every off-diagonal block included).  Routes:
 *   HSBA_COV_SCHUR  N^-1 through the Schur complement on the device:
 *                   S^-1 from the factored reduced system,
 *                   C_pr = -V^-1 W S^-1 and C_pp = V^-1 + V^-1 W S^-1 W^T V^-1
 *                   as matrix-core GEMMs.  Equal to the pseudo-inverse when
 *                   every singular value of N exceeds 1e-6.
 *   HSBA_COV_DENSE  the dense N eigen-decomposed on the device by the cyclic
 *                   Jacobi method (for symmetric N the SVD, as the reference
 *                   takes it), any N, e.g. a free gauge; O(x_size^3) per sweep,
 *                   x_size at most HSBA_COV_DENSE_MAX (HSBA_ERR_TOO_LARGE
 *                   beyond).
 *   HSBA_COV_AUTO   factors N - 1e-6 I (point blocks and reduced system); when
 *                   it is positive definite every singular value exceeds 1e-6
 *                   and the Schur route gives the pseudo-inverse, else the
 *                   dense route runs.
 * info (may be NULL) reports the route taken, the number of singular values
 * kept, the Jacobi sweeps and the smallest kept singular value (dense route).
 * y_obs may be NULL.  Dense-solver problems only (HSBA_ERR_UNSUPPORTED under
 * PCG). */
#define HSBA_COV_AUTO      0
#define HSBA_COV_SCHUR     1
#define HSBA_COV_DENSE     2
#define HSBA_COV_DENSE_MAX 4096
typedef struct hsba_covariance_info {
  int32_t method;        /* HSBA_COV_SCHUR or HSBA_COV_DENSE */
  int32_t rank;          /* singular values kept (x_size on the Schur route) */
  int32_t sweeps;        /* Jacobi sweeps of the dense route */
  int32_t reserved_;
  double smallest_kept;  /* smallest kept singular value (dense route), else 0 */
} hsba_covariance_info;
int hsba_x_covariance(hsba_problem *problem, const double *x, const double *y_obs,
                      const hsba_ycov_inverse *ycov, int method, double *x_covariance,
                      hsba_covariance_info *info);

/* LM scalar helpers on the last build:
 *   max diag(N)            camera_shared_normal_matrix_max_diagonal_value_calculator.hpp:26-38
 *   ||g||_inf              camera_shared_gradient_infinite_norm_calculator.hpp:27-39
 *   g^T delta              camera_shared_gradient_delta_x_dot_product_calculator.hpp:26-38
 *   r^T Sigma^-1 r at x    camera_shared_mahalanobis_distance_calculator.hpp:28-61 */
int hsba_max_diagonal(hsba_problem *problem, double *value);
int hsba_gradient_inf_norm(hsba_problem *problem, double *value);
int hsba_gradient_dot(hsba_problem *problem, const double *delta_x, double *value);
int hsba_mahalanobis(hsba_problem *problem, const double *x, const double *y_obs,
                     const hsba_ycov_inverse *ycov, double *value);

/* CameraSharedLevenbergMarquardtOptimizor::operator()
 * (camera_shared_levenberg_marquardt_optimizor.hpp:43-58) with the LM loop of
 * hs::optimizor::nllso::LevenbergMarquardt run entirely on the device. */
int hsba_lm_solve(hsba_problem *problem, const double *x0, const double *y_obs,
                  const hsba_ycov_inverse *ycov, const hsba_lm_options *lm,
                  double *x_out, hsba_lm_stats *stats);

/* Stepwise form of hsba_lm_solve: begin uploads x0 / y and runs the initial
 * cost, linearization and mu0; iterate runs `iterations` more LM iterations
 * (one captured hipGraph replay each, no host round trip); end copies x and
 * the statistics out.  No reference counterpart: used for timing and by the
 * serving loop. */
int hsba_lm_begin(hsba_problem *problem, const double *x0, const double *y_obs,
                  const hsba_ycov_inverse *ycov, const hsba_lm_options *lm);
/* Per-stage device time of one LM iteration, from HIP event nodes captured
 * around each stage on the library's stream (milliseconds, averaged over the
 * iterations of one hsba_lm_iterate call). */
typedef struct {
  double linearize_ms;    /* K1+K2 residual, Jacobian, point blocks */
  double image_reduce_ms; /* image / camera blocks, constraints, gradient stats */
  double point_factor_ms; /* K3 point factor + Z = L^-1 W */
  double pairs_ms;        /* K4 camera-pair Schur products */
  double assemble_ms;     /* reduced matrix and right-hand side */
  double cholesky_ms;     /* K5 factorization + triangular solves */
  double backsub_ms;      /* K6 point back-substitution */
  double step_cost_ms;    /* K7 x - delta, candidate cost, LM control (K8) */
  double iteration_ms;    /* whole iteration */
  int32_t iterations;
  int32_t reserved_;
} hsba_stage_times;
int hsba_lm_iterate(hsba_problem *problem, int32_t iterations, hsba_stage_times *times);
int hsba_lm_end(hsba_problem *problem, double *x_out, hsba_lm_stats *stats);

/* The autotuner's decisions so far (process-wide).  No reference
 * counterpart. */
typedef struct {
  int32_t precision;                /* HSBA_PRECISION_* */
  int32_t camera_block;             /* E: image + camera parameters of a reduced block row */
  int32_t bucket;                   /* floor(log2(mean incidences per camera pair)) of the windows */
  int32_t variant;                  /* chosen pair kernel: 0 wavefront per pair (VALU), 1 workgroup per
                                       pair (matrix cores), 2 wavefront per pair (matrix cores) */
  double  ms[3];                    /* measured time of each candidate (0 = not a candidate) */
  int64_t pairs;                    /* camera pairs the candidates were timed on */
} hsba_tune_entry;
int hsba_tuning_table(hsba_tune_entry *entries, int32_t max_entries, int32_t *count);
/* Forgets every decision (the next problem of each class is timed again). */
int hsba_tuning_reset(void);

/* ---- ba_naive family ---------------------------------------------------- */
/* 6-DoF cameras in normalised image coordinates, x = [cameras 6C | points 3P]
 * (ba_naive_vector_function.hpp:117-135); the same device engine as the
 * camera_shared entry points with one image per camera and identity
 * intrinsics. */
typedef struct {
  int64_t num_cameras;
  int64_t num_points;
  int64_t num_features;
  const int32_t *feature_maps;      /* [num_features][2] = (camera_id, point_id) */
} hsba_naive_desc;

typedef struct hsba_naive_problem hsba_naive_problem;

int hsba_naive_problem_create(const hsba_naive_desc *desc, const hsba_options *options,
                              hsba_naive_problem **problem);
int hsba_naive_problem_destroy(hsba_naive_problem *problem);
/* BANaiveVectorFunction::operator() (ba_naive_vector_function.hpp:137-196). */
int hsba_naive_evaluate(hsba_naive_problem *problem, const double *x, double *y);
/* BANaiveAnalyticalJacobianMatrixCalculator::operator()
 * (ba_naive_analytical_jacobian_matrix_calculator.hpp:75-220): camera blocks
 * [F][2][6], point blocks [F][2][3], feature order.  HSBA_ERR_INVALID at an
 * exactly zero camera rotation, where the reference returns -1 (:104-108). */
int hsba_naive_jacobian(hsba_naive_problem *problem, const double *x,
                        double *camera_blocks, double *point_blocks);
/* BANaiveNormalEquationBuilder::operator() (ba_naive_normal_equation_builder.hpp:78-152):
 * camera blocks [C][6][6], point blocks [P][3][3], gradient [6C + 3P];
 * feature_cov_inverse = [F][2][2] (BANaiveFeatureCovarianceInverse) or NULL. */
int hsba_naive_build_normal(hsba_naive_problem *problem, const double *x, const double *y_obs,
                            const double *feature_cov_inverse, double *camera_blocks,
                            double *point_blocks, double *gradient);
/* BANaiveNormalEquationSolver::operator() (ba_naive_normal_equation_solver.hpp:75-254)
 * on the last build, damped by mu; delta in the naive x layout. */
int hsba_naive_solve(hsba_naive_problem *problem, double mu, double *delta_x);
/* BANaiveLevenbergMarquardtOptimizor (ba_naive_levenberg_marquardt_optimizor.hpp). */
int hsba_naive_lm_solve(hsba_naive_problem *problem, const double *x0, const double *y_obs,
                        const double *feature_cov_inverse, const hsba_lm_options *lm,
                        double *x_out, hsba_lm_stats *stats);

/* ---- ba_gcp_constrained family ------------------------------------------ */
/* ba_naive plus ground-control-point priors: the last num_gcps points are
 * GCPs whose coordinates are also observed, so
 *   y = [features 2F | GCP coordinates 3G]
 * (ba_gcp_constrained_vector_function.hpp:73-113) with a 3x3 inverse
 * covariance per GCP (gcp_blocks, [G][3][3], symmetric).  The GCP rows of the
 * Jacobian are identity rows on the GCP points
 * (ba_gcp_constrained_jacobian_matrix.hpp:50-71); the builder adds each block
 * to its point block and B r to its gradient segment
 * (ba_gcp_constrained_normal_equation_builder.hpp:59-72).  x is the naive
 * layout [cameras 6C | points 3P]. */
typedef struct {
  int64_t num_cameras;
  int64_t num_points;               /* including the GCPs */
  int64_t num_features;
  const int32_t *feature_maps;      /* [num_features][2] = (camera_id, point_id) */
  int64_t num_gcps;                 /* the last num_gcps points */
} hsba_gcp_desc;

typedef struct hsba_gcp_problem hsba_gcp_problem;

/* BAGCPConstrainedVectorFunction ctor (ba_gcp_constrained_vector_function.hpp:43-50). */
int hsba_gcp_problem_create(const hsba_gcp_desc *desc, const hsba_options *options,
                            hsba_gcp_problem **problem);
int hsba_gcp_problem_destroy(hsba_gcp_problem *problem);
/* BAGCPConstrainedVectorFunction::operator() (ba_gcp_constrained_vector_function.hpp:73-113):
 * y [2F + 3G]. */
int hsba_gcp_evaluate(hsba_gcp_problem *problem, const double *x, double *y);
/* BAGCPConstrainedAnalyticalJacobianMatrixCalculator::operator()
 * (ba_gcp_constrained_analytical_jacobian_matrix_calculator.hpp:39-49): the
 * naive feature blocks ([F][2][6] camera, [F][2][3] point); the GCP rows are
 * the implicit identity rows.  HSBA_ERR_INVALID at a zero camera rotation. */
int hsba_gcp_jacobian(hsba_gcp_problem *problem, const double *x,
                      double *camera_blocks, double *point_blocks);
/* BAGCPConstrainedNormalEquationBuilder::operator()
 * (ba_gcp_constrained_normal_equation_builder.hpp:42-76): camera blocks
 * [C][6][6], point blocks [P][3][3], gradient [6C + 3P]. */
int hsba_gcp_build_normal(hsba_gcp_problem *problem, const double *x, const double *y_obs,
                          const double *feature_cov_inverse, const double *gcp_blocks,
                          double *camera_blocks, double *point_blocks, double *gradient);
/* BANaiveNormalEquationSolver on the last build (the GCP family reuses the
 * naive solver, ba_gcp_constrained_nllso_meta_levenberg_marquardt.hpp:21). */
int hsba_gcp_solve(hsba_gcp_problem *problem, double mu, double *delta_x);
/* BAGCPConstrainedMahalanobisDistanceCalculator::operator()
 * (ba_gcp_constrained_mahalanobis_distance_calculator.hpp:33-62) at x. */
int hsba_gcp_mahalanobis(hsba_gcp_problem *problem, const double *x, const double *y_obs,
                         const double *feature_cov_inverse, const double *gcp_blocks,
                         double *value);
/* BAGCPConstrainedLevenbergMarquardtOptimizor::operator()
 * (ba_gcp_constrained_levenberg_marquardt_optimizor.hpp:44-62). */
int hsba_gcp_lm_solve(hsba_gcp_problem *problem, const double *x0, const double *y_obs,
                      const double *feature_cov_inverse, const double *gcp_blocks,
                      const hsba_lm_options *lm, double *x_out, hsba_lm_stats *stats);

/* ---- problem I/O --------------------------------------------------------- */
/* A host-side problem read from files: the description, x0 and observations,
 * ready for hsba_problem_create / hsba_lm_solve, plus what is needed to write
 * an optimised x back.  Host memory only; no device call. */
typedef struct hsba_scene hsba_scene;

/* The hs_sfm text files (sfm_file_io/ *_loader.hpp), whitespace separated. */
typedef struct {
  const char *tracks;               /* tracks_loader.hpp:24-67: n, then per track n_views (image key blunder)* */
  const char *const *keysets;       /* [num_keysets], keyset_loader.hpp:23-40: one file per image */
  int64_t num_keysets;
  const char *intrinsics;           /* intrinsic_params_set_loader.hpp:24-56: f skew px py pr */
  const char *extrinsics;           /* extrinsic_params_set_loader.hpp:24-47: angle-axis, centre */
  const char *points;               /* points_loader.hpp:24-41 */
  const char *image_intrinsic_map;  /* object_index_loader.hpp:21-44: 0 = none, else id + 1 */
  const char *image_extrinsic_map;
  const char *track_point_map;
} hsba_sfm_files;

/* Assembles the camera_shared problem exactly as
 * BundleAdjustmentOptimizor::operator() does (bundle_adjustment_optimizor.hpp:59-203):
 * tracks with a point, views on images with an extrinsic and not flagged as
 * blunders, keys in track order, x = [points | (angle-axis, -R C) | intrinsics
 * [k1 k2 k3 d1 d2 f skew px py pr] restricted to the free mask], y = the keys.
 * intrinsic_free_mask as in hsba_problem_desc (0 = every parameter of the
 * enabled groups). */
int hsba_scene_load_sfm(const hsba_sfm_files *files, uint32_t intrinsic_computations_mask,
                        uint32_t intrinsic_free_mask, hsba_scene **scene);
/* "Bundle Adjustment in the Large" text format (Agarwal et al. 2010): one
 * image and one camera group per BAL camera [angle-axis, t, f, k1, k2],
 * projection -P / P_z, observations centred on the principal point.  Mapped
 * onto camera_shared by the camera-frame flip diag(1, -1, -1) (so that
 * P_z > 0) and a negated v observation; free intrinsics {f, k1, k2}. */
int hsba_scene_load_bal(const char *path, hsba_scene **scene);

typedef struct {
  hsba_problem_desc desc;           /* pointers stay valid while the scene lives */
  const double *x0;
  const double *y_obs;
  int64_t x_size;
  int64_t y_size;
} hsba_scene_view;
int hsba_scene_get(const hsba_scene *scene, hsba_scene_view *view);
/* Writes x back into the loaded entities and saves them in the formats of
 * points_saver.hpp, extrinsic_params_set_saver.hpp and
 * intrinsic_params_set_saver.hpp (all points / images / cameras of the input
 * files; entities outside the bundle keep their input values).  Values are
 * written with 17 significant digits (the reference streams with the default
 * 6).  Any path may be NULL to skip that file. */
int hsba_scene_store_sfm(hsba_scene *scene, const double *x, const char *intrinsics,
                         const char *extrinsics, const char *points);
/* Writes the BAL file of a scene loaded with hsba_scene_load_bal at x. */
int hsba_scene_store_bal(hsba_scene *scene, const double *x, const char *path);
int hsba_scene_destroy(hsba_scene *scene);

/* ---- batched independent problems (local-BA windows) ------------------- */
/* Many independent problems solved together: one set of launches per LM
 * iteration covers every window.  No reference counterpart (the reference
 * solves one problem per optimizer call); each window's arithmetic is the
 * single-problem path's. */
typedef struct hsba_batch hsba_batch;

int hsba_batch_create(const hsba_problem_desc *descs, int64_t count,
                      const hsba_options *options, hsba_batch **batch);
int hsba_batch_destroy(hsba_batch *batch);
/* x0s / y_obs / x_outs: arrays of `count` host pointers, one per window. */
int hsba_batch_lm_solve(hsba_batch *batch, const double *const *x0s,
                        const double *const *y_obs,
                        const hsba_ycov_inverse *ycovs, /* [count] or NULL */
                        const hsba_lm_options *lm,
                        double *const *x_outs, hsba_lm_stats *stats /* [count] */);

/* Stepwise form of hsba_batch_lm_solve (as hsba_lm_begin / _iterate / _end
 * for one problem): per-stage device times of the whole batch, for timing.
 * No reference counterpart. */
int hsba_batch_lm_begin(hsba_batch *batch, const double *const *x0s, const double *const *y_obs,
                        const hsba_ycov_inverse *ycovs, const hsba_lm_options *lm);
int hsba_batch_lm_iterate(hsba_batch *batch, int32_t iterations, hsba_stage_times *times);
int hsba_batch_lm_end(hsba_batch *batch, double *const *x_outs, hsba_lm_stats *stats);

/* ---- serving loop (continuous batching) -------------------------------- */
typedef struct hsba_server hsba_server;

typedef struct {
  int64_t hbm_budget_bytes;         /* admission limit on resident job state (0 = 16 GiB) */
  int32_t max_active_jobs;          /* jobs admitted at once (0 = 4096)      */
  int32_t precision;                /* HSBA_PRECISION_*                      */
  hsba_lm_options lm;               /* per-job LM settings                   */
  int32_t streams;                  /* concurrent batches, one HIP stream each (0 = 4) */
  int32_t max_batch_jobs;           /* jobs per batch (0 = 256)              */
  int32_t loss;                     /* HSBA_LOSS_* for every job             */
  int32_t reserved2_;
  double  loss_scale;               /* Huber scale (0 = 1.0)                 */
} hsba_server_options;

typedef struct {
  int64_t job_id;
  int32_t status;                   /* HSBA_OK or negative                   */
  int32_t reserved_;
  double  submit_time_s;            /* server clock                          */
  double  admit_time_s;
  double  finish_time_s;
  hsba_lm_stats stats;
} hsba_job_result;

int hsba_server_create(const hsba_server_options *options, hsba_server **server);
int hsba_server_destroy(hsba_server *server);
/* Copies the job in; returns immediately. x_out must stay valid until the
 * job's result is polled. */
int hsba_server_submit(hsba_server *server, const hsba_problem_desc *desc,
                       const double *x0, const double *y_obs,
                       const hsba_ycov_inverse *ycov, double *x_out,
                       int64_t *job_id);
/* Fetches up to max_results finished jobs; returns how many in *count. */
int hsba_server_poll(hsba_server *server, hsba_job_result *results,
                     int32_t max_results, int32_t *count);
/* Blocks until every submitted job has finished. */
int hsba_server_drain(hsba_server *server);
double hsba_server_clock(const hsba_server *server);

#ifdef __cplusplus
}
#endif

#endif /* HSBA_H_ */
