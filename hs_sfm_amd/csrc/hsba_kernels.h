// hsba_kernels.h — kernel-set interface between the host driver and the
// templated HIP kernels (one instantiation per precision and camera block).
#pragma once

#include <vector>

#include "hsba_device.h"

namespace hsba {

template <typename T>
struct KernelSet {
  virtual ~KernelSet() {}
  virtual int E_() const = 0;
  virtual void linearize(const Dev<T> &d, hipStream_t st, const LinSeg &seg, const int32_t *plist) = 0;
  virtual void image_reduce(const Dev<T> &d, hipStream_t st) = 0;
  virtual void camera_reduce(const Dev<T> &d, hipStream_t st) = 0;
  virtual void constraints(const Dev<T> &d, hipStream_t st) = 0;
  virtual void point_factor(const Dev<T> &d, hipStream_t st) = 0;
  virtual void key_z(const Dev<T> &d, hipStream_t st) = 0;
  // K4 over the pairs plist[0 .. count) (plist == nullptr: pairs 0 .. count;
  // entries < 0 are empty slots).  variant 0: one wavefront per pair, VALU
  // register tiles; variant 1: one workgroup per pair, LDS-staged incidences
  // on the matrix cores; variant 2: one wavefront per pair on the matrix
  // cores (variants 1 and 2 for E <= 15).
  virtual void pairs(const Dev<T> &d, hipStream_t st, int variant, const int32_t *plist, int count) = 0;
  virtual bool has_pair_variant(int variant) const = 0;
  virtual void assemble(const Dev<T> &d, hipStream_t st, int64_t lower_total, int rows) = 0;
  virtual void backsub(const Dev<T> &d, hipStream_t st, int glog2) = 0;
  virtual void step(const Dev<T> &d, hipStream_t st) = 0;
  virtual void cost(const Dev<T> &d, hipStream_t st, int use_new) = 0;
  virtual void grad_stats(const Dev<T> &d, hipStream_t st) = 0;
  virtual void lm_control(const Dev<T> &d, hipStream_t st, int phase) = 0;
  virtual void commit(const Dev<T> &d, hipStream_t st) = 0;
  // rotation quantities of the current image rotations (after x is set)
  virtual void image_rotations(const Dev<T> &d, hipStream_t st) = 0;
  virtual void export_keys(const Dev<T> &d, hipStream_t st, double *y, double *jp, double *ji,
                           double *jc, int want_j) = 0;
  // implicit-Schur PCG
  virtual void pcg_setup(const Dev<T> &d, hipStream_t st, int glog2) = 0;
  virtual void pcg_iteration(const Dev<T> &d, hipStream_t st, int glog2) = 0;
  virtual void pcg_finish(const Dev<T> &d, hipStream_t st) = 0;
  virtual void pcg_params(double rtol, int max_it) = 0;
  // posterior covariance of every point, [P][3][3]: L_p^-T (I + Z_p S^-1 Z_p^T) L_p^-1
  virtual void point_cov(const Dev<T> &d, hipStream_t st, const double *Sinv, int64_t ld, int r_off,
                         double *out) = 0;
  // Dense point rows of the coupling block, out[(3 p + a) * ld + j] for the
  // reduced columns j (window-local, r_off subtracted) of each key's image and
  // camera: which = 0 W_p = J_p^T Lambda [J_i | J_c], which = 1
  // V_p^-1 W_p = L_p^-T Z_p.  Camera columns accumulate, so out starts at zero.
  virtual void point_rows(const Dev<T> &d, hipStream_t st, int which, double *out, int64_t ld, int r_off) = 0;
  // Point diagonal blocks at out[(3 p + a) * ld + 3 p + b]: which = 0 stores
  // V_p, which = 1 adds V_p^-1 = L_p^-T L_p^-1.
  virtual void point_diag(const Dev<T> &d, hipStream_t st, int which, double *out, int64_t ld) = 0;
};

template <typename T>
KernelSet<T> *make_kernel_set(int img, int dcam);

// What the large-path Cholesky launches need of one large window, indexed
// by blockIdx.y: one 40-byte record instead of the window id and then its
// WinInfo (two dependent loads at the head of every launch).
struct CholWin {
  int64_t s_off, dinv_off;
  int32_t ld, nb, n, r_off, win, pad_;
};

// Dense FP64 Cholesky of every window's reduced system (hsba_cholesky.hip).
struct CholPlan {
  int nwin;
  const WinInfo *wi;         // device
  WinState *ws;              // device
  double *S, *rhs, *dred, *Dinv;
  double *Lf;                // factor of large windows (same layout as S)
  double *bwork, *ywork;     // right-hand side being reduced / forward solution
  const WinInfo *wi_host;    // host copy (large-window launch loop)
  const CholWin *large_wins; // device records of the large windows, in large-list order
};
void chol_init_padding(const CholPlan &plan, hipStream_t st);
void chol_factor_solve_impl(const CholPlan &plan, const int *small_list_dev, int num_small,
                            const int *large_host, int num_large, hipStream_t st);
// S^-1 of window w from its factor (after chol_factor_solve_impl): Y (ld x ld
// scratch at the window's offset) receives L^-T, Sinv the full symmetric inverse.
void chol_inverse_impl(const CholPlan &plan, int w, bool large, double *Y, double *Sinv, hipStream_t st);

// Dense FP64 building blocks (hsba_dense.hip).
// C[m][n] (m < M, n < N, pitch ldc) = alpha opA(m, k) opB(k, n) (+ C), with
// opA(m, k) = ta ? A[k * lda + m] : A[m * lda + k] and
// opB(k, n) = tb ? B[n * ldb + k] : B[k * ldb + n]; the operands must hold
// whole 64-row / 64-column tiles and K must be a multiple of 16 (zero padding).
void dense_gemm(bool ta, bool tb, int64_t M, int64_t N, int64_t K, double alpha, const double *A, int64_t lda,
                const double *B, int64_t ldb, double *C, int64_t ldc, bool accumulate, hipStream_t st);
// dst[i][j] for i < rows, j < cols: mode 0 src[i][j], mode 1 src[j][i],
// mode 2 the symmetric matrix of src's lower triangle.
void dense_copy(const double *src, int64_t lds, double *dst, int64_t ldd, int64_t rows, int64_t cols, int mode,
                hipStream_t st);
// Pseudo-inverse of the symmetric n x n matrix held in A (pitch ld, a
// multiple of 64, zero beyond n; n <= 4096, else HSBA_ERR_TOO_LARGE) by the
// cyclic Jacobi eigenvalue method: eigenvalues with |lambda| > thr are
// inverted, the others dropped (for a symmetric matrix this is the SVD
// pseudo-inverse with singular values |lambda|); out (pitch ldo) receives the
// result, lam the eigenvalues.  A is overwritten; Vt (ld x ld),
// rotations_dev (one int) and work_dev (n + 1 + n doubles) are work space.
// HSBA_ERR_GENERIC when max_sweeps sweeps do not converge.
int dense_pinv(double *A, double *Vt, int64_t n, int64_t ld, double thr, double *out, int64_t ldo, int max_sweeps,
               int *rotations_dev, double *work_dev, std::vector<double> &lam, int *sweeps, hipStream_t st);

// Fixed device scratch of the allocation-free covariance entry point
// (hsba_covariance_static): L^-T and S^-1 of one reduced system of at most
// kScratchMaxLd padded rows, then the point blocks of at most
// kScratchMaxPoints points.  8.5 MiB per GPU, held until the process ends.
constexpr int64_t kScratchMaxLd = 512;
constexpr int64_t kScratchMaxPoints = 65536;
constexpr int64_t kScratchDoubles = 2 * kScratchMaxLd * kScratchMaxLd + 9 * kScratchMaxPoints;
// Device address of the current GPU's scratch (looked up once per device and
// cached); 0 on success.
int scratch_region(double **ptr);

}  // namespace hsba
