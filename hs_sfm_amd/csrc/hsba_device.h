// hsba_device.h — device-side types and the per-key projection / Jacobian
// shared by every kernel of the MI355X bundle adjuster.
//
// Projection: CameraSharedVectorFunction::WorldPointToImageKey
//   (include/hs_sfm/bundle_adjustment/camera_shared_vector_function.hpp:634-744).
// Jacobian: CameraSharedAnalyticalJacobianMatrixCalculator
//   (include/hs_sfm/bundle_adjustment/camera_shared_analytical_jacobian_matrix_calculator.hpp:94-557),
//   evaluated in FP64 registers whatever the storage precision.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hsba {

constexpr int kFull = 10;  // k1 k2 k3 d1 d2 f skew px py pr
enum { IK1 = 0, IK2, IK3, ID1, ID2, IF, ISKEW, IPX, IPY, IPR };

// Per-window LM state, FP64; one record per independent problem.
struct WinState {
  double mu, nu, cost, cost_new, rho;
  double dnorm2, xnorm2, gain_den, ginf, maxdiag;
  double tau, eps1, eps2, cost0;
  int32_t iter, max_iter, accepted, status;
  int32_t stop_reason, active, need_lin, trial;
  int32_t take_step, step_accepted, chol_fail, pcg_iters;
  int32_t pcg_active, pad2_[3];
};

// Per-window state of the preconditioned conjugate gradient solver.
struct PcgState {
  double rz, pap, alpha, beta, bnorm2, rr, rtol2, pad_;
  int32_t it, done, max_it, pad2_;
};

// Static per-window layout (host-built).
struct WinInfo {
  int64_t key_off, s_off, dinv_off, blk_off, lo_off;
  int32_t pt_off, img_off, cam_off, num_pts;
  int32_t num_imgs, num_cams, num_keys, n;      // n = reduced size
  int32_t ld, nb, cb, r_off;                    // ld = padded n, nb = ld/64, cb = image rows, r_off = global row base
  int32_t con_begin, con_end, small, pad_;
};

struct ConstraintRow {  // y-order constraint row
  int32_t kind;         // 0 point, 1 image, 2 camera
  int32_t entity;       // global point / image / camera id
  int32_t param;        // index inside the entity's x block
  int32_t win;
  double weight;        // diagonal weight (block == 0)
  double y;             // observed value
  // block != 0: one of the three rows of a point constraint weighted by a
  // full 3x3 block; brow = row `param` of the block, the other two rows of
  // the record are the neighbouring entries (index - param + j).
  double brow[3];
  int32_t block;
  int32_t rec_mask;     // coordinate mask of the point constraint record this row belongs to
};

struct Contrib {        // one term of a reduced-system block
  int32_t kind;         // 0 diag image, 1 diag camera, 2 image-camera, 3 pair
  int32_t id;           // image / camera / pair id (global)
  int16_t ro, so;       // sub-block offsets inside the pair block
  int32_t trans;        // pair block used transposed
};

// Row stride of the per-key Z blocks: the 3 x E values of Z_k, then the
// point's q_p (3 values, read by the matrix-core pair kernels as the
// right-hand-side column), rounded up to whole 128-byte cache lines, so a
// gathered block touches as few lines as possible and loads as 16-byte
// vectors.
__host__ __device__ constexpr int z_stride(int e, int tbytes) {
  return ((3 * e + 3) * tbytes + 127) / 128 * (128 / tbytes);
}

// Per-image quantities of the angle-axis rotation (rot_quantities).
struct RotQ {
  double k0, k1, k2, ct, st, s, c, ds, dc, mst, theta, theta2;
};

// Keys per image chunk of the image-block reduction.
constexpr int kImgChunk = 512;

// One incidence of a camera pair: the keys a (first image) and b (second image)
// of one point, as a << 32 | b.  64 bits, so neither the track length nor the
// number of keys of a problem is limited by the code.
__host__ __device__ constexpr uint64_t inc_code(int64_t a, int64_t b) { return ((uint64_t)a << 32) | (uint64_t)b; }
__host__ __device__ constexpr int inc_a(uint64_t c) { return (int)(c >> 32); }
__host__ __device__ constexpr int inc_b(uint64_t c) { return (int)(c & 0xffffffffu); }

// Huber loss on a key's squared whitened residual s (scale a):
// rho(s) = s for s <= a^2, 2 a sqrt(s) - a^2 beyond; rho'(s) = min(1, a / sqrt(s)).
// rho'' <= 0 everywhere, so the robustified Gauss-Newton model is plain
// reweighting: residual and Jacobian rows times sqrt(rho'(s)).
__host__ __device__ inline double robust_cost(int loss, double a, double s) {
  if (loss == 1 /* HSBA_LOSS_HUBER */ && s > a * a) return 2.0 * a * sqrt(s) - a * a;
  return s;
}
__host__ __device__ inline double robust_row_scale(int loss, double a, double s) {
  if (loss == 1 /* HSBA_LOSS_HUBER */ && s > a * a) return sqrt(a / sqrt(s));
  return 1.0;
}

// Thread segments of one K1 launch: segment s covers threads [t0[s], t0[s+1])
// (multiples of 64, so a wave lies in one segment), 2^g[s] lanes per point,
// points plist[base[s] .. base[s] + cnt[s]) (plist null: point = index).
struct LinSeg {
  int32_t t0[4], g[3], base[3], cnt[3];
};

// All device buffers, passed by value to kernels.
template <typename T>
struct Dev {
  int32_t P, I, C, N, nwin, npairs, ncons, nchunks;
  int32_t pts_free, imgs_free, cams_free, has_cov;
  int32_t pcg_rows;
  int32_t loss;           // HSBA_LOSS_*: robust loss on each key's squared whitened residual
  int32_t fidx[kFull];    // free intrinsic indices (first D valid)
  double loss_scale;      // Huber a
  int64_t lower_max;      // the most lower-triangle entries of one window's reduced matrix
  // keys, point-major order
  const int32_t *key_img, *key_pt, *key_orig;
  const double *key_uv;   // [N][2]
  const double *key_cov;  // [N][3] = (l00, l10, l11): L^T whitening factor, or null
  // points
  const int32_t *pt_ptr, *pt_win;
  double *pts, *pts_new, *dpt;
  T *V, *gp, *Linv, *q;
  // images
  const int32_t *img_ptr, *img_keys, *img_win, *img_cam, *img_row;
  const int32_t *ich_img, *ich_begin, *ich_end, *img_ch_ptr;  // image-key chunks
  int32_t nich, pad_ich_;
  T *ipart;               // [chunks][E(E+1)/2 + E] partial image blocks
  double *imgs, *imgs_new;
  RotQ *imgq, *imgq_new;   // [I] rot_quantities of imgs / imgs_new
  T *Dimg, *gr, *rp;
  // cameras
  const int32_t *cam_ptr, *cam_imgs, *cam_win, *cam_row;
  double *cam_free, *cam_free_new, *cam_full, *cam_full_new;
  T *Uc, *gc;
  // per-key trial data
  T *W, *Z;
  // PCG mode: the Z rows again in image-major order (row t = key img_keys[t]);
  // key_ipos = each key's row there, ikey_pt = each such row's point
  T *Zi;
  const int32_t *key_ipos, *ikey_pt;
  // camera pairs
  const int64_t *pair_ptr;
  const uint64_t *inc;      // pair incidences: key a << 32 | key b
  const int32_t *pair_win, *pair_img;  // pair_img >= 0: diagonal pair of that image
  T *pair_out;
  // windows
  WinState *ws;
  const WinInfo *wi;
  const int32_t *win_of_chunk;
  // reduced systems
  double *S, *rhs, *dred, *Dinv;
  const int64_t *blk_ptr;
  const Contrib *contrib;
  const int32_t *row_blk;   // [rows] global: block index within window
  const int32_t *row_sub;   // [rows] offset inside block
  // constraints
  const ConstraintRow *cons;
  // reductions: chunk table and partials
  const int32_t *chunk_kind, *chunk_begin, *chunk_end, *chunk_win;
  double *partial;          // [nchunks][4]
  // PCG work vectors (global reduced space)
  double *pv, *pr, *pz, *pp, *pAp, *Minv;
  T *tpt;                   // unused
  double *tpd;              // [P][3] point-space intermediate of the implicit matvec
  double *pdot;             // [I][3] per-image partials of p.Ap, r.z, r.r (summed per window in image order)
  PcgState *pcg;
};

// ---------------------------------------------------------------- math
__device__ __forceinline__ double sel10(const double a[kFull], int j) {
  double v = a[0];
#pragma unroll
  for (int t = 1; t < kFull; ++t) v = (j == t) ? a[t] : v;
  return v;
}

// Per-image quantities of the angle-axis rotation r (theta = |r|): the unit
// axis k (zero for theta = 0), cos / sin of theta and the Rodrigues
// derivative coefficients s = sin(theta)/theta, c = (1 - cos(theta))/theta^2,
// ds, dc, mst = -sin(theta)/theta, with a Taylor branch below theta = 1e-4
// (the reference divides by theta).  Every key of the image uses the same
// values, so they are formed once per image (k_img_rot, k_step) and read by
// the per-key kernels (struct RotQ).
__device__ __forceinline__ RotQ rot_quantities(const double r[3]) {
  RotQ q;
  const double r0 = r[0], r1 = r[1], r2 = r[2];
  q.theta2 = r0 * r0 + r1 * r1 + r2 * r2;
  q.theta = sqrt(q.theta2);
  sincos(q.theta, &q.st, &q.ct);
  if (q.theta == 0.0) {
    q.k0 = q.k1 = q.k2 = 0.0;
  } else {
    q.k0 = r0 / q.theta; q.k1 = r1 / q.theta; q.k2 = r2 / q.theta;
  }
  const double theta = q.theta, theta2 = q.theta2, ct = q.ct, st = q.st;
  if (theta < 1e-4) {
    q.s = 1.0 - theta2 / 6.0;
    q.c = 0.5 - theta2 / 24.0;
    q.ds = -1.0 / 3.0 + theta2 / 30.0;
    q.dc = -1.0 / 12.0 + theta2 / 180.0;
    q.mst = -q.s;
  } else {
    q.s = st / theta;
    q.c = (1.0 - ct) / theta2;
    q.ds = (ct * theta - st) / theta2 / theta;
    q.dc = (st * theta - 2.0 * (1.0 - ct)) / (theta2 * theta2);
    q.mst = -st / theta;
  }
  return q;
}

// Residual and analytic Jacobian of one key, FP64.
//   X point, r/t image rotation & translation (q: rot_quantities of r), cam
//   full intrinsics, uv observation
// Outputs: res (f(x) - y), Jp [2][3], Ji [2][6], Jc [2][D] (free parameters).
// Rodrigues terms f = cos(theta) X, s v with s = sin(theta)/theta, v = r x X,
// c u with c = (1 - cos(theta))/theta^2, u = (r.X) r as in
// camera_shared_analytical_jacobian_matrix_calculator.hpp:334-429.
template <int D, bool WANT_J>
__device__ __forceinline__ void key_linearize_q(const double X[3], const double img[6], const RotQ &q,
                                                const double cam[kFull], const int fidx[kFull],
                                                double u_obs, double v_obs, double res[2],
                                                double Jp[2][3], double Ji[2][6],
                                                double Jc[2][D > 0 ? D : 1]) {
  const double r0 = img[0], r1 = img[1], r2 = img[2];
  const double ct = q.ct, st = q.st;
  double y[3];
  // projection exactly as camera_shared_vector_function.hpp:688-710
  if (q.theta == 0.0) {
    y[0] = X[0] + img[3]; y[1] = X[1] + img[4]; y[2] = X[2] + img[5];
  } else {
    const double k0 = q.k0, k1 = q.k1, k2 = q.k2;
    const double kx0 = k1 * X[2] - k2 * X[1], kx1 = k2 * X[0] - k0 * X[2], kx2 = k0 * X[1] - k1 * X[0];
    const double kd = k0 * X[0] + k1 * X[1] + k2 * X[2];
    const double omc = 1.0 - ct;
    y[0] = ct * X[0] + st * kx0 + omc * kd * k0 + img[3];
    y[1] = ct * X[1] + st * kx1 + omc * kd * k1 + img[4];
    y[2] = ct * X[2] + st * kx2 + omc * kd * k2 + img[5];
  }
  const double n0 = y[0] / y[2], n1 = y[1] / y[2];
  const double rr2 = n0 * n0 + n1 * n1, rr4 = rr2 * rr2, rr6 = rr2 * rr4;
  const double K1 = cam[IK1], K2 = cam[IK2], K3 = cam[IK3], D1 = cam[ID1], D2 = cam[ID2];
  const double coeff = K1 * rr2 + K2 * rr4 + K3 * rr6;
  const double dd0 = 2.0 * D1 * n0 * n1 + D2 * (rr2 + 2.0 * n0 * n0);
  const double dd1 = 2.0 * D2 * n0 * n1 + D1 * (rr2 + 2.0 * n1 * n1);
  const double m0 = n0 + (coeff * n0 + dd0), m1 = n1 + (coeff * n1 + dd1);
  const double F = cam[IF], SK = cam[ISKEW], PR = cam[IPR];
  res[0] = (F * m0 + SK * m1 + cam[IPX]) - u_obs;
  res[1] = (F * PR * m1 + cam[IPY]) - v_obs;
  if (!WANT_J) return;

  // derivative of the un-normalised camera point y w.r.t. r and X
  const double dot = r0 * X[0] + r1 * X[1] + r2 * X[2];
  const double s = q.s, c = q.c, ds = q.ds, dc = q.dc, mst = q.mst;
  const double rv[3] = {r0, r1, r2};
  const double v[3] = {r1 * X[2] - r2 * X[1], r2 * X[0] - r0 * X[2], r0 * X[1] - r1 * X[0]};
  const double dvdr[3][3] = {{0.0, X[2], -X[1]}, {-X[2], 0.0, X[0]}, {X[1], -X[0], 0.0}};
  const double dvdX[3][3] = {{0.0, -r2, r1}, {r2, 0.0, -r0}, {-r1, r0, 0.0}};
  // y recomputed from the derivative form (identical to the projection above
  // up to rounding) is not needed: the quotient rule uses the projected y.
  double pe[2][6], pp[2][3];
  const double iz2 = 1.0 / (y[2] * y[2]);
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    double dydr[3], dydX[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double dudr = (a == m ? dot : 0.0) + rv[a] * X[m];
      dydr[a] = X[a] * mst * rv[m] + v[a] * ds * rv[m] + s * dvdr[a][m] + dot * rv[a] * dc * rv[m] + c * dudr;
      dydX[a] = (a == m ? ct : 0.0) + s * dvdX[a][m] + c * rv[a] * rv[m];
    }
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      pe[a][m] = (dydr[a] * y[2] - y[a] * dydr[2]) * iz2;
      pe[a][3 + m] = ((a == m ? 1.0 : 0.0) * y[2] - y[a] * (m == 2 ? 1.0 : 0.0)) * iz2;
      pp[a][m] = (dydX[a] * y[2] - y[a] * dydX[2]) * iz2;
    }
  }
  // distortion layers (radial :431-470, decentering :472-503), both from the
  // undistorted derivative, then the intrinsic layer (:505-524)
#pragma unroll
  for (int m = 0; m < 9; ++m) {
    const double q0 = m < 3 ? pp[0][m] : pe[0][m - 3];
    const double q1 = m < 3 ? pp[1][m] : pe[1][m - 3];
    const double pr2 = 2.0 * n0 * q0 + 2.0 * n1 * q1;
    const double cd = K1 * pr2 + K2 * 2.0 * rr2 * pr2 + K3 * 3.0 * rr4 * pr2;
    const double cross = q0 * n1 + q1 * n0;
    const double t0 = q0 + (cd * n0 + coeff * q0) + (2.0 * D1 * cross + D2 * (pr2 + 4.0 * n0 * q0));
    const double t1 = q1 + (cd * n1 + coeff * q1) + (2.0 * D2 * cross + D1 * (pr2 + 4.0 * n1 * q1));
    const double j0 = F * t0 + SK * t1, j1 = F * PR * t1;
    if (m < 3) { Jp[0][m] = j0; Jp[1][m] = j1; }
    else { Ji[0][m - 3] = j0; Ji[1][m - 3] = j1; }
  }
  if (D > 0) {
    // camera block: [k1 k2 k3 d1 d2] through the intrinsic layer (:526-544),
    // [f skew px py pr] from IntrinsicIntrinsicDerivation (:546-557).  Only
    // the D free columns are formed; the column index is the same for every
    // thread, so the switch is a uniform branch (a table indexed by it would
    // be placed in scratch memory).
#pragma unroll
    for (int q2 = 0; q2 < (D > 0 ? D : 1); ++q2) {
      double nd0 = 0.0, nd1 = 0.0, c0 = 0.0, c1 = 0.0;
      bool distortion = true;
      switch (fidx[q2]) {
        case IK1: nd0 = rr2 * n0; nd1 = rr2 * n1; break;
        case IK2: nd0 = rr4 * n0; nd1 = rr4 * n1; break;
        case IK3: nd0 = rr6 * n0; nd1 = rr6 * n1; break;
        case ID1: nd0 = 2.0 * n0 * n1; nd1 = rr2 + 2.0 * n1 * n1; break;
        case ID2: nd0 = rr2 + 2.0 * n0 * n0; nd1 = 2.0 * n0 * n1; break;
        case IF: distortion = false; c0 = m0; c1 = PR * m1; break;
        case ISKEW: distortion = false; c0 = m1; c1 = 0.0; break;
        case IPX: distortion = false; c0 = 1.0; c1 = 0.0; break;
        case IPY: distortion = false; c0 = 0.0; c1 = 1.0; break;
        default: distortion = false; c0 = 0.0; c1 = F * m1; break;  // IPR
      }
      if (distortion) {
        c0 = F * nd0 + SK * nd1;
        c1 = F * PR * nd1;
      }
      Jc[0][q2] = c0;
      Jc[1][q2] = c1;
    }
  }
}

// The same with the rotation quantities formed here.
template <int D, bool WANT_J>
__device__ __forceinline__ void key_linearize(const double X[3], const double img[6],
                                              const double cam[kFull], const int fidx[kFull],
                                              double u_obs, double v_obs, double res[2],
                                              double Jp[2][3], double Ji[2][6],
                                              double Jc[2][D > 0 ? D : 1]) {
  const RotQ q = rot_quantities(img);
  key_linearize_q<D, WANT_J>(X, img, q, cam, fidx, u_obs, v_obs, res, Jp, Ji, Jc);
}

// Whitening by the 2x2 inverse covariance Lambda = L L^T: J^ = L^T J, r^ = L^T r,
// so that J^T Lambda J = J^^T J^ and r^T Lambda r = |r^|^2.
__device__ __forceinline__ void whiten2(const double *lt, double &a, double &b) {
  const double na = lt[0] * a + lt[1] * b;
  b = lt[2] * b;
  a = na;
}

}  // namespace hsba
