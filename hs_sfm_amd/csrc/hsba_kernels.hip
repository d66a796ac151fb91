// hsba_kernels.hip — per-key, per-point, per-image and per-pair kernels of
// the MI355X LM bundle adjuster (gfx950, wave64).
//
// Storage / accumulation precision T is float (mixed) or double (FP64 path);
// projections and Jacobians are always evaluated in FP64 registers.  Every
// reduction runs in a fixed order (segmented by point, image, camera or pair,
// then a fixed butterfly / tree), so results are bit-identical run to run and
// no kernel uses global atomics.
#include "hsba_kernels.h"

namespace hsba {

// ------------------------------------------------------------ small helpers
template <typename T>
__device__ __forceinline__ T shfl_xor(T v, int m) { return __shfl_xor(v, m, 64); }

__device__ __forceinline__ bool win_trial(const WinState &s) { return s.active && s.trial; }

template <typename T>
__device__ __forceinline__ void load_cam(const Dev<T> &d, int c, bool new_x, double cam[kFull]) {
  const double *src = (new_x ? d.cam_full_new : d.cam_full) + (int64_t)c * kFull;
#pragma unroll
  for (int j = 0; j < kFull; ++j) cam[j] = src[j];
}

// ------------------------------------------------------------ K1+K2 linearize
// One group of G lanes per point (points are contiguous per window; keys of a
// point contiguous); the launch is cut into up to three segments of whole
// waves with their own G (the host groups points by track length, so a
// point's lanes, and with them its order of summation, depend on the point
// alone).  Each lane linearizes keys lane, lane+G, ... of its point:
// residual and Jacobian in FP64, then W_k = Jp^T L Jr (3 x E) is stored and
// V_p = sum Jp^T L Jp, g_p = sum Jp^T L r are reduced across the group with a
// butterfly.  Replaces the per-key loops of
// camera_shared_analytical_jacobian_matrix_calculator.hpp:94-321 and
// camera_shared_normal_equation_builder.hpp:207-220, 249-287, 377-391.
template <typename T, int IMG, int D>
__global__ void __launch_bounds__(256) k_linearize(Dev<T> d, LinSeg seg, const int32_t *plist) {
  constexpr int E = IMG + D;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int sg = t < seg.t0[1] ? 0 : (t < seg.t0[2] ? 1 : 2);  // uniform in a wave
  const int glog2 = seg.g[sg];
  const int G = 1 << glog2;
  const int gid = t - seg.t0[sg];
  const int gi = gid >> glog2;
  const int p = (t < seg.t0[3] && gi < seg.cnt[sg]) ? (plist ? plist[seg.base[sg] + gi] : gi) : d.P;  // d.P: none
  const int lane = gid & (G - 1);
  bool act = false;
  if (p < d.P) {
    const WinState &s = d.ws[d.pt_win[p]];
    act = s.active && s.need_lin;
  }
  T v[6] = {0, 0, 0, 0, 0, 0}, g[3] = {0, 0, 0};
  if (act) {
    const double X[3] = {d.pts[3 * p], d.pts[3 * p + 1], d.pts[3 * p + 2]};
    const int kb = d.pt_ptr[p], ke = d.pt_ptr[p + 1];
    for (int k = kb + lane; k < ke; k += G) {
      const int im = d.key_img[k];
      double img[6], cam[kFull];
#pragma unroll
      for (int j = 0; j < 6; ++j) img[j] = d.imgs[6 * im + j];
      const RotQ rq = d.imgq[im];
      load_cam(d, d.img_cam[im], false, cam);
      double res[2], Jp[2][3], Ji[2][6], Jc[2][D > 0 ? D : 1];
      key_linearize_q<D, true>(X, img, rq, cam, d.fidx, d.key_uv[2 * k], d.key_uv[2 * k + 1], res, Jp, Ji, Jc);
      if (d.has_cov) {
        const double *lt = d.key_cov + 3 * k;
        whiten2(lt, res[0], res[1]);
#pragma unroll
        for (int m = 0; m < 3; ++m) whiten2(lt, Jp[0][m], Jp[1][m]);
#pragma unroll
        for (int m = 0; m < 6; ++m) whiten2(lt, Ji[0][m], Ji[1][m]);
#pragma unroll
        for (int m = 0; m < (D > 0 ? D : 1); ++m) whiten2(lt, Jc[0][m], Jc[1][m]);
      }
      if (d.loss) {
        const double sw = robust_row_scale(d.loss, d.loss_scale, res[0] * res[0] + res[1] * res[1]);
        res[0] *= sw; res[1] *= sw;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
#pragma unroll
          for (int m = 0; m < 3; ++m) Jp[r][m] *= sw;
#pragma unroll
          for (int m = 0; m < 6; ++m) Ji[r][m] *= sw;
#pragma unroll
          for (int m = 0; m < (D > 0 ? D : 1); ++m) Jc[r][m] *= sw;
        }
      }
      v[0] += T(Jp[0][0] * Jp[0][0] + Jp[1][0] * Jp[1][0]);
      v[1] += T(Jp[0][0] * Jp[0][1] + Jp[1][0] * Jp[1][1]);
      v[2] += T(Jp[0][0] * Jp[0][2] + Jp[1][0] * Jp[1][2]);
      v[3] += T(Jp[0][1] * Jp[0][1] + Jp[1][1] * Jp[1][1]);
      v[4] += T(Jp[0][1] * Jp[0][2] + Jp[1][1] * Jp[1][2]);
      v[5] += T(Jp[0][2] * Jp[0][2] + Jp[1][2] * Jp[1][2]);
#pragma unroll
      for (int a = 0; a < 3; ++a) g[a] += T(Jp[0][a] * res[0] + Jp[1][a] * res[1]);
      if (E > 0) {
        T *w = d.W + (int64_t)k * 3 * E;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
#pragma unroll
          for (int j = 0; j < E; ++j) {
            double jr0, jr1;
            if (j < IMG) { jr0 = Ji[0][j]; jr1 = Ji[1][j]; }
            else { jr0 = Jc[0][j - IMG]; jr1 = Jc[1][j - IMG]; }
            w[a * E + j] = T(Jp[0][a] * jr0 + Jp[1][a] * jr1);
          }
        }
      }
    }
  }
  for (int m = 1; m < G; m <<= 1) {
#pragma unroll
    for (int j = 0; j < 6; ++j) v[j] += shfl_xor(v[j], m);
#pragma unroll
    for (int j = 0; j < 3; ++j) g[j] += shfl_xor(g[j], m);
  }
  if (act && lane == 0) {
#pragma unroll
    for (int j = 0; j < 6; ++j) d.V[6 * p + j] = v[j];
#pragma unroll
    for (int j = 0; j < 3; ++j) d.gp[3 * p + j] = g[j];
  }
}

// ------------------------------------------------------------ image blocks
// D_i = sum Jr^T L Jr (E x E, upper triangle) and gr_i = sum Jr^T L r over the
// image's keys (image-major list), recomputing the Jacobian instead of reading
// a stored one.  One workgroup per chunk of at most kImgChunk of an image's
// keys writes the chunk's partial sums; k_image_combine adds an image's chunks
// in chunk order (deterministic).  Replaces
// camera_shared_normal_equation_builder.hpp:221-248, 288-307, 392-415.
template <typename T, int IMG, int D>
__global__ void __launch_bounds__(128) k_image_reduce(Dev<T> d) {
  constexpr int E = IMG + D;
  constexpr int NU = E * (E + 1) / 2;
  constexpr int NV = NU + E;
  const int ch = blockIdx.x;
  const int im = d.ich_img[ch];
  const WinState &s = d.ws[d.img_win[im]];
  if (!(s.active && s.need_lin)) return;
  T acc[NV > 0 ? NV : 1];
#pragma unroll
  for (int j = 0; j < NV; ++j) acc[j] = 0;
  double img[6], cam[kFull];
#pragma unroll
  for (int j = 0; j < 6; ++j) img[j] = d.imgs[6 * im + j];
  const RotQ rq = d.imgq[im];
  load_cam(d, d.img_cam[im], false, cam);
  const int kb = d.ich_begin[ch], ke = d.ich_end[ch];
  // the key and point indices of this thread's keys (at most kImgChunk / 128
  // of them) are loaded up front, so only the first key waits out the two
  // dependent index loads; keys are summed in the same order as one by one
  constexpr int KPT = kImgChunk / 128;
  int kk[KPT], pk[KPT];
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    const int t = kb + (int)threadIdx.x + 128 * i;
    kk[i] = t < ke ? d.img_keys[t] : 0;
  }
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    const int t = kb + (int)threadIdx.x + 128 * i;
    pk[i] = t < ke ? d.key_pt[kk[i]] : 0;
  }
#pragma unroll 1
  for (int i = 0; i < KPT; ++i) {
    if (kb + (int)threadIdx.x + 128 * i >= ke) break;
    const int k = kk[i];
    const int p = pk[i];
    const double X[3] = {d.pts[3 * p], d.pts[3 * p + 1], d.pts[3 * p + 2]};
    double res[2], Jp[2][3], Ji[2][6], Jc[2][D > 0 ? D : 1];
    key_linearize_q<D, true>(X, img, rq, cam, d.fidx, d.key_uv[2 * k], d.key_uv[2 * k + 1], res, Jp, Ji, Jc);
    if (d.has_cov) {
      const double *lt = d.key_cov + 3 * k;
      whiten2(lt, res[0], res[1]);
#pragma unroll
      for (int m = 0; m < 6; ++m) whiten2(lt, Ji[0][m], Ji[1][m]);
#pragma unroll
      for (int m = 0; m < (D > 0 ? D : 1); ++m) whiten2(lt, Jc[0][m], Jc[1][m]);
    }
    if (d.loss) {
      const double sw = robust_row_scale(d.loss, d.loss_scale, res[0] * res[0] + res[1] * res[1]);
      res[0] *= sw; res[1] *= sw;
#pragma unroll
      for (int r = 0; r < 2; ++r) {
#pragma unroll
        for (int m = 0; m < 6; ++m) Ji[r][m] *= sw;
#pragma unroll
        for (int m = 0; m < (D > 0 ? D : 1); ++m) Jc[r][m] *= sw;
      }
    }
    // products in the storage type (FP32 in mixed precision: the sums are
    // FP32 anyway, and the narrower operands keep the kernel out of scratch)
    T jr[2][E > 0 ? E : 1];
#pragma unroll
    for (int j = 0; j < E; ++j) {
      jr[0][j] = T(j < IMG ? Ji[0][j] : Jc[0][j - IMG]);
      jr[1][j] = T(j < IMG ? Ji[1][j] : Jc[1][j - IMG]);
    }
    const T r0 = T(res[0]), r1 = T(res[1]);
    int u = 0;
#pragma unroll
    for (int a = 0; a < E; ++a) {
#pragma unroll
      for (int b = a; b < E; ++b) acc[u++] += jr[0][a] * jr[0][b] + jr[1][a] * jr[1][b];
    }
#pragma unroll
    for (int a = 0; a < E; ++a) acc[NU + a] += jr[0][a] * r0 + jr[1][a] * r1;
  }
  // wave butterfly then fixed-order sum of the two waves
  for (int m = 1; m < 64; m <<= 1) {
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] += shfl_xor(acc[j], m);
  }
  __shared__ T red[2][NV > 0 ? NV : 1];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) red[wave][j] = acc[j];
  }
  __syncthreads();
  for (int j = threadIdx.x; j < NV; j += blockDim.x) d.ipart[(int64_t)ch * NV + j] = red[0][j] + red[1][j];
}

// Sum of an image's chunk partials in chunk order into D_i (both halves) and
// gr_i; one thread per (image, entry).
template <typename T, int IMG, int D>
__global__ void k_image_combine(Dev<T> d) {
  constexpr int E = IMG + D;
  constexpr int NU = E * (E + 1) / 2;
  constexpr int NV = NU + E;
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (NV == 0 || gid >= (int64_t)d.I * NV) return;
  const int im = (int)(gid / NV), j = (int)(gid - (int64_t)im * NV);
  const WinState &s = d.ws[d.img_win[im]];
  if (!(s.active && s.need_lin)) return;
  T v = 0;
  for (int ch = d.img_ch_ptr[im]; ch < d.img_ch_ptr[im + 1]; ++ch) v += d.ipart[(int64_t)ch * NV + j];
  if (j < NU) {
    // unpack upper-triangle index j -> (a, b) and write both halves
    int a = 0, rem = j;
    while (rem >= E - a) { rem -= E - a; ++a; }
    const int b = a + rem;
    d.Dimg[(int64_t)im * E * E + a * E + b] = v;
    d.Dimg[(int64_t)im * E * E + b * E + a] = v;
  } else {
    d.gr[(int64_t)im * E + (j - NU)] = v;
  }
}

// ------------------------------------------------------------ camera blocks
// U_c = sum over the camera's images of the camera part of D_i, g_c likewise,
// in image order (the reference accumulates U_c per key in key order,
// camera_shared_normal_equation_builder.hpp:235-248, 403-415).
template <typename T, int IMG, int D>
__global__ void __launch_bounds__(128) k_camera_reduce(Dev<T> d) {
  constexpr int E = IMG + D;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (D == 0 || c >= d.C) return;
  const WinState &s = d.ws[d.cam_win[c]];
  if (!(s.active && s.need_lin)) return;
  T u[D > 0 ? D * D : 1], g[D > 0 ? D : 1];
#pragma unroll
  for (int j = 0; j < (D > 0 ? D * D : 1); ++j) u[j] = 0;
#pragma unroll
  for (int j = 0; j < (D > 0 ? D : 1); ++j) g[j] = 0;
  for (int t = d.cam_ptr[c]; t < d.cam_ptr[c + 1]; ++t) {
    const int im = d.cam_imgs[t];
    const T *Di = d.Dimg + (int64_t)im * E * E;
#pragma unroll
    for (int a = 0; a < D; ++a) {
#pragma unroll
      for (int b = 0; b < D; ++b) u[a * D + b] += Di[(IMG + a) * E + IMG + b];
      g[a] += d.gr[(int64_t)im * E + IMG + a];
    }
  }
#pragma unroll
  for (int j = 0; j < D * D; ++j) d.Uc[(int64_t)c * D * D + j] = u[j];
#pragma unroll
  for (int j = 0; j < D; ++j) d.gc[(int64_t)c * D + j] = g[j];
}

// Constraint rows (identity rows of the vector function): weight on the
// diagonal and w * r in the gradient, in y order by a single thread so the
// order of additions matches the reference
// (camera_shared_normal_equation_builder.hpp:309-349, 418-459).
template <typename T, int IMG, int D>
__global__ void k_constraints(Dev<T> d) {
  constexpr int E = IMG + D;
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int q = 0; q < d.ncons; ++q) {
    const ConstraintRow cr = d.cons[q];
    const WinState &s = d.ws[cr.win];
    if (!(s.active && s.need_lin)) continue;
    if (cr.kind == 0 && cr.block) {
      // row `param` of a 3x3 block over the record's three rows: N gets the
      // block's upper triangle once (V holds 00 01 02 11 12 22), g gets B r
      // (ba_gcp_constrained_normal_equation_builder.hpp:64-71)
      const ConstraintRow *rec = d.cons + (q - cr.param);
      double t = 0.0;
      for (int j = 0; j < 3; ++j) t += cr.brow[j] * (d.pts[3 * cr.entity + j] - rec[j].y);
      for (int j = cr.param; j < 3; ++j) d.V[6 * cr.entity + (cr.param == 0 ? j : (cr.param == 1 ? j + 2 : 5))] += T(cr.brow[j]);
      d.gp[3 * cr.entity + cr.param] += T(t);
    } else if (cr.kind == 0) {
      const double r = d.pts[3 * cr.entity + cr.param] - cr.y;
      const int vi = cr.param == 0 ? 0 : (cr.param == 1 ? 3 : 5);
      d.V[6 * cr.entity + vi] += T(cr.weight);
      d.gp[3 * cr.entity + cr.param] += T(cr.weight * r);
    } else if (cr.kind == 1) {
      const double r = d.imgs[6 * cr.entity + cr.param] - cr.y;
      d.Dimg[(int64_t)cr.entity * E * E + cr.param * E + cr.param] += T(cr.weight);
      d.gr[(int64_t)cr.entity * E + cr.param] += T(cr.weight * r);
    } else {
      const double r = d.cam_free[(int64_t)cr.entity * (D > 0 ? D : 1) + cr.param] - cr.y;
      d.Uc[(int64_t)cr.entity * D * D + cr.param * D + cr.param] += T(cr.weight);
      d.gc[(int64_t)cr.entity * D + cr.param] += T(cr.weight * r);
    }
  }
}

// ------------------------------------------------------------ K3 point factor
// V* = V + mu I, V* = L L^T (FP64), store L^-1 (lower, 6 values) and
// q = L^-1 g_p.  Then W^T V*^-1 W' = (L^-1 W)^T (L^-1 W') — the inverse of
// camera_shared_normal_equation_solver.hpp:141-148 in factored form.
template <typename T>
__global__ void k_point_factor(Dev<T> d) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= d.P) return;
  WinState &s = d.ws[d.pt_win[p]];
  if (!win_trial(s)) return;
  const double mu = s.mu;
  const double a00 = double(d.V[6 * p]) + mu, a01 = d.V[6 * p + 1], a02 = d.V[6 * p + 2];
  const double a11 = double(d.V[6 * p + 3]) + mu, a12 = d.V[6 * p + 4], a22 = double(d.V[6 * p + 5]) + mu;
  const double l00 = sqrt(a00);
  const double l10 = a01 / l00, l20 = a02 / l00;
  const double t11 = a11 - l10 * l10;
  const double l11 = sqrt(t11);
  const double l21 = (a12 - l20 * l10) / l11;
  const double t22 = a22 - l20 * l20 - l21 * l21;
  const double l22 = sqrt(t22);
  if (!(a00 > 0.0) || !(t11 > 0.0) || !(t22 > 0.0)) {
    s.chol_fail = 1;  // benign race: every writer stores 1
    return;
  }
  // inverse of lower-triangular L
  const double i00 = 1.0 / l00, i11 = 1.0 / l11, i22 = 1.0 / l22;
  const double i10 = -l10 * i00 * i11;
  const double i21 = -l21 * i11 * i22;
  const double i20 = -(l20 * i00 + l21 * i10) * i22;
  T *L = d.Linv + 6 * p;
  L[0] = T(i00); L[1] = T(i10); L[2] = T(i11); L[3] = T(i20); L[4] = T(i21); L[5] = T(i22);
  const double g0 = d.gp[3 * p], g1 = d.gp[3 * p + 1], g2 = d.gp[3 * p + 2];
  d.q[3 * p] = T(i00 * g0);
  d.q[3 * p + 1] = T(i10 * g0 + i11 * g1);
  d.q[3 * p + 2] = T(i20 * g0 + i21 * g1 + i22 * g2);
}

// Lanes of one wave see each other's earlier LDS writes after this point (the
// LDS unit executes one wave's instructions in issue order; the fences keep
// the compiler from moving LDS accesses across the call).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Z rows staged through LDS when a row is one 128-byte line of FP32.
template <typename T, int E>
constexpr bool kKeyZStaged = sizeof(T) == 4 && z_stride(E, sizeof(T)) == 32;

// Z_k = L_p^-1 W_k and q_p, one thread per key.  The thread's W loads go out before
// the dependent point / window / factor lookups, so they overlap that chain;
// the Z row (padding included) leaves as whole 16-byte stores, so every
// 128-byte line of Z is written completely.
//
// FP32 rows of one line: each wave moves its 64 keys' W rows (contiguous in
// memory) and Z rows through a wave-private LDS region, so that every load
// and store instruction covers consecutive addresses (64 x 4 bytes of W,
// 64 x 16 bytes of Z) instead of one row per lane: per wave that is 2 lines
// per W load and 8 per Z store, against up to 64 for row-per-lane access.
// Rows of keys whose window is not in a trial are not written, as before.
template <typename T, int E>
__global__ void __launch_bounds__(256) k_key_z(Dev<T> d) {
  constexpr int BLK = 3 * E;
  constexpr int ZS = z_stride(E, sizeof(T));
  constexpr int VW = 16 / sizeof(T);  // elements per 16-byte store
  typedef T vec __attribute__((ext_vector_type(VW)));
  if constexpr (kKeyZStaged<T, E>) {
    constexpr int WP = BLK | 1;           // odd pitch: row reads by 64 lanes hit distinct banks
    constexpr int ZP = ZS + VW;           // 16-byte row writes of 16 lanes hit distinct banks
    constexpr int VPR = ZS / VW;          // 16-byte vectors per Z row
    constexpr int REG = 64 * (WP > ZP ? WP : ZP);
    __shared__ __align__(16) T stage[4][REG];
    __shared__ int actf[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t k0 = (int64_t)blockIdx.x * blockDim.x + 64 * wv;  // the wave's first key
    if (k0 >= d.N) return;                                           // whole wave past the end
    const int nk = (int)(d.N - k0 < 64 ? d.N - k0 : 64);
    const int k = (int)k0 + lane;
    T *sw = stage[wv];
    // W rows of keys k0 .. k0 + nk - 1: nk * BLK consecutive elements
    const T *wsrc = d.W + k0 * BLK;
    const int ne = nk * BLK;
    T wl[BLK];
#pragma unroll
    for (int i = 0; i < BLK; ++i) {
      const int e = 64 * i + lane;
      wl[i] = e < ne ? wsrc[e] : T(0);
    }
    const bool valid = lane < nk;
    const int p = valid ? d.key_pt[k] : 0;
    const bool act = valid && win_trial(d.ws[d.pt_win[p]]);
    T l[6], qp[3];
#pragma unroll
    for (int j = 0; j < 6; ++j) l[j] = act ? d.Linv[6 * p + j] : T(0);
#pragma unroll
    for (int j = 0; j < 3; ++j) qp[j] = act ? d.q[3 * p + j] : T(0);
#pragma unroll
    for (int i = 0; i < BLK; ++i) {
      const int e = 64 * i + lane, r = e / BLK;
      sw[r * WP + (e - r * BLK)] = wl[i];
    }
    actf[wv][lane] = act ? 1 : 0;
    wave_lds_sync();
    T w[BLK];
#pragma unroll
    for (int e = 0; e < BLK; ++e) w[e] = sw[lane * WP + e];
    wave_lds_sync();  // every W row read before the region takes the Z rows
    vec out[VPR];
#pragma unroll
    for (int e = 0; e < ZS; ++e) {
      T v = 0;
      if (e < E) v = l[0] * w[e];
      else if (e < 2 * E) v = l[1] * w[e - E] + l[2] * w[e];
      else if (e < BLK) v = l[3] * w[e - 2 * E] + l[4] * w[e - E] + l[5] * w[e];
      else if (e < BLK + 3) v = qp[e - BLK];
      out[e / VW][e % VW] = v;
    }
#pragma unroll
    for (int q = 0; q < VPR; ++q) *(vec *)(sw + lane * ZP + q * VW) = out[q];
    wave_lds_sync();
    vec *zdst = (vec *)(d.Z + k0 * ZS);
#pragma unroll
    for (int q = 0; q < VPR; ++q) {
      const int v = 64 * q + lane, r = v / VPR, c = v - r * VPR;
      if (r < nk && actf[wv][r]) zdst[v] = *(const vec *)(sw + r * ZP + c * VW);
    }
    if (d.Zi && act) {
      vec *zi = (vec *)(d.Zi + (int64_t)d.key_ipos[k] * ZS);
#pragma unroll
      for (int q = 0; q < VPR; ++q) zi[q] = out[q];
    }
    return;
  }
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d.N) return;
  // the W row does not depend on the point: its loads are issued first and
  // overlap the point / window lookups
  T w[BLK];
  const T *wk = d.W + (int64_t)k * BLK;
#pragma unroll
  for (int e = 0; e < BLK; ++e) w[e] = wk[e];
  const int p = d.key_pt[k];
  if (!win_trial(d.ws[d.pt_win[p]])) return;
  T l[6], qp[3];
#pragma unroll
  for (int j = 0; j < 6; ++j) l[j] = d.Linv[6 * p + j];
#pragma unroll
  for (int j = 0; j < 3; ++j) qp[j] = d.q[3 * p + j];
  // Z = L^-1 W (lower-triangular L^-1: l00 l10 l11 l20 l21 l22), then q_p,
  // padding zero, written as whole 16-byte vectors
  vec out[ZS / VW];
#pragma unroll
  for (int e = 0; e < ZS; ++e) {
    T v = 0;
    if (e < E) v = l[0] * w[e];
    else if (e < 2 * E) v = l[1] * w[e - E] + l[2] * w[e];
    else if (e < BLK) v = l[3] * w[e - 2 * E] + l[4] * w[e - E] + l[5] * w[e];
    else if (e < BLK + 3) v = qp[e - BLK];
    out[e / VW][e % VW] = v;
  }
  vec *z = (vec *)(d.Z + (int64_t)k * ZS);
#pragma unroll
  for (int q = 0; q < ZS / VW; ++q) z[q] = out[q];
  if (d.Zi) {
    vec *zi = (vec *)(d.Zi + (int64_t)d.key_ipos[k] * ZS);
#pragma unroll
    for (int q = 0; q < ZS / VW; ++q) zi[q] = out[q];
  }
}

// ------------------------------------------------------------ K4 camera pairs
// One wavefront per image pair (ia <= ib): P = sum over points seeing both of
// Z_a^T Z_b (E x E), plus for diagonal pairs rp_i = sum Z_a^T q_p.  Lanes
// stride the pair's incidence list; the 64 partial blocks are combined by a
// fixed butterfly.  Replaces the O(I^2 P) loops of
// camera_shared_normal_equation_solver.hpp:198-443.
template <typename T, int E>
__global__ void __launch_bounds__(256) k_pairs(Dev<T> d, const int32_t *plist, int count) {
  constexpr int ZS = z_stride(E, sizeof(T));
  constexpr int VW = 16 / sizeof(T), NV = (3 * E + VW - 1) / VW;
  typedef T vec __attribute__((ext_vector_type(VW)));
  const int slot = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (slot >= count) return;
  const int pair = plist ? plist[slot] : slot;
  if (!win_trial(d.ws[d.pair_win[pair]])) return;
  const bool diag = d.pair_img[pair] >= 0;
  T acc[E * E];
  T racc[E];
#pragma unroll
  for (int j = 0; j < E * E; ++j) acc[j] = 0;
#pragma unroll
  for (int j = 0; j < E; ++j) racc[j] = 0;
  const int64_t jb = d.pair_ptr[pair], je = d.pair_ptr[pair + 1];
  auto fetch = [&](uint64_t code, vec za[NV], vec zb[NV], T qv[3]) {
    const int a = inc_a(code);
    const int b = inc_b(code);
    const vec *ra = (const vec *)(d.Z + (int64_t)a * ZS);
    const vec *rb = (const vec *)(d.Z + (int64_t)b * ZS);
#pragma unroll
    for (int q = 0; q < NV; ++q) { za[q] = ra[q]; zb[q] = rb[q]; }
    if (diag) {
      const int p = d.key_pt[a];
      qv[0] = d.q[3 * p]; qv[1] = d.q[3 * p + 1]; qv[2] = d.q[3 * p + 2];
    }
  };
  auto mult = [&](const vec va[NV], const vec vb[NV], const T qc[3]) {
#pragma unroll
    for (int r = 0; r < E; ++r) {
      const T a0 = va[r / VW][r % VW], a1 = va[(E + r) / VW][(E + r) % VW];
      const T a2 = va[(2 * E + r) / VW][(2 * E + r) % VW];
#pragma unroll
      for (int c = 0; c < E; ++c)
        acc[r * E + c] += a0 * vb[c / VW][c % VW] + a1 * vb[(E + c) / VW][(E + c) % VW] +
                          a2 * vb[(2 * E + c) / VW][(2 * E + c) % VW];
      if (diag) racc[r] += a0 * qc[0] + a1 * qc[1] + a2 * qc[2];
    }
  };
  vec va[NV], vb[NV];
  T qc[3] = {0, 0, 0};
  int64_t j = jb + lane;
  if (j < je) fetch(d.inc[j], va, vb, qc);
  uint64_t cn = (j + 64 < je) ? d.inc[j + 64] : 0u;
  for (; j < je; j += 64) {
    vec na[NV], nb[NV];
    T qn[3] = {0, 0, 0};
    const bool more = j + 64 < je;
    const uint64_t cnn = (j + 128 < je) ? d.inc[j + 128] : 0u;
    if (more) fetch(cn, na, nb, qn);
    mult(va, vb, qc);
    if (more) {
#pragma unroll
      for (int q = 0; q < NV; ++q) { va[q] = na[q]; vb[q] = nb[q]; }
      qc[0] = qn[0]; qc[1] = qn[1]; qc[2] = qn[2];
    }
    cn = cnn;
  }
  for (int m = 1; m < 64; m <<= 1) {
#pragma unroll
    for (int j = 0; j < E * E; ++j) acc[j] += shfl_xor(acc[j], m);
    if (diag) {
#pragma unroll
      for (int j = 0; j < E; ++j) racc[j] += shfl_xor(racc[j], m);
    }
  }
  T *out = d.pair_out + (int64_t)pair * E * E;
#pragma unroll
  for (int j = 0; j < E * E; ++j)
    if ((j & 63) == lane) out[j] = acc[j];
  if (diag) {
    T *rp = d.rp + (int64_t)d.pair_img[pair] * E;
#pragma unroll
    for (int j = 0; j < E; ++j)
      if (j == lane) rp[j] = racc[j];
  }
}

// K4 on the matrix cores (camera blocks up to E = 15).  For a pair
// (a, b), P_ab = A^T B where A and B stack the Z rows of the shared points (3
// per point); column E of B carries q_p, so for a diagonal pair column E of
// the product is rp_i.  In v_mfma_{f32,f64}_16x16x4 lane i + 16k supplies
// A^T[i][k] = Z_a[row k][i] and B[k][i]; the product P[i][j] lands in lane
// j + 16 (i / 4), element i % 4 for the f32 form and in lane j + 16 (i % 4),
// element i / 4 for the f64 form (both checked on the device).
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mma;
template <>
struct Mma<float> {
  typedef f32x4 acc_t;
  static __device__ __forceinline__ int row(int kr, int e) { return 4 * kr + e; }
  static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mma<double> {
  typedef f64x4 acc_t;
  static __device__ __forceinline__ int row(int kr, int e) { return kr + 4 * e; }
  static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

// Destination in the K-major LDS stage (see k_pairs_mma) of element e of the
// concatenated [Z_a row | Z_b row] of one incidence (each row z_stride(E)
// values: Z's 3 x E, then q_p): row r = e' / E, column e' % E for Z values; q_p
// of the b row goes to column E of rows 0..2; everything else to the dump row.
// Offsets are in elements of an array [2][rows + 1][16] and relative to the
// incidence's first row, except the dump row, which is absolute (the caller
// passes the incidence base to add).
template <int E, int ZS>
__device__ __forceinline__ int kmajor_dst(int e, int rows, int base) {
  const int side = e < ZS ? 0 : 1;
  const int f = e - side * ZS;
  const int plane = side * (rows + 1) * 16;
  if (f < 3 * E) return plane + base + (f / E) * 16 + (f % E);
  if (side == 1 && f < 3 * E + 3) return plane + base + (f - 3 * E) * 16 + E;
  return plane + rows * 16;
}

// One workgroup per image pair (ia <= ib) on the matrix cores
// (v_mfma_f32_16x16x4_f32 / v_mfma_f64_16x16x4_f64): P = sum over incidences
// of Z_a^T Z_b (E x E) and, in column E of the same 16 x 16 tile, Z_a^T q_p
// (the pair's rhs part when ia == ib).  A stage of 64 incidences is laid out
// in LDS K-major: every incidence gives 3 rows of 16 values per side,
// [Z_a row r | 0...] and [Z_b row r | q_p[r] | 0...], so each v_mfma step is
// two LDS reads at a fixed offset from the lane's base and the product.  Wave
// w takes the stage's incidences 16 w .. 16 w + 15 (48 rows, 12 steps); its
// partial tile is added to the others at the end.  Loading: 4 lanes per
// incidence, each moving a fixed quarter of the two Z rows (16-byte loads,
// element stores to destinations computed once per lane); the next stage's
// rows are in flight while this one is multiplied, and the incidence codes
// of the stage after.  Replaces the O(I^2 P) loops of
// camera_shared_normal_equation_solver.hpp:198-443.
template <typename T, int E>
__global__ void __launch_bounds__(256) k_pairs_mma(Dev<T> d, const int32_t *plist) {
  static_assert(E >= 1 && E <= 15, "one 16-wide tile holds the E columns and the rhs column");
  constexpr int ZS = z_stride(E, sizeof(T));
  constexpr int VW = 16 / sizeof(T), NV = ZS / VW;
  constexpr int SI = 64;            // incidences per stage
  constexpr int KR = 3 * SI;        // K rows per stage
  constexpr int NU = 2 * NV / 4;    // vectors a lane moves per incidence
  typedef T vec __attribute__((ext_vector_type(VW)));
  typedef typename Mma<T>::acc_t acc_t;
  __shared__ T lds[2][KR + 1][16];  // K-major a and b sides; row KR takes the unread elements
  __shared__ T part[3][64][4];
  const int pair = plist ? plist[blockIdx.x] : blockIdx.x;
  if (pair < 0) return;  // an empty slot of the launch order (the whole workgroup)
  if (!win_trial(d.ws[d.pair_win[pair]])) return;
  const bool diag = d.pair_img[pair] >= 0;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int col = lane & 15, kr = lane >> 4;
  const int ls = t >> 2, lq = t & 3;  // loader: incidence, quarter of its 2 NV vectors
  T *flat = &lds[0][0][0];
  for (int e = t; e < 2 * (KR + 1) * 16; e += 256) flat[e] = 0;  // columns no stage writes stay zero
  int dst[NU * VW];
#pragma unroll
  for (int u = 0; u < NU; ++u)
#pragma unroll
    for (int c = 0; c < VW; ++c) dst[u * VW + c] = kmajor_dst<E, ZS>((lq + 4 * u) * VW + c, KR, 3 * ls * 16);
  const int64_t jb = d.pair_ptr[pair], je = d.pair_ptr[pair + 1];
  acc_t acc = {0, 0, 0, 0};
  vec v[NU];
  auto fetch = [&](uint64_t code) {
    const vec *ra = (const vec *)(d.Z + (int64_t)inc_a(code) * ZS);
    const vec *rb = (const vec *)(d.Z + (int64_t)inc_b(code) * ZS);
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int q = lq + 4 * u;
      v[u] = q < NV ? ra[q] : rb[q - NV];
    }
  };
  auto clear = [&]() {
#pragma unroll
    for (int u = 0; u < NU; ++u) v[u] = vec{0};
  };
  uint64_t code_next = 0;
  if (jb + ls < je) fetch(d.inc[jb + ls]);
  else clear();
  if (jb + SI + ls < je) code_next = d.inc[jb + SI + ls];
  const T *pa = &lds[0][48 * wave + kr][col], *pb = &lds[1][48 * wave + kr][col];
  __syncthreads();  // the zeroed rows before the first stage's stores
  for (int64_t base = jb; base < je; base += SI) {
    const int n = je - base < SI ? (int)(je - base) : SI;
    // stage the rows (zeros for the slots past the pair's last incidence)
#pragma unroll
    for (int u = 0; u < NU; ++u)
#pragma unroll
      for (int c = 0; c < VW; ++c) flat[dst[u * VW + c]] = v[u][c];
    __syncthreads();
    if (base + SI + ls < je) fetch(code_next);
    else clear();
    code_next = (base + 2 * SI + ls < je) ? d.inc[base + 2 * SI + ls] : 0;
    const int m = max(0, min(16, n - 16 * wave));  // this wave's incidences
    if (m == 16) {
#pragma unroll
      for (int k = 0; k < 12; ++k) acc = Mma<T>::op(pa[64 * k], pb[64 * k], acc);
    } else {
      for (int k = 0; 4 * k < 3 * m; ++k) acc = Mma<T>::op(pa[64 * k], pb[64 * k], acc);
    }
    __syncthreads();
  }
  if (wave > 0)
#pragma unroll
    for (int e = 0; e < 4; ++e) part[wave - 1][lane][e] = acc[e];
  __syncthreads();
  if (wave > 0) return;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const T v = acc[e] + part[0][lane][e] + part[1][lane][e] + part[2][lane][e];
    const int row = Mma<T>::row(kr, e);
    if (row < E) {
      if (col < E) d.pair_out[(int64_t)pair * E * E + row * E + col] = v;
      else if (diag && col == E) d.rp[(int64_t)d.pair_img[pair] * E + row] = v;
    }
  }
}

// One wavefront per image pair on the matrix cores, four independent pairs
// per workgroup and no workgroup barrier: for short incidence lists, where
// k_pairs_mma leaves most of its loaders and waves idle and still pays the
// cross-wave combine.  A stage is 16 incidences in the wave's own K-major LDS
// rows (as in k_pairs_mma; 4 lanes load each incidence); while it is
// multiplied (48 rows, 12 v_mfma steps) the next stage's rows are in flight
// and the incidence codes of the one after.  One accumulator per pair: the
// incidences are summed in list order.
template <typename T, int E>
__global__ void __launch_bounds__(256) k_pairs_mma_wave(Dev<T> d, const int32_t *plist, int count) {
  static_assert(E >= 1 && E <= 15, "one 16-wide tile holds the E columns and the rhs column");
  constexpr int ZS = z_stride(E, sizeof(T));
  constexpr int VW = 16 / sizeof(T), NV = ZS / VW;
  constexpr int SI = 16;            // incidences per stage
  constexpr int KR = 3 * SI;        // K rows per stage
  constexpr int NU = 2 * NV / 4;    // vectors a lane moves per incidence
  typedef T vec __attribute__((ext_vector_type(VW)));
  typedef typename Mma<T>::acc_t acc_t;
  __shared__ T lds[4][2][KR + 1][16];  // per wave: K-major a and b sides, row KR for unread elements
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int slot = blockIdx.x * 4 + wave;
  if (slot >= count) return;  // the whole wave
  const int pair = plist ? plist[slot] : slot;
  if (pair < 0) return;  // an empty slot of the launch order
  if (!win_trial(d.ws[d.pair_win[pair]])) return;
  const bool diag = d.pair_img[pair] >= 0;
  const int col = lane & 15, kr = lane >> 4;
  const int ls = lane >> 2, lq = lane & 3;  // loader: incidence, quarter of its 2 NV vectors
  T *flat = &lds[wave][0][0][0];
  for (int e = lane; e < 2 * (KR + 1) * 16; e += 64) flat[e] = 0;
  int dst[NU * VW];
#pragma unroll
  for (int u = 0; u < NU; ++u)
#pragma unroll
    for (int c = 0; c < VW; ++c) dst[u * VW + c] = kmajor_dst<E, ZS>((lq + 4 * u) * VW + c, KR, 3 * ls * 16);
  const int64_t jb = d.pair_ptr[pair], je = d.pair_ptr[pair + 1];
  acc_t acc = {0, 0, 0, 0};
  vec v[NU];
  auto fetch = [&](uint64_t code) {
    const vec *ra = (const vec *)(d.Z + (int64_t)inc_a(code) * ZS);
    const vec *rb = (const vec *)(d.Z + (int64_t)inc_b(code) * ZS);
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int q = lq + 4 * u;
      v[u] = q < NV ? ra[q] : rb[q - NV];
    }
  };
  auto clear = [&]() {
#pragma unroll
    for (int u = 0; u < NU; ++u) v[u] = vec{0};
  };
  uint64_t code_next = 0;
  if (jb + ls < je) fetch(d.inc[jb + ls]);
  else clear();
  if (jb + SI + ls < je) code_next = d.inc[jb + SI + ls];
  const T *pa = &lds[wave][0][kr][col], *pb = &lds[wave][1][kr][col];
  wave_lds_sync();  // the zeroed rows before the first stage's stores
  for (int64_t base = jb; base < je; base += SI) {
    const int n = je - base < SI ? (int)(je - base) : SI;
#pragma unroll
    for (int u = 0; u < NU; ++u)
#pragma unroll
      for (int c = 0; c < VW; ++c) flat[dst[u * VW + c]] = v[u][c];
    wave_lds_sync();
    if (base + SI + ls < je) fetch(code_next);
    else clear();
    code_next = (base + 2 * SI + ls < je) ? d.inc[base + 2 * SI + ls] : 0;
    if (n == SI) {
#pragma unroll
      for (int k = 0; k < 12; ++k) acc = Mma<T>::op(pa[64 * k], pb[64 * k], acc);
    } else {
      for (int k = 0; 4 * k < 3 * n; ++k) acc = Mma<T>::op(pa[64 * k], pb[64 * k], acc);
    }
    wave_lds_sync();  // the stage's rows are read before the next stage overwrites them
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int row = Mma<T>::row(kr, e);
    if (row < E) {
      if (col < E) d.pair_out[(int64_t)pair * E * E + row * E + col] = acc[e];
      else if (diag && col == E) d.rp[(int64_t)d.pair_img[pair] * E + row] = acc[e];
    }
  }
}

// ------------------------------------------------------------ assembly
// One thread per lower-triangle entry of a window's reduced matrix (FP64):
// S = U + mu I (+ image-camera coupling) - sum of pair blocks, by the
// per-block contribution lists built at problem creation.
// Entry u (row-major lower-triangle order) of window w.
template <typename T, int IMG, int D>
__device__ __forceinline__ void assemble_entry(const Dev<T> &d, int w, int64_t u) {
  constexpr int E = IMG + D;
  const WinInfo &wi = d.wi[w];
  const WinState &s = d.ws[w];
  if (!win_trial(s) || s.chol_fail) return;
  int64_t R = (int64_t)((sqrt(8.0 * (double)u + 1.0) - 1.0) * 0.5);
  while (R * (R + 1) / 2 > u) --R;
  while ((R + 1) * (R + 2) / 2 <= u) ++R;
  const int64_t Cc = u - R * (R + 1) / 2;
  const int bR = d.row_blk[wi.r_off + R], bC = d.row_blk[wi.r_off + Cc];
  const int oR = d.row_sub[wi.r_off + R], oC = d.row_sub[wi.r_off + Cc];
  const int64_t blk = (int64_t)bR * (bR + 1) / 2 + bC;
  double v = (R == Cc) ? s.mu : 0.0;
  for (int64_t j = d.blk_ptr[wi.blk_off + blk]; j < d.blk_ptr[wi.blk_off + blk + 1]; ++j) {
    const Contrib cb = d.contrib[j];
    if (cb.kind == 0) {
      v += (double)d.Dimg[(int64_t)cb.id * E * E + oR * E + oC];
    } else if (cb.kind == 1) {
      v += (double)d.Uc[(int64_t)cb.id * D * D + oR * D + oC];
    } else if (cb.kind == 2) {
      // camera row block x image column block: W_ic^T = D_i[IMG + oR][oC]
      v += (double)d.Dimg[(int64_t)cb.id * E * E + (IMG + oR) * E + oC];
    } else {
      const T *P = d.pair_out + (int64_t)cb.id * E * E;
      const int r = cb.ro + oR, c = cb.so + oC;
      v -= cb.trans ? (double)P[c * E + r] : (double)P[r * E + c];
    }
  }
  d.S[wi.s_off + R * wi.ld + Cc] = v;
}

// One thread per lower-triangle entry of all windows, the window found by
// binary search over the windows' entry offsets.
template <typename T, int IMG, int D>
__global__ void k_assemble(Dev<T> d, int64_t total) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  int lo = 0, hi = d.nwin - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d.wi[mid].lo_off <= t) lo = mid; else hi = mid - 1;
  }
  assemble_entry<T, IMG, D>(d, lo, t - d.wi[lo].lo_off);
}

// Batches of windows: blockIdx.y is the window, so no thread walks the
// dependent loads of a search (the grid covers the largest window; threads
// past a smaller window's entries exit at once).
template <typename T, int IMG, int D>
__global__ void k_assemble_win(Dev<T> d) {
  const int w = blockIdx.y;
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = d.wi[w].n;
  if (u >= n * (n + 1) / 2) return;
  assemble_entry<T, IMG, D>(d, w, u);
}

// Reduced right-hand side: g_r - sum of rp over the row's images.
template <typename T, int IMG, int D>
__global__ void k_assemble_rhs(Dev<T> d, int rows) {
  constexpr int E = IMG + D;
  const int R = blockIdx.x * blockDim.x + threadIdx.x;
  if (R >= rows) return;
  // row -> window via row_blk sign: windows store row ranges contiguously
  int lo = 0, hi = d.nwin - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d.wi[mid].r_off <= R) lo = mid; else hi = mid - 1;
  }
  const WinInfo &wi = d.wi[lo];
  const WinState &s = d.ws[lo];
  if (!win_trial(s) || s.chol_fail) return;
  const int r = R - wi.r_off;
  if (r >= wi.n) return;  // padding rows: the solves read zeros there
  const int blk = d.row_blk[R], sub = d.row_sub[R];
  double v;
  if (IMG > 0 && r < wi.cb) {
    const int im = wi.img_off + blk;
    v = (double)d.gr[(int64_t)im * E + sub];
    if (d.pts_free) v -= (double)d.rp[(int64_t)im * E + sub];
  } else {
    const int c = wi.cam_off + (blk - (IMG > 0 ? wi.num_imgs : 0));
    v = (double)d.gc[(int64_t)c * D + sub];
    if (d.pts_free)
      for (int t = d.cam_ptr[c]; t < d.cam_ptr[c + 1]; ++t)
        v -= (double)d.rp[(int64_t)d.cam_imgs[t] * E + IMG + sub];
  }
  d.rhs[R] = v;
}

// ------------------------------------------------------------ K6 back-substitution
// delta_p = L^-T (q_p - sum_a Z_a delta_r(a)), one lane group per point
// (camera_shared_normal_equation_solver.hpp:468-538 in factored form).
template <typename T, int IMG, int D>
__global__ void __launch_bounds__(256) k_backsub(Dev<T> d, int glog2) {
  constexpr int E = IMG + D;
  const int G = 1 << glog2;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = gid >> glog2;
  const int lane = gid & (G - 1);
  bool act = false;
  if (p < d.P) {
    const WinState &s = d.ws[d.pt_win[p]];
    act = win_trial(s) && !s.chol_fail;
  }
  double t[3] = {0.0, 0.0, 0.0};
  if (act && E > 0) {
    for (int k = d.pt_ptr[p] + lane; k < d.pt_ptr[p + 1]; k += G) {
      const int im = d.key_img[k];
      const T *z = d.Z + (int64_t)k * z_stride(E, sizeof(T));
      double dr[E > 0 ? E : 1];
#pragma unroll
      for (int j = 0; j < E; ++j)
        dr[j] = j < IMG ? d.dred[d.img_row[im] + j] : d.dred[d.cam_row[d.img_cam[im]] + (j - IMG)];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < E; ++j) acc += (double)z[a * E + j] * dr[j];
        t[a] += acc;
      }
    }
  }
  for (int m = 1; m < G; m <<= 1) {
#pragma unroll
    for (int a = 0; a < 3; ++a) t[a] += shfl_xor(t[a], m);
  }
  if (act && lane == 0) {
    const T *L = d.Linv + 6 * p;
    const double u0 = (double)d.q[3 * p] - t[0], u1 = (double)d.q[3 * p + 1] - t[1],
                 u2 = (double)d.q[3 * p + 2] - t[2];
    // delta = L^-T u with L^-1 = [[i00,0,0],[i10,i11,0],[i20,i21,i22]]
    const double i00 = L[0], i10 = L[1], i11 = L[2], i20 = L[3], i21 = L[4], i22 = L[5];
    d.dpt[3 * p] = i00 * u0 + i10 * u1 + i20 * u2;
    d.dpt[3 * p + 1] = i11 * u1 + i21 * u2;
    d.dpt[3 * p + 2] = i22 * u2;
  }
}

// ------------------------------------------------------------ step statistics
// Chunked reductions over parameters, one workgroup per chunk (a chunk never
// straddles windows): x_new = x - delta and partial sums of |delta|^2, |x|^2
// and delta . (mu delta + g) in FP64.  kind 0 points, 1 images, 2 cameras.
template <typename T, int IMG, int D>
__global__ void __launch_bounds__(256) k_step(Dev<T> d) {
  constexpr int E = IMG + D;
  const int ch = blockIdx.x;
  const int kind = d.chunk_kind[ch];
  const int w = d.chunk_win[ch];
  const WinState &s = d.ws[w];
  __shared__ double red[3][256];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  const bool act = win_trial(s) && !s.chol_fail;
  if (act && kind < 3) {
    const double mu = s.mu;
    for (int e = d.chunk_begin[ch] + threadIdx.x; e < d.chunk_end[ch]; e += blockDim.x) {
      if (kind == 0) {
        const int p = e / 3, j = e - 3 * p;
        const double dl = d.dpt[e], x = d.pts[e];
        const double g = d.gp[3 * p + j];
        d.pts_new[e] = x - dl;
        a0 += dl * dl; a1 += x * x; a2 += dl * (mu * dl + g);
      } else if (kind == 1) {
        const int im = e / 6, j = e - 6 * im;
        const double dl = d.dred[d.img_row[im] + j], x = d.imgs[e];
        const double g = d.gr[(int64_t)im * E + j];
        d.imgs_new[e] = x - dl;
        a0 += dl * dl; a1 += x * x; a2 += dl * (mu * dl + g);
        if (j == 0) {  // the image's new rotation, formed as the three threads above form it
          double rn[3];
#pragma unroll
          for (int a = 0; a < 3; ++a) rn[a] = d.imgs[e + a] - d.dred[d.img_row[im] + a];
          d.imgq_new[im] = rot_quantities(rn);
        }
      } else {
        const int c = e / (D > 0 ? D : 1), j = e - c * (D > 0 ? D : 1);
        const double dl = d.dred[d.cam_row[c] + j], x = d.cam_free[e];
        const double g = d.gc[(int64_t)c * D + j];
        d.cam_free_new[e] = x - dl;
        d.cam_full_new[(int64_t)c * kFull + d.fidx[j]] = x - dl;
        a0 += dl * dl; a1 += x * x; a2 += dl * (mu * dl + g);
      }
    }
  }
  red[0][threadIdx.x] = a0; red[1][threadIdx.x] = a1; red[2][threadIdx.x] = a2;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      red[0][threadIdx.x] += red[0][threadIdx.x + h];
      red[1][threadIdx.x] += red[1][threadIdx.x + h];
      red[2][threadIdx.x] += red[2][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && kind < 3) {
    d.partial[4 * ch] = red[0][0];
    d.partial[4 * ch + 1] = red[1][0];
    d.partial[4 * ch + 2] = red[2][0];
  }
}

// K7 candidate cost: chunks of keys (kind 3), FP64 residuals at x_new,
// partial sums of r^T Lambda r; camera_shared_mahalanobis_distance_calculator.hpp:28-61.
// `use_new` = 0 evaluates at the current x (initial cost).
template <typename T, int IMG, int D>
__global__ void __launch_bounds__(256) k_cost(Dev<T> d, int use_new) {
  const int ch = blockIdx.x;
  const int kind = d.chunk_kind[ch];
  const WinState &s = d.ws[d.chunk_win[ch]];
  __shared__ double red[256];
  double acc = 0.0;
  const bool act = use_new ? (win_trial(s) && s.take_step) : (s.active != 0);
  if (act && kind == 3) {
    const double *pts = use_new && d.pts_free ? d.pts_new : d.pts;
    const double *imgs = use_new && d.imgs_free ? d.imgs_new : d.imgs;
    const RotQ *imgq = use_new && d.imgs_free ? d.imgq_new : d.imgq;
    const double *camf = use_new && d.cams_free ? d.cam_full_new : d.cam_full;
    for (int k = d.chunk_begin[ch] + threadIdx.x; k < d.chunk_end[ch]; k += blockDim.x) {
      const int p = d.key_pt[k], im = d.key_img[k];
      const double X[3] = {pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]};
      double img[6], cam[kFull];
#pragma unroll
      for (int j = 0; j < 6; ++j) img[j] = imgs[6 * im + j];
      const int c = d.img_cam[im];
#pragma unroll
      for (int j = 0; j < kFull; ++j) cam[j] = camf[(int64_t)c * kFull + j];
      double res[2], Jp[2][3], Ji[2][6], Jc[2][1];
      key_linearize_q<0, false>(X, img, imgq[im], cam, d.fidx, d.key_uv[2 * k], d.key_uv[2 * k + 1], res, Jp, Ji,
                                Jc);
      if (d.has_cov) whiten2(d.key_cov + 3 * k, res[0], res[1]);
      acc += robust_cost(d.loss, d.loss_scale, res[0] * res[0] + res[1] * res[1]);
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0 && kind == 3) d.partial[4 * ch + 3] = red[0];
}

// Gradient / diagonal statistics after a linearization: chunk partials of
// max |g| and max diag(N) (kinds 0..2).
template <typename T, int IMG, int D>
__global__ void __launch_bounds__(256) k_grad_stats(Dev<T> d) {
  constexpr int E = IMG + D;
  const int ch = blockIdx.x;
  const int kind = d.chunk_kind[ch];
  const WinState &s = d.ws[d.chunk_win[ch]];
  __shared__ double red[2][256];
  double gm = 0.0, dm = -1e308;
  if (s.active && s.need_lin && kind < 3) {
    for (int e = d.chunk_begin[ch] + threadIdx.x; e < d.chunk_end[ch]; e += blockDim.x) {
      if (kind == 0) {
        const int p = e / 3, j = e - 3 * p;
        gm = fmax(gm, fabs((double)d.gp[e]));
        dm = fmax(dm, (double)d.V[6 * p + (j == 0 ? 0 : (j == 1 ? 3 : 5))]);
      } else if (kind == 1) {
        const int im = e / 6, j = e - 6 * im;
        gm = fmax(gm, fabs((double)d.gr[(int64_t)im * E + j]));
        dm = fmax(dm, (double)d.Dimg[(int64_t)im * E * E + j * E + j]);
      } else {
        const int c = e / (D > 0 ? D : 1), j = e - c * (D > 0 ? D : 1);
        gm = fmax(gm, fabs((double)d.gc[(int64_t)c * D + j]));
        dm = fmax(dm, (double)d.Uc[(int64_t)c * D * D + j * D + j]);
      }
    }
  }
  red[0][threadIdx.x] = gm; red[1][threadIdx.x] = dm;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][threadIdx.x + h]);
      red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && kind < 3) {
    d.partial[4 * ch] = red[0][0];
    d.partial[4 * ch + 1] = red[1][0];
  }
}

// ------------------------------------------------------------ K8 LM control
// One thread per window.  Madsen–Nielsen–Tingleff Algorithm 3.16 with the
// reference's conventions: F = r^T W r, (N + mu I) delta = g, x_new = x - delta,
// rho = (F - F_new) / (delta^T (mu delta + g)).
//   phase 0: after the initial cost + linearization: F0, mu0 = tau max diag N,
//            found = ||g||_inf <= eps1
//   phase 1: start of an iteration: stop at max_iter, mark the trial
//   phase 2: after the solve: ||delta|| <= eps2 (||x|| + eps2) -> stop
//   phase 3: after the candidate cost: gain ratio, accept / reject, mu update
//   phase 4: after a relinearization: ||g||_inf <= eps1 -> stop
template <typename T>
__global__ void __launch_bounds__(256) k_lm_control(Dev<T> d, int phase) {
  __shared__ double red[4][6];
  const int w = blockIdx.x;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  WinState &s = d.ws[w];
  double F = 0.0, gm = 0.0, dm = -1e308, dn = 0.0, xn = 0.0, gd = 0.0;
  if (phase != 1) {  // phase 1 reads no sums
    // chunk range of this window: chunks are sorted by window
    int cb, ce;
    {
      int lo = 0, hi = d.nchunks;
      while (lo < hi) { const int m = (lo + hi) >> 1; if (d.chunk_win[m] < w) lo = m + 1; else hi = m; }
      cb = lo; hi = d.nchunks;
      while (lo < hi) { const int m = (lo + hi) >> 1; if (d.chunk_win[m] <= w) lo = m + 1; else hi = m; }
      ce = lo;
    }
    // thread partial sums / maxima over the window's chunks; then a fixed
    // butterfly in each wave and the four waves combined in order
    // (selects instead of branches on the chunk kind, so that the loads of
    // several passes are issued together; a select's neutral value leaves
    // the sum or maximum exactly as it was)
#pragma unroll 4
    for (int c = cb + t; c < ce; c += 256) {
      const bool key_chunk = d.chunk_kind[c] == 3;
      const double p0 = d.partial[4 * c], p1 = d.partial[4 * c + 1], p2 = d.partial[4 * c + 2],
                   p3 = d.partial[4 * c + 3];
      F += key_chunk ? p3 : 0.0;
      if (phase == 0 || phase == 4) {
        gm = fmax(gm, key_chunk ? 0.0 : p0);
        dm = fmax(dm, key_chunk ? -1e308 : p1);
      }
      if (phase == 2) {
        dn += key_chunk ? 0.0 : p0;
        xn += key_chunk ? 0.0 : p1;
        gd += key_chunk ? 0.0 : p2;
      }
    }
    for (int m = 1; m < 64; m <<= 1) {
      F += __shfl_xor(F, m, 64);
      gm = fmax(gm, __shfl_xor(gm, m, 64));
      dm = fmax(dm, __shfl_xor(dm, m, 64));
      dn += __shfl_xor(dn, m, 64);
      xn += __shfl_xor(xn, m, 64);
      gd += __shfl_xor(gd, m, 64);
    }
    if (lane == 0) {
      red[wave][0] = F; red[wave][1] = gm; red[wave][2] = dm;
      red[wave][3] = dn; red[wave][4] = xn; red[wave][5] = gd;
    }
    __syncthreads();
    if (t != 0) return;
    F = (red[0][0] + red[1][0]) + (red[2][0] + red[3][0]);
    gm = fmax(fmax(red[0][1], red[1][1]), fmax(red[2][1], red[3][1]));
    dm = fmax(fmax(red[0][2], red[1][2]), fmax(red[2][2], red[3][2]));
    dn = (red[0][3] + red[1][3]) + (red[2][3] + red[3][3]);
    xn = (red[0][4] + red[1][4]) + (red[2][4] + red[3][4]);
    gd = (red[0][5] + red[1][5]) + (red[2][5] + red[3][5]);
  }
  if (t != 0) return;
  if (phase == 0) {
    if (!s.active) return;
    s.cost = F + s.cost_new;  // cost_new holds the constraint part
    s.cost0 = s.cost;
    s.ginf = gm; s.maxdiag = dm;
    s.mu = s.tau * dm;
    s.nu = 2.0;
    s.need_lin = 0;
    if (!(F == F) || !(gm == gm)) { s.status = -4; s.stop_reason = 3; s.active = 0; return; }
    if (gm <= s.eps1) { s.stop_reason = 1; s.active = 0; }
    if (s.max_iter <= 0) { s.active = 0; }
  } else if (phase == 1) {
    s.trial = 0;
    s.take_step = 0;
    s.step_accepted = 0;
    s.chol_fail = 0;
    if (!s.active) return;
    if (s.iter >= s.max_iter) { s.active = 0; s.stop_reason = 0; return; }
    s.iter += 1;
    s.trial = 1;
  } else if (phase == 2) {
    if (!(s.active && s.trial)) return;
    if (s.chol_fail) {
      // (N + mu I) not factorable (in mixed precision the FP32 Schur
      // complement of an ill-conditioned problem loses definiteness as mu
      // shrinks): a rejected step, more damping next time
      s.mu *= s.nu;
      s.nu *= 2.0;
      s.need_lin = 0;
      s.trial = 0;
      if (!(s.mu < 1.0e300)) { s.status = -3; s.stop_reason = 3; s.active = 0; }
      return;
    }
    s.dnorm2 = dn; s.xnorm2 = xn; s.gain_den = gd;
    if (!(dn == dn)) { s.status = -4; s.stop_reason = 3; s.active = 0; s.trial = 0; return; }
    if (sqrt(dn) <= s.eps2 * (sqrt(xn) + s.eps2)) {
      s.stop_reason = 2; s.active = 0; s.trial = 0; return;
    }
    s.take_step = 1;
  } else if (phase == 3) {
    if (!(s.active && s.trial && s.take_step)) return;
    F += s.cost_new;  // constraint part, written by k_constraint_cost
    s.cost_new = F;
    const double rho = (s.cost - F) / s.gain_den;
    s.rho = rho;
    if (rho > 0.0 && F == F) {
      s.cost = F;
      const double t = 2.0 * rho - 1.0;
      s.mu *= fmax(1.0 / 3.0, 1.0 - t * t * t);
      s.nu = 2.0;
      s.accepted += 1;
      s.step_accepted = 1;
      s.need_lin = 1;
    } else {
      s.mu *= s.nu;
      s.nu *= 2.0;
      s.need_lin = 0;
    }
  } else if (phase == 4) {
    if (!(s.active && s.need_lin)) return;
    s.ginf = gm;
    s.need_lin = 0;
    if (gm <= s.eps1) { s.stop_reason = 1; s.active = 0; }
  }
}

// Constraint-row part of the cost (single thread, y order) into cost_new.
template <typename T>
__global__ void k_constraint_cost(Dev<T> d, int use_new) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int w = 0; w < d.nwin; ++w) d.ws[w].cost_new = 0.0;
  for (int q = 0; q < d.ncons; ++q) {
    const ConstraintRow cr = d.cons[q];
    WinState &s = d.ws[cr.win];
    const bool act = use_new ? (s.active && s.trial && s.take_step) : (s.active != 0);
    if (!act) continue;
    if (cr.kind == 0 && cr.block) {
      // r_param * (B r)_param: summed over the record's rows this is r^T B r
      // (ba_gcp_constrained_mahalanobis_distance_calculator.hpp:50-58)
      const double *X = (use_new ? d.pts_new : d.pts) + 3 * cr.entity;
      const ConstraintRow *rec = d.cons + (q - cr.param);
      double t = 0.0;
      for (int j = 0; j < 3; ++j) t += cr.brow[j] * (X[j] - rec[j].y);
      s.cost_new += (X[cr.param] - cr.y) * t;
      continue;
    }
    double x;
    if (cr.kind == 0) x = (use_new ? d.pts_new : d.pts)[3 * cr.entity + cr.param];
    else if (cr.kind == 1) x = (use_new ? d.imgs_new : d.imgs)[6 * cr.entity + cr.param];
    else x = (use_new ? d.cam_full_new : d.cam_full)[(int64_t)cr.entity * kFull + d.fidx[cr.param]];
    const double r = x - cr.y;
    s.cost_new += r * cr.weight * r;
  }
}

// Rotation quantities of every image's current rotation (after x is set
// from the host); one thread per image.
template <typename T>
__global__ void k_img_rot(Dev<T> d) {
  const int im = blockIdx.x * blockDim.x + threadIdx.x;
  if (im >= d.I) return;
  d.imgq[im] = rot_quantities(d.imgs + 6 * (int64_t)im);
}

// Commit x_new -> x for windows whose step was accepted.
template <typename T, int IMG, int D>
__global__ void k_commit(Dev<T> d) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t np = d.pts_free ? 3LL * d.P : 0;
  const int64_t ni = d.imgs_free ? 6LL * d.I : 0;
  const int64_t nc = d.cams_free ? (int64_t)kFull * d.C : 0;
  if (e < np) {
    if (d.ws[d.pt_win[e / 3]].step_accepted) d.pts[e] = d.pts_new[e];
  } else if (e < np + ni) {
    const int64_t j = e - np;
    if (d.ws[d.img_win[j / 6]].step_accepted) {
      d.imgs[j] = d.imgs_new[j];
      if (j % 6 == 0) d.imgq[j / 6] = d.imgq_new[j / 6];
    }
  } else if (e < np + ni + nc) {
    const int64_t j = e - np - ni;
    const int c = (int)(j / kFull), q = (int)(j - (int64_t)c * kFull);
    if (d.ws[d.cam_win[c]].step_accepted) {
      d.cam_full[j] = d.cam_full_new[j];
      if (q < D) d.cam_free[(int64_t)c * D + q] = d.cam_free_new[(int64_t)c * D + q];
    }
  }
}

// ------------------------------------------------------------ posterior covariance
// Point block of N^-1 (N undamped): with V_p = L L^T and Z_a = L^-1 W_a,
//   (N^-1)_pp = V^-1 + V^-1 W S^-1 W^T V^-1 = L^-T (I + sum_{a,b} Z_a S^-1_ab Z_b^T) L^-1
// over the point's keys a, b (S^-1_ab = the E x E block of S^-1 at the rows of
// key a's image / camera and key b's).  One wavefront per point: lane a forms
// U_a = sum_b S^-1_ab Z_b^T, then the lanes sum Z_a U_a.  Replaces the point
// blocks of the dense pseudo-inverse of
// camera_shared_analytical_x_covariance_calculator.hpp:64-87 (equal when N is
// nonsingular).
template <typename T, int IMG, int D>
__global__ void __launch_bounds__(256) k_point_cov(Dev<T> d, const double *Sinv, int64_t ld, int r_off, double *out) {
  constexpr int E = IMG + D;
  constexpr int ZS = z_stride(E > 0 ? E : 1, sizeof(T));
  const int p = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (p >= d.P) return;  // the whole wavefront
  double m[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) m[q] = 0.0;
  if (E > 0) {
    const int kb = d.pt_ptr[p], ke = d.pt_ptr[p + 1];
    for (int a = kb + lane; a < ke; a += 64) {
      const int ia = d.key_img[a];
      int ra[E > 0 ? E : 1];
#pragma unroll
      for (int j = 0; j < E; ++j) ra[j] = (j < IMG ? d.img_row[ia] + j : d.cam_row[d.img_cam[ia]] + (j - IMG)) - r_off;
      double u[E > 0 ? E : 1][3];
#pragma unroll
      for (int i = 0; i < E; ++i) u[i][0] = u[i][1] = u[i][2] = 0.0;
      for (int b = kb; b < ke; ++b) {
        const int ib = d.key_img[b];
        const T *zb = d.Z + (int64_t)b * ZS;
        double z[3][E > 0 ? E : 1];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int j = 0; j < E; ++j) z[c][j] = (double)zb[c * E + j];
#pragma unroll
        for (int j = 0; j < E; ++j) {
          const int rb = (j < IMG ? d.img_row[ib] + j : d.cam_row[d.img_cam[ib]] + (j - IMG)) - r_off;
#pragma unroll
          for (int i = 0; i < E; ++i) {
            const double g = Sinv[(int64_t)ra[i] * ld + rb];
            u[i][0] += g * z[0][j];
            u[i][1] += g * z[1][j];
            u[i][2] += g * z[2][j];
          }
        }
      }
      const T *za = d.Z + (int64_t)a * ZS;
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int i = 0; i < E; ++i) {
          const double zv = (double)za[r * E + i];
          m[3 * r] += zv * u[i][0];
          m[3 * r + 1] += zv * u[i][1];
          m[3 * r + 2] += zv * u[i][2];
        }
    }
  }
  for (int s = 1; s < 64; s <<= 1) {
#pragma unroll
    for (int q = 0; q < 9; ++q) m[q] += __shfl_xor(m[q], s, 64);
  }
  if (lane != 0) return;
  m[0] += 1.0; m[4] += 1.0; m[8] += 1.0;
  // C = Li^T M Li, Li = L^-1 lower
  const T *Lp = d.Linv + 6 * p;
  const double Li[3][3] = {{(double)Lp[0], 0.0, 0.0}, {(double)Lp[1], (double)Lp[2], 0.0},
                           {(double)Lp[3], (double)Lp[4], (double)Lp[5]}};
  double t[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) t[r][c] = m[3 * r] * Li[0][c] + m[3 * r + 1] * Li[1][c] + m[3 * r + 2] * Li[2][c];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      out[9 * (int64_t)p + 3 * r + c] = Li[0][r] * t[0][c] + Li[1][r] * t[1][c] + Li[2][r] * t[2][c];
}

// Dense point rows of the coupling block (full posterior covariance): one
// thread per (point p, row a) writes out[(3 p + a) * ld + j] over the reduced
// columns j of each of p's keys — which = 0: W_k row a; which = 1: row a of
// L_p^-T Z_k = sum_{c >= a} (L_p^-1)[c][a] Z_k[c].  Image columns are written
// (a point has one key per image), camera columns accumulated over the keys
// whose images share the camera, in key order.
template <typename T, int IMG, int D>
__global__ void __launch_bounds__(256) k_point_rows(Dev<T> d, int which, double *out, int64_t ld, int r_off) {
  constexpr int E = IMG + D;
  constexpr int ZS = z_stride(E > 0 ? E : 1, sizeof(T));
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * (int64_t)d.P) return;
  const int p = (int)(t / 3), a = (int)(t - 3 * (int64_t)p);
  const T *Lp = d.Linv + 6 * p;
  // column a of L_p^-1 (rows c >= a): lower-triangular storage l00 l10 l11 l20 l21 l22
  const double lc[3] = {a == 0 ? (double)Lp[0] : 0.0, a == 0 ? (double)Lp[1] : (a == 1 ? (double)Lp[2] : 0.0),
                        a == 0 ? (double)Lp[3] : (a == 1 ? (double)Lp[4] : (double)Lp[5])};
  double *row = out + t * ld;
  for (int s = d.pt_ptr[p]; s < d.pt_ptr[p + 1]; ++s) {
    const int im = d.key_img[s];
    double v[E > 0 ? E : 1];
    if (which == 0) {
#pragma unroll
      for (int j = 0; j < E; ++j) v[j] = (double)d.W[((int64_t)3 * s + a) * E + j];
    } else {
      const T *z = d.Z + (int64_t)s * ZS;
#pragma unroll
      for (int j = 0; j < E; ++j)
        v[j] = lc[0] * (double)z[j] + lc[1] * (double)z[E + j] + lc[2] * (double)z[2 * E + j];
    }
    if (IMG) {
      const int r = d.img_row[im] - r_off;
#pragma unroll
      for (int j = 0; j < IMG; ++j) row[r + j] = v[j];
    }
    if (D) {
      const int r = d.cam_row[d.img_cam[im]] - r_off;
#pragma unroll
      for (int j = 0; j < D; ++j) row[r + j] += v[IMG + j];
    }
  }
}

// Point diagonal blocks of the dense matrices: which = 0 stores V_p, which = 1
// adds V_p^-1 = L_p^-T L_p^-1, at out[(3 p + a) * ld + 3 p + b].
template <typename T>
__global__ void __launch_bounds__(256) k_point_diag(Dev<T> d, int which, double *out, int64_t ld) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= d.P) return;
  double m[3][3];
  if (which == 0) {
    const T *v = d.V + 6 * p;
    const int sym[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) m[a][b] = (double)v[sym[a][b]];
  } else {
    const T *Lp = d.Linv + 6 * p;
    const double Li[3][3] = {{(double)Lp[0], 0.0, 0.0}, {(double)Lp[1], (double)Lp[2], 0.0},
                             {(double)Lp[3], (double)Lp[4], (double)Lp[5]}};
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) m[a][b] = Li[0][a] * Li[0][b] + Li[1][a] * Li[1][b] + Li[2][a] * Li[2][b];
  }
  double *o = out + (int64_t)3 * p * ld + 3 * p;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) o[a * ld + b] = which == 0 ? m[a][b] : o[a * ld + b] + m[a][b];
}

// ------------------------------------------------------------ export kernels
// Per-key residuals and Jacobian blocks in the caller's key order (FP64),
// used by hsba_evaluate / hsba_jacobian.
template <typename T, int IMG, int D>
__global__ void __launch_bounds__(256) k_export_keys(Dev<T> d, double *y, double *jp, double *ji, double *jc,
                                                     int want_j) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d.N) return;
  const int p = d.key_pt[k], im = d.key_img[k];
  const double X[3] = {d.pts[3 * p], d.pts[3 * p + 1], d.pts[3 * p + 2]};
  double img[6], cam[kFull];
#pragma unroll
  for (int j = 0; j < 6; ++j) img[j] = d.imgs[6 * im + j];
  load_cam(d, d.img_cam[im], false, cam);
  double res[2], Jp[2][3], Ji[2][6], Jc[2][D > 0 ? D : 1];
  // observation 0: res is the projection itself
  key_linearize<D, true>(X, img, cam, d.fidx, 0.0, 0.0, res, Jp, Ji, Jc);
  const int64_t o = d.key_orig[k];  // global original key index
  if (y) { y[2 * o] = res[0]; y[2 * o + 1] = res[1]; }
  if (!want_j) return;
  if (jp)
    for (int a = 0; a < 2; ++a)
      for (int m = 0; m < 3; ++m) jp[o * 6 + a * 3 + m] = Jp[a][m];
  if (ji)
    for (int a = 0; a < 2; ++a)
      for (int m = 0; m < 6; ++m) ji[o * 12 + a * 6 + m] = Ji[a][m];
  if (jc && D > 0)
    for (int a = 0; a < 2; ++a)
      for (int m = 0; m < D; ++m) jc[o * 2 * D + a * D + m] = Jc[a][m];
}


// ------------------------------------------------------------ K9 PCG
// Block-Jacobi preconditioned conjugate gradients on the implicit reduced
// system S = U + mu I - sum_p Z_p^T Z_p (no reference counterpart: the
// reference only has the direct SimplicialLDLT; Ceres' ITERATIVE_SCHUR is
// commented out at camera_shared_ceres_optimizor.hpp:476).  Every image owns
// one super-block of E rows (its 6 image rows and the D rows of its camera;
// one image per camera), which is also the preconditioner block.  FP32 Z,
// FP64 vectors, dot products and recurrences.

__device__ __forceinline__ bool pcg_live(const WinState &s, const PcgState &g) {
  return s.active && s.trial && !s.chol_fail && !g.done;
}

template <int IMG, int D>
__device__ __forceinline__ int sup_row(const int *img_row, const int *cam_row, int cam, int im, int j) {
  return j < IMG ? img_row[im] + j : cam_row[cam] + (j - IMG);
}

// Preconditioner blocks M_i = [D_i with the camera block of U_c] + mu I - P_ii,
// factored M_i = L L^T in FP64 registers (one thread per image).
template <typename T, int IMG, int D>
__global__ void __launch_bounds__(64) k_pcg_precond(Dev<T> d) {
  constexpr int E = IMG + D;
  const int im = blockIdx.x * blockDim.x + threadIdx.x;
  if (im >= d.I) return;
  WinState &s = d.ws[d.img_win[im]];
  if (!(s.active && s.trial) || s.chol_fail) return;
  const int c = d.img_cam[im];
  double M[E > 0 ? E : 1][E > 0 ? E : 1];
  const T *Di = d.Dimg + (int64_t)im * E * E;
  const T *P = d.pair_out + (int64_t)im * E * E;  // diagonal pair of image im
#pragma unroll
  for (int a = 0; a < E; ++a)
#pragma unroll
    for (int b = 0; b < E; ++b) {
      double v;
      if (a >= IMG && b >= IMG) v = (double)d.Uc[(int64_t)c * (D > 0 ? D : 1) * D + (a - IMG) * D + (b - IMG)];
      else v = (double)Di[a * E + b];
      M[a][b] = v - (d.pts_free ? (double)P[a * E + b] : 0.0) + (a == b ? s.mu : 0.0);
    }
  bool ok = true;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    double djj = M[j][j];
#pragma unroll
    for (int m = 0; m < j; ++m) djj -= M[j][m] * M[j][m];
    ok = ok && djj > 0.0;
    const double ljj = sqrt(djj > 0.0 ? djj : 1.0);
    M[j][j] = ljj;
#pragma unroll
    for (int i = j + 1; i < E; ++i) {
      double v = M[i][j];
#pragma unroll
      for (int m = 0; m < j; ++m) v -= M[i][m] * M[j][m];
      M[i][j] = v / ljj;
    }
  }
  if (!ok) s.chol_fail = 1;
  double *out = d.Minv + (int64_t)im * E * E;
#pragma unroll
  for (int a = 0; a < E; ++a)
#pragma unroll
    for (int b = 0; b < E; ++b) out[a * E + b] = b <= a ? M[a][b] : 0.0;
}

// Per super-block (image and its own camera, one thread per image):
//   init = 1:  r = rhs, x = 0, z = M^-1 r, p = z
//   init = 0:  x += alpha p, r -= alpha Ap, z = M^-1 r
// and the block's partials of r.z and r.r for the window's dot product.
template <typename T, int IMG, int D>
__global__ void __launch_bounds__(64) k_pcg_apply(Dev<T> d, int init) {
  constexpr int E = IMG + D;
  const int im = blockIdx.x * blockDim.x + threadIdx.x;
  if (im >= d.I) return;
  const int w = d.img_win[im];
  const PcgState &g = d.pcg[w];
  if (!pcg_live(d.ws[w], g)) return;
  const int c = d.img_cam[im];
  const double *L = d.Minv + (int64_t)im * E * E;
  const double alpha = g.alpha;
  double v[E > 0 ? E : 1], rv[E > 0 ? E : 1];
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int R = sup_row<IMG, D>(d.img_row, d.cam_row, c, im, j);
    double r;
    if (init) {
      r = d.rhs[R];
      d.dred[R] = 0.0;
    } else {
      d.dred[R] += alpha * d.pp[R];
      r = d.pr[R] - alpha * d.pAp[R];
    }
    d.pr[R] = r;
    rv[j] = r;
    v[j] = r;
  }
#pragma unroll
  for (int i = 0; i < E; ++i) {
#pragma unroll
    for (int m = 0; m < i; ++m) v[i] -= L[i * E + m] * v[m];
    v[i] /= L[i * E + i];
  }
#pragma unroll
  for (int i = E - 1; i >= 0; --i) {
#pragma unroll
    for (int m = i + 1; m < E; ++m) v[i] -= L[m * E + i] * v[m];
    v[i] /= L[i * E + i];
  }
  double rz = 0.0, rr = 0.0;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int R = sup_row<IMG, D>(d.img_row, d.cam_row, c, im, j);
    d.pz[R] = v[j];
    if (init) d.pp[R] = v[j];
    rz += rv[j] * v[j];
    rr += rv[j] * rv[j];
  }
  d.pdot[3 * im + 1] = rz;
  d.pdot[3 * im + 2] = rr;
}

// t_p = sum_{b in p} Z_b v_sup(b), one lane group per point.
template <typename T, int IMG, int D>
__global__ void __launch_bounds__(256) k_pcg_points(Dev<T> d, int glog2) {
  constexpr int E = IMG + D;
  const int G = 1 << glog2;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = gid >> glog2;
  const int lane = gid & (G - 1);
  bool act = false;
  if (p < d.P) {
    const int w = d.pt_win[p];
    act = pcg_live(d.ws[w], d.pcg[w]);
  }
  double t[3] = {0.0, 0.0, 0.0};
  if (act) {
    constexpr int ZS = z_stride(E > 0 ? E : 1, sizeof(T));
    constexpr int VW = 16 / sizeof(T), NV = (3 * E + VW - 1) / VW;
    typedef T vec __attribute__((ext_vector_type(VW)));
    for (int k = d.pt_ptr[p] + lane; k < d.pt_ptr[p + 1]; k += G) {
      const int im = d.key_img[k], c = d.img_cam[im];
      const vec *zr = (const vec *)(d.Z + (int64_t)k * ZS);
      vec zv[NV > 0 ? NV : 1];
#pragma unroll
      for (int q = 0; q < NV; ++q) zv[q] = zr[q];
      double v[E > 0 ? E : 1];
#pragma unroll
      for (int j = 0; j < E; ++j) v[j] = d.pp[sup_row<IMG, D>(d.img_row, d.cam_row, c, im, j)];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int j = 0; j < E; ++j) t[a] += (double)zv[(a * E + j) / VW][(a * E + j) % VW] * v[j];
    }
  }
  for (int m = 1; m < G; m <<= 1)
#pragma unroll
    for (int a = 0; a < 3; ++a) t[a] += __shfl_xor(t[a], m, 64);
  if (act && lane == 0)
#pragma unroll
    for (int a = 0; a < 3; ++a) d.tpd[3 * p + a] = t[a];
}

// Ap_i = (U_i + mu I) p_i - sum_{a in i} Z_a^T t_p(a), one workgroup per image.
// The image's Z rows are contiguous in the image-major copy (coalesced 128-byte
// rows); t_p is a 24-byte gather from a small, cache-resident vector.
template <typename T, int IMG, int D>
__global__ void __launch_bounds__(128) k_pcg_images(Dev<T> d) {
  constexpr int E = IMG + D;
  constexpr int ZS = z_stride(E, sizeof(T));
  constexpr int VW = 16 / sizeof(T), NV = (3 * E + VW - 1) / VW;
  typedef T vec __attribute__((ext_vector_type(VW)));
  const int im = blockIdx.x;
  const int w = d.img_win[im];
  if (!pcg_live(d.ws[w], d.pcg[w])) return;
  const int c = d.img_cam[im];
  double acc[E > 0 ? E : 1];
#pragma unroll
  for (int j = 0; j < E; ++j) acc[j] = 0.0;
  for (int t = d.img_ptr[im] + threadIdx.x; t < d.img_ptr[im + 1]; t += blockDim.x) {
    const int p = d.ikey_pt[t];
    const vec *zr = (const vec *)(d.Zi + (int64_t)t * ZS);
    vec zv[NV > 0 ? NV : 1];
#pragma unroll
    for (int q = 0; q < NV; ++q) zv[q] = zr[q];
    const double t0 = d.tpd[3 * p], t1 = d.tpd[3 * p + 1], t2 = d.tpd[3 * p + 2];
#pragma unroll
    for (int j = 0; j < E; ++j)
      acc[j] += (double)zv[j / VW][j % VW] * t0 + (double)zv[(E + j) / VW][(E + j) % VW] * t1 +
                (double)zv[(2 * E + j) / VW][(2 * E + j) % VW] * t2;
  }
  for (int m = 1; m < 64; m <<= 1)
#pragma unroll
    for (int j = 0; j < E; ++j) acc[j] += __shfl_xor(acc[j], m, 64);
  __shared__ double red[2][E > 0 ? E : 1];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0)
#pragma unroll
    for (int j = 0; j < E; ++j) red[wave][j] = acc[j];
  __syncthreads();
  if (threadIdx.x < 64) {  // wave 0: rows of the super-block, then p.Ap over them
    double prod = 0.0;
    if (threadIdx.x < E) {
      const int a = threadIdx.x;
      const T *Di = d.Dimg + (int64_t)im * E * E;
      double y = 0.0;
      for (int b = 0; b < E; ++b) {
        double u;
        if (a >= IMG && b >= IMG) u = (double)d.Uc[(int64_t)c * (D > 0 ? D : 1) * D + (a - IMG) * D + (b - IMG)];
        else u = (double)Di[a * E + b];
        y += u * d.pp[sup_row<IMG, D>(d.img_row, d.cam_row, c, im, b)];
      }
      const int r = sup_row<IMG, D>(d.img_row, d.cam_row, c, im, a);
      y += d.ws[w].mu * d.pp[r];
      const double ap = y - (d.pts_free ? (red[0][a] + red[1][a]) : 0.0);
      d.pAp[r] = ap;
      prod = d.pp[r] * ap;
    }
    for (int m = 1; m < 64; m <<= 1) prod += __shfl_xor(prod, m, 64);
    if (threadIdx.x == 0) d.pdot[3 * im] = prod;
  }
}

// Per-window dot products (one workgroup per window, fixed tree):
//   mode 0: rz = r.z, bnorm2 = rr = r.r (start)      mode 1: alpha = rz / p.Ap
//   mode 2: rz' = r.z, rr = r.r, beta, convergence, it++
template <typename T>
__global__ void __launch_bounds__(256) k_pcg_dot(Dev<T> d, int mode) {
  const int w = blockIdx.x;
  const WinInfo &wi = d.wi[w];
  PcgState &g = d.pcg[w];
  if (!pcg_live(d.ws[w], g)) return;
  // every row of the window lies in exactly one image's super-block (one
  // image per camera), so the per-image partials add up to the dot products
  double a0 = 0.0, a1 = 0.0;
  for (int i = threadIdx.x; i < wi.num_imgs; i += blockDim.x) {
    const double *pd = d.pdot + 3 * (int64_t)(wi.img_off + i);
    if (mode == 1) a0 += pd[0];
    else { a0 += pd[1]; a1 += pd[2]; }
  }
  __shared__ double red[2][256];
  red[0][threadIdx.x] = a0; red[1][threadIdx.x] = a1;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) { red[0][threadIdx.x] += red[0][threadIdx.x + h]; red[1][threadIdx.x] += red[1][threadIdx.x + h]; }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (mode == 0) {
    g.rz = red[0][0]; g.rr = red[1][0]; g.bnorm2 = red[1][0];
    if (!(g.bnorm2 > 0.0)) g.done = 1;
  } else if (mode == 1) {
    g.pap = red[0][0];
    g.alpha = g.pap > 0.0 ? g.rz / g.pap : 0.0;
    if (!(g.pap > 0.0)) g.done = 1;  // exact solution reached or loss of definiteness
  } else {
    g.rr = red[1][0];
    g.beta = g.rz > 0.0 ? red[0][0] / g.rz : 0.0;
    g.rz = red[0][0];
    g.it += 1;
    if (g.rr <= g.rtol2 * g.bnorm2 || g.it >= g.max_it) g.done = 1;
  }
}

// mode 0: x += alpha p, r -= alpha Ap        mode 1: p = z + beta p
// mode 2 (start): r = rhs
template <typename T>
__global__ void k_pcg_vec(Dev<T> d, int mode, int rows) {
  const int R = blockIdx.x * blockDim.x + threadIdx.x;
  if (R >= rows) return;
  int lo = 0, hi = d.nwin - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d.wi[mid].r_off <= R) lo = mid; else hi = mid - 1;
  }
  const WinInfo &wi = d.wi[lo];
  if (R - wi.r_off >= wi.n) return;
  const PcgState &g = d.pcg[lo];
  if (!pcg_live(d.ws[lo], g)) return;
  if (mode == 0) {
    d.dred[R] += g.alpha * d.pp[R];
    d.pr[R] -= g.alpha * d.pAp[R];
  } else if (mode == 1) {
    d.pp[R] = d.pz[R] + g.beta * d.pp[R];
  } else {
    d.pr[R] = d.rhs[R];
  }
}

// Reset the solver state of every window at the start of a trial.
template <typename T>
__global__ void k_pcg_reset(Dev<T> d, double rtol2, int max_it) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= d.nwin) return;
  PcgState &g = d.pcg[w];
  g.rz = g.pap = g.alpha = g.beta = g.bnorm2 = g.rr = 0.0;
  g.rtol2 = rtol2;
  g.it = 0;
  g.max_it = max_it;
  const WinState &s = d.ws[w];
  g.done = (s.active && s.trial && !s.chol_fail) ? 0 : 1;
}

// Adds the PCG iteration count of the trial to the window statistics.
template <typename T>
__global__ void k_pcg_finish(Dev<T> d) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= d.nwin) return;
  WinState &s = d.ws[w];
  if (s.active && s.trial) s.pcg_iters += d.pcg[w].it;
}

// ------------------------------------------------------------ launchers
template <typename T, int IMG, int D>
struct KernelSetImpl final : KernelSet<T> {
  static constexpr int E = IMG + D;
  int E_() const override { return E; }
  void linearize(const Dev<T> &d, hipStream_t st, const LinSeg &seg, const int32_t *plist) override {
    if (!d.pts_free || d.P == 0 || seg.t0[3] <= 0) return;
    hipLaunchKernelGGL((k_linearize<T, IMG, D>), dim3((seg.t0[3] + 255) / 256), dim3(256), 0, st, d, seg, plist);
  }
  void image_reduce(const Dev<T> &d, hipStream_t st) override {
    if (E == 0 || d.I == 0) return;
    if (d.nich > 0) hipLaunchKernelGGL((k_image_reduce<T, IMG, D>), dim3(d.nich), dim3(128), 0, st, d);
    {
      constexpr int NV = E * (E + 1) / 2 + E;
      const int64_t n = (int64_t)d.I * NV;
      if (n > 0) hipLaunchKernelGGL((k_image_combine<T, IMG, D>), dim3((n + 255) / 256), dim3(256), 0, st, d);
    }
  }
  void camera_reduce(const Dev<T> &d, hipStream_t st) override {
    if (D == 0 || d.C == 0) return;
    hipLaunchKernelGGL((k_camera_reduce<T, IMG, D>), dim3((d.C + 127) / 128), dim3(128), 0, st, d);
  }
  void constraints(const Dev<T> &d, hipStream_t st) override {
    if (d.ncons == 0) return;
    hipLaunchKernelGGL((k_constraints<T, IMG, D>), dim3(1), dim3(64), 0, st, d);
  }
  void point_factor(const Dev<T> &d, hipStream_t st) override {
    if (!d.pts_free || d.P == 0) return;
    hipLaunchKernelGGL((k_point_factor<T>), dim3((d.P + 255) / 256), dim3(256), 0, st, d);
  }
  void key_z(const Dev<T> &d, hipStream_t st) override {
    if (!d.pts_free || E == 0 || d.N == 0) return;
    hipLaunchKernelGGL((k_key_z<T, (E > 0 ? E : 1)>), dim3((d.N + 255) / 256), dim3(256), 0, st, d);
  }
  bool has_pair_variant(int variant) const override { return variant == 0 ? E > 0 : (E >= 1 && E <= 15); }
  void pairs(const Dev<T> &d, hipStream_t st, int variant, const int32_t *plist, int count) override {
    if (!d.pts_free || E == 0 || count <= 0) return;
    if constexpr (E >= 1 && E <= 15) {
      if (variant == 1) {
        hipLaunchKernelGGL((k_pairs_mma<T, E>), dim3(count), dim3(256), 0, st, d, plist);
        return;
      }
      if (variant == 2) {
        hipLaunchKernelGGL((k_pairs_mma_wave<T, E>), dim3((count + 3) / 4), dim3(256), 0, st, d, plist, count);
        return;
      }
    }
    hipLaunchKernelGGL((k_pairs<T, (E > 0 ? E : 1)>), dim3((count + 3) / 4), dim3(256), 0, st, d, plist, count);
  }
  void assemble(const Dev<T> &d, hipStream_t st, int64_t lower_total, int rows) override {
    if (E == 0 || rows == 0) return;
    if (lower_total > 0) {
      if (d.nwin > 1 && d.nwin <= 65535)
        hipLaunchKernelGGL((k_assemble_win<T, IMG, D>), dim3((unsigned)((d.lower_max + 255) / 256), d.nwin), dim3(256),
                           0, st, d);
      else
        hipLaunchKernelGGL((k_assemble<T, IMG, D>), dim3((lower_total + 255) / 256), dim3(256), 0, st, d,
                           lower_total);
    }
    hipLaunchKernelGGL((k_assemble_rhs<T, IMG, D>), dim3((rows + 255) / 256), dim3(256), 0, st, d, rows);
  }
  void backsub(const Dev<T> &d, hipStream_t st, int glog2) override {
    if (!d.pts_free || d.P == 0) return;
    const int64_t threads = (int64_t)d.P << glog2;
    hipLaunchKernelGGL((k_backsub<T, IMG, D>), dim3((threads + 255) / 256), dim3(256), 0, st, d, glog2);
  }
  void step(const Dev<T> &d, hipStream_t st) override {
    hipLaunchKernelGGL((k_step<T, IMG, D>), dim3(d.nchunks), dim3(256), 0, st, d);
  }
  void cost(const Dev<T> &d, hipStream_t st, int use_new) override {
    hipLaunchKernelGGL((k_constraint_cost<T>), dim3(1), dim3(64), 0, st, d, use_new);
    hipLaunchKernelGGL((k_cost<T, IMG, D>), dim3(d.nchunks), dim3(256), 0, st, d, use_new);
  }
  void grad_stats(const Dev<T> &d, hipStream_t st) override {
    hipLaunchKernelGGL((k_grad_stats<T, IMG, D>), dim3(d.nchunks), dim3(256), 0, st, d);
  }
  void lm_control(const Dev<T> &d, hipStream_t st, int phase) override {
    hipLaunchKernelGGL((k_lm_control<T>), dim3(d.nwin), dim3(256), 0, st, d, phase);
  }
  void image_rotations(const Dev<T> &d, hipStream_t st) override {
    if (d.I > 0) hipLaunchKernelGGL((k_img_rot<T>), dim3((d.I + 255) / 256), dim3(256), 0, st, d);
  }
  void commit(const Dev<T> &d, hipStream_t st) override {
    const int64_t n = (d.pts_free ? 3LL * d.P : 0) + (d.imgs_free ? 6LL * d.I : 0) +
                      (d.cams_free ? (int64_t)kFull * d.C : 0);
    if (n == 0) return;
    hipLaunchKernelGGL((k_commit<T, IMG, D>), dim3((n + 255) / 256), dim3(256), 0, st, d);
  }
  double pcg_rtol = 1e-6;
  int pcg_max_it = 100;
  void pcg_setup(const Dev<T> &d, hipStream_t st, int glog2) override {
    const int rows = (int)d.pcg_rows;
    hipLaunchKernelGGL((k_pcg_reset<T>), dim3((d.nwin + 63) / 64), dim3(64), 0, st, d, pcg_rtol * pcg_rtol,
                       pcg_max_it);
    if (E == 0 || rows == 0) return;
    hipLaunchKernelGGL((k_pcg_precond<T, IMG, D>), dim3((d.I + 63) / 64), dim3(64), 0, st, d);
    hipLaunchKernelGGL((k_pcg_reset<T>), dim3((d.nwin + 63) / 64), dim3(64), 0, st, d, pcg_rtol * pcg_rtol,
                       pcg_max_it);
    hipLaunchKernelGGL((k_pcg_apply<T, IMG, D>), dim3((d.I + 63) / 64), dim3(64), 0, st, d, 1);
    hipLaunchKernelGGL((k_pcg_dot<T>), dim3(d.nwin), dim3(256), 0, st, d, 0);
  }
  void pcg_iteration(const Dev<T> &d, hipStream_t st, int glog2) override {
    const int rows = (int)d.pcg_rows;
    if (E == 0 || rows == 0) return;
    const int64_t threads = (int64_t)d.P << glog2;
    if (d.pts_free && d.P)
      hipLaunchKernelGGL((k_pcg_points<T, IMG, D>), dim3((threads + 255) / 256), dim3(256), 0, st, d, glog2);
    hipLaunchKernelGGL((k_pcg_images<T, IMG, D>), dim3(d.I), dim3(128), 0, st, d);
    hipLaunchKernelGGL((k_pcg_dot<T>), dim3(d.nwin), dim3(256), 0, st, d, 1);
    hipLaunchKernelGGL((k_pcg_apply<T, IMG, D>), dim3((d.I + 63) / 64), dim3(64), 0, st, d, 0);
    hipLaunchKernelGGL((k_pcg_dot<T>), dim3(d.nwin), dim3(256), 0, st, d, 2);
    hipLaunchKernelGGL((k_pcg_vec<T>), dim3((rows + 255) / 256), dim3(256), 0, st, d, 1, rows);
  }
  void pcg_finish(const Dev<T> &d, hipStream_t st) override {
    hipLaunchKernelGGL((k_pcg_finish<T>), dim3((d.nwin + 63) / 64), dim3(64), 0, st, d);
  }
  void pcg_params(double rtol, int max_it) override {
    pcg_rtol = rtol;
    pcg_max_it = max_it;
  }
  void point_cov(const Dev<T> &d, hipStream_t st, const double *Sinv, int64_t ld, int r_off, double *out) override {
    if (d.P == 0) return;
    hipLaunchKernelGGL((k_point_cov<T, IMG, D>), dim3((d.P + 3) / 4), dim3(256), 0, st, d, Sinv, ld, r_off, out);
  }
  void point_rows(const Dev<T> &d, hipStream_t st, int which, double *out, int64_t ld, int r_off) override {
    if (d.P == 0 || IMG + D == 0) return;
    hipLaunchKernelGGL((k_point_rows<T, IMG, D>), dim3((3 * d.P + 255) / 256), dim3(256), 0, st, d, which, out, ld,
                       r_off);
  }
  void point_diag(const Dev<T> &d, hipStream_t st, int which, double *out, int64_t ld) override {
    if (d.P == 0) return;
    hipLaunchKernelGGL((k_point_diag<T>), dim3((d.P + 255) / 256), dim3(256), 0, st, d, which, out, ld);
  }
  void export_keys(const Dev<T> &d, hipStream_t st, double *y, double *jp, double *ji, double *jc,
                   int want_j) override {
    if (d.N == 0) return;
    hipLaunchKernelGGL((k_export_keys<T, IMG, D>), dim3((d.N + 255) / 256), dim3(256), 0, st, d, y, jp, ji,
                       jc, want_j);
  }
};

template <typename T>
KernelSet<T> *make_kernel_set(int img, int dcam) {
#define HSBA_CASE(I_, D_) \
  if (img == I_ && dcam == D_) return new KernelSetImpl<T, I_, D_>();
  HSBA_CASE(6, 0) HSBA_CASE(6, 1) HSBA_CASE(6, 2) HSBA_CASE(6, 3) HSBA_CASE(6, 4) HSBA_CASE(6, 5)
  HSBA_CASE(6, 6) HSBA_CASE(6, 7) HSBA_CASE(6, 8) HSBA_CASE(6, 9) HSBA_CASE(6, 10)
  HSBA_CASE(0, 0) HSBA_CASE(0, 1) HSBA_CASE(0, 2) HSBA_CASE(0, 3) HSBA_CASE(0, 4) HSBA_CASE(0, 5)
  HSBA_CASE(0, 6) HSBA_CASE(0, 7) HSBA_CASE(0, 8) HSBA_CASE(0, 9) HSBA_CASE(0, 10)
#undef HSBA_CASE
  return nullptr;
}

template KernelSet<float> *make_kernel_set<float>(int, int);
template KernelSet<double> *make_kernel_set<double>(int, int);

}  // namespace hsba
