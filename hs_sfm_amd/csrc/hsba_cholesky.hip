// hsba_cholesky.hip — dense FP64 Cholesky and triangular solves of the
// reduced camera systems, replacing Eigen::SimplicialLDLT
// (include/hs_sfm/bundle_adjustment/camera_shared_normal_equation_solver.hpp:448-466).
//
// Storage: each window's reduced matrix is row-major with leading dimension
// ld (a multiple of 64), lower triangle valid, padding rows / columns set to
// the identity once.  64 x 64 tiles are staged through LDS.
//
// Diagonal tiles: a blocked factorization with 16-wide panels (each 16 x 16
// diagonal block eliminated by one wavefront, four columns per step on the
// matrix cores, while the other three waves form the panel, trailing and
// inverse blocks around it) yields L_kk^-1, the only form of the diagonal
// factor tiles that the solves and the inverse read; panel tiles are then
// GEMMs with L_kk^-1.
//
// Large windows (more than 4 tiles): one fused launch per panel step k.
// Every workgroup of step k owns one trailing tile (i, j), k < j <= i: it forms
// L_ik = A_ik L_kk^-T and L_jk itself and applies A_ij -= L_ik L_jk^T; the
// workgroup of tile (k+1, k+1) then factors that tile (look-ahead), so each
// step is a single launch.  The right-hand side rides along: step k also
// forms y_k = L_kk^-1 b_k and updates b_j -= L_jk y_k.  The factor goes to a
// separate buffer (column k of A is read by every workgroup of step k); its
// diagonal tiles are not written.  The backward solve L^T x = y is one
// multi-workgroup launch per four block rows.
//
// Small windows: one workgroup (256 threads) per window runs the whole
// blocked factorization and both solves; a batch of windows is one launch.
//
// A non-positive pivot marks the window (WinState::chol_fail); the reference
// reports the same condition as SimplicialLDLT's info() != Success.
#include <algorithm>

#include "hsba_kernels.h"

namespace hsba {

constexpr int NB = 64;
constexpr int LDT = NB + 1;  // LDS tile pitch (bank-conflict padding)

// Tile <-> LDS with 256 threads: every global load of the tile is issued
// before the first LDS store, so the 16 loads of a thread overlap instead of
// each waiting out the memory latency.
__device__ __forceinline__ void fetch_tile(double v[16], const double *A, int64_t ld, int r0, int c0) {
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int e = threadIdx.x + 256 * q;
    v[q] = A[(int64_t)(r0 + (e >> 6)) * ld + c0 + (e & 63)];
  }
}
__device__ __forceinline__ void put_tile(double (*t)[LDT], const double v[16]) {
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int e = threadIdx.x + 256 * q;
    t[e >> 6][e & 63] = v[q];
  }
}
__device__ __forceinline__ void load_tile(double (*t)[LDT], const double *A, int64_t ld, int r0, int c0) {
  double v[16];
  fetch_tile(v, A, ld, r0, c0);
  put_tile(t, v);
}
__device__ __forceinline__ void store_tile(const double (*t)[LDT], double *A, int64_t ld, int r0, int c0) {
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int e = threadIdx.x + 256 * q;
    A[(int64_t)(r0 + (e >> 6)) * ld + c0 + (e & 63)] = t[e >> 6][e & 63];
  }
}

// 1/d for d > 0: hardware reciprocal estimate refined by two Newton–Raphson
// steps r <- r (2 - d r) (quadratic convergence, so ~1 ulp).
__device__ __forceinline__ double recip_nr(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = r * (2.0 - d * r);
  r = r * (2.0 - d * r);
  return r;
}

// Lanes of one wave see each other's earlier LDS writes after this point: the
// LDS unit executes one wave's instructions in issue order, and the fences
// keep the compiler from moving LDS accesses across the call.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int PB = 16;  // panel width inside a 64 x 64 tile

typedef double dbl4 __attribute__((ext_vector_type(4)));

// 16 x 16 x 16 block products of one wave on the matrix cores
// (v_mfma_f64_16x16x4_f64), operands read from LDS tiles by block offset:
//   blk_nt: acc += A B^T,  blk_nn: acc += A B,
// A the 16 x 16 block of a at (ar, ac), B that of b at (br, bc).  Operand
// lanes: A[i][k] from lane i + 16 k, B[k][j] from lane j + 16 k; the result
// D[i][j] is in lane j + 16 (i % 4), element i / 4 (blk_put / blk_sub).
__device__ __forceinline__ dbl4 blk_nt(const double (*a)[LDT], int ar, int ac, const double (*b)[LDT], int br,
                                       int bc, dbl4 acc) {
  const int l = threadIdx.x & 63, li = l & 15, lk = l >> 4;
#pragma unroll
  for (int k = 0; k < 16; k += 4)
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ar + li][ac + k + lk], b[br + li][bc + k + lk], acc, 0, 0, 0);
  return acc;
}
__device__ __forceinline__ dbl4 blk_nn(const double (*a)[LDT], int ar, int ac, const double (*b)[LDT], int br,
                                       int bc, dbl4 acc) {
  const int l = threadIdx.x & 63, li = l & 15, lk = l >> 4;
#pragma unroll
  for (int k = 0; k < 16; k += 4)
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ar + li][ac + k + lk], b[br + k + lk][bc + li], acc, 0, 0, 0);
  return acc;
}
// out block at (r, c) = s * acc  (s = +1 or -1)
__device__ __forceinline__ void blk_put(double (*out)[LDT], int r, int c, dbl4 acc, double s) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int e = 0; e < 4; ++e) out[r + (l >> 4) + 4 * e][c + (l & 15)] = s * acc[e];
}
// out block at (r, c) -= acc
__device__ __forceinline__ void blk_sub(double (*out)[LDT], int r, int c, dbl4 acc) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int e = 0; e < 4; ++e) out[r + (l >> 4) + 4 * e][c + (l & 15)] -= acc[e];
}

// One wavefront factors the 16 x 16 diagonal block of wa at (c0, c0) (lower
// triangle, SPD) and writes L_p^-1 (zero upper) to the same block of wx; the
// block of wa is only read (nothing reads the diagonal factor blocks L_p).
// A (symmetric, both triangles) and X (starting at I) live in matrix-core
// result layout: lane jj + 16 q, element e holds row q + 4 e, column jj.
// Elimination by 4 x 4 pivot blocks J (rows / columns p .. p+3): the block's
// four columns of A are published through LDS, and every lane forms
// P = L_P D_P L_P^T (unit lower L_P, four reciprocals, the same in every
// lane), L_P^-1 and, for its row i, y_i = A_iJ L_P^-T, z_i = y_i D_P^-1 and
// G_i = z_i L_P^-1.  Then one v_mfma_f64_16x16x4_f64 per matrix applies
//   A_i -= G_i A_J           (rows i below the block: A_ic -= z_i . y_c)
//   X_i -= G_i X_J           (rows below the block)
//   X_J += (L_P^-1 - I) X_J  (the block's rows: X_J <- L_P^-1 X_J)
// with the product's k = 4 dimension the pivot block's rows A_J, X_J as they
// stand, which every lane holds as its own accumulator element J.  This is
// four steps of LDL^T at once; columns left of the block are not kept.  At
// the end X = Lt^-1 (A = Lt D Lt^T, Lt unit lower) and L^-1 = D^-1/2 X.  A
// non-positive (or NaN) pivot sets bad and continues with pivot 1.
struct WaveScratch16 {
  double pan[PB][4];  // A[i][p .. p+3]: the pivot block's columns
};
// a, b, c or d by q = 0..3 as selects; the operands are pinned to registers
// first so that the compiler does not turn the choice into branches.
__device__ __forceinline__ double sel4(int q, double a, double b, double c, double d) {
  asm volatile("" : "+v"(a), "+v"(b), "+v"(c), "+v"(d));
  double v = a;
  v = (q == 1) ? b : v;
  v = (q == 2) ? c : v;
  v = (q == 3) ? d : v;
  return v;
}
__device__ __forceinline__ void diag16_wave(double (*wa)[LDT], double (*wx)[LDT], int c0, WaveScratch16 &sc,
                                            bool &bad) {
  const int l = threadIdx.x & 63, jj = l & 15, q = l >> 4;
  dbl4 accA, accX;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int i = q + 4 * e;
    accA[e] = (i >= jj) ? wa[c0 + i][c0 + jj] : wa[c0 + jj][c0 + i];
    accX[e] = (i == jj) ? 1.0 : 0.0;
  }
  double dv[PB];
#pragma unroll
  for (int J = 0; J < 4; ++J) {
    const int p = 4 * J;
    if (jj >= p && jj < p + 4) {
#pragma unroll
      for (int e = 0; e < 4; ++e) sc.pan[q + 4 * e][jj - p] = accA[e];
    }
    wave_sync();
    const double p00 = sc.pan[p][0], p10 = sc.pan[p + 1][0], p11 = sc.pan[p + 1][1];
    const double p20 = sc.pan[p + 2][0], p21 = sc.pan[p + 2][1], p22 = sc.pan[p + 2][2];
    const double p30 = sc.pan[p + 3][0], p31 = sc.pan[p + 3][1], p32 = sc.pan[p + 3][2], p33 = sc.pan[p + 3][3];
    double ai[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) ai[r] = sc.pan[jj][r];
    wave_sync();  // the next block's writes come after every lane's reads
    // P = L_P D_P L_P^T (t_rs = l_rs d_s)
    const double d0 = p00;
    const double r0 = recip_nr(d0 > 0.0 ? d0 : 1.0);
    const double l10 = p10 * r0, l20 = p20 * r0, l30 = p30 * r0;
    const double d1 = p11 - l10 * p10;
    const double r1 = recip_nr(d1 > 0.0 ? d1 : 1.0);
    const double t21 = p21 - l20 * p10, t31 = p31 - l30 * p10;
    const double l21 = t21 * r1, l31 = t31 * r1;
    const double d2 = p22 - l20 * p20 - l21 * t21;
    const double r2 = recip_nr(d2 > 0.0 ? d2 : 1.0);
    const double t32 = p32 - l30 * p20 - l31 * t21;
    const double l32 = t32 * r2;
    const double d3 = p33 - l30 * p30 - l31 * t31 - l32 * t32;
    const double r3 = recip_nr(d3 > 0.0 ? d3 : 1.0);
    bad = bad || !(d0 > 0.0 && d1 > 0.0 && d2 > 0.0 && d3 > 0.0);
    dv[p] = d0;
    dv[p + 1] = d1;
    dv[p + 2] = d2;
    dv[p + 3] = d3;
    // M = L_P^-1 (unit lower)
    const double m10 = -l10, m21 = -l21, m32 = -l32;
    const double m20 = -(l20 + l21 * m10), m31 = -(l31 + l32 * m21);
    const double m30 = -(l30 + l31 * m10 + l32 * m20);
    // row jj: y = a L_P^-T, z = y D_P^-1, G = z M
    double y[4];
    y[0] = ai[0];
    y[1] = ai[1] - l10 * y[0];
    y[2] = ai[2] - l20 * y[0] - l21 * y[1];
    y[3] = ai[3] - l30 * y[0] - l31 * y[1] - l32 * y[2];
    const double z0 = y[0] * r0, z1 = y[1] * r1, z2 = y[2] * r2, z3 = y[3] * r3;
    const double g3 = z3;
    const double g2 = z2 + z3 * m32;
    const double g1 = z1 + z2 * m21 + z3 * m31;
    const double g0 = z0 + z1 * m10 + z2 * m20 + z3 * m30;
    const double gk = sel4(q, g0, g1, g2, g3);  // this lane's k = q entry of G
    const bool below = jj >= p + 4;
    // the block's rows: (M - I) at (jj - p, q)
    const int rr = jj - p;
    double hb = 0.0;
    hb = (rr == 1 && q == 0) ? m10 : hb;
    hb = (rr == 2 && q == 0) ? m20 : hb;
    hb = (rr == 2 && q == 1) ? m21 : hb;
    hb = (rr == 3 && q == 0) ? m30 : hb;
    hb = (rr == 3 && q == 1) ? m31 : hb;
    hb = (rr == 3 && q == 2) ? m32 : hb;
    // operand A[i = jj][k = q]; operand B[k = q][c = jj] = row p + q of A or X, element J of this lane
    const double opA = below ? -gk : 0.0;
    const double opX = below ? -gk : hb;
    const double bA = accA[J], bX = accX[J];
    accA = __builtin_amdgcn_mfma_f64_16x16x4f64(opA, bA, accA, 0, 0, 0);
    accX = __builtin_amdgcn_mfma_f64_16x16x4f64(opX, bX, accX, 0, 0, 0);
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int i = q + 4 * e;
    const double di = sel4(q, dv[4 * e], dv[4 * e + 1], dv[4 * e + 2], dv[4 * e + 3]);
    wx[c0 + i][c0 + jj] = (jj <= i) ? accX[e] * rsqrt(di > 0.0 ? di : 1.0) : 0.0;
  }
}

// Factor a 64 x 64 SPD tile and invert the factor, in LDS: wa holds the tile
// (lower 16 x 16 blocks valid, synchronised); with pa != null the tile is
// first updated, T = wa - pa pa^T, fused into the schedule below.  Writes
// L^-1 (zero upper) to dst; a non-positive pivot sets *fail.  Only L^-1 is
// kept: the solves and the inverse use L_kk^-1 for diagonal tiles and never
// read L_kk.  wa ends with the off-diagonal blocks of L (its diagonal
// blocks hold the updated T_pp), wx (scratch, must not alias wa or pa) with
// L^-1 in its lower blocks.  256 threads.
//
// Right-looking blocked Cholesky with 16-wide panels p = 0..3 and blocked
// inversion of the factor:
//   L_pp = chol(T_pp), M_p = L_pp^-1        (wave 0, diag16_wave)
//   L_rp = T_rp M_p^T                       (r > p)
//   T_rc -= L_rp L_cp^T                     (p < c <= r)
//   X_rc = -M_r S_rc, S_rc = sum_{k=c}^{r-1} L_rk X_kc   (X = L^-1, X_pp = M_p)
// Wave 0 runs the chain of diagonal blocks (update of T_pp, then its factor);
// waves 1-3 do every other block product (16 x 16 x 16 on the matrix cores)
// in the stage where its operands are ready, so that between two diagonal
// blocks wave 0 waits for one panel block only.  Stages (barriers between):
//   S1   w0: T_00 update, diag(0)      w1-3: T updates of the other 9 blocks
//   Sa1  w1-3: L_10 L_20 L_30
//   Sb1  w0: T_11 -= L_10 L_10^T, diag(1)
//        w1: T_21 T_31   w2: T_22 T_32   w3: T_33, S_10 = L_10 M_0
//   Sa2  w1: L_21   w2: L_31   w3: X_10
//   Sb2  w0: T_22 -= L_21 L_21^T, diag(2)
//        w1: T_32, S_21 = L_21 M_1   w2: T_33   w3: S_20 = L_20 M_0 + L_21 X_10
//   Sa3  w1: L_32   w2: X_21   w3: X_20
//   Sb3  w0: T_33 -= L_32 L_32^T, diag(3)   w1-3: S_32, S_31, S_30
//   end  w1-3: X_32, X_31, X_30
// The S blocks are staged in the unused upper blocks (c, r) of wx.
//
// nreal = the tile's rows that belong to the system (the rest is padding,
// identity on the diagonal and zero elsewhere, and stays so under every
// update).  A tile with at most 16 real rows (the last tile of a system whose
// size leaves up to 16 rows in it) has L^-1 = [M_0 0; 0 I]: only block 0 is
// updated and factored.
__device__ __forceinline__ void tile_factor_inverse_lds(double *dst, int64_t ld, int *fail, double (*wa)[LDT],
                                                        const double (*pa)[LDT], double (*wx)[LDT],
                                                        WaveScratch16 &ws, int nreal) {
  const int tid = threadIdx.x, w = tid >> 6;
  const dbl4 zero = dbl4{0.0, 0.0, 0.0, 0.0};
  bool bad = false;
  auto sub_pp = [&](int r, int c) {  // T_rc -= sum_k P_rk P_ck^T
    dbl4 acc = zero;
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) acc = blk_nt(pa, PB * r, PB * kb, pa, PB * c, PB * kb, acc);
    blk_sub(wa, PB * r, PB * c, acc);
  };
  if (nreal <= PB) {
    if (w == 0) {
      if (pa) {
        sub_pp(0, 0);
        wave_sync();
      }
      diag16_wave(wa, wx, 0, ws, bad);
    }
    __syncthreads();
    if (bad && tid == 0) *fail = 1;
#pragma unroll 4
    for (int q = 0; q < 16; ++q) {
      const int e = tid + 256 * q, i = e >> 6, j = e & 63;
      dst[(int64_t)i * ld + j] = (i < PB) ? ((j <= i) ? wx[i][j] : 0.0) : (i == j ? 1.0 : 0.0);
    }
    return;
  }
  auto lpanel = [&](int r, int c) {  // L_rc = T_rc M_c^T, in place (reads precede the writes)
    blk_put(wa, PB * r, PB * c, blk_nt(wa, PB * r, PB * c, wx, PB * c, PB * c, zero), 1.0);
  };
  auto sub_ll = [&](int r, int c, int k) {  // T_rc -= L_rk L_ck^T
    blk_sub(wa, PB * r, PB * c, blk_nt(wa, PB * r, PB * k, wa, PB * c, PB * k, zero));
  };
  auto xblk = [&](int r, int c) {  // X_rc = -M_r S_rc (S_rc in upper block (c, r) of wx)
    blk_put(wx, PB * r, PB * c, blk_nn(wx, PB * r, PB * r, wx, PB * c, PB * r, zero), -1.0);
  };
  // S1
  if (w == 0) {
    if (pa) {
      sub_pp(0, 0);
      wave_sync();
    }
    diag16_wave(wa, wx, 0, ws, bad);
  } else if (pa) {
    if (w == 1) { sub_pp(1, 0); sub_pp(1, 1); sub_pp(2, 0); }
    else if (w == 2) { sub_pp(2, 1); sub_pp(2, 2); sub_pp(3, 0); }
    else { sub_pp(3, 1); sub_pp(3, 2); sub_pp(3, 3); }
  }
  __syncthreads();
  // Sa1
  if (w >= 1) lpanel(w, 0);
  __syncthreads();
  // Sb1
  if (w == 0) {
    sub_ll(1, 1, 0);
    wave_sync();
    diag16_wave(wa, wx, PB, ws, bad);
  } else if (w == 1) { sub_ll(2, 1, 0); sub_ll(3, 1, 0); }
  else if (w == 2) { sub_ll(2, 2, 0); sub_ll(3, 2, 0); }
  else {
    sub_ll(3, 3, 0);
    blk_put(wx, 0, PB, blk_nn(wa, PB, 0, wx, 0, 0, zero), 1.0);  // S_10
  }
  __syncthreads();
  // Sa2
  if (w == 1) lpanel(2, 1);
  else if (w == 2) lpanel(3, 1);
  else if (w == 3) xblk(1, 0);
  __syncthreads();
  // Sb2
  if (w == 0) {
    sub_ll(2, 2, 1);
    wave_sync();
    diag16_wave(wa, wx, 2 * PB, ws, bad);
  } else if (w == 1) {
    sub_ll(3, 2, 1);
    blk_put(wx, PB, 2 * PB, blk_nn(wa, 2 * PB, PB, wx, PB, PB, zero), 1.0);  // S_21
  } else if (w == 2) {
    sub_ll(3, 3, 1);
  } else {
    dbl4 t = blk_nn(wa, 2 * PB, 0, wx, 0, 0, zero);
    t = blk_nn(wa, 2 * PB, PB, wx, PB, 0, t);
    blk_put(wx, 0, 2 * PB, t, 1.0);  // S_20
  }
  __syncthreads();
  // Sa3
  if (w == 1) lpanel(3, 2);
  else if (w == 2) xblk(2, 1);
  else if (w == 3) xblk(2, 0);
  __syncthreads();
  // Sb3
  if (w == 0) {
    sub_ll(3, 3, 2);
    wave_sync();
    diag16_wave(wa, wx, 3 * PB, ws, bad);
  } else {
    const int c = 3 - w;  // w1: S_32, w2: S_31, w3: S_30
    dbl4 t = zero;
    for (int k = c; k < 3; ++k) t = blk_nn(wa, 3 * PB, PB * k, wx, PB * k, PB * c, t);
    blk_put(wx, PB * c, 3 * PB, t, 1.0);
  }
  __syncthreads();
  if (w >= 1) xblk(3, 3 - w);
  __syncthreads();
  if (bad && tid == 0) *fail = 1;
#pragma unroll 4
  for (int q = 0; q < 16; ++q) {
    const int e = tid + 256 * q, i = e >> 6, j = e & 63;
    dst[(int64_t)i * ld + j] = (j <= i) ? wx[i][j] : 0.0;
  }
}

__device__ __forceinline__ void tile_factor_inverse(const double *src, int64_t lds, double *dst, int64_t ld,
                                                    int *fail, double (*wa)[LDT], double (*wx)[LDT],
                                                    WaveScratch16 &ws, int nreal) {
  load_tile(wa, src, lds, 0, 0);
  __syncthreads();
  tile_factor_inverse_lds(dst, ld, fail, wa, nullptr, wx, ws, nreal);
}

// The same, not inlined: inside k_chol_small's loop over tiles the inlined
// factor's LDS addresses are hoisted out of the loop and the kernel needs
// more than 256 registers per lane (one wave per SIMD: one workgroup per CU
// instead of two); called, it needs 243.
__device__ __noinline__ void tile_factor_inverse_call(const double *src, int64_t lds, double *dst, int64_t ld,
                                                      int *fail, double (*wa)[LDT], double (*wx)[LDT],
                                                      WaveScratch16 &ws, int nreal) {
  tile_factor_inverse(src, lds, dst, ld, fail, wa, wx, ws, nreal);
}

// 64 x 64 x 64 products on the matrix cores (v_mfma_f64_16x16x4_f64).  Wave
// w owns the 32 x 32 quadrant (rows 32 (w / 2), columns 32 (w % 2)) as 2 x 2
// blocks of 16 x 16.  Operand lanes: A[i][k] from lane i + 16 k, B[k][j] from
// lane j + 16 k; the result D[i][j] lands in lane j + 16 (i % 4), element
// i / 4.  acc = a * b^T with a, b in LDS.

__device__ __forceinline__ void mfma_nt(const double (*a)[LDT], const double (*b)[LDT], dbl4 acc[2][2]) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int r0 = 32 * (w >> 1), c0 = 32 * (w & 1);
  const int li = l & 15, lk = l >> 4;
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v) acc[u][v] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int k = 0; k < NB; k += 4) {
    const double a0 = a[r0 + li][k + lk], a1 = a[r0 + 16 + li][k + lk];
    const double b0 = b[c0 + li][k + lk], b1 = b[c0 + 16 + li][k + lk];
    acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
  }
}

// out (LDS) = a * b^T for LDS tiles (256 threads); out may alias a or b.
__device__ __forceinline__ void tile_mul_nt_lds(double (*out)[LDT], const double (*a)[LDT], const double (*b)[LDT]) {
  dbl4 acc[2][2];
  mfma_nt(a, b, acc);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int r0 = 32 * (w >> 1) + (l >> 4), c0 = 32 * (w & 1) + (l & 15);
  __syncthreads();  // a or b may alias out
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v)
#pragma unroll
      for (int e = 0; e < 4; ++e) out[r0 + 16 * u + 4 * e][c0 + 16 * v] = acc[u][v][e];
  __syncthreads();
}

// C (global 64 x 64 at r0, c0) -= a b^T; the tile of C is fetched before the
// products so its loads overlap them.
__device__ __forceinline__ void tile_sub_nt_global(double *C, int64_t ld, int r0, int c0, const double (*a)[LDT],
                                                   const double (*b)[LDT]) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int rr = r0 + 32 * (w >> 1) + (l >> 4), cc = c0 + 32 * (w & 1) + (l & 15);
  double cv[2][2][4];
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v)
#pragma unroll
      for (int e = 0; e < 4; ++e) cv[u][v][e] = C[(int64_t)(rr + 16 * u + 4 * e) * ld + cc + 16 * v];
  dbl4 acc[2][2];
  mfma_nt(a, b, acc);
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v)
#pragma unroll
      for (int e = 0; e < 4; ++e) C[(int64_t)(rr + 16 * u + 4 * e) * ld + cc + 16 * v] = cv[u][v][e] - acc[u][v][e];
}

// out = a b^T for LDS tiles with b lower triangular (b[j][k] = 0 for k > j,
// as every stored L_kk^-1 is); out may alias a.  Wave w forms row blocks
// 2 (w / 2), 2 (w / 2) + 1 of column blocks {0, 3} (w even) or {1, 2} (w odd)
// and skips the zero part of b: 40 block products per wave instead of 64,
// with the same result (only exact zeros are left out).
__device__ __forceinline__ void tile_mul_nt_tril_lds(double (*out)[LDT], const double (*a)[LDT],
                                                     const double (*b)[LDT]) {
  const int w = threadIdx.x >> 6;
  const int r0 = 32 * (w >> 1);
  const int cb0 = (w & 1) ? 1 : 0, cb1 = (w & 1) ? 2 : 3;
  dbl4 acc[2][2];
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v) acc[u][v] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    if (kb <= cb0) {
      acc[0][0] = blk_nt(a, r0, PB * kb, b, PB * cb0, PB * kb, acc[0][0]);
      acc[1][0] = blk_nt(a, r0 + PB, PB * kb, b, PB * cb0, PB * kb, acc[1][0]);
    }
    if (kb <= cb1) {
      acc[0][1] = blk_nt(a, r0, PB * kb, b, PB * cb1, PB * kb, acc[0][1]);
      acc[1][1] = blk_nt(a, r0 + PB, PB * kb, b, PB * cb1, PB * kb, acc[1][1]);
    }
  }
  __syncthreads();  // out may alias a
  blk_put(out, r0, PB * cb0, acc[0][0], 1.0);
  blk_put(out, r0 + PB, PB * cb0, acc[1][0], 1.0);
  blk_put(out, r0, PB * cb1, acc[0][1], 1.0);
  blk_put(out, r0 + PB, PB * cb1, acc[1][1], 1.0);
  __syncthreads();
}

// y = M v (transposed = false) or M^T v for a 64 x 64 global tile M with row
// pitch ld and a 64-vector v in LDS; result to LDS out (256 threads).
__device__ __forceinline__ void tile_gemv(const double *M, int64_t ld, const double *v, double *out, bool transposed,
                                          double (*scratch)[NB]) {
  const int r = threadIdx.x & 63, sl = threadIdx.x >> 6;  // 4 slices of 16
  double mv[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int k = sl * 16 + q;
    mv[q] = transposed ? M[(int64_t)k * ld + r] : M[(int64_t)r * ld + k];
  }
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < 16; ++q) acc[q & 3] += mv[q] * v[sl * 16 + q];
  scratch[sl][r] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  __syncthreads();
  if (threadIdx.x < NB) out[threadIdx.x] = scratch[0][threadIdx.x] + scratch[1][threadIdx.x] +
                                           scratch[2][threadIdx.x] + scratch[3][threadIdx.x];
  __syncthreads();
}

__device__ __forceinline__ bool chol_active(const WinState &s) { return s.active && s.trial && !s.chol_fail; }

// ---------------------------------------------------------------- large path
// Every large-path launch covers all large windows of the batch: the window
// of a workgroup is large window blockIdx.y (its CholWin record), and each
// window runs exactly the work (and order of work) it runs alone, so its
// results do not depend on the batch.  Windows with fewer tiles finish their steps earlier; their
// workgroups of later launches return at once.

// Diagonal tile 0 (before the first step).
// Workgroup 0: diagonal tile 0.  Workgroups 1..: b = rhs with zero padding
// (the step launches then update b in place).
__global__ void __launch_bounds__(256) k_chol_first(const CholWin *cws, WinState *ws, const double *S, double *Dinv,
                                                    const double *rhs, double *b) {
  __shared__ double wa[NB][LDT];
  __shared__ double wx[NB][LDT];
  __shared__ WaveScratch16 fs;
  __shared__ int fail;
  const CholWin cw = cws[blockIdx.y];
  WinState &s = ws[cw.win];
  if (blockIdx.x > 0) {
    if (!chol_active(s)) return;
    for (int i = (blockIdx.x - 1) * blockDim.x + threadIdx.x; i < cw.ld; i += (gridDim.x - 1) * blockDim.x)
      b[cw.r_off + i] = i < cw.n ? rhs[cw.r_off + i] : 0.0;
    return;
  }
  double v[16];
  fetch_tile(v, S + cw.s_off, cw.ld, 0, 0);  // in flight while the window's state is read
  if (!chol_active(s)) return;
  if (threadIdx.x == 0) fail = 0;
  put_tile(wa, v);
  __syncthreads();
  tile_factor_inverse_lds(Dinv + cw.dinv_off, NB, &fail, wa, nullptr, wx, fs, cw.n);
  __syncthreads();
  if (threadIdx.x == 0 && fail) s.chol_fail = 1;
}

// Step k of large window blockIdx.y: trailing tiles (i, j),
// k < j <= i < nb, in row order, then the right-hand-side workgroups
// j = k+1 .. nb-1.
__global__ void __launch_bounds__(256) k_chol_step(const CholWin *cws, WinState *ws, double *S, double *Lf,
                                                   double *Dinv, double *b, double *y, int k) {
  __shared__ double li[NB][LDT];
  __shared__ double pa[NB][LDT];
  __shared__ double pb[NB][LDT];
  __shared__ double vsh[3][NB];
  __shared__ double red[4][NB];
  __shared__ WaveScratch16 fs;
  __shared__ int fail;
  const CholWin wi = cws[blockIdx.y];
  if (k + 1 >= wi.nb) return;  // this window's factorization is complete
  const int rest = wi.nb - k - 1, ntiles = rest * (rest + 1) / 2;
  if ((int)blockIdx.x >= ntiles + rest) return;
  WinState &s = ws[wi.win];
  const int64_t ld = wi.ld;
  double *A = S + wi.s_off;
  double *L = Lf + wi.s_off;
  const double *Li = Dinv + wi.dinv_off + (int64_t)k * NB * NB;
  if (blockIdx.x >= ntiles) {
    if (!chol_active(s)) return;
    // right-hand side: y_k = Linv_k b_k, w = Linv_k^T y_k, b_j -= A_jk w
    const int jb = k + 1 + (blockIdx.x - ntiles);
    for (int e = threadIdx.x; e < NB; e += blockDim.x) vsh[0][e] = b[wi.r_off + k * NB + e];
    __syncthreads();
    tile_gemv(Li, NB, vsh[0], vsh[1], false, red);       // y_k
    tile_gemv(Li, NB, vsh[1], vsh[2], true, red);        // w
    if (jb == k + 1 && threadIdx.x < NB) y[wi.r_off + k * NB + threadIdx.x] = vsh[1][threadIdx.x];
    tile_gemv(A + (int64_t)jb * NB * ld + k * NB, ld, vsh[2], vsh[0], false, red);
    if (threadIdx.x < NB) b[wi.r_off + jb * NB + threadIdx.x] -= vsh[0][threadIdx.x];
    return;
  }
  // tile index -> (i, j) relative to k + 1
  const int t = blockIdx.x;
  int r = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while (r * (r + 1) / 2 > t) --r;
  while ((r + 1) * (r + 2) / 2 <= t) ++r;
  const int ib = k + 1 + r, jb = k + 1 + (t - r * (r + 1) / 2);
  // look-ahead: tile (k+1, k+1) is final after this update, so it is formed
  // and factored in LDS by its workgroup
  const bool look = ib == k + 1 && jb == k + 1;
  {
    // all three tiles' loads in flight together, and with the read of the
    // window's state (the look-ahead workgroup's third tile is the diagonal
    // tile it updates)
    double v0[16], v1[16], v2[16];
    fetch_tile(v0, Li, NB, 0, 0);
    fetch_tile(v1, A, ld, ib * NB, k * NB);
    if (jb != ib) fetch_tile(v2, A, ld, jb * NB, k * NB);
    else if (look) fetch_tile(v2, A, ld, ib * NB, jb * NB);
    if (!chol_active(s)) return;
    put_tile(li, v0);
    put_tile(pa, v1);
    if (jb != ib || look) put_tile(pb, v2);
  }
  __syncthreads();
  tile_mul_nt_tril_lds(pa, pa, li);                     // L_ik = A_ik Linv^T
  if (jb != ib) tile_mul_nt_tril_lds(pb, pb, li);       // L_jk
  if (jb == k + 1) store_tile(pa, L, ld, ib * NB, k * NB);
  if (look) {
    // T = A_{k+1,k+1} - L_{k+1,k} L_{k+1,k}^T is fused into the factorization;
    // li (read only by the product above) becomes its scratch
    if (threadIdx.x == 0) fail = 0;
    __syncthreads();
    tile_factor_inverse_lds(Dinv + wi.dinv_off + (int64_t)ib * NB * NB, NB, &fail, pb, pa, li, fs, wi.n - ib * NB);
    __syncthreads();
    if (threadIdx.x == 0 && fail) s.chol_fail = 1;
  } else {
    tile_sub_nt_global(A, ld, ib * NB, jb * NB, pa, jb != ib ? pb : pa);
  }
}

// Products of a 64 x 64 global tile with a 64-vector by 1024 threads: thread
// t holds the four entries M'[r][4 s + q] (q = 0..3) of M' = M^T
// (transposed) or M, r = t % 64, s = t / 64 (16 slices of four columns).
constexpr int BT = 1024;
__device__ __forceinline__ void fetch4(double m[4], const double *M, int64_t ld, bool transposed) {
  const int r = threadIdx.x & 63, sl = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    m[q] = transposed ? M[(int64_t)(4 * sl + q) * ld + r] : M[(int64_t)r * ld + 4 * sl + q];
}
// this thread's share of sum_k M'[r][k] v[k] (v in LDS)
__device__ __forceinline__ double part4(const double m[4], const double *v) {
  const int sl = threadIdx.x >> 6;
  return (m[0] * v[4 * sl] + m[1] * v[4 * sl + 1]) + (m[2] * v[4 * sl + 2] + m[3] * v[4 * sl + 3]);
}
// sum of the 16 slices' shares; the result is valid in threads 0..63 (row r = t)
__device__ __forceinline__ double reduce16(double part, double (*red)[NB]) {
  red[threadIdx.x >> 6][threadIdx.x & 63] = part;
  __syncthreads();
  double v = 0.0;
  if (threadIdx.x < NB) {
    double q[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int sl = 0; sl < 16; ++sl) q[sl & 3] += red[sl][threadIdx.x];
    v = (q[0] + q[1]) + (q[2] + q[3]);
  }
  return v;
}

// Backward solve of the block rows k_hi, k_hi - 1, ..., k_lo (at most four)
// in one launch of 1024-thread workgroups.  Every workgroup forms that short
// chain itself,
//   v = y_k - sum_{m = k+1..k_hi} L_mk^T x_m   (y_k = Linv_k b_k for the last block),
//   x_k = Linv_k^T v,
// with every tile of the chain fetched into registers before the first
// product (the chain is then bound by its reductions, not by memory latency);
// workgroup ib < k_lo then applies y_ib -= sum_{m = k_lo..k_hi} L_m,ib^T x_m,
// its tiles of that update fetched while the chain runs (each into the
// registers of chain tiles already used, so the peak register count stays
// that of the chain).
__global__ void __launch_bounds__(BT) k_chol_back(const CholWin *cws, WinState *ws, const double *Lf,
                                                  const double *Dinv, const double *b, double *y, double *x, int q) {
  __shared__ double xs[4][NB];
  __shared__ double vs[NB];
  __shared__ double red[16][NB];
  const CholWin wi = cws[blockIdx.y];
  // launch q solves block rows nb - 1 - 4 q down to (at most) nb - 4 - 4 q
  const int k_hi = wi.nb - 1 - 4 * q;
  if (k_hi < 0) return;
  const int k_lo = k_hi >= 3 ? k_hi - 3 : 0;
  if ((int)blockIdx.x >= (k_lo > 0 ? k_lo : 1)) return;
  const WinState &s = ws[wi.win];
  const int64_t ld = wi.ld;
  const double *L = Lf + wi.s_off;
  const double *Dw = Dinv + wi.dinv_off;
  const int t = threadIdx.x;
  const int nk = k_hi - k_lo + 1;  // 1..4; step a is block row k_hi - a
  const bool last = k_hi == wi.nb - 1;
  double di[4][4], lt[6][4], dl[4], yv[4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
    if (a < nk) fetch4(di[a], Dw + (int64_t)(k_hi - a) * NB * NB, NB, true);
  // L_{k_hi - c, k_hi - a} for c < a, at lt[a (a - 1) / 2 + c]
#pragma unroll
  for (int a = 1; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < a; ++c)
      if (a < nk) fetch4(lt[a * (a - 1) / 2 + c], L + (int64_t)(k_hi - c) * NB * ld + (k_hi - a) * NB, ld, true);
  if (last) fetch4(dl, Dw + (int64_t)k_hi * NB * NB, NB, false);
#pragma unroll
  for (int a = 0; a < 4; ++a)
    yv[a] = (t < NB && a < nk) ? ((last && a == 0) ? b[wi.r_off + k_hi * NB + t] : y[wi.r_off + (k_hi - a) * NB + t])
                               : 0.0;
  if (!chol_active(s)) return;  // read while the tiles above are in flight
  const int ib = blockIdx.x;      // < k_lo when k_lo > 0
  double mu[4][4];                // L_{k_hi - a, ib}^T for the update of y_ib
  auto fetch_upd = [&](int a) { fetch4(mu[a], L + (int64_t)(k_hi - a) * NB * ld + ib * NB, ld, true); };
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    if (a < nk) {
      double v;
      if (a == 0 && last) {
        if (t < NB) vs[t] = yv[0];
        __syncthreads();
        v = reduce16(part4(dl, vs), red);  // Linv_k b_k
      } else if (a == 0) {
        v = yv[0];
      } else {
        double part = 0.0;
#pragma unroll
        for (int c = 0; c < a; ++c) part += part4(lt[a * (a - 1) / 2 + c], xs[c]);
        v = yv[a] - reduce16(part, red);
      }
      __syncthreads();
      if (t < NB) vs[t] = v;
      __syncthreads();
      const double xk = reduce16(part4(di[a], vs), red);
      if (t < NB) xs[a][t] = xk;
      // step a has used its chain tiles: the update tiles take their place
      if (k_lo > 0) {
        fetch_upd(a);
        if (a + 2 == nk) fetch_upd(a + 1);
      }
      __syncthreads();
    }
  }
  if (blockIdx.x == 0 && t < NB)
    for (int a = 0; a < nk; ++a) x[wi.r_off + (k_hi - a) * NB + t] = xs[a][t];
  if (k_lo == 0) return;
  double part = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a)
    if (a < nk) part += part4(mu[a], xs[a]);
  const double u = reduce16(part, red);
  if (t < NB) y[wi.r_off + ib * NB + t] -= u;
}

// ---------------------------------------------------------------- small path
// Whole blocked factorization and both triangular solves of one window per
// workgroup (nb <= 4).  b is the right-hand side, x receives the solution.
__global__ void __launch_bounds__(256) k_chol_small(const WinInfo *wis, WinState *ws, const int *win_list,
                                                    double *S, double *Dinv, const double *rhs, double *dred) {
  __shared__ double a[NB][LDT];
  __shared__ double c[NB][LDT];
  __shared__ double vsh[3][NB];
  __shared__ double red[4][NB];
  __shared__ WaveScratch16 fs;
  __shared__ int fail;
  const int w = win_list[blockIdx.x];
  const WinInfo wi = wis[w];
  WinState &s = ws[w];
  if (!chol_active(s)) return;
  if (threadIdx.x == 0) fail = 0;
  double *A = S + wi.s_off;
  const int64_t ld = wi.ld;
  double *Dw = Dinv + wi.dinv_off;
  for (int kb = 0; kb < wi.nb; ++kb) {
    __syncthreads();
    __threadfence_block();
    tile_factor_inverse_call(A + (int64_t)kb * NB * ld + kb * NB, ld, Dw + (int64_t)kb * NB * NB, NB, &fail, a, c,
                             fs, wi.n - kb * NB);
    __threadfence_block();
    __syncthreads();
    // panel: L_ib = A_ib Linv^T.  c holds Linv in its lower blocks (the
    // triangular product reads no other block of c); a tile with at most 16
    // real rows is the window's last, so no panel follows its short factor.
    for (int ib = kb + 1; ib < wi.nb; ++ib) {
      load_tile(a, A, ld, ib * NB, kb * NB);
      __syncthreads();
      tile_mul_nt_tril_lds(a, a, c);
      store_tile(a, A, ld, ib * NB, kb * NB);
    }
    __syncthreads();
    for (int ib = kb + 1; ib < wi.nb; ++ib)
      for (int jb = kb + 1; jb <= ib; ++jb) {
        load_tile(a, A, ld, ib * NB, kb * NB);
        load_tile(c, A, ld, jb * NB, kb * NB);
        __syncthreads();
        tile_sub_nt_global(A, ld, ib * NB, jb * NB, a, c);
        __syncthreads();
      }
  }
  __syncthreads();
  if (threadIdx.x == 0 && fail) s.chol_fail = 1;
  if (fail) return;
  // forward: x = rhs; for kb: x_kb = Linv_kb x_kb; x_i -= L_i,kb x_kb (i > kb)
  double *x = dred + wi.r_off;
  for (int i = threadIdx.x; i < ld; i += blockDim.x) x[i] = i < wi.n ? rhs[wi.r_off + i] : 0.0;
  __syncthreads();
  for (int kb = 0; kb < wi.nb; ++kb) {
    for (int e = threadIdx.x; e < NB; e += blockDim.x) vsh[0][e] = x[kb * NB + e];
    __syncthreads();
    tile_gemv(Dw + (int64_t)kb * NB * NB, NB, vsh[0], vsh[1], false, red);
    if (threadIdx.x < NB) x[kb * NB + threadIdx.x] = vsh[1][threadIdx.x];
    for (int ib = kb + 1; ib < wi.nb; ++ib) {
      tile_gemv(A + (int64_t)ib * NB * ld + kb * NB, ld, vsh[1], vsh[2], false, red);
      if (threadIdx.x < NB) x[ib * NB + threadIdx.x] -= vsh[2][threadIdx.x];
      __syncthreads();
    }
    __syncthreads();
  }
  for (int kb = wi.nb - 1; kb >= 0; --kb) {
    for (int e = threadIdx.x; e < NB; e += blockDim.x) vsh[0][e] = x[kb * NB + e];
    __syncthreads();
    tile_gemv(Dw + (int64_t)kb * NB * NB, NB, vsh[0], vsh[1], true, red);
    if (threadIdx.x < NB) x[kb * NB + threadIdx.x] = vsh[1][threadIdx.x];
    for (int ib = 0; ib < kb; ++ib) {
      tile_gemv(A + (int64_t)kb * NB * ld + ib * NB, ld, vsh[1], vsh[2], true, red);
      if (threadIdx.x < NB) x[ib * NB + threadIdx.x] -= vsh[2][threadIdx.x];
      __syncthreads();
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- inverse
// S^-1 = L^-T L^-1 from the factor, for the posterior covariance.
// Y = L^-T is upper block-triangular: Y_ii = Linv_i^T and, for j < i,
//   Y_ji = -(sum_{k=j..i-1} Y_jk L_ik^T) Linv_i^T,
// one launch per block column i (its workgroups j = 0..i are independent);
// then S^-1_ab = sum_{i >= max(a, b)} Y_ai Y_bi^T, one workgroup per tile.

// acc += a * b^T (LDS tiles, 256 threads, MFMA result layout of mfma_nt)
__device__ __forceinline__ void mfma_nt_acc(const double (*a)[LDT], const double (*b)[LDT], dbl4 acc[2][2]) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int r0 = 32 * (w >> 1), c0 = 32 * (w & 1);
  const int li = l & 15, lk = l >> 4;
#pragma unroll 4
  for (int k = 0; k < NB; k += 4) {
    const double a0 = a[r0 + li][k + lk], a1 = a[r0 + 16 + li][k + lk];
    const double b0 = b[c0 + li][k + lk], b1 = b[c0 + 16 + li][k + lk];
    acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
  }
}

// element (r, c) of the 64 x 64 result held by this thread in the mfma_nt layout
#define HSBA_ACC_FOR(BODY)                                                            \
  {                                                                                   \
    const int w_ = threadIdx.x >> 6, l_ = threadIdx.x & 63;                           \
    const int rb_ = 32 * (w_ >> 1) + (l_ >> 4), cb_ = 32 * (w_ & 1) + (l_ & 15);      \
    _Pragma("unroll") for (int u = 0; u < 2; ++u)                                     \
    _Pragma("unroll") for (int v = 0; v < 2; ++v)                                     \
    _Pragma("unroll") for (int e = 0; e < 4; ++e) {                                   \
      const int r = rb_ + 16 * u + 4 * e, c = cb_ + 16 * v;                           \
      BODY;                                                                           \
    }                                                                                 \
  }

__global__ void __launch_bounds__(256) k_chol_inv_col(WinInfo wi, const double *Lsrc, const double *Dinv, double *Y,
                                                      int i) {
  __shared__ double ta[NB][LDT];
  __shared__ double tb[NB][LDT];
  const int j = blockIdx.x;  // 0..i
  const int64_t ld = wi.ld;
  const double *L = Lsrc + wi.s_off;
  double *Yw = Y + wi.s_off;
  const double *Di = Dinv + wi.dinv_off + (int64_t)i * NB * NB;
  if (j == i) {
    load_tile(ta, Di, NB, 0, 0);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int e = threadIdx.x + 256 * q;
      Yw[(int64_t)(i * NB + (e >> 6)) * ld + i * NB + (e & 63)] = ta[e & 63][e >> 6];
    }
    return;
  }
  dbl4 acc[2][2];
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v) acc[u][v] = dbl4{0.0, 0.0, 0.0, 0.0};
  for (int k = j; k < i; ++k) {
    {
      double v0[16], v1[16];
      fetch_tile(v0, Yw, ld, j * NB, k * NB);
      fetch_tile(v1, L, ld, i * NB, k * NB);
      put_tile(ta, v0);
      put_tile(tb, v1);
    }
    __syncthreads();
    mfma_nt_acc(ta, tb, acc);  // += Y_jk L_ik^T
    __syncthreads();
  }
  HSBA_ACC_FOR(ta[r][c] = acc[u][v][e]);
  load_tile(tb, Di, NB, 0, 0);
  __syncthreads();
  mfma_nt(ta, tb, acc);  // T Linv_i^T
  HSBA_ACC_FOR(Yw[(int64_t)(j * NB + r) * ld + i * NB + c] = -acc[u][v][e]);
}

__global__ void __launch_bounds__(256) k_chol_inv_gram(WinInfo wi, const double *Y, double *Sinv) {
  __shared__ double ta[NB][LDT];
  __shared__ double tb[NB][LDT];
  const int t = blockIdx.x;
  int a = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while (a * (a + 1) / 2 > t) --a;
  while ((a + 1) * (a + 2) / 2 <= t) ++a;
  const int b = t - a * (a + 1) / 2;  // b <= a
  const int64_t ld = wi.ld;
  const double *Yw = Y + wi.s_off;
  double *Sw = Sinv + wi.s_off;
  dbl4 acc[2][2];
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v) acc[u][v] = dbl4{0.0, 0.0, 0.0, 0.0};
  for (int i = a; i < wi.nb; ++i) {
    {
      double v0[16], v1[16];
      fetch_tile(v0, Yw, ld, a * NB, i * NB);
      fetch_tile(v1, Yw, ld, b * NB, i * NB);
      put_tile(ta, v0);
      put_tile(tb, v1);
    }
    __syncthreads();
    mfma_nt_acc(ta, tb, acc);  // += Y_ai Y_bi^T
    __syncthreads();
  }
  // both triangles from the same element (the diagonal tile from its lower half)
  HSBA_ACC_FOR(if (a != b || r >= c) {
    Sw[(int64_t)(a * NB + r) * ld + b * NB + c] = acc[u][v][e];
    Sw[(int64_t)(b * NB + c) * ld + a * NB + r] = acc[u][v][e];
  });
}
#undef HSBA_ACC_FOR

void chol_inverse_impl(const CholPlan &plan, int w, bool large, double *Y, double *Sinv, hipStream_t st) {
  const WinInfo wi = plan.wi_host[w];
  const double *Lsrc = large ? plan.Lf : plan.S;
  for (int i = 0; i < wi.nb; ++i)
    hipLaunchKernelGGL(k_chol_inv_col, dim3(i + 1), dim3(256), 0, st, wi, Lsrc, plan.Dinv, Y, i);
  hipLaunchKernelGGL(k_chol_inv_gram, dim3(wi.nb * (wi.nb + 1) / 2), dim3(256), 0, st, wi, Y, Sinv);
}

// Identity on the padding rows / columns (once, after allocation).
__global__ void k_pad_identity(const WinInfo *wis, int nwin, double *S) {
  const int w = blockIdx.y;
  if (w >= nwin) return;
  const WinInfo wi = wis[w];
  double *A = S + wi.s_off;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < (int64_t)wi.ld * wi.ld;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / wi.ld), j = (int)(e - (int64_t)i * wi.ld);
    if (i >= wi.n || j >= wi.n) A[e] = (i == j) ? 1.0 : 0.0;
  }
}

void chol_init_padding(const CholPlan &plan, hipStream_t st) {
  if (plan.nwin == 0) return;
  hipLaunchKernelGGL(k_pad_identity, dim3(64, plan.nwin), dim3(256), 0, st, plan.wi, plan.nwin, plan.S);
}

void chol_factor_solve_impl(const CholPlan &plan, const int *small_list_dev, int num_small,
                            const int *large_host, int num_large, hipStream_t st) {
  if (num_small > 0)
    hipLaunchKernelGGL(k_chol_small, dim3(num_small), dim3(256), 0, st, plan.wi, plan.ws, small_list_dev,
                       plan.S, plan.Dinv, plan.rhs, plan.dred);
  if (num_large > 0) {
    // every launch covers all large windows (blockIdx.y); the grid is sized
    // for the window with the most work in that launch
    int max_nb = 0;
    for (int q = 0; q < num_large; ++q) max_nb = std::max(max_nb, plan.wi_host[large_host[q]].nb);
    hipLaunchKernelGGL(k_chol_first, dim3(1 + 8, num_large), dim3(256), 0, st, plan.large_wins, plan.ws, plan.S,
                       plan.Dinv, plan.rhs, plan.bwork);
    for (int k = 0; k + 1 < max_nb; ++k) {
      int slots = 0;
      for (int q = 0; q < num_large; ++q) {
        const int rest = plan.wi_host[large_host[q]].nb - k - 1;
        if (rest > 0) slots = std::max(slots, rest * (rest + 1) / 2 + rest);
      }
      hipLaunchKernelGGL(k_chol_step, dim3(slots, num_large), dim3(256), 0, st, plan.large_wins, plan.ws, plan.S,
                         plan.Lf, plan.Dinv, plan.bwork, plan.ywork, k);
    }
    for (int q = 0; 4 * q < max_nb; ++q) {
      int blocks = 1;
      for (int u = 0; u < num_large; ++u) {
        const int k_hi = plan.wi_host[large_host[u]].nb - 1 - 4 * q;
        if (k_hi >= 4) blocks = std::max(blocks, k_hi - 3);
      }
      hipLaunchKernelGGL(k_chol_back, dim3(blocks, num_large), dim3(BT), 0, st, plan.large_wins, plan.ws, plan.Lf,
                         plan.Dinv, plan.bwork, plan.ywork, plan.dred, q);
    }
  }
}

}  // namespace hsba
