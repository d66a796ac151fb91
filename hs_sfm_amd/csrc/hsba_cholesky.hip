// hsba_cholesky.hip — dense FP64 Cholesky and triangular solves of the
// reduced camera systems, replacing Eigen::SimplicialLDLT
// (include/hs_sfm/bundle_adjustment/camera_shared_normal_equation_solver.hpp:448-466).
//
// Storage: each window's reduced matrix is row-major with leading dimension
// ld (a multiple of 64), lower triangle valid, padding rows/columns set to the
// identity once.  64 x 64 tiles are staged through LDS.
//   * large windows: one launch per stage (diag factor, panel solve, trailing
//     update) so every CU works on the trailing update;
//   * small windows: one workgroup per window runs the whole blocked
//     factorization and both triangular solves (a batch of windows is one
//     launch).
// A non-positive pivot marks the window (WinState::chol_fail); the reference
// reports the same condition as SimplicialLDLT's info() != Success.
#include "hsba_kernels.h"

namespace hsba {

constexpr int NB = 64;
constexpr int LDT = NB + 1;  // LDS tile pitch (bank-conflict padding)

__device__ __forceinline__ void load_tile(double (*t)[LDT], const double *A, int64_t ld, int r0, int c0) {
  for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) {
    const int i = e / NB, j = e - i * NB;
    t[i][j] = A[(int64_t)(r0 + i) * ld + c0 + j];
  }
}
__device__ __forceinline__ void store_tile(const double (*t)[LDT], double *A, int64_t ld, int r0, int c0) {
  for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) {
    const int i = e / NB, j = e - i * NB;
    A[(int64_t)(r0 + i) * ld + c0 + j] = t[i][j];
  }
}

// Unblocked right-looking Cholesky of the lower triangle of an LDS tile.
__device__ void tile_potrf(double (*t)[LDT], int *fail) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int j = 0; j < NB; ++j) {
    if (tid == 0) {
      const double djj = t[j][j];
      if (!(djj > 0.0)) { *fail = 1; t[j][j] = 1.0; }
      else t[j][j] = sqrt(djj);
    }
    __syncthreads();
    const double ljj = t[j][j];
    for (int i = j + 1 + tid; i < NB; i += nt) t[i][j] /= ljj;
    __syncthreads();
    // trailing lower triangle: thread owns column c, rows c.. stepping by nt/64
    const int c = j + 1 + (tid & 63);
    if (c < NB) {
      const double lcj = t[c][j];
      for (int i = c + (tid >> 6); i < NB; i += (nt >> 6)) t[i][c] -= t[i][j] * lcj;
    }
    __syncthreads();
  }
}

// Inverse of the lower-triangular LDS tile l into li (thread j solves column j).
__device__ void tile_trtri(const double (*l)[LDT], double (*li)[LDT]) {
  const int j = threadIdx.x;
  if (j < NB) {
    for (int i = 0; i < j; ++i) li[i][j] = 0.0;
    li[j][j] = 1.0 / l[j][j];
    for (int i = j + 1; i < NB; ++i) {
      double s = 0.0;
      for (int c = j; c < i; ++c) s += l[i][c] * li[c][j];
      li[i][j] = -s / l[i][i];
    }
  }
  __syncthreads();
}

// C (global, 64 x 64 at r0, c0) = alpha * C + beta * (a b^T) with a, b LDS tiles.
// 256 threads, each 4 rows x 4 strided columns.
__device__ void tile_gemm_nt(double *C, int64_t ld, int r0, int c0, const double (*a)[LDT],
                             const double (*b)[LDT], bool accumulate_negative) {
  const int tid = threadIdx.x;
  const int ti = (tid >> 4) * 4, tc = tid & 15;
  double acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
  for (int k = 0; k < NB; ++k) {
    double av[4], bv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) av[u] = a[ti + u][k];
#pragma unroll
    for (int v = 0; v < 4; ++v) bv[v] = b[tc + 16 * v][k];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[u][v] += av[u] * bv[v];
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      double *dst = C + (int64_t)(r0 + ti + u) * ld + c0 + tc + 16 * v;
      if (accumulate_negative) *dst -= acc[u][v];
      else *dst = acc[u][v];
    }
}

// ---------------------------------------------------------------- large path
__global__ void __launch_bounds__(256) k_chol_diag(WinInfo wi, WinState *ws, int w, double *S, double *Dinv, int kb) {
  __shared__ double t[NB][LDT];
  __shared__ double li[NB][LDT];
  __shared__ int fail;
  WinState &s = ws[w];
  if (!(s.active && s.trial) || s.chol_fail) return;
  if (threadIdx.x == 0) fail = 0;
  double *A = S + wi.s_off;
  load_tile(t, A, wi.ld, kb * NB, kb * NB);
  __syncthreads();
  tile_potrf(t, &fail);
  tile_trtri(t, li);
  store_tile(t, A, wi.ld, kb * NB, kb * NB);
  store_tile(li, Dinv + wi.dinv_off + (int64_t)kb * NB * NB, NB, 0, 0);
  if (threadIdx.x == 0 && fail) s.chol_fail = 1;
}

__global__ void __launch_bounds__(256) k_chol_panel(WinInfo wi, WinState *ws, int w, double *S, const double *Dinv,
                                                    int kb) {
  __shared__ double a[NB][LDT];
  __shared__ double li[NB][LDT];
  const WinState &s = ws[w];
  if (!(s.active && s.trial) || s.chol_fail) return;
  const int ib = kb + 1 + blockIdx.x;
  double *A = S + wi.s_off;
  load_tile(a, A, wi.ld, ib * NB, kb * NB);
  load_tile(li, Dinv + wi.dinv_off + (int64_t)kb * NB * NB, NB, 0, 0);
  __syncthreads();
  __shared__ double out_unused;
  (void)out_unused;
  tile_gemm_nt(A, wi.ld, ib * NB, kb * NB, a, li, false);  // L_ib = A_ib Linv^T
}

__global__ void __launch_bounds__(256) k_chol_update(WinInfo wi, WinState *ws, int w, double *S, int kb) {
  __shared__ double a[NB][LDT];
  __shared__ double b[NB][LDT];
  const WinState &s = ws[w];
  if (!(s.active && s.trial) || s.chol_fail) return;
  // tile index -> (ib, jb), kb < jb <= ib < nb, lower triangle in row order
  const int t = blockIdx.x;
  int r = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while (r * (r + 1) / 2 > t) --r;
  while ((r + 1) * (r + 2) / 2 <= t) ++r;
  const int ib = kb + 1 + r, jb = kb + 1 + (t - r * (r + 1) / 2);
  double *A = S + wi.s_off;
  load_tile(a, A, wi.ld, ib * NB, kb * NB);
  load_tile(b, A, wi.ld, jb * NB, kb * NB);
  __syncthreads();
  tile_gemm_nt(A, wi.ld, ib * NB, jb * NB, a, b, true);
}

// Blocked forward (L y = b) and backward (L^T x = y) substitution of one window
// with the diagonal-tile inverses; x overwrites dred.  1024 threads.
__device__ void trsv_window(const WinInfo &wi, const double *S, const double *Dinv, const double *rhs,
                            double *x) {
  const double *A = S + wi.s_off;
  const int64_t ld = wi.ld;
  const int n = wi.n, nb = wi.nb;
  __shared__ double yb[NB];
  __shared__ double part[16][NB];
  // y = rhs (padding rows are zero)
  for (int i = threadIdx.x; i < ld; i += blockDim.x) x[i] = i < n ? rhs[i] : 0.0;
  __syncthreads();
  for (int kb = 0; kb < nb; ++kb) {
    const double *Li = Dinv + wi.dinv_off + (int64_t)kb * NB * NB;
    // z = Linv_kk y_kb : row r = tid & 63, partial over 16 slices of k
    {
      const int r = threadIdx.x & 63, sl = threadIdx.x >> 6;  // sl < 16
      double acc = 0.0;
      for (int k = sl * 4; k < sl * 4 + 4; ++k) acc += Li[r * NB + k] * x[kb * NB + k];
      part[sl][r] = acc;
    }
    __syncthreads();
    if (threadIdx.x < NB) {
      double z = 0.0;
      for (int sl = 0; sl < 16; ++sl) z += part[sl][threadIdx.x];
      yb[threadIdx.x] = z;
      x[kb * NB + threadIdx.x] = z;
    }
    __syncthreads();
    for (int i = (kb + 1) * NB + threadIdx.x; i < ld; i += blockDim.x) {
      const double *row = A + (int64_t)i * ld + kb * NB;
      double acc = 0.0;
      for (int k = 0; k < NB; ++k) acc += row[k] * yb[k];
      x[i] -= acc;
    }
    __syncthreads();
  }
  for (int kb = nb - 1; kb >= 0; --kb) {
    const double *Li = Dinv + wi.dinv_off + (int64_t)kb * NB * NB;
    // z = Linv_kk^T y_kb
    {
      const int r = threadIdx.x & 63, sl = threadIdx.x >> 6;
      double acc = 0.0;
      for (int k = sl * 4; k < sl * 4 + 4; ++k) acc += Li[k * NB + r] * x[kb * NB + k];
      part[sl][r] = acc;
    }
    __syncthreads();
    if (threadIdx.x < NB) {
      double z = 0.0;
      for (int sl = 0; sl < 16; ++sl) z += part[sl][threadIdx.x];
      yb[threadIdx.x] = z;
      x[kb * NB + threadIdx.x] = z;
    }
    __syncthreads();
    // x_r -= sum_{c in block kb} L[c][r] z_c for r < kb*NB
    for (int r = threadIdx.x; r < kb * NB; r += blockDim.x) {
      double acc = 0.0;
      for (int k = 0; k < NB; ++k) acc += A[(int64_t)(kb * NB + k) * ld + r] * yb[k];
      x[r] -= acc;
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(1024) k_trsv_large(WinInfo wi, WinState *ws, int w, const double *S,
                                                     const double *Dinv, const double *rhs, double *dred) {
  const WinState &s = ws[w];
  if (!(s.active && s.trial) || s.chol_fail) return;
  trsv_window(wi, S, Dinv, rhs + wi.r_off, dred + wi.r_off);
}

// ---------------------------------------------------------------- small path
// One workgroup (256 threads) per window: the whole blocked factorization.
__global__ void __launch_bounds__(256) k_chol_small(const WinInfo *wis, WinState *ws, const int *win_list,
                                                    double *S, double *Dinv) {
  __shared__ double a[NB][LDT];
  __shared__ double b[NB][LDT];
  __shared__ int fail;
  const int w = win_list[blockIdx.x];
  const WinInfo wi = wis[w];
  WinState &s = ws[w];
  if (!(s.active && s.trial) || s.chol_fail) return;
  if (threadIdx.x == 0) fail = 0;
  double *A = S + wi.s_off;
  const int64_t ld = wi.ld;
  for (int kb = 0; kb < wi.nb; ++kb) {
    __syncthreads();
    load_tile(a, A, ld, kb * NB, kb * NB);
    __syncthreads();
    tile_potrf(a, &fail);
    tile_trtri(a, b);
    store_tile(a, A, ld, kb * NB, kb * NB);
    store_tile(b, Dinv + wi.dinv_off + (int64_t)kb * NB * NB, NB, 0, 0);
    __syncthreads();
    for (int ib = kb + 1; ib < wi.nb; ++ib) {
      load_tile(a, A, ld, ib * NB, kb * NB);
      __syncthreads();
      tile_gemm_nt(A, ld, ib * NB, kb * NB, a, b, false);
      __syncthreads();
    }
    __threadfence_block();
    for (int ib = kb + 1; ib < wi.nb; ++ib) {
      for (int jb = kb + 1; jb <= ib; ++jb) {
        __syncthreads();
        load_tile(a, A, ld, ib * NB, kb * NB);
        load_tile(b, A, ld, jb * NB, kb * NB);
        __syncthreads();
        tile_gemm_nt(A, ld, ib * NB, jb * NB, a, b, true);
      }
    }
    __threadfence_block();
  }
  __syncthreads();
  if (threadIdx.x == 0 && fail) s.chol_fail = 1;
}

__global__ void __launch_bounds__(1024) k_trsv_small(const WinInfo *wis, WinState *ws, const int *win_list,
                                                     const double *S, const double *Dinv, const double *rhs,
                                                     double *dred) {
  const int w = win_list[blockIdx.x];
  const WinInfo wi = wis[w];
  const WinState &s = ws[w];
  if (!(s.active && s.trial) || s.chol_fail) return;
  trsv_window(wi, S, Dinv, rhs + wi.r_off, dred + wi.r_off);
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

// small_list / large windows are described by the host copy of WinInfo.
void chol_factor_solve_impl(const CholPlan &plan, const int *small_list_dev, int num_small,
                            const int *large_host, int num_large, hipStream_t st) {
  if (num_small > 0) {
    hipLaunchKernelGGL(k_chol_small, dim3(num_small), dim3(256), 0, st, plan.wi, plan.ws, small_list_dev,
                       plan.S, plan.Dinv);
    hipLaunchKernelGGL(k_trsv_small, dim3(num_small), dim3(1024), 0, st, plan.wi, plan.ws, small_list_dev,
                       plan.S, plan.Dinv, plan.rhs, plan.dred);
  }
  for (int q = 0; q < num_large; ++q) {
    const int w = large_host[q];
    const WinInfo wi = plan.wi_host[w];
    for (int kb = 0; kb < wi.nb; ++kb) {
      hipLaunchKernelGGL(k_chol_diag, dim3(1), dim3(256), 0, st, wi, plan.ws, w, plan.S, plan.Dinv, kb);
      const int rest = wi.nb - kb - 1;
      if (rest > 0) {
        hipLaunchKernelGGL(k_chol_panel, dim3(rest), dim3(256), 0, st, wi, plan.ws, w, plan.S, plan.Dinv, kb);
        hipLaunchKernelGGL(k_chol_update, dim3(rest * (rest + 1) / 2), dim3(256), 0, st, wi, plan.ws, w, plan.S,
                           kb);
      }
    }
    hipLaunchKernelGGL(k_trsv_large, dim3(1), dim3(1024), 0, st, wi, plan.ws, w, plan.S, plan.Dinv, plan.rhs,
                       plan.dred);
  }
}

}  // namespace hsba
