// hsba_dense.hip — dense FP64 building blocks of the full posterior
// covariance (CameraSharedAnalyticalXCovarianceCalculator,
// camera_shared_analytical_x_covariance_calculator.hpp:45-90): a matrix-core
// GEMM, block copies, and the pseudo-inverse of a dense symmetric matrix by
// the cyclic Jacobi eigenvalue method.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/hsba.h"
#include "hsba_kernels.h"

namespace hsba {
namespace {

typedef double gdbl4 __attribute__((ext_vector_type(4)));
constexpr int GT = 64;        // output tile (rows and columns)
constexpr int GK = 16;        // depth of one LDS stage
constexpr int GLD = GK + 1;   // LDS row stride (doubles)

// C[m][n] (m < M, n < N, row pitch ldc) = alpha * sum_k opA(m, k) opB(k, n)
// (+ C when accumulate), with opA(m, k) = TA ? A[k * lda + m] : A[m * lda + k]
// and opB(k, n) = TB ? B[n * ldb + k] : B[k * ldb + n].  Operands are read in
// whole tiles: rows / columns up to the next multiple of 64 of M and N and K
// (a multiple of 16) must exist in A and B (the callers pad with zeros).
// One workgroup per 64 x 64 tile of C; wave w owns the 32 x 32 quadrant
// (rows 32 (w / 2), columns 32 (w % 2)) as 2 x 2 blocks of 16 x 16 on
// v_mfma_f64_16x16x4_f64 (A[i][k] from lane i + 16 k, B[k][j] from lane
// j + 16 k, D[i][j] in lane j + 16 (i % 4), element i / 4).  The next stage's
// operands are loaded into registers while the current stage multiplies.
template <bool TA, bool TB>
__global__ void __launch_bounds__(256) k_gemm(int64_t M, int64_t N, int64_t K, double alpha, const double *A,
                                              int64_t lda, const double *B, int64_t ldb, double *C, int64_t ldc,
                                              int accumulate) {
  __shared__ double sa[GT][GLD];
  __shared__ double sb[GT][GLD];
  const int t = threadIdx.x;
  const int64_t bm = (int64_t)blockIdx.y * GT, bn = (int64_t)blockIdx.x * GT;
  double ra[4], rb[4];
  // untransposed operand (row-major over k): row t / 4, k = 4 (t % 4) .. +3;
  // transposed operand (k-major): k = t / 16, rows 4 (t % 16) .. +3
  auto fetch = [&](int64_t k0) {
    const double *pa = TA ? A + (k0 + (t >> 4)) * lda + bm + (t & 15) * 4 : A + (bm + (t >> 2)) * lda + k0 + (t & 3) * 4;
    const double *pb = TB ? B + (bn + (t >> 2)) * ldb + k0 + (t & 3) * 4 : B + (k0 + (t >> 4)) * ldb + bn + (t & 15) * 4;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      ra[e] = pa[e];
      rb[e] = pb[e];
    }
  };
  auto put = [&]() {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (TA) sa[(t & 15) * 4 + e][t >> 4] = ra[e];
      else sa[t >> 2][(t & 3) * 4 + e] = ra[e];
      if (TB) sb[t >> 2][(t & 3) * 4 + e] = rb[e];
      else sb[(t & 15) * 4 + e][t >> 4] = rb[e];
    }
  };
  const int w = t >> 6, l = t & 63, li = l & 15, lk = l >> 4;
  const int r0 = 32 * (w >> 1), c0 = 32 * (w & 1);
  gdbl4 acc[2][2];
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v) acc[u][v] = gdbl4{0.0, 0.0, 0.0, 0.0};
  fetch(0);
  for (int64_t k0 = 0; k0 < K; k0 += GK) {
    put();
    __syncthreads();
    if (k0 + GK < K) fetch(k0 + GK);
#pragma unroll
    for (int k = 0; k < GK; k += 4) {
      const double a0 = sa[r0 + li][k + lk], a1 = sa[r0 + 16 + li][k + lk];
      const double b0 = sb[c0 + li][k + lk], b1 = sb[c0 + 16 + li][k + lk];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
  const int64_t rb_ = bm + r0 + (l >> 4), cb_ = bn + c0 + (l & 15);
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t r = rb_ + 16 * u + 4 * e, c = cb_ + 16 * v;
        if (r < M && c < N) {
          double *o = C + r * ldc + c;
          *o = accumulate ? *o + alpha * acc[u][v][e] : alpha * acc[u][v][e];
        }
      }
}

// dst[i][j] (i < rows, j < cols) from src: mode 0 src[i][j], mode 1 src[j][i]
// (through a 32 x 32 LDS tile, so both sides are read and written along
// rows), mode 2 the symmetric matrix whose lower triangle src holds.
__global__ void __launch_bounds__(256) k_copy(const double *src, int64_t lds, double *dst, int64_t ldd, int64_t rows,
                                              int64_t cols, int mode) {
  __shared__ double tile[32][33];
  const int64_t i0 = (int64_t)blockIdx.y * 32, j0 = (int64_t)blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  if (mode == 1) {
    // tile[jj][ii] = src[j0 + jj][i0 + ii]
    for (int jj = ty; jj < 32; jj += 8)
      if (j0 + jj < cols && i0 + tx < rows) tile[jj][tx] = src[(j0 + jj) * lds + i0 + tx];
    __syncthreads();
    for (int ii = ty; ii < 32; ii += 8)
      if (i0 + ii < rows && j0 + tx < cols) dst[(i0 + ii) * ldd + j0 + tx] = tile[tx][ii];
    return;
  }
  for (int ii = ty; ii < 32; ii += 8) {
    const int64_t i = i0 + ii, j = j0 + tx;
    if (i >= rows || j >= cols) continue;
    dst[i * ldd + j] = (mode == 2 && j > i) ? src[j * lds + i] : src[i * lds + j];
  }
}

__global__ void k_fill_int(int *p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// V = I on rows / columns < n of an ld x ld matrix, zero elsewhere.
__global__ void k_identity(double *V, int64_t ld, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ld * ld; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / ld, j = e - i * ld;
    V[e] = (i == j && i < n) ? 1.0 : 0.0;
  }
}

// Sum of v over the 256 threads, the same value in every thread; fixed order
// (butterfly inside each wave, then the four waves in order).
__device__ __forceinline__ void block_sum3(double v[3], double (*red)[3]) {
#pragma unroll
  for (int s = 1; s < 64; s <<= 1)
#pragma unroll
    for (int q = 0; q < 3; ++q) v[q] += __shfl_xor(v[q], s, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int q = 0; q < 3; ++q) red[w][q] = v[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 3; ++q) v[q] = (red[0][q] + red[1][q]) + (red[2][q] + red[3][q]);
}

// Cyclic two-sided Jacobi eigenvalue method for the symmetric n x n matrix A
// (row-major, pitch ld), in the parallel round-robin order: in round r of a
// sweep over m = n rounded up to even indices, position 0 holds index 0 and
// position k > 0 index (k - 1 + r) mod (m - 1) + 1; slot i pairs the indices
// at positions i and m - 1 - i (p < q; q >= n is a bye).
__device__ __forceinline__ void round_pair(int slot, int m, int r, int *p, int *q) {
  const int a = slot == 0 ? 0 : (slot - 1 + r) % (m - 1) + 1;
  const int pos = m - 1 - slot;
  const int b = pos == 0 ? 0 : (pos - 1 + r) % (m - 1) + 1;
  *p = a < b ? a : b;
  *q = a < b ? b : a;
}

// Rotation of each pair of round r (one thread per slot): for
// |a_pq| > max(tol sqrt|a_pp a_qq|, floor), tau = (a_qq - a_pp) / (2 a_pq),
// t = sign(tau) / (|tau| + sqrt(1 + tau^2)), c = 1 / sqrt(1 + t^2), s = t c —
// the rotation J (J_pp = J_qq = c, J_pq = s, J_qp = -s) for which J^T A J has
// a zero (p, q) entry (Golub & Van Loan, Matrix Computations, Alg. 8.5.1).
// Other slots get (1, 0).  Each rotation counts into *rotations.
__global__ void k_jac_angles(const double *A, int64_t ld, int n, int m, int r, double tol, double floor_,
                             double *rot, int *rotations) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m / 2) return;
  int p, q;
  round_pair(i, m, r, &p, &q);
  double c = 1.0, s = 0.0;
  if (q < n) {
    const double apq = A[(int64_t)p * ld + q], app = A[(int64_t)p * ld + p], aqq = A[(int64_t)q * ld + q];
    if (fabs(apq) > fmax(tol * sqrt(fabs(app * aqq)), floor_)) {
      const double tau = (aqq - app) / (2.0 * apq);
      const double at = fabs(tau);
      double t = at > 1e150 ? 0.5 / at : 1.0 / (at + sqrt(1.0 + at * at));
      if (tau < 0.0) t = -t;
      c = 1.0 / sqrt(1.0 + t * t);
      s = t * c;
      atomicAdd(rotations, 1);
    }
  }
  rot[2 * i] = c;
  rot[2 * i + 1] = s;
}

constexpr int kJacMaxN = 4096;  // both rows of a pair in LDS (64 KiB)

// A <- J^T A J and Vt <- J^T Vt for all rotations J of round r: workgroup i
// owns rows p, q of slot i (no other workgroup reads or writes them), applies
// its own rotation to the two rows, then every slot's rotation to their
// columns (slots are disjoint, so each LDS entry is updated by one thread).
__global__ void __launch_bounds__(256) k_jac_apply(double *A, double *Vt, int64_t ld, int n, int m, int r,
                                                   const double *rot) {
  __shared__ double rp[kJacMaxN];
  __shared__ double rq[kJacMaxN];
  const int i = blockIdx.x;
  int p, q;
  round_pair(i, m, r, &p, &q);
  const bool two = q < n;
  const double c = rot[2 * i], s = rot[2 * i + 1];
  double *ap = A + (int64_t)p * ld, *aq = two ? A + (int64_t)q * ld : nullptr;
  for (int e = threadIdx.x; e < n; e += 256) {
    const double x = ap[e], y = two ? aq[e] : 0.0;
    rp[e] = c * x - s * y;
    if (two) rq[e] = s * x + c * y;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < m / 2; j += 256) {
    int pj, qj;
    round_pair(j, m, r, &pj, &qj);
    const double sj = rot[2 * j + 1];
    if (qj >= n || sj == 0.0) continue;
    const double cj = rot[2 * j];
    const double x = rp[pj], y = rp[qj];
    rp[pj] = cj * x - sj * y;
    rp[qj] = sj * x + cj * y;
    if (two) {
      const double u = rq[pj], v = rq[qj];
      rq[pj] = cj * u - sj * v;
      rq[qj] = sj * u + cj * v;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n; e += 256) {
    ap[e] = rp[e];
    if (two) aq[e] = rq[e];
  }
  if (two && s != 0.0) {
    double *vp = Vt + (int64_t)p * ld, *vq = Vt + (int64_t)q * ld;
    for (int e = threadIdx.x; e < n; e += 256) {
      const double x = vp[e], y = vq[e];
      vp[e] = c * x - s * y;
      vq[e] = s * x + c * y;
    }
  }
}

// Sum of squares of each row (the Frobenius norm, invariant under the
// rotations, sets the noise floor of the off-diagonal test).
__global__ void __launch_bounds__(256) k_row_sq(const double *A, int64_t ld, int n, double *out) {
  __shared__ double red[4][3];
  const double *a = A + (int64_t)blockIdx.x * ld;
  double s3[3] = {0.0, 0.0, 0.0};
  for (int e = threadIdx.x; e < n; e += 256) s3[0] += a[e] * a[e];
  block_sum3(s3, red);
  if (threadIdx.x == 0) out[blockIdx.x] = s3[0];
}

// Eigenvalue lambda_j = A[j][j] after convergence; row j of A becomes
// w_j Vt[j] with w_j = 1 / lambda_j when |lambda_j| > thr, else 0, so that
// Vt^T A = sum over kept j of v_j v_j^T / lambda_j = V D U^T of the SVD
// (u_j = sign(lambda_j) v_j, d_j = |lambda_j|).
__global__ void __launch_bounds__(256) k_eig_scale(double *A, const double *Vt, int64_t ld, int n, double thr,
                                                   double *lam) {
  __shared__ double lj;
  const int j = blockIdx.x;
  if (threadIdx.x == 0) lj = A[(int64_t)j * ld + j];
  __syncthreads();
  const double l = lj;
  const double w = fabs(l) > thr ? 1.0 / l : 0.0;
  for (int e = threadIdx.x; e < n; e += 256) A[(int64_t)j * ld + e] = w * Vt[(int64_t)j * ld + e];
  if (threadIdx.x == 0) lam[j] = l;
}

}  // namespace

void dense_gemm(bool ta, bool tb, int64_t M, int64_t N, int64_t K, double alpha, const double *A, int64_t lda,
                const double *B, int64_t ldb, double *C, int64_t ldc, bool accumulate, hipStream_t st) {
  if (M <= 0 || N <= 0) return;
  const dim3 grid((unsigned)((N + GT - 1) / GT), (unsigned)((M + GT - 1) / GT));
  const int acc = accumulate ? 1 : 0;
  if (!ta && !tb) hipLaunchKernelGGL((k_gemm<false, false>), grid, dim3(256), 0, st, M, N, K, alpha, A, lda, B, ldb, C, ldc, acc);
  else if (!ta && tb) hipLaunchKernelGGL((k_gemm<false, true>), grid, dim3(256), 0, st, M, N, K, alpha, A, lda, B, ldb, C, ldc, acc);
  else if (ta && !tb) hipLaunchKernelGGL((k_gemm<true, false>), grid, dim3(256), 0, st, M, N, K, alpha, A, lda, B, ldb, C, ldc, acc);
  else hipLaunchKernelGGL((k_gemm<true, true>), grid, dim3(256), 0, st, M, N, K, alpha, A, lda, B, ldb, C, ldc, acc);
}

void dense_copy(const double *src, int64_t lds, double *dst, int64_t ldd, int64_t rows, int64_t cols, int mode,
                hipStream_t st) {
  if (rows <= 0 || cols <= 0) return;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32)), dim3(256), 0, st, src,
                     lds, dst, ldd, rows, cols, mode);
}

int dense_pinv(double *A, double *Vt, int64_t n, int64_t ld, double thr, double *out, int64_t ldo, int max_sweeps,
               int *rotations_dev, double *work_dev, std::vector<double> &lam, int *sweeps, hipStream_t st) {
  lam.clear();
  if (n <= 0) return HSBA_OK;
  if (n > kJacMaxN) return HSBA_ERR_TOO_LARGE;
  const int nn = (int)n, m = nn + (nn & 1);
  double *rot = work_dev, *aux = work_dev + m;
  const int64_t cells = ld * ld;
  hipLaunchKernelGGL(k_identity, dim3((unsigned)std::min<int64_t>((cells + 255) / 256, 65536)), dim3(256), 0, st, Vt,
                     ld, n);
  hipLaunchKernelGGL(k_row_sq, dim3(nn), dim3(256), 0, st, A, ld, nn, aux);
  std::vector<double> sq((size_t)n);
  if (hipMemcpyAsync(sq.data(), aux, sizeof(double) * sq.size(), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return HSBA_ERR_DEVICE;
  double fro2 = 0.0;
  for (double v : sq) fro2 += v;
  const double eps = 2.220446049250313e-16;
  const double floor_ = eps * std::sqrt(fro2);
  // one sweep (the rotation counter cleared by a kernel, then every round's
  // two launches) captured as a graph and replayed until a sweep rotates
  // nothing
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return HSBA_ERR_DEVICE;
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, rotations_dev, 1, 0);
  for (int r = 0; r < m - 1; ++r) {
    hipLaunchKernelGGL(k_jac_angles, dim3((m / 2 + 255) / 256), dim3(256), 0, st, A, ld, nn, m, r, eps, floor_, rot,
                       rotations_dev);
    hipLaunchKernelGGL(k_jac_apply, dim3(m / 2), dim3(256), 0, st, A, Vt, ld, nn, m, r, rot);
  }
  const hipError_t ec = hipStreamEndCapture(st, &graph);
  if (ec != hipSuccess || hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return HSBA_ERR_DEVICE;
  }
  int done = 0, rc = HSBA_OK;
  for (int sweep = 0; sweep < max_sweeps && rc == HSBA_OK; ++sweep) {
    int count = 0;
    if (hipGraphLaunch(exec, st) != hipSuccess ||
        hipMemcpyAsync(&count, rotations_dev, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      rc = HSBA_ERR_DEVICE;
      break;
    }
    *sweeps = sweep + 1;
    if (count == 0) done = 1;
    if (done) break;
  }
  (void)hipGraphExecDestroy(exec);
  (void)hipGraphDestroy(graph);
  if (rc) return rc;
  if (!done) return HSBA_ERR_GENERIC;
  hipLaunchKernelGGL(k_eig_scale, dim3(nn), dim3(256), 0, st, A, Vt, ld, nn, thr, aux);
  // out = Vt^T (w Vt): element (a, b) = sum_j Vt[j][a] w_j Vt[j][b]
  dense_gemm(true, false, n, n, ld, 1.0, Vt, ld, A, ld, out, ldo, false, st);
  lam.resize((size_t)n);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(lam.data(), aux, sizeof(double) * lam.size(), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return HSBA_ERR_DEVICE;
  return HSBA_OK;
}

}  // namespace hsba
