// Dense float64 Cholesky of the reduced camera system with MFMA trailing updates.
//
// Replaces torch.linalg.cholesky / cholesky_solve on the dense Schur matrix
// (solvers.py:516-518).  Blocked right-looking algorithm on the lower
// triangle, panel width 64:
//   1. panel kernel: every workgroup factors the 64x64 diagonal block in LDS
//      (redundantly, no inter-workgroup sync); workgroup 0 writes L_kk, the
//      others solve one 64-row block of the panel below: A_ik <- A_ik L_kk^-T;
//   2. trailing kernel: one workgroup per lower 64x64 block (i >= j) of the
//      trailing matrix computes A_ij -= L_ik L_jk^T with
//      v_mfma_f64_16x16x4_f64 (4 waves x 4 16x16 tiles, K = 64).  Operand
//      layout: lane l supplies A[l % 16][k0 + l / 16] and B[k0 + l / 16][l % 16];
//      accumulator element v of lane l is C[4 v + l / 16][l % 16].
// Non-positive or non-finite pivots set *status (LinAlgError analogue).
#include "common.hpp"

using namespace tlh;

namespace {

constexpr int NB = 64;
typedef double v4d __attribute__((ext_vector_type(4)));

__device__ void factor_diag(double (*D)[NB + 1], int kb, bool& ok) {
  __shared__ double pivot;
  for (int j = 0; j < kb; ++j) {
    if (threadIdx.x == 0) {
      const double d = D[j][j];
      if (!(d > 0.0) || !isfinite(d)) ok = false;
      pivot = sqrt(d > 0.0 ? d : 1.0);
      D[j][j] = pivot;
    }
    __syncthreads();
    const double p = pivot;
    for (int i = j + 1 + threadIdx.x; i < kb; i += TLH_THREADS) D[i][j] /= p;
    __syncthreads();
    const int m = kb - j - 1;
    for (int e = threadIdx.x; e < m * m; e += TLH_THREADS) {
      const int ii = j + 1 + e / m, ll = j + 1 + e % m;
      if (ll <= ii) D[ii][ll] -= D[ii][j] * D[ll][j];
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(TLH_THREADS)
    chol_panel_kernel(double* __restrict__ A, int64_t n, int64_t ld, int64_t k, int32_t* status) {
  __shared__ double D[NB][NB + 1];
  __shared__ double P[NB][NB + 1];
  __shared__ int okflag;
  const int kb = (int)min((int64_t)NB, n - k);
  for (int e = threadIdx.x; e < NB * NB; e += TLH_THREADS) {
    const int i = e / NB, j = e % NB;
    D[i][j] = (i < kb && j <= i) ? A[(k + i) * ld + k + j] : 0.0;
  }
  if (threadIdx.x == 0) okflag = 1;
  __syncthreads();
  bool ok = true;
  factor_diag(D, kb, ok);
  if (threadIdx.x == 0 && !ok) okflag = 0;
  __syncthreads();
  if (blockIdx.x == 0) {
    for (int e = threadIdx.x; e < kb * kb; e += TLH_THREADS) {
      const int i = e / kb, j = e % kb;
      if (j <= i) A[(k + i) * ld + k + j] = D[i][j];
    }
    if (threadIdx.x == 0 && !okflag) __hip_atomic_fetch_or(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  const int64_t r0 = k + kb + (int64_t)(blockIdx.x - 1) * NB;
  const int rows = (int)min((int64_t)NB, n - r0);
  if (rows <= 0) return;
  for (int e = threadIdx.x; e < rows * kb; e += TLH_THREADS) {
    const int i = e / kb, j = e % kb;
    P[i][j] = A[(r0 + i) * ld + k + j];
  }
  __syncthreads();
  if ((int)threadIdx.x < rows) {
    const int t = threadIdx.x;
    for (int j = 0; j < kb; ++j) {
      double s = P[t][j];
      for (int l = 0; l < j; ++l) s -= P[t][l] * D[j][l];
      P[t][j] = s / D[j][j];
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < rows * kb; e += TLH_THREADS) {
    const int i = e / kb, j = e % kb;
    A[(r0 + i) * ld + k + j] = P[i][j];
  }
}

// Trailing update of the lower blocks (bi >= bj) right of panel k.
__global__ void __launch_bounds__(TLH_THREADS)
    chol_trailing_kernel(double* __restrict__ A, int64_t n, int64_t ld, int64_t k) {
  __shared__ double Pi[NB][NB + 1];
  __shared__ double Pj[NB][NB + 1];
  // triangular block index: blockIdx.x -> (bi, bj), bi >= bj
  const int64_t b = blockIdx.x;
  int64_t bi = (int64_t)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
  while ((bi + 1) * (bi + 2) / 2 <= b) ++bi;
  while (bi * (bi + 1) / 2 > b) --bi;
  const int64_t bj = b - bi * (bi + 1) / 2;
  const int64_t base = k + NB;
  const int64_t ri = base + bi * NB, rj = base + bj * NB;
  const int kb = (int)min((int64_t)NB, n - k);
  for (int e = threadIdx.x; e < NB * NB; e += TLH_THREADS) {
    const int i = e / NB, j = e % NB;
    Pi[i][j] = (ri + i < n && j < kb) ? A[(ri + i) * ld + k + j] : 0.0;
    Pj[i][j] = (rj + i < n && j < kb) ? A[(rj + i) * ld + k + j] : 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x % 64;
  const int w = threadIdx.x / 64;
  v4d acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < NB; k0 += 4) {
    const double a = Pi[16 * w + (lane % 16)][k0 + lane / 16];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const double bb = Pj[16 * t + (lane % 16)][k0 + lane / 16];
      acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bb, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t row = ri + 16 * w + 4 * v + lane / 16;
      const int64_t col = rj + 16 * t + (lane % 16);
      if (row < n && col < n && col <= row) A[row * ld + col] -= acc[t][v];
    }
  }
}

// Forward (L y = b) then backward (L^T x = y) substitution, one workgroup.
__global__ void __launch_bounds__(1024) chol_solve_kernel(const double* __restrict__ L, int64_t n, int64_t ld,
                                                          double* __restrict__ b) {
  __shared__ double xs;
  for (int64_t j = 0; j < n; ++j) {
    if (threadIdx.x == 0) {
      const double v = b[j] / L[j * ld + j];
      b[j] = v;
      xs = v;
    }
    __syncthreads();
    const double v = xs;
    for (int64_t i = j + 1 + threadIdx.x; i < n; i += blockDim.x) b[i] -= L[i * ld + j] * v;
    __syncthreads();
  }
  for (int64_t j = n - 1; j >= 0; --j) {
    if (threadIdx.x == 0) {
      const double v = b[j] / L[j * ld + j];
      b[j] = v;
      xs = v;
    }
    __syncthreads();
    const double v = xs;
    for (int64_t i = threadIdx.x; i < j; i += blockDim.x) b[i] -= L[j * ld + i] * v;
    __syncthreads();
  }
}

}  // namespace

extern "C" int tlh_cholesky_f64(double* A, int64_t n, int64_t ld, int32_t* status, tlh_stream_t stream) {
  if (!A || !status || n < 0 || ld < n) return TLH_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  for (int64_t k = 0; k < n; k += NB) {
    const int64_t below = n - k - NB > 0 ? ceil_div(n - k - NB, NB) : 0;
    hipLaunchKernelGGL(chol_panel_kernel, dim3((unsigned)(1 + below)), dim3(TLH_THREADS), 0, s, A, n, ld, k, status);
    TLH_RETURN_IF_LAUNCH_FAILED();
    if (below > 0) {
      const int64_t blocks = below * (below + 1) / 2;
      hipLaunchKernelGGL(chol_trailing_kernel, dim3((unsigned)blocks), dim3(TLH_THREADS), 0, s, A, n, ld, k);
      TLH_RETURN_IF_LAUNCH_FAILED();
    }
  }
  return 0;
}

extern "C" int tlh_cholesky_solve_f64(const double* L, int64_t n, int64_t ld, double* b, tlh_stream_t stream) {
  if (!L || !b || n < 0 || ld < n) return TLH_EINVAL;
  if (n == 0) return 0;
  hipLaunchKernelGGL(chol_solve_kernel, dim3(1), dim3(1024), 0, reinterpret_cast<hipStream_t>(stream), L, n, ld, b);
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}
