// Dense float64 Cholesky of the reduced camera system with MFMA trailing updates.
//
// Replaces torch.linalg.cholesky / cholesky_solve on the dense Schur matrix
// (solvers.py:516-518).  Blocked right-looking algorithm on the lower
// triangle, panel width 64:
//   1. panel kernel (eight wavefronts per workgroup, a row per lane): every
//      workgroup factors the 64x64 diagonal block (redundantly, no
//      inter-workgroup sync); workgroup 0 writes L_kk, the others solve one
//      64-row block of the panel below: A_ik <- A_ik L_kk^-T;
//   2. trailing kernel: one workgroup per lower 64x64 block (i >= j) of the
//      trailing matrix computes A_ij -= L_ik L_jk^T with
//      v_mfma_f64_16x16x4_f64 (4 waves x 4 16x16 tiles, K = 64).  Operand
//      layout: lane l supplies A[l % 16][k0 + l / 16] and B[k0 + l / 16][l % 16];
//      accumulator element v of lane l is C[4 v + l / 16][l % 16].
// Non-positive or non-finite pivots set *status (LinAlgError analogue).
#include <mutex>

#include "common.hpp"

using namespace tlh;

namespace {

constexpr int NB = 64;
typedef double v4d __attribute__((ext_vector_type(4)));

// Panel kernel: one wavefront per workgroup, lane r owning row r of a 64-wide
// block kept in LDS; columns are processed in sub-panels of 16 held in
// registers (v[16]), so every register index is a constant.  Inside a
// sub-panel the pivot and the column entries a lane needs from other rows are
// read with lane reads (no LDS round trip, no barrier); the update of the
// later sub-panels reads the finished sub-panel rows from LDS.
constexpr int PT = 64;
constexpr int SUB = 16;

__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// sum_l v[l] w[l] over a sub-panel, as four independent partial sums (the
// chains overlap instead of waiting on each other's FMA latency).
__device__ __forceinline__ double dot_sub(const double (&v)[SUB], const double* w) {
  double p[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int l = 0; l < SUB; ++l) p[l & 3] += v[l] * w[l];
  return (p[0] + p[1]) + (p[2] + p[3]);
}

// Panel workgroups hold eight wavefronts: wavefront 0 runs the pivot (and
// triangular) steps of each 16-column sub-panel, lane r owning row r, and all
// eight then share the update of the columns right of the sub-panel, wavefront
// w taking columns c0 + 16 + w, c0 + 24 + w, ... (every (row, column) update is
// the same dot_sub as with one wavefront, so the factor is unchanged).
constexpr int PANEL_WAVES = 8;
constexpr int PANEL_THREADS = 64 * PANEL_WAVES;

// Cholesky of the 64x64 block in W (row r = lane; lower triangle; padded rows
// hold the identity).  Right-looking by sub-panels of 16 columns.
__device__ __forceinline__ void factor_block(double (*W)[NB + 1], int wave, int lane, bool& ok) {
#pragma unroll
  for (int sp = 0; sp < NB / SUB; ++sp) {
    const int c0 = sp * SUB;
    if (wave == 0) {
      double v[SUB];
#pragma unroll
      for (int c = 0; c < SUB; ++c) v[c] = W[lane][c0 + c];
#pragma unroll
      for (int j = 0; j < SUB; ++j) {
        const int pr = c0 + j;
        const double dj = readlane_f64(v[j], pr);
        const bool good = dj > 0.0 && isfinite(dj);
        if (!good) ok = false;
        const double inv = rsqrt(good ? dj : 1.0);  // one reciprocal square root, no divide
        const double p = (good ? dj : 1.0) * inv;
        if (lane == pr) v[j] = p;
        else if (lane > pr) v[j] *= inv;
#pragma unroll
        for (int c = j + 1; c < SUB; ++c) {
          const double lc = readlane_f64(v[j], c0 + c);  // L[c0 + c][pr]
          if (lane >= c0 + c) v[c] -= v[j] * lc;
        }
      }
#pragma unroll
      for (int c = 0; c < SUB; ++c) W[lane][c0 + c] = v[c];
    }
    __syncthreads();
    if (c0 + SUB < NB) {
      double v[SUB];
#pragma unroll
      for (int c = 0; c < SUB; ++c) v[c] = W[lane][c0 + c];
      for (int c = c0 + SUB + wave; c < NB; c += PANEL_WAVES) {
        if (lane >= c) W[lane][c] -= dot_sub(v, &W[c][c0]);
      }
      __syncthreads();
    }
  }
}

// Panel k, job wg: every job factors the diagonal block A_kk itself (no
// inter-workgroup sync).  Job 0 writes L_kk over A_kk and reports a bad pivot;
// job w >= 1 solves X L_kk^T = A_ik for the w-th 64-row block below, lane r
// forming row r of X independently (sub-panel by sub-panel, L_kk read from LDS
// by every lane at the same address).  Job 0 never runs in the same launch as
// jobs w >= 1 of its panel, which read A_kk.  The rows of A_ik are loaded
// before the factorisation so those loads overlap it.  W and X are the two
// 64x65 LDS tiles of the caller.
__device__ __forceinline__ void panel_body(double* __restrict__ A, int64_t n, int64_t ld, int64_t k, int32_t* status,
                                           int wg, double (*W)[NB + 1], double (*X)[NB + 1], double* inv_diag) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int kb = (int)min((int64_t)NB, n - k);
  const int64_t r0 = k + kb + (int64_t)(wg - 1) * NB;
  const int rows = wg == 0 ? 0 : (int)max((int64_t)0, min((int64_t)NB, n - r0));
  if (wg != 0 && rows == 0) return;  // uniform over the workgroup
  double gw[NB / PANEL_WAVES], gx[NB / PANEL_WAVES];  // all loads in flight before the LDS stores
#pragma unroll
  for (int i = 0; i < NB / PANEL_WAVES; ++i) {  // coalesced: lane = column
    const int r = wave + PANEL_WAVES * i;
    gw[i] = (r < kb && lane <= r) ? A[(k + r) * ld + k + lane] : (r == lane ? 1.0 : 0.0);
    gx[i] = (rows > 0 && r < rows && lane < kb) ? A[(r0 + r) * ld + k + lane] : 0.0;
  }
#pragma unroll
  for (int i = 0; i < NB / PANEL_WAVES; ++i) {
    const int r = wave + PANEL_WAVES * i;
    W[r][lane] = gw[i];
    X[r][lane] = gx[i];
  }
  __syncthreads();
  bool ok = true;
  factor_block(W, wave, lane, ok);
  if (wg == 0) {
    const bool all_ok = __builtin_amdgcn_readfirstlane((int)__any(!ok)) == 0;
#pragma unroll
    for (int i = 0; i < NB / PANEL_WAVES; ++i) {
      const int r = wave + PANEL_WAVES * i;
      if (r < kb && lane <= r) A[(k + r) * ld + k + lane] = W[r][lane];
    }
    if (threadIdx.x == 0 && !all_ok) __hip_atomic_fetch_or(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  if (wave == 0) inv_diag[lane] = 1.0 / W[lane][lane];
  __syncthreads();
#pragma unroll
  for (int sp = 0; sp < NB / SUB; ++sp) {
    const int c0 = sp * SUB;
    if (wave == 0) {
      double v[SUB];
#pragma unroll
      for (int c = 0; c < SUB; ++c) v[c] = X[lane][c0 + c];
#pragma unroll
      for (int j = 0; j < SUB; ++j) {
        v[j] *= inv_diag[c0 + j];
#pragma unroll
        for (int c = j + 1; c < SUB; ++c) v[c] -= v[j] * W[c0 + c][c0 + j];
      }
#pragma unroll
      for (int c = 0; c < SUB; ++c) X[lane][c0 + c] = v[c];
    }
    __syncthreads();
    if (c0 + SUB < NB) {
      double v[SUB];
#pragma unroll
      for (int c = 0; c < SUB; ++c) v[c] = X[lane][c0 + c];
      for (int c = c0 + SUB + wave; c < NB; c += PANEL_WAVES) X[lane][c] -= dot_sub(v, &W[c][c0]);
      __syncthreads();
    }
  }
#pragma unroll
  for (int i = 0; i < NB / PANEL_WAVES; ++i) {
    const int r = wave + PANEL_WAVES * i;
    if (r < rows && lane < kb) A[(r0 + r) * ld + k + lane] = X[r][lane];
  }
}

// Trailing update of one lower tile (bi >= bj) right of panel k:
// A_ij -= L_ik L_jk^T with v_mfma_f64_16x16x4_f64, the first four wavefronts
// each computing 16 rows of the 64x64 tile (any further wavefronts only help
// load the two panel tiles).
template <int NTH>
__device__ __forceinline__ void trailing_tile(double* __restrict__ A, int64_t n, int64_t ld, int64_t k, int64_t bi,
                                              int64_t bj, double (*Pi)[NB + 1], double (*Pj)[NB + 1]) {
  const int64_t base = k + NB;
  const int64_t ri = base + bi * NB, rj = base + bj * NB;
  const int kb = (int)min((int64_t)NB, n - k);
  const int lane = threadIdx.x % 64;
  const int w = threadIdx.x / 64;
  // every load is issued before the first LDS store (one memory round trip),
  // including the tile being updated, held in the MFMA output layout
  constexpr int PER = NB * NB / NTH;
  double ga[PER], gb[PER], cold[4][4];
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    const int e = (int)threadIdx.x + m * NTH;
    const int i = e / NB, j = e % NB;
    ga[m] = (ri + i < n && j < kb) ? A[(ri + i) * ld + k + j] : 0.0;
    gb[m] = (rj + i < n && j < kb) ? A[(rj + i) * ld + k + j] : 0.0;
  }
  if (w < 4) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t row = ri + 16 * w + 4 * v + lane / 16;
        const int64_t col = rj + 16 * t + (lane % 16);
        cold[t][v] = (row < n && col < n && col <= row) ? A[row * ld + col] : 0.0;
      }
    }
  }
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    const int e = (int)threadIdx.x + m * NTH;
    const int i = e / NB, j = e % NB;
    Pi[i][j] = ga[m];
    Pj[i][j] = gb[m];
  }
  __syncthreads();
  if (w >= 4) return;
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
      if (row < n && col < n && col <= row) A[row * ld + col] = cold[t][v] - acc[t][v];
    }
  }
}

// Tile b of a lower block triangle numbered row by row: b -> (bi, bj), bi >= bj.
__device__ __forceinline__ void triangle_index(int64_t b, int64_t& bi, int64_t& bj) {
  bi = (int64_t)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
  while ((bi + 1) * (bi + 2) / 2 <= b) ++bi;
  while (bi * (bi + 1) / 2 > b) --bi;
  bj = b - bi * (bi + 1) / 2;
}

// Panel k: jobs first_job + blockIdx.x.
__global__ void __launch_bounds__(PANEL_THREADS)
    chol_panel_kernel(double* __restrict__ A, int64_t n, int64_t ld, int64_t k, int first_job, int32_t* status) {
  __shared__ double S[2][NB][NB + 1];
  __shared__ double inv_diag[NB];
  panel_body(A, n, ld, k, status, first_job + (int)blockIdx.x, S[0], S[1], inv_diag);
}

// Look-ahead, part 1 of step k: the trailing update of the next panel's column
// block only (tiles (bi, 0)).
__global__ void __launch_bounds__(TLH_THREADS) chol_trailing_col_kernel(double* __restrict__ A, int64_t n, int64_t ld,
                                                                        int64_t k) {
  __shared__ double S[2][NB][NB + 1];
  trailing_tile<TLH_THREADS>(A, n, ld, k, (int64_t)blockIdx.x, 0, S[0], S[1]);
}

// Look-ahead, part 2 of step k, by workgroup:
//   0                 job 0 of panel k (writes L_kk; A_kk is read by no other job here)
//   1 .. npanel       panel k + 64: its jobs 1 .. npanel (the row blocks below it), or
//                     its job 0 when it is the last panel (no row blocks below)
//   npanel + 1 ..     the rest of step k's trailing update (tiles with bj >= 1)
// Part 1 finished panel k + 64's columns.  The sets touch disjoint tiles: job 0
// reads and writes the diagonal block of rows [k, k + 64); the panel jobs read
// its diagonal block and write rows >= k + 128 of columns [k + 64, k + 128); the
// trailing tiles read columns [k, k + 64) of rows >= k + 128 and write columns
// >= k + 128.
__global__ void __launch_bounds__(PANEL_THREADS)
    chol_step_kernel(double* __restrict__ A, int64_t n, int64_t ld, int64_t k, int npanel, int last, int32_t* status) {
  __shared__ double S[2][NB][NB + 1];
  __shared__ double inv_diag[NB];
  const int g = (int)blockIdx.x;
  if (g == 0) {
    panel_body(A, n, ld, k, status, 0, S[0], S[1], inv_diag);
    return;
  }
  if (g <= npanel) {
    panel_body(A, n, ld, k + NB, status, last ? 0 : g, S[0], S[1], inv_diag);
    return;
  }
  int64_t bi, bj;
  triangle_index(g - 1 - npanel, bi, bj);
  trailing_tile<PANEL_THREADS>(A, n, ld, k, bi + 1, bj + 1, S[0], S[1]);
}

// Forward (L y = b) then backward (L^T x = y) substitution, one workgroup of
// 1024 threads, x in LDS.  Blocked by 64: the off-diagonal part of a block row
// (column, going backward) is a coalesced GEMV reduced in a fixed order; the
// 64x64 diagonal triangle is staged in LDS and solved by one wavefront, lane r
// holding row r, the solved value broadcast by a lane read.
constexpr int SOLVE_THREADS = 1024;

__global__ void __launch_bounds__(SOLVE_THREADS)
    chol_solve_kernel(const double* __restrict__ L, int64_t n, int64_t ld, double* __restrict__ b) {
  extern __shared__ double xs[];  // [n]
  __shared__ double Ld[NB][NB + 1];
  __shared__ double inv_d[NB];
  __shared__ double part[SOLVE_THREADS / NB][NB];
  const int t = threadIdx.x;
  for (int64_t i = t; i < n; i += SOLVE_THREADS) xs[i] = b[i];
  __syncthreads();
  // forward: block rows in order
  for (int64_t i0 = 0; i0 < n; i0 += NB) {
    const int m = (int)min((int64_t)NB, n - i0);
    {
      const int row = t >> 4, g = t & 15;  // 16 consecutive lanes per row
      double p4[4] = {0.0, 0.0, 0.0, 0.0};
      if (row < m) {
        const double* lr = L + (i0 + row) * ld;
        int64_t j = g;
        for (; j + 112 < i0; j += 128) {
#pragma unroll
          for (int u = 0; u < 8; ++u) p4[u & 3] += lr[j + 16 * u] * xs[j + 16 * u];
        }
        for (; j < i0; j += 16) p4[0] += lr[j] * xs[j];
      }
      double sacc = (p4[0] + p4[1]) + (p4[2] + p4[3]);
#pragma unroll
      for (int off = 8; off > 0; off >>= 1) sacc += __shfl_xor(sacc, off, 16);
      for (int e = t; e < NB * NB; e += SOLVE_THREADS) {
        const int r = e / NB, c = e % NB;
        Ld[r][c] = (r < m && c <= r) ? L[(i0 + r) * ld + i0 + c] : 0.0;
      }
      if (t < NB) inv_d[t] = t < m ? 1.0 / L[(i0 + t) * ld + i0 + t] : 1.0;
      __syncthreads();
      if (g == 0 && row < m) xs[i0 + row] -= sacc;
      __syncthreads();
    }
    if (t < 64) {
      double v = t < m ? xs[i0 + t] : 0.0;
      for (int j = 0; j < m; ++j) {
        const double xj = readlane_f64(v, j) * inv_d[j];
        if (t == j) v = xj;
        if (t > j) v -= Ld[t][j] * xj;
      }
      if (t < m) xs[i0 + t] = v;
    }
    __syncthreads();
  }
  // backward: block columns in reverse order
  const int64_t nblocks = (n + NB - 1) / NB;
  for (int64_t bi = nblocks - 1; bi >= 0; --bi) {
    const int64_t i0 = bi * NB;
    const int m = (int)min((int64_t)NB, n - i0);
    {
      const int c = t & (NB - 1), g = t / NB;  // 64 consecutive lanes read one row of L
      constexpr int G = SOLVE_THREADS / NB;
      double p4[4] = {0.0, 0.0, 0.0, 0.0};
      if (c < m) {
        int64_t j = i0 + m + g;
        for (; j + 7 * G < n; j += 8 * G) {
#pragma unroll
          for (int u = 0; u < 8; ++u) p4[u & 3] += L[(j + u * G) * ld + i0 + c] * xs[j + u * G];
        }
        for (; j < n; j += G) p4[0] += L[j * ld + i0 + c] * xs[j];
      }
      part[g][c] = (p4[0] + p4[1]) + (p4[2] + p4[3]);
      for (int e = t; e < NB * NB; e += SOLVE_THREADS) {
        const int r = e / NB, cc = e % NB;
        Ld[r][cc] = (r < m && cc <= r) ? L[(i0 + r) * ld + i0 + cc] : 0.0;
      }
      if (t < NB) inv_d[t] = t < m ? 1.0 / L[(i0 + t) * ld + i0 + t] : 1.0;
      __syncthreads();
      if (t < m) {
        double tot = 0.0;
        for (int gg = 0; gg < SOLVE_THREADS / NB; ++gg) tot += part[gg][t];
        xs[i0 + t] -= tot;
      }
      __syncthreads();
    }
    if (t < 64) {
      double v = t < m ? xs[i0 + t] : 0.0;
      for (int j = m - 1; j >= 0; --j) {
        const double xj = readlane_f64(v, j) * inv_d[j];
        if (t == j) v = xj;
        if (t < j) v -= Ld[j][t] * xj;
      }
      if (t < m) xs[i0 + t] = v;
    }
    __syncthreads();
  }
  for (int64_t i = t; i < n; i += SOLVE_THREADS) b[i] = xs[i];
}

// Blocked substitution over many workgroups.  Workgroup w owns one 64-row block;
// a block's off-diagonal updates need the solutions of the blocks before it
// (forward) or after it (backward).  Each solved block is also stored, value by
// value with agent-scope atomic stores, into a workspace that the launch first
// fills with a signalling-NaN pattern; a waiting workgroup polls those values
// with agent-scope atomic loads until none holds the pattern, so the values are
// their own ready flags (no fences, no separate flag round trip); a solved
// value never carries the pattern (see publish_value).
// Each update loads its L slice before waiting, so the serial chain is the
// 64-step diagonal solves and the polls.  Workgroups are dispatched in index
// order and a block waits only on blocks with lower workgroup indices, so every
// wait is on a workgroup that is already running; a wait still gives up after
// SOLVE_SPIN_MAX polls (or the limit set by tlh_set_wait_limit), setting
// *status |= 2.
constexpr int SOLVE_WG = 256;
constexpr int SOLVE_SPIN_MAX = 1 << 20;  // about a second of polls; a real wait is the solve's length
constexpr long long SOLVE_EMPTY = 0x7ff4dead5eed1234ll;  // a signalling NaN

__global__ void __launch_bounds__(TLH_THREADS) solve_work_reset(long long* work, int64_t count) {
  const int64_t i = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x;
  if (i < count) work[i] = SOLVE_EMPTY;
}

// Wavefront 0 polls the cnt values at src until all are solved and stages them
// in xs (zero past cnt); the barrier also orders every earlier read of xs
// before the new values land.
__device__ __forceinline__ void wait_block(const double* src, int cnt, double* xs, int32_t* status, int spin) {
  __syncthreads();
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    const long long* p = reinterpret_cast<const long long*>(src) + lane;
    long long bits = 0;
    int polls = 0;
    while (true) {
      bits = lane < cnt ? __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
      if (__all(bits != SOLVE_EMPTY)) break;
      if (++polls > spin) {
        if (lane == 0) __hip_atomic_fetch_or(status, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    xs[lane] = lane < cnt ? __longlong_as_double(bits) : 0.0;
  }
  __syncthreads();
}

// A value equal to the pattern (possible only if a signalling NaN in b passed
// through unquieted) is published as a quiet NaN, so every wait ends.
__device__ __forceinline__ void publish_value(double* dst, double v) {
  long long bits = __double_as_longlong(v);
  if (bits == SOLVE_EMPTY) bits = 0x7ff8000000000000ll;
  __hip_atomic_store(reinterpret_cast<long long*>(dst), bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Forward: block i solves L_ii y_i = b_i - sum_{j<i} L_ij y_j.
__global__ void __launch_bounds__(SOLVE_WG)
    chol_forward_blocks(const double* __restrict__ L, int64_t n, int64_t ld, double* b, double* ywork,
                        int32_t* status, int spin) {
  __shared__ double D[NB][NB + 1];
  __shared__ double xs[NB], acc[NB];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int i = blockIdx.x;
  const int64_t i0 = (int64_t)i * NB;
  const int m = (int)min((int64_t)NB, n - i0);
  for (int e = t; e < NB * NB; e += SOLVE_WG) {
    const int r = e / NB, c = e % NB;
    D[r][c] = (r < m && c <= r) ? L[(i0 + r) * ld + i0 + c] : (r == c ? 1.0 : 0.0);
  }
  if (t < NB) acc[t] = t < m ? b[i0 + t] : 0.0;
  __syncthreads();
  // wave 0 keeps row `lane` of the diagonal block in registers for the substitution
  double drow[NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) drow[c] = D[lane][c];
  const double dinv = 1.0 / D[lane][lane];
  // four lanes per row, 16 columns each, read straight into registers; the next
  // block's slice is loaded before waiting on the current one
  const int r = t >> 2, q = t & 3;
  auto load = [&](int j, double (&g)[16]) {
#pragma unroll
    for (int c = 0; c < 16; ++c) g[c] = r < m ? L[(i0 + r) * ld + (int64_t)j * NB + q * 16 + c] : 0.0;
  };
  double cur[16];
  if (i > 0) load(0, cur);
  for (int j = 0; j < i; ++j) {
    double nxt[16];
    if (j + 1 < i) load(j + 1, nxt);
    wait_block(ywork + (int64_t)j * NB, NB, xs, status, spin);
    double sacc = 0.0;
#pragma unroll
    for (int c = 0; c < 16; ++c) sacc += cur[c] * xs[q * 16 + c];
    sacc += __shfl_xor(sacc, 1);
    sacc += __shfl_xor(sacc, 2);
    if (q == 0 && r < m) acc[r] -= sacc;
#pragma unroll
    for (int c = 0; c < 16; ++c) cur[c] = nxt[c];
  }
  __syncthreads();
  if (wave == 0) {
    double v = lane < m ? acc[lane] : 0.0;
#pragma unroll
    for (int jj = 0; jj < NB; ++jj) {
      if (jj < m) {
        const double xj = readlane_f64(v, jj) * readlane_f64(dinv, jj);
        v = lane == jj ? xj : (lane > jj ? v - drow[jj] * xj : v);
      }
    }
    if (lane < m) {
      b[i0 + lane] = v;
      publish_value(ywork + i0 + lane, v);
    }
  }
}

// Backward: workgroup w solves block k = nb - 1 - w: L_kk^T x_k = y_k - sum_{j>k} L_jk^T x_j.
__global__ void __launch_bounds__(SOLVE_WG)
    chol_backward_blocks(const double* __restrict__ L, int64_t n, int64_t ld, double* b, double* xwork,
                         int32_t* status, int spin) {
  __shared__ double D[NB][NB + 1];
  __shared__ double xs[NB], acc[NB];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int nb = (int)gridDim.x;
  const int k = nb - 1 - (int)blockIdx.x;
  const int64_t k0 = (int64_t)k * NB;
  const int m = (int)min((int64_t)NB, n - k0);
  for (int e = t; e < NB * NB; e += SOLVE_WG) {
    const int rr = e / NB, c = e % NB;
    D[rr][c] = (rr < m && c <= rr) ? L[(k0 + rr) * ld + k0 + c] : (rr == c ? 1.0 : 0.0);
  }
  if (t < NB) acc[t] = t < m ? b[k0 + t] : 0.0;
  __syncthreads();
  // wave 0 keeps column `lane` of the diagonal block in registers for the substitution
  double dcol[NB];
#pragma unroll
  for (int rr = 0; rr < NB; ++rr) dcol[rr] = D[rr][lane];
  const double dinv = 1.0 / D[lane][lane];
  // four lanes per column, 16 rows each, read straight into registers; the next
  // block's slice is loaded before waiting on the current one
  const int c = t >> 2, q = t & 3;
  auto load = [&](int jb, double (&g)[16]) {
    const int64_t j0 = (int64_t)jb * NB;
    const int mj = (int)min((int64_t)NB, n - j0);
#pragma unroll
    for (int rr = 0; rr < 16; ++rr)
      g[rr] = (q * 16 + rr < mj && c < m) ? L[(j0 + q * 16 + rr) * ld + k0 + c] : 0.0;
  };
  double cur[16];
  if (k < nb - 1) load(nb - 1, cur);
  for (int jb = nb - 1; jb > k; --jb) {
    const int64_t j0 = (int64_t)jb * NB;
    const int mj = (int)min((int64_t)NB, n - j0);
    double nxt[16];
    if (jb - 1 > k) load(jb - 1, nxt);
    wait_block(xwork + j0, mj, xs, status, spin);
    double sacc = 0.0;
#pragma unroll
    for (int rr = 0; rr < 16; ++rr) sacc += cur[rr] * xs[q * 16 + rr];
    sacc += __shfl_xor(sacc, 1);
    sacc += __shfl_xor(sacc, 2);
    if (q == 0 && c < m) acc[c] -= sacc;
#pragma unroll
    for (int rr = 0; rr < 16; ++rr) cur[rr] = nxt[rr];
  }
  __syncthreads();
  if (wave == 0) {
    double v = lane < m ? acc[lane] : 0.0;
#pragma unroll
    for (int jj = NB - 1; jj >= 0; --jj) {
      if (jj < m) {
        const double xj = readlane_f64(v, jj) * readlane_f64(dinv, jj);
        v = lane == jj ? xj : (lane < jj ? v - dcol[jj] * xj : v);
      }
    }
    if (lane < m) {
      b[k0 + lane] = v;
      publish_value(xwork + k0 + lane, v);
    }
  }
}

}  // namespace

extern "C" int tlh_cholesky_f64(double* A, int64_t n, int64_t ld, int32_t* status, tlh_stream_t stream) {
  if (!A || !status || n < 0 || ld < n) return TLH_EINVAL;
  if (n == 0) return 0;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  // panel 0's row blocks (or, for n <= 64, its diagonal block), then per step k:
  // the next panel's columns, then L_kk beside the next panel and the rest of
  // step k's trailing update
  const int64_t below0 = n > NB ? ceil_div(n - NB, NB) : 0;
  hipLaunchKernelGGL(chol_panel_kernel, dim3((unsigned)(below0 > 0 ? below0 : 1)), dim3(PANEL_THREADS), 0, s, A, n, ld,
                     (int64_t)0, below0 > 0 ? 1 : 0, status);
  TLH_RETURN_IF_LAUNCH_FAILED();
  for (int64_t k = 0; k + NB < n; k += NB) {
    const int64_t below = ceil_div(n - k - NB, NB);  // block rows below panel k (>= 1)
    hipLaunchKernelGGL(chol_trailing_col_kernel, dim3((unsigned)below), dim3(TLH_THREADS), 0, s, A, n, ld, k);
    TLH_RETURN_IF_LAUNCH_FAILED();
    const int last = below == 1;                    // panel k + 64 has no row blocks below
    const int64_t npanel = last ? 1 : below - 1;
    const int64_t rest = (below - 1) * below / 2;   // tiles 1 <= bj <= bi <= below - 1
    hipLaunchKernelGGL(chol_step_kernel, dim3((unsigned)(1 + npanel + rest)), dim3(PANEL_THREADS), 0, s, A, n, ld, k,
                       (int)npanel, last, status);
    TLH_RETURN_IF_LAUNCH_FAILED();
  }
  return 0;
}

extern "C" int tlh_cholesky_solve_f64(const double* L, int64_t n, int64_t ld, double* b, tlh_stream_t stream) {
  if (!L || !b || n < 0 || ld < n || n > 8192) return TLH_EINVAL;
  if (n == 0) return 0;
  const int64_t lds = n * (int64_t)sizeof(double);  // x; the diagonal block and GEMV partials are static
  TLH_RETURN_IF_ERROR(hipFuncSetAttribute((const void*)chol_solve_kernel,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(chol_solve_kernel, dim3(1), dim3(SOLVE_THREADS), (size_t)lds, reinterpret_cast<hipStream_t>(stream),
                     L, n, ld, b);
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

namespace {
// Largest system of the blocks solve: the dense Schur path caps the reduced
// dimension at 8192, and the bounded waits of a 2 x 128-step chain stay far
// inside their poll limit.
constexpr int64_t SOLVE_MAX_N = 8192;

int solve_blocks(const double* L, int64_t n, int64_t ld, double* b, double* work, int32_t* status, hipStream_t s) {
  const int64_t nb = ceil_div(n, NB);
  hipLaunchKernelGGL(solve_work_reset, dim3((unsigned)ceil_div(2 * n, TLH_THREADS)), dim3(TLH_THREADS), 0, s,
                     reinterpret_cast<long long*>(work), 2 * n);
  TLH_RETURN_IF_LAUNCH_FAILED();
  const int spin = wait_limit > 0 ? (int)std::min<int64_t>(wait_limit, SOLVE_SPIN_MAX) : SOLVE_SPIN_MAX;
  hipLaunchKernelGGL(chol_forward_blocks, dim3((unsigned)nb), dim3(SOLVE_WG), 0, s, L, n, ld, b, work, status,
                     spin);
  TLH_RETURN_IF_LAUNCH_FAILED();
  hipLaunchKernelGGL(chol_backward_blocks, dim3((unsigned)nb), dim3(SOLVE_WG), 0, s, L, n, ld, b, work + n, status,
                     spin);
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

// Workspace of tlh_cholesky_solve_blocks_f64_noalloc: one copy per device, held
// until the process ends.  Host code never uses this name as an address; it
// uses the per-device pointer resolved by hipGetSymbolAddress below.
__device__ __align__(256) double scratch[2 * SOLVE_MAX_N];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mutex;
double* scratch_ptr[MAX_DEVICES] = {};

int scratch_for_current_device(double** out) {
  int dev = 0;
  TLH_RETURN_IF_ERROR(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return TLH_EINVAL;
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    TLH_RETURN_IF_ERROR(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return TLH_EINVAL;
    scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = scratch_ptr[dev];
  return 0;
}
}  // namespace

extern "C" int tlh_cholesky_solve_blocks_f64(const double* L, int64_t n, int64_t ld, double* b, double* work,
                                             int32_t* status, tlh_stream_t stream) {
  if (!L || !b || !work || !status || n < 0 || ld < n || n > SOLVE_MAX_N) return TLH_EINVAL;
  if (n == 0) return 0;
  return solve_blocks(L, n, ld, b, work, status, reinterpret_cast<hipStream_t>(stream));
}

extern "C" int tlh_cholesky_solve_blocks_f64_noalloc(const double* L, int64_t n, int64_t ld, double* b,
                                                     int32_t* status, tlh_stream_t stream) {
  if (!L || !b || !status || n < 0 || ld < n || n > SOLVE_MAX_N) return TLH_EINVAL;
  if (n == 0) return 0;
  double* work = nullptr;
  if (int rc = scratch_for_current_device(&work)) return rc;
  return solve_blocks(L, n, ld, b, work, status, reinterpret_cast<hipStream_t>(stream));
}

namespace tlh {
int64_t wait_limit = 0;
}  // namespace tlh

extern "C" int tlh_set_wait_limit(int64_t polls) {
  if (polls < 0) return TLH_EINVAL;
  tlh::wait_limit = polls;
  return 0;
}
