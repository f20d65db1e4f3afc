// Dense float64 Cholesky as one persistent launch over 64x64 tiles.
//
// Replaces torch.linalg.cholesky on the dense Schur matrix (solvers.py:516-518).
// The lower triangle is cut into T x T tiles (T = ceil(n / 64); padding rows and
// columns act as the identity).  One cooperative launch keeps every workgroup
// resident; work moves between workgroups through per-tile states in device
// memory (agent-scope release after a tile is stored, acquire before it is read):
//
//   workgroup 0 ("chain") runs the diagonal: for s = 1 .. T-1 it takes the
//     subdiagonal tile A_{s,s-1} and the diagonal tile A_ss once their owners
//     have applied every update k <= s - 2, solves L_{s,s-1} = A_{s,s-1}
//     L_{s-1,s-1}^-T, subtracts L_{s,s-1} L_{s,s-1}^T from A_ss and factors it;
//     it publishes each L tile and the four inverse 16x16 diagonal blocks of
//     every L_ss (used by the other solves as small matrix products);
//   workgroups 1 .. W ("workers") own the tiles round-robin in column-major
//     order (column j ascending, row i ascending within a column: a
//     topological order of the dependencies, so no workgroup waits on a tile
//     it has not reached).  A worker keeps its tile in registers (MFMA
//     accumulator layout) through every update A_ij -= L_ik L_jk^T (k < j,
//     diagonal tiles k < i - 1), then either hands it to the chain (diagonal
//     and subdiagonal tiles) or solves L_ij = A_ij L_jj^-T with the published
//     inverse blocks.
//
// Products use v_mfma_f64_16x16x4_f64 on operands staged in LDS: lane l
// supplies A[l % 16][k0 + l / 16] and B[k0 + l / 16][l % 16]; accumulator
// element v of lane l is C[4 v + l / 16][l % 16].  Wave w of a workgroup owns
// rows 16 w .. 16 w + 15 of a tile.
//
// The 64x64 diagonal factorisation is blocked by 16: wavefront 0 factors a
// 16x16 block with lane reads (a row per lane) and inverts it (a column per
// lane), the block column below is multiplied by that inverse and the trailing
// blocks are updated with MFMA.
//
// Every wait is bounded; a wait that gives up sets *status |= 4 and an abort
// flag that ends every other wait at once, so the launch always drains (the
// caller then factors again with tlh_cholesky_f64).  A non-positive or
// non-finite pivot sets *status |= 1 (the pivot is replaced by 1 and the
// factorisation continues, as tlh_cholesky_f64 does).
#include <math.h>

#include "common.hpp"

using namespace tlh;

namespace {

constexpr int NB = 64;          // tile size
constexpr int SB = 16;          // diagonal sub-block
constexpr int CT = 256;         // threads per workgroup: four wavefronts
constexpr int LDT = NB + 1;     // LDS row stride of a tile (doubles)
constexpr int LDI = SB + 1;     // LDS row stride of an inverse block
constexpr int TILE_LDS = NB * LDT;
constexpr int INV_LDS = 4 * SB * LDI;
constexpr int INV_GLOBAL = 4 * SB * SB;  // doubles per published set of inverse blocks
constexpr size_t LDS_BYTES = (size_t)(3 * TILE_LDS + 2 * INV_LDS) * sizeof(double);

typedef double v4d __attribute__((ext_vector_type(4)));

struct TileArgs {
  double* A;
  int64_t n;
  int64_t ld;
  int32_t T;
  int32_t nworkers;
  int32_t* state;   // [T (T + 1) / 2]: 0 initial, 1 owner's updates applied (chain tiles), 2 L final
  double* dinv;     // [T][4][16][16] inverse diagonal blocks of each L_ss
  int32_t* status;
  int32_t* abort;
  int64_t spin;
};

__device__ __forceinline__ int64_t tile_id(int i, int j) { return (int64_t)i * (i + 1) / 2 + j; }

__device__ __forceinline__ v4d mfma(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Orders a wavefront's LDS writes before its later LDS reads of other lanes'
// values (LDS executes one wavefront's instructions in order; this keeps the
// compiler from moving the reads up).
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double gload(const TileArgs& a, int64_t R, int64_t C) {
  return (R < a.n && C < a.n) ? a.A[R * a.ld + C] : (R == C ? 1.0 : 0.0);
}

// Tile (ti, tj) of A into LDS S (every thread; all loads in flight before the stores).
__device__ __forceinline__ void load_tile(const TileArgs& a, int ti, int tj, double* S) {
  double g[NB * NB / CT];
#pragma unroll
  for (int m = 0; m < NB * NB / CT; ++m) {
    const int e = (int)threadIdx.x + CT * m;
    g[m] = gload(a, (int64_t)NB * ti + (e >> 6), (int64_t)NB * tj + (e & 63));
  }
#pragma unroll
  for (int m = 0; m < NB * NB / CT; ++m) {
    const int e = (int)threadIdx.x + CT * m;
    S[(e >> 6) * LDT + (e & 63)] = g[m];
  }
}

// LDS tile S into A at tile (ti, tj); lower_only: entries on or below the diagonal.
__device__ __forceinline__ void store_tile(const TileArgs& a, int ti, int tj, const double* S, bool lower_only) {
#pragma unroll
  for (int m = 0; m < NB * NB / CT; ++m) {
    const int e = (int)threadIdx.x + CT * m;
    const int r = e >> 6, c = e & 63;
    const int64_t R = (int64_t)NB * ti + r, C = (int64_t)NB * tj + c;
    if (R < a.n && C < a.n && (!lower_only || c <= r)) a.A[R * a.ld + C] = S[r * LDT + c];
  }
}

// Accumulator layout <-> LDS / global.
__device__ __forceinline__ void acc_load_global(const TileArgs& a, int ti, int tj, v4d (&acc)[4]) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      acc[t][v] = gload(a, (int64_t)NB * ti + 16 * w + 4 * v + (l >> 4), (int64_t)NB * tj + 16 * t + (l & 15));
}

__device__ __forceinline__ void acc_to_lds(const v4d (&acc)[4], double* S) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) S[(16 * w + 4 * v + (l >> 4)) * LDT + 16 * t + (l & 15)] = acc[t][v];
}

// acc -= Li Lj^T over the wave's row block (lower_only: column blocks t <= w).
__device__ __forceinline__ void tile_update(v4d (&acc)[4], const double* Li, const double* Lj, bool lower_only) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const double* ai = Li + (16 * w + (l & 15)) * LDT + (l >> 4);
  const double* bj = Lj + (l & 15) * LDT + (l >> 4);
  for (int k0 = 0; k0 < NB; k0 += 4) {
    const double x = -ai[k0];
#pragma unroll
    for (int t = 0; t < 4; ++t)
      if (!lower_only || t <= w) acc[t] = mfma(x, bj[16 * t * LDT + k0], acc[t]);
  }
}

// X <- X L^-T for the wave's 16 rows of X (LDS), L lower 64x64 (LDS) with the
// inverses Dinv of its 16x16 diagonal blocks (LDS, [4][16][LDI]): column block p
// is (X_p - sum_{q<p} X_q L_pq^T) Dinv_p^T.  No barrier: the rows are the wave's own.
__device__ __forceinline__ void solve_rows(double* X, const double* L, const double* Dinv) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int row_a = 16 * w + (l & 15);  // operand row of this lane
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    v4d c;
#pragma unroll
    for (int v = 0; v < 4; ++v) c[v] = X[(16 * w + 4 * v + (l >> 4)) * LDT + 16 * p + (l & 15)];
    for (int q = 0; q < p; ++q)
#pragma unroll
      for (int k0 = 0; k0 < SB; k0 += 4)
        c = mfma(-X[row_a * LDT + 16 * q + k0 + (l >> 4)], L[(16 * p + (l & 15)) * LDT + 16 * q + k0 + (l >> 4)], c);
#pragma unroll
    for (int v = 0; v < 4; ++v) X[(16 * w + 4 * v + (l >> 4)) * LDT + 16 * p + (l & 15)] = c[v];
    wave_lds_fence();
    v4d d = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k0 = 0; k0 < SB; k0 += 4)
      d = mfma(X[row_a * LDT + 16 * p + k0 + (l >> 4)], Dinv[(SB * p + (l & 15)) * LDI + k0 + (l >> 4)], d);
    wave_lds_fence();
#pragma unroll
    for (int v = 0; v < 4; ++v) X[(16 * w + 4 * v + (l >> 4)) * LDT + 16 * p + (l & 15)] = d[v];
    wave_lds_fence();
  }
}

// Lane c of each 16-lane row of the wavefront, broadcast to that row (DPP
// row_newbcast; c must fold to a constant).
__device__ __forceinline__ int row_bcast_i32(int x, int c) {
  switch (c) {
    case 0: return __builtin_amdgcn_update_dpp(0, x, 0x150, 0xf, 0xf, false);
    case 1: return __builtin_amdgcn_update_dpp(0, x, 0x151, 0xf, 0xf, false);
    case 2: return __builtin_amdgcn_update_dpp(0, x, 0x152, 0xf, 0xf, false);
    case 3: return __builtin_amdgcn_update_dpp(0, x, 0x153, 0xf, 0xf, false);
    case 4: return __builtin_amdgcn_update_dpp(0, x, 0x154, 0xf, 0xf, false);
    case 5: return __builtin_amdgcn_update_dpp(0, x, 0x155, 0xf, 0xf, false);
    case 6: return __builtin_amdgcn_update_dpp(0, x, 0x156, 0xf, 0xf, false);
    case 7: return __builtin_amdgcn_update_dpp(0, x, 0x157, 0xf, 0xf, false);
    case 8: return __builtin_amdgcn_update_dpp(0, x, 0x158, 0xf, 0xf, false);
    case 9: return __builtin_amdgcn_update_dpp(0, x, 0x159, 0xf, 0xf, false);
    case 10: return __builtin_amdgcn_update_dpp(0, x, 0x15a, 0xf, 0xf, false);
    case 11: return __builtin_amdgcn_update_dpp(0, x, 0x15b, 0xf, 0xf, false);
    case 12: return __builtin_amdgcn_update_dpp(0, x, 0x15c, 0xf, 0xf, false);
    case 13: return __builtin_amdgcn_update_dpp(0, x, 0x15d, 0xf, 0xf, false);
    case 14: return __builtin_amdgcn_update_dpp(0, x, 0x15e, 0xf, 0xf, false);
    default: return __builtin_amdgcn_update_dpp(0, x, 0x15f, 0xf, 0xf, false);
  }
}

__device__ __forceinline__ double row_bcast(double v, int c) {
  const long long b = __double_as_longlong(v);
  const int lo = row_bcast_i32((int)(b & 0xffffffffll), c);
  const int hi = row_bcast_i32((int)(b >> 32), c);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Cholesky of the 64x64 tile D (LDS, lower triangle) in place; the inverses of
// its four 16x16 diagonal blocks go to Dinv (LDS).  Returns false (every
// thread) on a non-positive or non-finite pivot.  Blocked by 8 columns; per
// 8-column block p (rows and columns o = 8 p .. o + 7):
//   A. wavefront 0 factors block (p, p), lane r of each 8-lane group holding row
//      r (the pivot column reaches the other rows by DPP row broadcasts; entries
//      above the diagonal are updated too and never read); its rows go to LDS (P);
//   B. wavefront 1 solves the rows below, x L_pp^T = a, a row per lane (row c of
//      L_pp read from P), while wavefront 2 inverts L_pp (a column per lane);
//   C. the trailing part (rows and columns >= o + 8) takes the rank-8 update with
//      MFMA on the 16x16 tiles that cover it (rows of a tile above o + 8 get a
//      zero operand); the tile holding block (p+1, p+1) goes to wavefront 0,
//      which goes on to step A of p+1 without waiting for the other tiles.
// Finally each wavefront q forms the inverse of the 16x16 diagonal block q from
// its two 8x8 inverses: [[A^-1, 0], [-C^-1 B A^-1, C^-1]].
constexpr int FB = 8;  // factorisation block

// 1 / sqrt(d) for a positive finite d: the hardware estimate (about 2^-29 relative
// error) refined by one Newton step, y (3 - d y^2) / 2, to float64 accuracy.
__device__ __forceinline__ double rsqrt_refined(double d) {
  const double y = __builtin_amdgcn_rsq(d);
  return y * fma(-0.5 * d * y, y, 1.5);
}

__device__ __forceinline__ bool factor_tile(double* D, double* Dinv, int* flag) {
  __shared__ double P[FB][FB];      // L_pp rows
  __shared__ double idiag[FB];      // 1 / L[r][r] of the current block
  __shared__ double I8[NB / FB][FB][FB + 1];  // inverses of the 8x8 diagonal blocks
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  bool ok = true;
#pragma unroll 1
  for (int p = 0; p < NB / FB; ++p) {
    const int o = FB * p;
    if (w == 0) {
      const int r = l & 7;
      double v[FB];
#pragma unroll
      for (int c = 0; c < FB; ++c) v[c] = D[(o + r) * LDT + o + c];
      double invd = 1.0;
#pragma unroll
      for (int j = 0; j < FB; ++j) {
        const double dj = row_bcast(v[j], j);
        const bool good = dj > 0.0 && isfinite(dj);
        if (!good) ok = false;
        const double inv = rsqrt_refined(good ? dj : 1.0);
        const double piv = (good ? dj : 1.0) * inv;
        v[j] = r == j ? piv : v[j] * inv;
        invd = r == j ? inv : invd;
#pragma unroll
        for (int c = j + 1; c < FB; ++c) v[c] = fma(-v[j], row_bcast(v[j], c), v[c]);
      }
      if (l < FB) {
#pragma unroll
        for (int c = 0; c < FB; ++c) {
          P[l][c] = v[c];
          if (c <= l) D[(o + l) * LDT + o + c] = v[c];
        }
        idiag[l] = invd;
      }
    }
    __syncthreads();
    if (w == 1 && o + FB + l < NB) {
      // row o + 8 + l below the block: x L^T = a, column by column
      const int R = o + FB + l;
      double x[FB];
#pragma unroll
      for (int c = 0; c < FB; ++c) x[c] = D[R * LDT + o + c];
#pragma unroll
      for (int c = 0; c < FB; ++c) {
        double t = x[c];
#pragma unroll
        for (int m = 0; m < c; ++m) t = fma(-x[m], P[c][m], t);
        x[c] = t * idiag[c];
      }
#pragma unroll
      for (int c = 0; c < FB; ++c) D[R * LDT + o + c] = x[c];
    } else if (w == 2 && l < FB) {
      // column l of L_pp^-1: x[c] = -(sum_{m<c} L[c][m] x[m]) / L[c][c], x[l] = 1 / L[l][l]
      double x[FB];
#pragma unroll
      for (int c = 0; c < FB; ++c) {
        double t = c == l ? 1.0 : 0.0;
#pragma unroll
        for (int m = 0; m < c; ++m) t = fma(-x[m], P[c][m], t);
        x[c] = t * idiag[c];
      }
#pragma unroll
      for (int c = 0; c < FB; ++c) I8[p][c][l] = x[c];
    }
    __syncthreads();
    if (p < NB / FB - 1) {
      // trailing update of the rows and columns >= e = o + 8 on the 16x16 tiles (a2, b2),
      // t0 <= b2 <= a2 < 4, t0 = e / 16; operand rows below e are zero
      const int e = o + FB, t0 = e >> 4;
      int job = 0;
#pragma unroll 1
      for (int a2 = t0; a2 < 4; ++a2) {
#pragma unroll 1
        for (int b2 = t0; b2 <= a2; ++b2) {
          const bool first = a2 == t0 && b2 == t0;
          const int owner = first ? 0 : 1 + (job++ % 3);
          if (owner != w) continue;
          const int ra = SB * a2 + (l & 15), rb = SB * b2 + (l & 15);
          v4d c;
#pragma unroll
          for (int v = 0; v < 4; ++v) c[v] = D[(SB * a2 + 4 * v + (l >> 4)) * LDT + SB * b2 + (l & 15)];
#pragma unroll
          for (int k0 = 0; k0 < FB; k0 += 4) {
            const double xa = ra >= e ? D[ra * LDT + o + k0 + (l >> 4)] : 0.0;
            const double xb = rb >= e ? D[rb * LDT + o + k0 + (l >> 4)] : 0.0;
            c = mfma(-xa, xb, c);
          }
#pragma unroll
          for (int v = 0; v < 4; ++v) D[(SB * a2 + 4 * v + (l >> 4)) * LDT + SB * b2 + (l & 15)] = c[v];
        }
      }
      // No barrier: wavefront 0 updated the tile holding block (p+1, p+1), which is all
      // that step A of p+1 reads (and only wavefront 0 writes there); the other
      // wavefronts' tiles are read after the barrier that ends that step A.
      wave_lds_fence();
    }
  }
  __syncthreads();
  // 16x16 inverses: wavefront q, lane (i, j) = (l / 8, l % 8) forms entry (8 + i, j) of
  // block q's inverse, -(C^-1 B A^-1)[i][j], with A = L_{2q}, B = L rows 16q + 8.., columns
  // 16q.., C = L_{2q+1}; the two diagonal 8x8 inverses are copied as they are
  {
    const int q = w;
    const int i = l >> 3, j = l & 7;
    const int r0 = SB * q;
    double bt = 0.0;  // (B A^-1)[i][j]
#pragma unroll
    for (int m = 0; m < FB; ++m) bt = fma(D[(r0 + FB + i) * LDT + r0 + m], I8[2 * q][m][j], bt);
    // -(C^-1 (B A^-1))[i][j] needs row i of C^-1 and column j of B A^-1: exchange through LDS
    __shared__ double BT[4][FB][FB + 1];
    BT[q][i][j] = bt;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int m = 0; m < FB; ++m) t = fma(I8[2 * q + 1][i][m], BT[q][m][j], t);
    double* inv = Dinv + (SB * q) * LDI;
    inv[(FB + i) * LDI + j] = -t;
    inv[i * LDI + j] = I8[2 * q][i][j];
    inv[(FB + i) * LDI + FB + j] = I8[2 * q + 1][i][j];
    inv[i * LDI + FB + j] = 0.0;
  }
  if (threadIdx.x == 0) *flag = 0;
  __syncthreads();
  if (!ok) *flag = 1;  // wavefront 0 only
  __syncthreads();
  const bool all_ok = *flag == 0;
  __syncthreads();
  return all_ok;
}

// Waits until state[t0] >= v0 and (t1 < 0 or state[t1] >= v1); false when the
// wait gave up or another workgroup aborted.  After it returns true, the
// workgroup's loads see the data published with those states.
__device__ __forceinline__ bool wait_states(const TileArgs& a, int64_t t0, int v0, int64_t t1, int v1, int* flag) {
  if (threadIdx.x == 0) {
    int ok = 1;
    for (int64_t it = 0;; ++it) {
      const int s0 = __hip_atomic_load(&a.state[t0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const int s1 = t1 < 0 ? v1 : __hip_atomic_load(&a.state[t1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (s0 >= v0 && s1 >= v1) break;
      if (__hip_atomic_load(a.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
        ok = 0;
        break;
      }
      if (it >= a.spin) {
        __hip_atomic_fetch_or(a.status, 4, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(a.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    *flag = ok;
  }
  __syncthreads();
  const bool ok = *flag != 0;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  __syncthreads();
  return ok;
}

// Every thread's stores of the tile are made visible at agent scope, then
// state[t] = v.
__device__ __forceinline__ void publish(const TileArgs& a, int64_t t, int v) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(&a.state[t], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void store_dinv(const TileArgs& a, int s, const double* Dinv) {
  for (int e = threadIdx.x; e < INV_GLOBAL; e += CT) {
    const int blk = e >> 8, r = (e >> 4) & 15, c = e & 15;
    a.dinv[(int64_t)s * INV_GLOBAL + e] = Dinv[(SB * blk + r) * LDI + c];
  }
}

__device__ __forceinline__ void load_dinv(const TileArgs& a, int s, double* Dinv) {
  double g[INV_GLOBAL / CT];
#pragma unroll
  for (int m = 0; m < INV_GLOBAL / CT; ++m) g[m] = a.dinv[(int64_t)s * INV_GLOBAL + threadIdx.x + CT * m];
#pragma unroll
  for (int m = 0; m < INV_GLOBAL / CT; ++m) {
    const int e = threadIdx.x + CT * m;
    Dinv[(SB * (e >> 8) + ((e >> 4) & 15)) * LDI + (e & 15)] = g[m];
  }
}

// D -= X X^T on the lower 16x16 blocks (a2, b2), round-robin over the waves.
__device__ __forceinline__ void chain_syrk(double* D, const double* X) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  int job = 0;
#pragma unroll 1
  for (int a2 = 0; a2 < 4; ++a2) {
#pragma unroll 1
    for (int b2 = 0; b2 <= a2; ++b2, ++job) {
      if ((job & 3) != w) continue;
      v4d c;
#pragma unroll
      for (int v = 0; v < 4; ++v) c[v] = D[(SB * a2 + 4 * v + (l >> 4)) * LDT + SB * b2 + (l & 15)];
      for (int k0 = 0; k0 < NB; k0 += 4)
        c = mfma(-X[(SB * a2 + (l & 15)) * LDT + k0 + (l >> 4)], X[(SB * b2 + (l & 15)) * LDT + k0 + (l >> 4)], c);
#pragma unroll
      for (int v = 0; v < 4; ++v) D[(SB * a2 + 4 * v + (l >> 4)) * LDT + SB * b2 + (l & 15)] = c[v];
    }
  }
}

// The diagonal.  Step s: X = A_{s,s-1} solved against L_{s-1,s-1} and published
// (its stores overlap the update of A_ss), A_ss -= X X^T, then the factor of A_ss;
// the next step's two tiles are loaded while L_ss is being stored.
__device__ void chain(const TileArgs& a, double* lds, int* flag) {
  double* Lp = lds;                  // L_{s-1,s-1}
  double* D = lds + TILE_LDS;        // A_ss -> L_ss
  double* X = lds + 2 * TILE_LDS;    // A_{s,s-1} -> L_{s,s-1}
  double* Ip = lds + 3 * TILE_LDS;   // inverse blocks of L_{s-1,s-1}
  double* In = Ip + INV_LDS;         // inverse blocks of L_ss
  load_tile(a, 0, 0, D);
  __syncthreads();
  if (!factor_tile(D, In, flag) && threadIdx.x == 0)
    __hip_atomic_fetch_or(a.status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  store_tile(a, 0, 0, D, true);
  store_dinv(a, 0, In);
  if (a.T > 1) {  // step 1's tiles need no owner updates
    load_tile(a, 1, 0, X);
    load_tile(a, 1, 1, Lp);
  }
  publish(a, tile_id(0, 0), 2);
#pragma unroll 1
  for (int s = 1; s < a.T; ++s) {
    { double* t = Lp; Lp = D; D = t; }  // D holds A_ss (loaded at the end of step s - 1)
    { double* t = Ip; Ip = In; In = t; }
    __syncthreads();
    solve_rows(X, Lp, Ip);
    __syncthreads();
    store_tile(a, s, s - 1, X, false);
    chain_syrk(D, X);
    publish(a, tile_id(s, s - 1), 2);
    if (!factor_tile(D, In, flag) && threadIdx.x == 0)
      __hip_atomic_fetch_or(a.status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    store_tile(a, s, s, D, true);
    store_dinv(a, s, In);
    if (s + 1 < a.T) {
      // the owners of (s+1, s) and (s+1, s+1) apply the updates k <= s - 1
      if (!wait_states(a, tile_id(s + 1, s), 1, tile_id(s + 1, s + 1), 1, flag)) return;
      load_tile(a, s + 1, s, X);
      load_tile(a, s + 1, s + 1, Lp);  // L_{s-1,s-1} is no longer needed
    }
    publish(a, tile_id(s, s), 2);
  }
}

// Tile seq (column-major, row ascending inside a column) -> (i, j).
__device__ __forceinline__ void tile_of(int T, int64_t seq, int& i, int& j) {
  int col = 0;
  int64_t start = 0;
  while (start + (T - col) <= seq) {
    start += T - col;
    ++col;
  }
  j = col;
  i = col + (int)(seq - start);
}

__device__ void worker(const TileArgs& a, int wid, double* lds, int* flag) {
  double* Li = lds;
  double* Lj = lds + TILE_LDS;
  double* X = lds + 2 * TILE_LDS;
  double* Inv = lds + 3 * TILE_LDS;
  const int64_t ntiles = (int64_t)a.T * (a.T + 1) / 2;
#pragma unroll 1
  for (int64_t seq = wid; seq < ntiles; seq += a.nworkers) {
    int i, j;
    tile_of(a.T, seq, i, j);
    // tiles of the chain with no update for their owner (rows 0 and 1): the chain reads A
    if (i <= 1 && j + 1 >= i) continue;
    v4d acc[4];
    acc_load_global(a, i, j, acc);
    const bool diag = i == j;
    const int kend = diag ? i - 1 : j;  // updates k < kend (the chain applies the diagonal's last)
#pragma unroll 1
    for (int k = 0; k < kend; ++k) {
      if (!wait_states(a, tile_id(i, k), 2, diag ? -1 : tile_id(j, k), 2, flag)) return;
      load_tile(a, i, k, Li);
      if (!diag) load_tile(a, j, k, Lj);
      __syncthreads();
      tile_update(acc, Li, diag ? Li : Lj, diag);
      __syncthreads();
    }
    if (diag || i == j + 1) {  // handed to the chain
      acc_to_lds(acc, X);
      __syncthreads();
      store_tile(a, i, j, X, diag);
      publish(a, tile_id(i, j), 1);
      continue;
    }
    if (!wait_states(a, tile_id(j, j), 2, -1, 0, flag)) return;
    load_tile(a, j, j, Li);
    load_dinv(a, j, Inv);
    acc_to_lds(acc, X);
    __syncthreads();
    solve_rows(X, Li, Inv);
    __syncthreads();
    store_tile(a, i, j, X, false);
    publish(a, tile_id(i, j), 2);
  }
}

__global__ void __launch_bounds__(CT) chol_tiles_kernel(TileArgs a) {
  extern __shared__ double lds[];
  __shared__ int flag;
  if (blockIdx.x == 0)
    chain(a, lds, &flag);
  else
    worker(a, (int)blockIdx.x - 1, lds, &flag);
}

__global__ void __launch_bounds__(CT) chol_tiles_reset(int32_t* state, int64_t count, int32_t* abort) {
  const int64_t i = (int64_t)blockIdx.x * CT + threadIdx.x;
  if (i < count) state[i] = 0;
  if (i == 0) *abort = 0;
}

int64_t tiles_of(int64_t n) { return ceil_div(n, NB); }

int64_t state_bytes(int64_t T) { return ceil_div(T * (T + 1) / 2 * (int64_t)sizeof(int32_t) + 256, 256) * 256; }

}  // namespace

extern "C" int64_t tlh_cholesky_tiles_workspace(int64_t n) {
  if (n < 0) return -1;
  const int64_t T = tiles_of(n);
  return state_bytes(T) + T * INV_GLOBAL * (int64_t)sizeof(double);
}

extern "C" int tlh_cholesky_tiles_f64(double* A, int64_t n, int64_t ld, void* work, int64_t work_bytes,
                                      int32_t* status, tlh_stream_t stream) {
  if (!A || !status || n < 0 || ld < n || n > 8192) return TLH_EINVAL;
  if (n == 0) return 0;
  if (!work || work_bytes < tlh_cholesky_tiles_workspace(n)) return TLH_EINVAL;
  const int64_t T = tiles_of(n);
  const int64_t ntiles = T * (T + 1) / 2;
  char* base = static_cast<char*>(work);
  TileArgs a;
  a.A = A;
  a.n = n;
  a.ld = ld;
  a.T = (int32_t)T;
  a.state = reinterpret_cast<int32_t*>(base);
  a.abort = a.state + ntiles;
  a.dinv = reinterpret_cast<double*>(base + state_bytes(T));
  a.status = status;
  a.spin = wait_limit > 0 ? std::min<int64_t>(wait_limit, (int64_t)1 << 18) : (int64_t)1 << 18;
  TLH_RETURN_IF_ERROR(hipFuncSetAttribute((const void*)chol_tiles_kernel,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)LDS_BYTES));
  int dev = 0, cus = 0, per_cu = 0;
  TLH_RETURN_IF_ERROR(hipGetDevice(&dev));
  TLH_RETURN_IF_ERROR(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  TLH_RETURN_IF_ERROR(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)chol_tiles_kernel, CT,
                                                                   LDS_BYTES));
  const int64_t resident = (int64_t)per_cu * cus;
  if (resident < 2 && T > 1) return TLH_ELDS;
  const int64_t grid = T == 1 ? 1 : (1 + ntiles < resident ? 1 + ntiles : resident);
  a.nworkers = (int32_t)(grid - 1);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(chol_tiles_reset, dim3((unsigned)ceil_div(ntiles, CT)), dim3(CT), 0, s, a.state, ntiles,
                     a.abort);
  TLH_RETURN_IF_LAUNCH_FAILED();
  void* args[] = {&a};
  TLH_RETURN_IF_ERROR(hipLaunchCooperativeKernel((const void*)chol_tiles_kernel, dim3((unsigned)grid), dim3(CT), args,
                                                 LDS_BYTES, s));
  return 0;
}
