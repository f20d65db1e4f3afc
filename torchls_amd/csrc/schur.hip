// Implicit Schur complement of one independent variable type (points) on gfx950.
//
// Replaces DampedSchurSystem (schur.py:263-397) without storing the cross
// blocks W or any CSR matrix.  With e the eliminated type, C_p = D_p + lambda I
// and its float64 inverse precomputed per variable:
//
//   matvec   S x  = J_c^T J_c x + lambda x - W C^-1 W^T x            (schur.py:303-328)
//   rhs      b_c' = b_c - W C^-1 b_p                                 (schur.py:285-290)
//   backsub  x_p  = C^-1 (b_p - W^T x_c)                             (schur.py:292-301)
//   precond  M_c  = D_c + lambda I - sum_edges W_e C_p^-1 W_e^T      (schur.py:353-378)
//   dense    M    = blockdiag(D_c + lambda I) - sum_p W_p C_p^-1 W_p^T (schur.py:380-397)
//
// The W-group (the residual group holding a retained slot and the eliminated
// slot, e.g. reprojection) is sorted by the eliminated variable with every
// variable's observations inside one tile, so one pass over its Jacobian
// evaluates W C^-1 W^T x: per observation u = J_c x_c; per point (one
// segment owner) t_p = sum J_p^T u and s_p = C_p^-1 t_p in float64; per
// observation q = u - J_p s_p and the camera contribution J_c^T q, reduced
// deterministically exactly like the normal matvec (DENSE LDS accumulator or
// SPARSE partials + finish).  Groups holding only retained slots are added by
// the normal tile kernels of csrc/normal.hip.
#include <algorithm>

#include "common.hpp"

using namespace tlh;

namespace {

enum SchurMode { SCHUR_MATVEC = 0, SCHUR_RHS = 1, SCHUR_BACKSUB = 2 };

template <typename Real>
struct SchurArgs {
  const Real* packed;
  int64_t stride;
  const int32_t* tile_start;
  int32_t ntiles;
  int32_t tiles_per_wg;
  const int32_t* idx_c;
  const int32_t* idx_p;
  const uint16_t* rank_c;
  const int32_t* seg_ptr_c;
  const int32_t* seg_start_c;
  const int32_t* seg_target_c;
  int32_t mode_c;
  int32_t dense_block_c;
  int32_t count_c;
  Real* part_c;
  const int32_t* seg_ptr_p;
  const int32_t* seg_start_p;
  const Real* xc;      // retained input (x_c), offset to the retained type
  const Real* bp;      // eliminated right-hand side, offset to the eliminated type
  Real* outp;          // back-substituted eliminated update, offset to the eliminated type
  const double* cinv;  // [count_p][DP*(DP+1)/2] upper triangle of (D_p + lambda I)^-1
};

template <int D>
__device__ __forceinline__ double sym(const double* m, int i, int j) {
  const int a = i < j ? i : j, b = i < j ? j : i;
  return m[a * D - a * (a - 1) / 2 + (b - a)];
}

// One observation per thread.  The point phase is a segmented sum over the
// tile (sorted by point): each wavefront scans its 64 observations with DPP
// (flag = first observation of a point), lane 63 publishes the wave's tail,
// and a lane whose point began in an earlier wavefront adds the tails of the
// waves before it, nearest first, up to the wave where the point starts.  The
// point's last observation then holds t_p = sum J_p^T u and forms
// s_p = C_p^-1 (...) in float64; the scan also carries the point's first
// position, where s_p is published for the point's other observations.  XL:
// the retained x slice is copied into LDS once per workgroup (after the
// accumulator), so the per-observation camera gathers are LDS reads.
template <typename Real, int R, int DC, int DP, bool CFIRST, int ITEMS, int MODE, int NTH, bool XL>
__global__ void __launch_bounds__(NTH) schur_tile_kernel(SchurArgs<Real> a) {
  constexpr int OC = CFIRST ? 0 : R * DP;
  constexpr int OP = CFIRST ? R * DC : 0;
  constexpr int T = NTH * ITEMS;
  constexpr int NM = DP * (DP + 1) / 2;
  constexpr int NW = NTH / 64;
  static_assert(ITEMS == 1, "one observation per thread");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* tail = reinterpret_cast<double*>(smem_raw);  // [NW][DP + 1]
  int* tail_flag = reinterpret_cast<int*>(tail + NW * (DP + 1));
  Real* stage = reinterpret_cast<Real*>(tail_flag + NW * 2);
  Real* sstage = stage + T * DC;
  Real* acc = sstage + T * DP;
  Real* xl = acc + (int64_t)a.count_c * DC;
  const bool dense = MODE != SCHUR_BACKSUB && a.mode_c == TLH_REDUCE_DENSE;
  if (dense)
    for (int i = threadIdx.x; i < a.count_c * DC; i += NTH) acc[i] = Real(0);
  if constexpr (XL) {
    for (int i = threadIdx.x; i < a.count_c * DC; i += NTH) xl[i] = a.xc[i];
  }
  if (dense || XL) __syncthreads();

  const int pos = (int)threadIdx.x;
  const int wave = pos >> 6, lane = pos & 63;
  const int t0 = blockIdx.x * a.tiles_per_wg;
  const int t1 = min(t0 + a.tiles_per_wg, a.ntiles);
  for (int t = t0; t < t1; ++t) {
    const int start = a.tile_start[t];
    const int len = a.tile_start[t + 1] - start;
    const bool valid = pos < len;
    const int64_t n = start + min(pos, len - 1);
    Real jc[R][DC], jp[R][DP];
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
      for (int c = 0; c < DC; ++c) jc[r][c] = a.packed[(int64_t)(OC + r * DC + c) * a.stride + n];
#pragma unroll
      for (int c = 0; c < DP; ++c) jp[r][c] = a.packed[(int64_t)(OP + r * DP + c) * a.stride + n];
    }
    const int rk = a.rank_c ? (int)a.rank_c[n] : min(pos, len - 1);
    // this thread's camera segment (a tile of at most NTH observations has at most NTH),
    // loaded here so the reduction after the barriers does not wait on memory
    int cseg = -1, clo = 0, chi = 0;
    if (MODE != SCHUR_BACKSUB) {
      const int cb = a.seg_ptr_c[t], ce = a.seg_ptr_c[t + 1];
      if (cb + pos < ce) {
        cseg = a.seg_target_c[cb + pos];
        clo = a.seg_start_c[cb + pos];
        chi = (cb + pos + 1 < ce) ? a.seg_start_c[cb + pos + 1] : len;
      }
    }
    // point boundaries; a lane past the tile's end is a segment of its own
    const int vp = a.idx_p[n];
    const bool first = !valid || pos == 0 || a.idx_p[n - 1] != vp;
    const bool last = valid && (pos == len - 1 || a.idx_p[n + 1] != vp);
    double mk[NM], bpv[DP];
    if (last) {
#pragma unroll
      for (int e = 0; e < NM; ++e) mk[e] = a.cinv[(int64_t)vp * NM + e];
      if (MODE != SCHUR_MATVEC) {
#pragma unroll
        for (int c = 0; c < DP; ++c) bpv[c] = (double)a.bp[(int64_t)vp * DP + c];
      }
    }
    Real u[R];
    if (MODE != SCHUR_RHS) {
      const int ic = a.idx_c[n];
      const Real* x = XL ? xl + ic * DC : a.xc + (int64_t)ic * DC;
      Real xv[DC];
#pragma unroll
      for (int c = 0; c < DC; ++c) xv[c] = x[c];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        Real s = Real(0);
#pragma unroll
        for (int c = 0; c < DC; ++c) s += jc[r][c] * xv[c];
        u[r] = s;
      }
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) u[r] = Real(0);
    }
    // w = (J_p^T u, first position of the point), summed per point in float64
    double w[DP + 1];
#pragma unroll
    for (int c = 0; c < DP; ++c) {
      Real s = Real(0);
      if (MODE != SCHUR_RHS) {
#pragma unroll
        for (int r = 0; r < R; ++r) s += jp[r][c] * u[r];
      }
      w[c] = valid ? (double)s : 0.0;
    }
    w[DP] = first ? (double)pos : 0.0;
    int flag = first ? 1 : 0;
    wave_segmented_scan<double, DP + 1>(flag, w);
    if (lane == 63) {
#pragma unroll
      for (int c = 0; c <= DP; ++c) tail[wave * (DP + 1) + c] = w[c];
      tail_flag[wave] = flag;
    }
    __syncthreads();
    if (!flag) {
      // the point began in an earlier wavefront (wave 0 always holds a start)
      for (int k = wave - 1; k >= 0; --k) {
#pragma unroll
        for (int c = 0; c <= DP; ++c) w[c] += tail[k * (DP + 1) + c];
        if (tail_flag[k]) break;
      }
    }
    const int fp = (int)w[DP];
    if (last) {
      double tp[DP];
#pragma unroll
      for (int c = 0; c < DP; ++c) {
        if (MODE == SCHUR_MATVEC) tp[c] = w[c];
        if (MODE == SCHUR_RHS) tp[c] = bpv[c];
        if (MODE == SCHUR_BACKSUB) tp[c] = bpv[c] - w[c];  // b_p - W^T x
      }
      double sp[DP];
#pragma unroll
      for (int i = 0; i < DP; ++i) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < DP; ++j) s += sym<DP>(mk, i, j) * tp[j];
        sp[i] = s;
      }
      if (MODE == SCHUR_BACKSUB) {
#pragma unroll
        for (int c = 0; c < DP; ++c) a.outp[(int64_t)vp * DP + c] = (Real)sp[c];
      } else {
#pragma unroll
        for (int c = 0; c < DP; ++c) sstage[fp * DP + c] = (Real)sp[c];
      }
    }
    if (MODE == SCHUR_BACKSUB) {
      __syncthreads();  // the tails are rewritten by the next tile
      continue;
    }
    __syncthreads();
    if (valid) {
      Real sv[DP];
#pragma unroll
      for (int c = 0; c < DP; ++c) sv[c] = sstage[fp * DP + c];
      Real q[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        Real s = u[r];
#pragma unroll
        for (int c = 0; c < DP; ++c) s -= jp[r][c] * sv[c];
        q[r] = s;
      }
#pragma unroll
      for (int c = 0; c < DC; ++c) {
        Real s = Real(0);
#pragma unroll
        for (int r = 0; r < R; ++r) s += jc[r][c] * q[r];
        stage[rk * DC + c] = s;
      }
    }
    __syncthreads();
    if (cseg >= 0) {
      Real sum[DC];
#pragma unroll
      for (int c = 0; c < DC; ++c) sum[c] = stage[clo * DC + c];
      for (int i = clo + 1; i < chi; ++i)
#pragma unroll
        for (int c = 0; c < DC; ++c) sum[c] += stage[i * DC + c];
      const int target = cseg;
      if (dense) {
#pragma unroll
        for (int c = 0; c < DC; ++c) acc[target * DC + c] += sum[c];
      } else {
#pragma unroll
        for (int c = 0; c < DC; ++c) a.part_c[(int64_t)target * DC + c] = sum[c];
      }
    }
  }
  if (dense) {
    __syncthreads();
    const int64_t nn = (int64_t)a.count_c * DC;
    Real* out = a.part_c + (int64_t)(a.dense_block_c + blockIdx.x) * nn;
    for (int64_t i = threadIdx.x; i < nn; i += NTH) out[i] = acc[i];
  }
}

// ------------------------------------------- pipelined matvec (dense cameras)
// The matvec of the tile kernel above (same arithmetic, same order: identical
// results) for a DENSE camera slot, restructured so that memory latency is
// hidden: the next tile's block, id, segment and inverse loads are issued before
// the current tile is processed, into a second register buffer.  Every load is
// issued on every path (indices clamped, values selected afterwards) and the tile
// loop stores nothing to global memory, so the compiler's load waits stay exact
// and two tiles of loads are in flight.  The workgroup's tile starts and camera
// segment ranges are read into LDS once.
template <typename Real, int R, int DC, int DP, bool XL>
struct SchurTileRegs {
  static constexpr int NM = DP * (DP + 1) / 2;
  Real jc[R][DC], jp[R][DP];
  int vp, vprev, vnext, ic, rk, cseg, clo, chi;
  double mk[NM];
};

template <typename Real, int R, int DC, int DP, bool CFIRST, int NTH, bool XL>
__device__ __forceinline__ void schur_load_tile(const SchurArgs<Real>& a, const int* tstart, const int* tseg, int t,
                                                int j, int64_t count, SchurTileRegs<Real, R, DC, DP, XL>& d) {
  constexpr int OC = CFIRST ? 0 : R * DP;
  constexpr int OP = CFIRST ? R * DC : 0;
  constexpr int NM = DP * (DP + 1) / 2;
  const int pos = (int)threadIdx.x;
  const int start = tstart[j], len = tstart[j + 1] - start;
  const uint32_t k = (uint32_t)max(0, min(pos, len - 1));
  const Real* tile = a.packed + start;
#pragma unroll
  for (int r = 0; r < R; ++r) {
#pragma unroll
    for (int c = 0; c < DC; ++c) d.jc[r][c] = tile[(int64_t)(OC + r * DC + c) * a.stride + k];
#pragma unroll
    for (int c = 0; c < DP; ++c) d.jp[r][c] = tile[(int64_t)(OP + r * DP + c) * a.stride + k];
  }
  const int64_t n = (int64_t)start + k;
  d.vp = a.idx_p[n];
  d.vprev = a.idx_p[n > 0 ? n - 1 : 0];
  d.vnext = a.idx_p[n + 1 < count ? n + 1 : n];
  d.ic = a.idx_c[n];
  const bool ranked = a.rank_c != nullptr;
  const int rv = (ranked ? a.rank_c : reinterpret_cast<const uint16_t*>(a.idx_c))[n];
  d.rk = ranked ? rv : (int)k;
  // this thread's camera segment (a tile of at most NTH observations has at most NTH)
  const int cb = tseg[j], ce = tseg[j + 1];
  const int si = max(cb, min(cb + pos, ce - 1));
  const int tg = a.seg_target_c[si], lo = a.seg_start_c[si], hi = a.seg_start_c[min(si + 1, ce - 1)];
  d.cseg = cb + pos < ce ? tg : -1;
  d.clo = lo;
  d.chi = cb + pos + 1 < ce ? hi : len;
#pragma unroll
  for (int e = 0; e < NM; ++e) d.mk[e] = a.cinv[(int64_t)d.vp * NM + e];
  (void)t;
}

template <typename Real, int R, int DC, int DP, bool CFIRST, int NTH, bool XL>
__global__ void __launch_bounds__(NTH) schur_matvec_pipe_kernel(SchurArgs<Real> a, int64_t count) {
  constexpr int T = NTH;
  constexpr int NW = NTH / 64;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* tail = reinterpret_cast<double*>(smem_raw);  // [NW][DP + 1]
  int* tail_flag = reinterpret_cast<int*>(tail + NW * (DP + 1));
  Real* stage = reinterpret_cast<Real*>(tail_flag + NW * 2);
  Real* sstage = stage + T * DC;
  Real* acc = sstage + T * DP;
  Real* xl = acc + (int64_t)a.count_c * DC;
  int* tstart = reinterpret_cast<int*>(XL ? xl + (int64_t)a.count_c * DC : xl);  // [tiles_per_wg + 2]
  int* tseg = tstart + a.tiles_per_wg + 2;                                        // [tiles_per_wg + 2]
  const int pos = (int)threadIdx.x;
  const int wave = pos >> 6, lane = pos & 63;
  const int t0 = blockIdx.x * a.tiles_per_wg;
  const int t1 = min(t0 + a.tiles_per_wg, a.ntiles);
  for (int i = threadIdx.x; i < a.count_c * DC; i += NTH) acc[i] = Real(0);
  if constexpr (XL) {
    for (int i = threadIdx.x; i < a.count_c * DC; i += NTH) xl[i] = a.xc[i];
  }
  // tile starts and camera segment ranges of tiles t0 .. t1 (entry t1 - t0 + 1
  // repeats the last tile's end: the padding half of an odd tile count)
  for (int i = threadIdx.x; i <= t1 - t0 + 1; i += NTH) {
    const int tt = min(t0 + i, t1);
    tstart[i] = a.tile_start[tt];
    tseg[i] = a.seg_ptr_c[tt];
  }
  __syncthreads();

  auto process = [&](SchurTileRegs<Real, R, DC, DP, XL>& cur, SchurTileRegs<Real, R, DC, DP, XL>& nxt, int t) {
    // next tile's loads first (a re-read of this one, discarded, after the last tile)
    const int jn = t + 1 < t1 ? t + 1 - t0 : t - t0;
    schur_load_tile<Real, R, DC, DP, CFIRST, NTH, XL>(a, tstart, tseg, t + 1, min(jn, t1 - 1 - t0), count, nxt);
    const int j = t - t0;
    const int len = t < t1 ? tstart[j + 1] - tstart[j] : 0;
    const bool valid = pos < len;
    const int vp = cur.vp;
    const bool first = !valid || pos == 0 || cur.vprev != vp;
    const bool last = valid && (pos == len - 1 || cur.vnext != vp);
    Real u[R];
    {
      const Real* x = XL ? xl + cur.ic * DC : a.xc + (int64_t)cur.ic * DC;
      Real xv[DC];
#pragma unroll
      for (int c = 0; c < DC; ++c) xv[c] = x[c];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        Real s = Real(0);
#pragma unroll
        for (int c = 0; c < DC; ++c) s += cur.jc[r][c] * xv[c];
        u[r] = s;
      }
    }
    double w[DP + 1];
#pragma unroll
    for (int c = 0; c < DP; ++c) {
      Real s = Real(0);
#pragma unroll
      for (int r = 0; r < R; ++r) s += cur.jp[r][c] * u[r];
      w[c] = valid ? (double)s : 0.0;
    }
    w[DP] = first ? (double)pos : 0.0;
    int flag = first ? 1 : 0;
    wave_segmented_scan<double, DP + 1>(flag, w);
    if (lane == 63) {
#pragma unroll
      for (int c = 0; c <= DP; ++c) tail[wave * (DP + 1) + c] = w[c];
      tail_flag[wave] = flag;
    }
    __syncthreads();
    if (!flag) {
      for (int k = wave - 1; k >= 0; --k) {
#pragma unroll
        for (int c = 0; c <= DP; ++c) w[c] += tail[k * (DP + 1) + c];
        if (tail_flag[k]) break;
      }
    }
    const int fp = (int)w[DP];
    if (last) {
      double sp[DP];
#pragma unroll
      for (int i = 0; i < DP; ++i) {
        double s = 0.0;
#pragma unroll
        for (int jj = 0; jj < DP; ++jj) s += sym<DP>(cur.mk, i, jj) * w[jj];
        sp[i] = s;
      }
#pragma unroll
      for (int c = 0; c < DP; ++c) sstage[fp * DP + c] = (Real)sp[c];
    }
    __syncthreads();
    if (valid) {
      Real sv[DP];
#pragma unroll
      for (int c = 0; c < DP; ++c) sv[c] = sstage[fp * DP + c];
      Real q[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        Real s = u[r];
#pragma unroll
        for (int c = 0; c < DP; ++c) s -= cur.jp[r][c] * sv[c];
        q[r] = s;
      }
#pragma unroll
      for (int c = 0; c < DC; ++c) {
        Real s = Real(0);
#pragma unroll
        for (int r = 0; r < R; ++r) s += cur.jc[r][c] * q[r];
        stage[cur.rk * DC + c] = s;
      }
    }
    __syncthreads();
    if (t < t1 && cur.cseg >= 0) {
      Real sum[DC];
#pragma unroll
      for (int c = 0; c < DC; ++c) sum[c] = stage[cur.clo * DC + c];
      for (int i = cur.clo + 1; i < cur.chi; ++i)
#pragma unroll
        for (int c = 0; c < DC; ++c) sum[c] += stage[i * DC + c];
#pragma unroll
      for (int c = 0; c < DC; ++c) acc[cur.cseg * DC + c] += sum[c];
    }
  };
  SchurTileRegs<Real, R, DC, DP, XL> ta, tb;
  if (t0 < t1) schur_load_tile<Real, R, DC, DP, CFIRST, NTH, XL>(a, tstart, tseg, t0, 0, count, ta);
  // pairs of tiles with no branch between the halves (a conditional half would
  // move loop-carried registers and force waits for loads in flight)
  for (int t = t0; t < t1; t += 2) {
    process(ta, tb, t);
    process(tb, ta, t + 1);
  }
  __syncthreads();
  const int64_t nn = (int64_t)a.count_c * DC;
  Real* out = a.part_c + (int64_t)(a.dense_block_c + blockIdx.x) * nn;
  for (int64_t i = threadIdx.x; i < nn; i += NTH) out[i] = acc[i];
}

// ------------------------------------------------ preconditioner edge blocks
// Per tile (all-SPARSE plan): each (camera, point) edge head forms
// W_e = sum_{obs in edge} J_c^T J_p (float64), then Q = W_e C_p^-1 W_e^T
// (upper triangle); values are staged by camera rank and summed per segment
// into float64 partials [P][DC*(DC+1)/2].
template <typename Real>
__global__ void __launch_bounds__(TLH_THREADS)
    schur_precond_kernel(SchurArgs<Real> a, int R, int DC, int DP, int cfirst, int chunk) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* stage = reinterpret_cast<double*>(smem_raw);
  const int OC = cfirst ? 0 : R * DP;
  const int OP = cfirst ? R * DC : 0;
  const int W = DC * (DC + 1) / 2;
  double* part = reinterpret_cast<double*>(a.part_c);
  const int t = blockIdx.x;
  if (t >= a.ntiles) return;
  const int start = a.tile_start[t];
  const int len = a.tile_start[t + 1] - start;
  for (int c0 = 0; c0 < W; c0 += chunk) {
    const int cw = min(chunk, W - c0);
    __syncthreads();
    for (int pos = threadIdx.x; pos < len; pos += TLH_THREADS) {
      const int64_t n = start + pos;
      const int rk = a.rank_c ? (int)a.rank_c[n] : pos;
      const bool head = pos == 0 || a.idx_c[n] != a.idx_c[n - 1] || a.idx_p[n] != a.idx_p[n - 1];
      double vals[8];
      for (int k = 0; k < cw; ++k) vals[k] = 0.0;
      if (head) {
        double w[8][8];
        for (int i = 0; i < DC; ++i)
          for (int j = 0; j < DP; ++j) w[i][j] = 0.0;
        for (int m = pos; m < len; ++m) {
          const int64_t nm = start + m;
          if (m > pos && (a.idx_c[nm] != a.idx_c[n] || a.idx_p[nm] != a.idx_p[n])) break;
          for (int r = 0; r < R; ++r)
            for (int i = 0; i < DC; ++i) {
              const double ci = (double)a.packed[(int64_t)(OC + r * DC + i) * a.stride + nm];
              for (int j = 0; j < DP; ++j) w[i][j] += ci * (double)a.packed[(int64_t)(OP + r * DP + j) * a.stride + nm];
            }
        }
        const double* m = a.cinv + (int64_t)a.idx_p[n] * (DP * (DP + 1) / 2);
        // wm = W C^-1
        double wm[8][8];
        for (int i = 0; i < DC; ++i)
          for (int j = 0; j < DP; ++j) {
            double s = 0.0;
            for (int k = 0; k < DP; ++k) {
              const int lo = k < j ? k : j, hi = k < j ? j : k;
              s += w[i][k] * m[lo * DP - lo * (lo - 1) / 2 + (hi - lo)];
            }
            wm[i][j] = s;
          }
        int e = 0, k = 0;
        for (int i = 0; i < DC; ++i)
          for (int j = i; j < DC; ++j, ++e) {
            if (e < c0 || e >= c0 + cw) continue;
            double s = 0.0;
            for (int l = 0; l < DP; ++l) s += wm[i][l] * w[j][l];
            vals[e - c0] = s;
            (void)k;
          }
      }
      for (int k = 0; k < cw; ++k) stage[rk * cw + k] = vals[k];
    }
    __syncthreads();
    const int sb = a.seg_ptr_c[t], se = a.seg_ptr_c[t + 1];
    for (int k = sb + (int)threadIdx.x; k < se; k += TLH_THREADS) {
      const int lo = a.seg_start_c[k];
      const int hi = (k + 1 < se) ? a.seg_start_c[k + 1] : len;
      double* dst = part + (int64_t)a.seg_target_c[k] * W + c0;
      for (int c = 0; c < cw; ++c) {
        double s = stage[lo * cw + c];
        for (int i = lo + 1; i < hi; ++i) s += stage[i * cw + c];
        dst[c] = s;
      }
    }
  }
}

// Same result as schur_precond_kernel for fixed block shapes: every edge value
// lives in registers (no local arrays), all W = DC(DC+1)/2 components are
// computed once per pass, and the LDS staging takes ceil(W / chunk) passes.
template <typename Real, int R, int DC, int DP, bool CFIRST>
__global__ void __launch_bounds__(TLH_THREADS)
    schur_precond_fixed(SchurArgs<Real> a, int chunk) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* stage = reinterpret_cast<double*>(smem_raw);
  constexpr int OC = CFIRST ? 0 : R * DP;
  constexpr int OP = CFIRST ? R * DC : 0;
  constexpr int W = DC * (DC + 1) / 2;
  constexpr int WP = DP * (DP + 1) / 2;
  double* part = reinterpret_cast<double*>(a.part_c);
  const int t = blockIdx.x;
  if (t >= a.ntiles) return;
  const int start = a.tile_start[t];
  const int len = a.tile_start[t + 1] - start;
  for (int c0 = 0; c0 < W; c0 += chunk) {
    const int cw = min(chunk, W - c0);
    __syncthreads();
    for (int pos = threadIdx.x; pos < len; pos += TLH_THREADS) {
      const int64_t n = start + pos;
      const int rk = a.rank_c ? (int)a.rank_c[n] : pos;
      const int32_t ic = a.idx_c[n], ip = a.idx_p[n];
      const bool head = pos == 0 || ic != a.idx_c[n - 1] || ip != a.idx_p[n - 1];
      double vals[W];
#pragma unroll
      for (int k = 0; k < W; ++k) vals[k] = 0.0;
      if (head) {
        double w[DC][DP];
#pragma unroll
        for (int i = 0; i < DC; ++i)
#pragma unroll
          for (int j = 0; j < DP; ++j) w[i][j] = 0.0;
        for (int m = pos; m < len; ++m) {
          const int64_t nm = start + m;
          if (m > pos && (a.idx_c[nm] != ic || a.idx_p[nm] != ip)) break;
#pragma unroll
          for (int r = 0; r < R; ++r) {
            double jp[DP];
#pragma unroll
            for (int j = 0; j < DP; ++j) jp[j] = (double)a.packed[(int64_t)(OP + r * DP + j) * a.stride + nm];
#pragma unroll
            for (int i = 0; i < DC; ++i) {
              const double ci = (double)a.packed[(int64_t)(OC + r * DC + i) * a.stride + nm];
#pragma unroll
              for (int j = 0; j < DP; ++j) w[i][j] += ci * jp[j];
            }
          }
        }
        double cm[WP];
        const double* mp = a.cinv + (int64_t)ip * WP;
#pragma unroll
        for (int k = 0; k < WP; ++k) cm[k] = mp[k];
        double wm[DC][DP];
#pragma unroll
        for (int i = 0; i < DC; ++i)
#pragma unroll
          for (int j = 0; j < DP; ++j) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < DP; ++k) {
              const int lo = k < j ? k : j, hi = k < j ? j : k;
              acc += w[i][k] * cm[lo * DP - lo * (lo - 1) / 2 + (hi - lo)];
            }
            wm[i][j] = acc;
          }
        int e = 0;
#pragma unroll
        for (int i = 0; i < DC; ++i)
#pragma unroll
          for (int j = i; j < DC; ++j, ++e) {
            double acc = 0.0;
#pragma unroll
            for (int l = 0; l < DP; ++l) acc += wm[i][l] * w[j][l];
            vals[e] = acc;
          }
      }
#pragma unroll
      for (int k = 0; k < W; ++k)
        if (k >= c0 && k < c0 + cw) stage[rk * cw + (k - c0)] = vals[k];
    }
    __syncthreads();
    const int sb = a.seg_ptr_c[t], se = a.seg_ptr_c[t + 1];
    for (int k = sb + (int)threadIdx.x; k < se; k += TLH_THREADS) {
      const int lo = a.seg_start_c[k];
      const int hi = (k + 1 < se) ? a.seg_start_c[k + 1] : len;
      double* dst = part + (int64_t)a.seg_target_c[k] * W + c0;
      for (int c = 0; c < cw; ++c) {
        double acc = stage[lo * cw + c];
        for (int i = lo + 1; i < hi; ++i) acc += stage[i * cw + c];
        dst[c] = acc;
      }
    }
  }
}

// M_c = D_c + lambda I - sum of the camera's edge partials (full symmetric).
__global__ void __launch_bounds__(TLH_THREADS)
    schur_precond_finish(double* __restrict__ M, const double* __restrict__ D, double damping,
                         const double* __restrict__ partials, const int32_t* __restrict__ csr_ptr,
                         const int32_t* __restrict__ csr_idx, int64_t count, int dim) {
  const int W = dim * (dim + 1) / 2;
  const int64_t e = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x;
  if (e >= count * W) return;
  const int64_t v = e / W;
  int k = (int)(e - v * W), i = 0, row = dim;
  while (k >= row) {
    k -= row;
    ++i;
    --row;
  }
  const int j = i + k;
  const int kk = (int)(e - v * W);
  double s = 0.0;
  for (int p = csr_ptr[v]; p < csr_ptr[v + 1]; ++p) s += partials[(int64_t)csr_idx[p] * W + kk];
  const double val = D[v * dim * dim + i * dim + j] + (i == j ? damping : 0.0) - s;
  M[v * dim * dim + i * dim + j] = val;
  M[v * dim * dim + j * dim + i] = val;
}

__global__ void __launch_bounds__(TLH_THREADS)
    schur_precond_finish_wave(double* __restrict__ M, const double* __restrict__ D, double damping,
                              const double* __restrict__ partials, const int32_t* __restrict__ csr_ptr,
                              const int32_t* __restrict__ csr_idx, int64_t count, int dim) {
  const int W = dim * (dim + 1) / 2;
  const int lane = threadIdx.x & 63;
  const int64_t e = (int64_t)blockIdx.x * (TLH_THREADS / 64) + (threadIdx.x >> 6);
  if (e >= count * W) return;
  const int64_t v = e / W;
  const int k = (int)(e - v * W);
  const double s = csr_wave_sum<double>(partials, csr_idx, csr_ptr[v], csr_ptr[v + 1], W, k, lane);
  if (lane == 0) {
    int i, j;
    upper_index(k, dim, i, j);
    const double val = D[v * dim * dim + i * dim + j] + (i == j ? damping : 0.0) - s;
    M[v * dim * dim + i * dim + j] = val;
    M[v * dim * dim + j * dim + i] = val;
  }
}

// ------------------------------------------------------- dense reduced matrix
// One thread per unique camera pair (ci <= cj): sum over the pair's
// observation pairs (i in edge(ci,p), j in edge(cj,p)) of
// (J_c^T J_p)_i C_p^-1 (J_c^T J_p)_j^T, subtract from block (ci, cj) of M
// (and its transpose).  M must hold blockdiag(D_c + lambda I) beforehand.
// FR/FDC/FDP > 0 fix the block shape at compile time (bundle adjustment: 2, 6,
// 3), so the per-thread blocks live in registers; 0 reads it from the
// arguments (the arrays then live in private memory).
template <typename Real, int FR, int FDC, int FDP>
__global__ void __launch_bounds__(TLH_THREADS)
    schur_dense_kernel(SchurArgs<Real> a, int R_, int DC_, int DP_, int cfirst, const int32_t* __restrict__ pair_ptr,
                       const int32_t* __restrict__ pair_i, const int32_t* __restrict__ pair_j,
                       const int32_t* __restrict__ pair_key_ci, const int32_t* __restrict__ pair_key_cj,
                       int64_t npairs, double* __restrict__ M, int64_t ld) {
  constexpr int MC = FDC ? FDC : 8, MP = FDP ? FDP : 8;
  const int R = FR ? FR : R_, DC = FDC ? FDC : DC_, DP = FDP ? FDP : DP_;
  const int64_t key = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x;
  if (key >= npairs) return;
  const int OC = cfirst ? 0 : R * DP;
  const int OP = cfirst ? R * DC : 0;
  double acc[MC][MC];
#pragma unroll
  for (int i = 0; i < MC; ++i)
#pragma unroll
    for (int j = 0; j < MC; ++j) acc[i][j] = 0.0;
  for (int e = pair_ptr[key]; e < pair_ptr[key + 1]; ++e) {
    const int64_t ni = pair_i[e], nj = pair_j[e];
    double wi[MC][MP], wj[MC][MP];
#pragma unroll
    for (int c = 0; c < MC; ++c)
#pragma unroll
      for (int p = 0; p < MP; ++p) {
        if (c >= DC || p >= DP) continue;
        double si = 0.0, sj = 0.0;
#pragma unroll
        for (int r = 0; r < (FR ? FR : 16); ++r) {
          if (r >= R) break;
          si += (double)a.packed[(int64_t)(OC + r * DC + c) * a.stride + ni] *
                (double)a.packed[(int64_t)(OP + r * DP + p) * a.stride + ni];
          sj += (double)a.packed[(int64_t)(OC + r * DC + c) * a.stride + nj] *
                (double)a.packed[(int64_t)(OP + r * DP + p) * a.stride + nj];
        }
        wi[c][p] = si;
        wj[c][p] = sj;
      }
    const double* m = a.cinv + (int64_t)a.idx_p[ni] * (DP * (DP + 1) / 2);
#pragma unroll
    for (int c = 0; c < MC; ++c) {
      if (c >= DC) continue;
      double t[MP];
#pragma unroll
      for (int p = 0; p < MP; ++p) {
        if (p >= DP) continue;
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < MP; ++q) {
          if (q >= DP) continue;
          const int lo = q < p ? q : p, hi = q < p ? p : q;
          s += wi[c][q] * m[lo * DP - lo * (lo - 1) / 2 + (hi - lo)];
        }
        t[p] = s;
      }
#pragma unroll
      for (int d = 0; d < MC; ++d) {
        if (d >= DC) continue;
        double s = 0.0;
#pragma unroll
        for (int p = 0; p < MP; ++p)
          if (p < DP) s += t[p] * wj[d][p];
        acc[c][d] += s;
      }
    }
  }
  const int64_t ci = pair_key_ci[key], cj = pair_key_cj[key];
#pragma unroll
  for (int c = 0; c < MC; ++c)
#pragma unroll
    for (int d = 0; d < MC; ++d) {
      if (c >= DC || d >= DC) continue;
      M[(ci * DC + c) * ld + cj * DC + d] -= acc[c][d];
      if (ci != cj) M[(cj * DC + d) * ld + ci * DC + c] -= acc[c][d];
    }
}

// Fixed-shape dense path in two passes: schur_edge_kernel writes, per
// observation, W_o = J_c^T J_p (DC x DP) and U_o = W_o C_p^-1 (float64);
// schur_dense_pairs then sums U_i W_j^T over each camera pair's observation
// pairs (the same products, in the same order, as schur_dense_kernel, without
// re-reading and re-multiplying the blocks of an observation for every pair),
// one wavefront per camera pair.
template <typename Real, int R, int DC, int DP>
__global__ void __launch_bounds__(TLH_THREADS)
    schur_edge_kernel(SchurArgs<Real> a, int cfirst, int64_t count, double* __restrict__ work) {
  const int64_t n = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x;
  if (n >= count) return;
  const int OC = cfirst ? 0 : R * DP;
  const int OP = cfirst ? R * DC : 0;
  double w[DC][DP];
#pragma unroll
  for (int c = 0; c < DC; ++c)
#pragma unroll
    for (int p = 0; p < DP; ++p) {
      double s = 0.0;
#pragma unroll
      for (int r = 0; r < R; ++r)
        s += (double)a.packed[(int64_t)(OC + r * DC + c) * a.stride + n] *
             (double)a.packed[(int64_t)(OP + r * DP + p) * a.stride + n];
      w[c][p] = s;
    }
  const double* m = a.cinv + (int64_t)a.idx_p[n] * (DP * (DP + 1) / 2);
  double* out = work + n * (2 * DC * DP);
#pragma unroll
  for (int c = 0; c < DC; ++c) {
#pragma unroll
    for (int p = 0; p < DP; ++p) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < DP; ++q) s += w[c][q] * sym<DP>(m, q, p);
      out[c * DP + p] = w[c][p];
      out[DC * DP + c * DP + p] = s;
    }
  }
}

template <int DC, int DP>
__global__ void __launch_bounds__(TLH_THREADS)
    schur_dense_pairs(const double* __restrict__ work, const int32_t* __restrict__ pair_ptr,
                      const int32_t* __restrict__ pair_i, const int32_t* __restrict__ pair_j,
                      const int32_t* __restrict__ pair_key_ci, const int32_t* __restrict__ pair_key_cj,
                      int64_t npairs, double* __restrict__ M, int64_t ld) {
  // one wavefront per camera pair: lane l sums entries l, l + 64, ... (a camera's
  // own pair lists every observation of it), then a fixed xor tree adds the lanes
  const int64_t key = ((int64_t)blockIdx.x * TLH_THREADS + threadIdx.x) / 64;
  const int lane = threadIdx.x & 63;
  if (key >= npairs) return;
  double acc[DC][DC];
#pragma unroll
  for (int i = 0; i < DC; ++i)
#pragma unroll
    for (int j = 0; j < DC; ++j) acc[i][j] = 0.0;
  for (int e = pair_ptr[key] + lane; e < pair_ptr[key + 1]; e += 64) {
    const double* ui = work + (int64_t)pair_i[e] * (2 * DC * DP) + DC * DP;
    const double* wj = work + (int64_t)pair_j[e] * (2 * DC * DP);
    double u[DC][DP], v[DC][DP];
#pragma unroll
    for (int c = 0; c < DC; ++c)
#pragma unroll
      for (int p = 0; p < DP; ++p) {
        u[c][p] = ui[c * DP + p];
        v[c][p] = wj[c * DP + p];
      }
#pragma unroll
    for (int c = 0; c < DC; ++c)
#pragma unroll
      for (int d = 0; d < DC; ++d) {
        double s = 0.0;
#pragma unroll
        for (int p = 0; p < DP; ++p) s += u[c][p] * v[d][p];
        acc[c][d] += s;
      }
  }
#pragma unroll
  for (int c = 0; c < DC; ++c)
#pragma unroll
    for (int d = 0; d < DC; ++d)
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) acc[c][d] += __shfl_xor(acc[c][d], off);
  // every lane holds all DC*DC sums; lane l < DC*DC writes entry (l / DC, l % DC)
  // and its mirror, both loads issued before either store
  if (lane >= DC * DC) return;
  double mine = 0.0;
#pragma unroll
  for (int cd = 0; cd < DC * DC; ++cd)
    if (lane == cd) mine = acc[cd / DC][cd % DC];
  const int64_t ci = pair_key_ci[key], cj = pair_key_cj[key];
  const int c = lane / DC, d = lane % DC;
  const int64_t p0 = (ci * DC + c) * ld + cj * DC + d;
  const int64_t p1 = (cj * DC + d) * ld + ci * DC + c;
  const bool mirror = ci != cj;
  const double m0 = M[p0];
  const double m1 = mirror ? M[p1] : 0.0;
  M[p0] = m0 - mine;
  if (mirror) M[p1] = m1 - mine;
}

// ------------------------------------------------------------------ host side
template <typename Real>
SchurArgs<Real> make_schur_args(const tlh_operator_t* op, const tlh_group_t& g, int sc, int sp, const void* xc,
                                const void* bp, void* outp, const double* cinv) {
  SchurArgs<Real> a{};
  a.packed = static_cast<const Real*>(g.packed);
  a.stride = g.stride;
  a.tile_start = g.tile_start;
  a.ntiles = g.ntiles;
  a.tiles_per_wg = g.tiles_per_wg;
  a.idx_c = g.slot[sc].index;
  a.idx_p = g.slot[sp].index;
  a.rank_c = g.slot[sc].rank;
  a.seg_ptr_c = g.slot[sc].seg_ptr;
  a.seg_start_c = g.slot[sc].seg_start;
  a.seg_target_c = g.slot[sc].seg_target;
  a.mode_c = g.slot[sc].mode;
  a.dense_block_c = g.slot[sc].dense_block;
  const tlh_type_t& tc = op->types[g.slot[sc].type];
  const tlh_type_t& tp = op->types[g.slot[sp].type];
  a.count_c = (int32_t)tc.count;
  a.part_c = static_cast<Real*>(tc.partials);
  a.seg_ptr_p = g.slot[sp].seg_ptr;
  a.seg_start_p = g.slot[sp].seg_start;
  a.xc = xc ? static_cast<const Real*>(xc) + tc.offset : nullptr;
  a.bp = static_cast<const Real*>(bp);
  a.outp = static_cast<Real*>(outp);
  a.cinv = cinv;
  (void)tp;
  return a;
}

// The W-group tile kernel runs one observation per thread: 1,024 threads over a
// float32 tile (512 for float64), so a compute unit holding one workgroup (the
// dense camera accumulator fills most of its LDS) still has 16 (8) wavefronts.
template <typename Real>
constexpr int schur_threads() {
  return sizeof(Real) == 4 ? 1024 : 512;
}

template <typename Real, int R, int DC, int DP, bool CFIRST, int MODE>
int launch_schur_t(const tlh_operator_t* op, const tlh_group_t& g, int sc, int sp, const void* xc, const void* bp,
                   void* outp, const double* cinv, hipStream_t s) {
  constexpr int ITEMS = 1;
  constexpr int NTH = schur_threads<Real>();
  if (g.tile_cap > NTH * ITEMS) return TLH_EINVAL;
  constexpr int T = NTH * ITEMS;
  // wave tails (float64 sums and flags), camera stage [T][DC], point stage [T][DP]
  int64_t lds = (int64_t)(NTH / 64) * ((DP + 1) * sizeof(double) + 2 * sizeof(int)) +
                (int64_t)T * (DC + DP) * sizeof(Real);
  const int64_t slice = op->types[g.slot[sc].type].count * DC * (int64_t)sizeof(Real);
  const bool dense = MODE != SCHUR_BACKSUB && g.slot[sc].mode == TLH_REDUCE_DENSE;
  if (dense) lds += slice;
  if (MODE == SCHUR_MATVEC && dense) {
    // pipelined matvec: its tile starts and camera segment ranges also go to LDS
    const int64_t info = 2 * (int64_t)(g.tiles_per_wg + 2) * sizeof(int);
    const bool pxl = lds + slice + info <= device_lds_max();
    const int64_t plds = lds + (pxl ? slice : 0) + info;
    auto pk = pxl ? schur_matvec_pipe_kernel<Real, R, DC, DP, CFIRST, NTH, true>
                  : schur_matvec_pipe_kernel<Real, R, DC, DP, CFIRST, NTH, false>;
    TLH_RETURN_IF_ERROR(allow_lds((const void*)pk, plds));
    hipLaunchKernelGGL(pk, dim3((unsigned)g.nwg), dim3(NTH), (size_t)plds, s,
                       make_schur_args<Real>(op, g, sc, sp, xc, bp, outp, cinv), (int64_t)g.count);
    TLH_RETURN_IF_LAUNCH_FAILED();
    return 0;
  }
  // the retained x slice goes to LDS (after the accumulator region) when it fits
  const bool xl = MODE != SCHUR_RHS && dense && lds + slice <= device_lds_max();
  if (xl) lds += slice;
  auto kernel = xl ? schur_tile_kernel<Real, R, DC, DP, CFIRST, ITEMS, MODE, NTH, (MODE != SCHUR_RHS)>
                   : schur_tile_kernel<Real, R, DC, DP, CFIRST, ITEMS, MODE, NTH, false>;
  TLH_RETURN_IF_ERROR(allow_lds((const void*)kernel, lds));
  const unsigned grid = MODE == SCHUR_BACKSUB ? (unsigned)g.ntiles : (unsigned)g.nwg;
  SchurArgs<Real> args = make_schur_args<Real>(op, g, sc, sp, xc, bp, outp, cinv);
  if (MODE == SCHUR_BACKSUB) args.tiles_per_wg = 1;
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(NTH), (size_t)lds, s, args);
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

template <typename Real, int MODE>
int launch_schur(const tlh_operator_t* op, const tlh_group_t& g, int sc, int sp, const void* xc, const void* bp,
                 void* outp, const double* cinv, hipStream_t s) {
  const int R = g.rows, DC = g.slot[sc].dim, DP = g.slot[sp].dim;
  const bool cfirst = sc < sp;
  if (R == 2 && DC == 6 && DP == 3)
    return cfirst ? launch_schur_t<Real, 2, 6, 3, true, MODE>(op, g, sc, sp, xc, bp, outp, cinv, s)
                  : launch_schur_t<Real, 2, 6, 3, false, MODE>(op, g, sc, sp, xc, bp, outp, cinv, s);
  if (R == 2 && DC == 7 && DP == 3)
    return cfirst ? launch_schur_t<Real, 2, 7, 3, true, MODE>(op, g, sc, sp, xc, bp, outp, cinv, s)
                  : launch_schur_t<Real, 2, 7, 3, false, MODE>(op, g, sc, sp, xc, bp, outp, cinv, s);
  if (R == 3 && DC == 6 && DP == 3)
    return cfirst ? launch_schur_t<Real, 3, 6, 3, true, MODE>(op, g, sc, sp, xc, bp, outp, cinv, s)
                  : launch_schur_t<Real, 3, 6, 3, false, MODE>(op, g, sc, sp, xc, bp, outp, cinv, s);
  return TLH_ENOTFOUND;
}

template <typename Real>
int schur_dispatch(const tlh_operator_t* op, int wgroup, int eslot, int mode, const void* xc, const void* bp, void* outp,
                   const double* cinv, hipStream_t s) {
  const tlh_group_t& g = op->groups[wgroup];
  // the eliminated slot is the group's primary slot: sorted, identity rank
  if (g.nslots != 2 || eslot < 0 || eslot > 1 || g.slot[eslot].rank != nullptr) return TLH_EINVAL;
  const int se = eslot, sc = 1 - eslot;
  if (g.count == 0) return 0;
  switch (mode) {
    case SCHUR_MATVEC:
      return launch_schur<Real, SCHUR_MATVEC>(op, g, sc, se, xc, bp, outp, cinv, s);
    case SCHUR_RHS:
      return launch_schur<Real, SCHUR_RHS>(op, g, sc, se, xc, bp, outp, cinv, s);
    case SCHUR_BACKSUB:
      return launch_schur<Real, SCHUR_BACKSUB>(op, g, sc, se, xc, bp, outp, cinv, s);
  }
  return TLH_EINVAL;
}

}  // namespace

extern "C" int tlh_schur_apply(const tlh_operator_t* op, int wgroup, int eslot, int mode, const void* xc, const void* bp,
                               void* outp, const double* cinv, tlh_stream_t stream) {
  if (!op || wgroup < 0 || wgroup >= op->ngroups || !cinv) return TLH_EINVAL;
  if (mode != SCHUR_RHS && !xc) return TLH_EINVAL;
  if (mode != SCHUR_MATVEC && !bp) return TLH_EINVAL;
  if (mode == SCHUR_BACKSUB && !outp) return TLH_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (op->dtype == TLH_F32) return schur_dispatch<float>(op, wgroup, eslot, mode, xc, bp, outp, cinv, s);
  if (op->dtype == TLH_F64) return schur_dispatch<double>(op, wgroup, eslot, mode, xc, bp, outp, cinv, s);
  return TLH_EINVAL;
}

extern "C" int tlh_schur_precond(const tlh_operator_t* op, int wgroup, int eslot, const double* cinv, const double* D,
                                 double damping, double* M, int64_t lds_limit, tlh_stream_t stream) {
  if (!op || wgroup < 0 || wgroup >= op->ngroups || !cinv || !D || !M) return TLH_EINVAL;
  const tlh_group_t& g = op->groups[wgroup];
  if (g.nslots != 2 || eslot < 0 || eslot > 1 || g.slot[eslot].rank != nullptr) return TLH_EINVAL;
  const int se = eslot, sc = 1 - eslot;
  const int R = g.rows, DC = g.slot[sc].dim, DP = g.slot[se].dim;
  if (DC > 8 || DP > 8 || g.slot[sc].mode != TLH_REDUCE_SPARSE) return TLH_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const int W = DC * (DC + 1) / 2;
  const bool fixed = R == 2 && DC == 6 && DP == 3;
  const int64_t fit = lds_limit / ((int64_t)g.tile_cap * 8);
  int chunk = (int)std::min<int64_t>(std::min(W, fixed ? W : 8), fit);
  if (chunk < 1) return TLH_ELDS;
  if (fixed) chunk = (int)ceil_div(W, ceil_div(W, chunk));  // equal passes
  const int64_t lds = (int64_t)g.tile_cap * chunk * 8;
  const tlh_type_t& tc = op->types[g.slot[sc].type];
  if (g.count > 0 && fixed) {
    const bool cfirst = sc < se;
    if (op->dtype == TLH_F32) {
      SchurArgs<float> a = make_schur_args<float>(op, g, sc, se, nullptr, nullptr, nullptr, cinv);
      const void* k = cfirst ? (const void*)schur_precond_fixed<float, 2, 6, 3, true>
                             : (const void*)schur_precond_fixed<float, 2, 6, 3, false>;
      TLH_RETURN_IF_ERROR(allow_lds(k, lds));
      if (cfirst)
        hipLaunchKernelGGL((schur_precond_fixed<float, 2, 6, 3, true>), dim3(g.ntiles), dim3(TLH_THREADS),
                           (size_t)lds, s, a, chunk);
      else
        hipLaunchKernelGGL((schur_precond_fixed<float, 2, 6, 3, false>), dim3(g.ntiles), dim3(TLH_THREADS),
                           (size_t)lds, s, a, chunk);
    } else {
      SchurArgs<double> a = make_schur_args<double>(op, g, sc, se, nullptr, nullptr, nullptr, cinv);
      const void* k = cfirst ? (const void*)schur_precond_fixed<double, 2, 6, 3, true>
                             : (const void*)schur_precond_fixed<double, 2, 6, 3, false>;
      TLH_RETURN_IF_ERROR(allow_lds(k, lds));
      if (cfirst)
        hipLaunchKernelGGL((schur_precond_fixed<double, 2, 6, 3, true>), dim3(g.ntiles), dim3(TLH_THREADS),
                           (size_t)lds, s, a, chunk);
      else
        hipLaunchKernelGGL((schur_precond_fixed<double, 2, 6, 3, false>), dim3(g.ntiles), dim3(TLH_THREADS),
                           (size_t)lds, s, a, chunk);
    }
    TLH_RETURN_IF_LAUNCH_FAILED();
  } else if (g.count > 0) {
    if (op->dtype == TLH_F32) {
      SchurArgs<float> a = make_schur_args<float>(op, g, sc, se, nullptr, nullptr, nullptr, cinv);
      TLH_RETURN_IF_ERROR(allow_lds((const void*)schur_precond_kernel<float>, lds));
      hipLaunchKernelGGL(schur_precond_kernel<float>, dim3(g.ntiles), dim3(TLH_THREADS), (size_t)lds, s, a, R, DC,
                         DP, (int)(sc < se), chunk);
    } else {
      SchurArgs<double> a = make_schur_args<double>(op, g, sc, se, nullptr, nullptr, nullptr, cinv);
      TLH_RETURN_IF_ERROR(allow_lds((const void*)schur_precond_kernel<double>, lds));
      hipLaunchKernelGGL(schur_precond_kernel<double>, dim3(g.ntiles), dim3(TLH_THREADS), (size_t)lds, s, a, R,
                         DC, DP, (int)(sc < se), chunk);
    }
    TLH_RETURN_IF_LAUNCH_FAILED();
  }
  const int64_t total = tc.count * W;
  if (total) {
    if (csr_heavy(tc.csr_entries, tc.count))
      hipLaunchKernelGGL(schur_precond_finish_wave, dim3((unsigned)ceil_div(total, TLH_THREADS / 64)),
                         dim3(TLH_THREADS), 0, s, M, D, damping, static_cast<const double*>(tc.partials), tc.csr_ptr,
                         tc.csr_idx, tc.count, DC);
    else
      hipLaunchKernelGGL(schur_precond_finish, dim3((unsigned)ceil_div(total, TLH_THREADS)), dim3(TLH_THREADS), 0, s,
                         M, D, damping, static_cast<const double*>(tc.partials), tc.csr_ptr, tc.csr_idx, tc.count, DC);
    TLH_RETURN_IF_LAUNCH_FAILED();
  }
  return 0;
}

extern "C" int tlh_schur_dense(const tlh_operator_t* op, int wgroup, int eslot, const double* cinv, const int32_t* pair_ptr,
                               const int32_t* pair_i, const int32_t* pair_j, const int32_t* key_ci,
                               const int32_t* key_cj, int64_t npairs, double* M, int64_t ld, double* work,
                               tlh_stream_t stream) {
  if (!op || wgroup < 0 || wgroup >= op->ngroups || !cinv || !M) return TLH_EINVAL;
  const tlh_group_t& g = op->groups[wgroup];
  if (g.nslots != 2 || eslot < 0 || eslot > 1 || g.slot[eslot].rank != nullptr) return TLH_EINVAL;
  const int se = eslot, sc = 1 - eslot;
  const int R = g.rows, DC = g.slot[sc].dim, DP = g.slot[se].dim;
  if (DC > 8 || DP > 8) return TLH_EINVAL;
  if (npairs == 0) return 0;
  // camera-pair blocks sit at the W-group retained type's rows and columns: its
  // offset in the retained layout (other retained types may come first)
  const int64_t roff = op->types[g.slot[sc].type].offset;
  if (roff < 0 || roff + op->types[g.slot[sc].type].count * DC > ld) return TLH_EINVAL;
  M += roff * ld + roff;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid((unsigned)ceil_div(npairs, TLH_THREADS)), block(TLH_THREADS);
  const bool ba = R == 2 && DC == 6 && DP == 3;
  if (ba && work) {
    const dim3 egrid((unsigned)ceil_div(g.count, TLH_THREADS));
    if (op->dtype == TLH_F32) {
      SchurArgs<float> a = make_schur_args<float>(op, g, sc, se, nullptr, nullptr, nullptr, cinv);
      hipLaunchKernelGGL((schur_edge_kernel<float, 2, 6, 3>), egrid, block, 0, s, a, (int)(sc < se), g.count, work);
    } else {
      SchurArgs<double> a = make_schur_args<double>(op, g, sc, se, nullptr, nullptr, nullptr, cinv);
      hipLaunchKernelGGL((schur_edge_kernel<double, 2, 6, 3>), egrid, block, 0, s, a, (int)(sc < se), g.count, work);
    }
    TLH_RETURN_IF_LAUNCH_FAILED();
    const dim3 wgrid((unsigned)ceil_div(npairs, TLH_THREADS / 64));
    hipLaunchKernelGGL((schur_dense_pairs<6, 3>), wgrid, block, 0, s, work, pair_ptr, pair_i, pair_j, key_ci, key_cj,
                       npairs, M, ld);
    TLH_RETURN_IF_LAUNCH_FAILED();
    return 0;
  }
  if (op->dtype == TLH_F32) {
    SchurArgs<float> a = make_schur_args<float>(op, g, sc, se, nullptr, nullptr, nullptr, cinv);
    auto kernel = ba ? schur_dense_kernel<float, 2, 6, 3> : schur_dense_kernel<float, 0, 0, 0>;
    hipLaunchKernelGGL(kernel, grid, block, 0, s, a, R, DC, DP, (int)(sc < se), pair_ptr, pair_i, pair_j, key_ci,
                       key_cj, npairs, M, ld);
  } else {
    SchurArgs<double> a = make_schur_args<double>(op, g, sc, se, nullptr, nullptr, nullptr, cinv);
    auto kernel = ba ? schur_dense_kernel<double, 2, 6, 3> : schur_dense_kernel<double, 0, 0, 0>;
    hipLaunchKernelGGL(kernel, grid, block, 0, s, a, R, DC, DP, (int)(sc < se), pair_ptr, pair_i, pair_j, key_ci,
                       key_cj, npairs, M, ld);
  }
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}
