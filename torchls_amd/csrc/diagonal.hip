// Deterministic variable diagonal blocks D_t[v] = sum J_s^T J_s (float64).
//
// Replaces warp_diagonal.py (_gram_kernel 14-47, _finish 50-66) and the torch
// reference BlockJacobian.diagonal_blocks (sparse.py:104-121): for every slot
// the upper triangle of J_s^T J_s is formed per observation in float64 (from
// the packed float32 or float64 blocks), plus the alias cross terms
// J_s^T J_o + J_o^T J_s for earlier slots o of the same type whose variable is
// the same in that observation (sparse.py:113-119).  Values are summed with
// the tile/segment order of an all-SPARSE plan (one partial per segment), then
// a finish kernel adds each variable's partials in plan order and mirrors the
// upper triangle.  Runs once per linearization, so it favours generality
// (runtime dims, J re-read from L1/L2 per component) over register blocking.
#include <algorithm>

#include "common.hpp"

using namespace tlh;

namespace {

__device__ __forceinline__ void upper_ij(int k, int d, int& i, int& j) {
  i = 0;
  int row = d;
  while (k >= row) {
    k -= row;
    ++i;
    --row;
  }
  j = i + k;
}

template <typename Real>
__global__ void __launch_bounds__(TLH_THREADS)
    diag_tile_kernel(GroupArgs<Real> g, int chunk, int alias0, int alias1, int alias2) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* stage = reinterpret_cast<double*>(smem_raw);
  const int R = g.rows;
  int off[TLH_MAX_SLOTS + 1];
  off[0] = 0;
  for (int s = 0; s < TLH_MAX_SLOTS; ++s) off[s + 1] = off[s] + (s < g.nslots ? g.slot[s].dim : 0);
  const int alias[3] = {alias0, alias1, alias2};
  const int t = blockIdx.x;
  if (t >= g.ntiles) return;
  const int start = g.tile_start[t];
  const int len = g.tile_start[t + 1] - start;
  for (int s = 0; s < g.nslots; ++s) {
    const SlotArgs<Real>& sl = g.slot[s];
    const int D = sl.dim;
    const int W = D * (D + 1) / 2;
    double* out = reinterpret_cast<double*>(sl.partials);
    for (int c0 = 0; c0 < W; c0 += chunk) {
      const int cw = min(chunk, W - c0);
      __syncthreads();
      for (int pos = threadIdx.x; pos < len; pos += TLH_THREADS) {
        const int64_t n = start + pos;
        const int rk = sl.rank ? (int)sl.rank[n] : pos;
        const int32_t vs = sl.index[n];
        for (int k = 0; k < cw; ++k) {
          int i, j;
          upper_ij(c0 + k, D, i, j);
          double v = 0.0;
          for (int r = 0; r < R; ++r) {
            const double a = (double)g.packed[(int64_t)(R * off[s] + r * D + i) * g.stride + n];
            const double b = (double)g.packed[(int64_t)(R * off[s] + r * D + j) * g.stride + n];
            v += a * b;
          }
          for (int o = 0; o < s; ++o) {
            if (!((alias[s] >> o) & 1) || g.slot[o].index[n] != vs) continue;
            double cross = 0.0;
            for (int r = 0; r < R; ++r) {
              const double si = (double)g.packed[(int64_t)(R * off[s] + r * D + i) * g.stride + n];
              const double sj = (double)g.packed[(int64_t)(R * off[s] + r * D + j) * g.stride + n];
              const double oi = (double)g.packed[(int64_t)(R * off[o] + r * D + i) * g.stride + n];
              const double oj = (double)g.packed[(int64_t)(R * off[o] + r * D + j) * g.stride + n];
              cross += si * oj + oi * sj;
            }
            v += cross;
          }
          stage[rk * cw + k] = v;
        }
      }
      __syncthreads();
      const int sb = sl.seg_ptr[t];
      const int se = sl.seg_ptr[t + 1];
      for (int k = sb + (int)threadIdx.x; k < se; k += TLH_THREADS) {
        const int a = sl.seg_start[k];
        const int b = (k + 1 < se) ? sl.seg_start[k + 1] : len;
        double* dst = out + (int64_t)sl.seg_target[k] * W + c0;
        for (int c = 0; c < cw; ++c) {
          double sum = stage[a * cw + c];
          for (int p = a + 1; p < b; ++p) sum += stage[p * cw + c];
          dst[c] = sum;
        }
      }
    }
  }
}

// Fixed-shape version for two slots of different types (no alias terms): each
// observation's block rows are loaded once into registers and all upper-
// triangle products of both slots are formed there; they are staged to LDS in
// ceil(W / chunk) passes for the same segment sums as diag_tile_kernel.
template <typename Real, int R, int D0, int D1>
__global__ void __launch_bounds__(TLH_THREADS) diag_pair_kernel(GroupArgs<Real> g, int chunk) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* stage = reinterpret_cast<double*>(smem_raw);
  constexpr int W0 = D0 * (D0 + 1) / 2, W1 = D1 * (D1 + 1) / 2;
  const int t = blockIdx.x;
  if (t >= g.ntiles) return;
  const int start = g.tile_start[t];
  const int len = g.tile_start[t + 1] - start;
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const SlotArgs<Real>& sl = g.slot[s];
    const int D = s == 0 ? D0 : D1;
    const int W = s == 0 ? W0 : W1;
    const int off = s == 0 ? 0 : R * D0;
    double* out = reinterpret_cast<double*>(sl.partials);
    if (!out) continue;  // type computed elsewhere (uniform over the workgroup)
    for (int c0 = 0; c0 < W; c0 += chunk) {
      const int cw = min(chunk, W - c0);
      __syncthreads();
      for (int pos = threadIdx.x; pos < len; pos += TLH_THREADS) {
        const int64_t n = start + pos;
        const int rk = sl.rank ? (int)sl.rank[n] : pos;
        double j[R][D0 > D1 ? D0 : D1];
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
          for (int c = 0; c < (D0 > D1 ? D0 : D1); ++c)
            j[r][c] = c < D ? (double)g.packed[(int64_t)(off + r * D + c) * g.stride + n] : 0.0;
        int e = 0;
#pragma unroll
        for (int a = 0; a < (D0 > D1 ? D0 : D1); ++a)
#pragma unroll
          for (int b = a; b < (D0 > D1 ? D0 : D1); ++b) {
            if (a < D && b < D) {
              if (e >= c0 && e < c0 + cw) {
                double v = 0.0;
#pragma unroll
                for (int r = 0; r < R; ++r) v += j[r][a] * j[r][b];
                stage[rk * cw + (e - c0)] = v;
              }
              ++e;
            }
          }
      }
      __syncthreads();
      const int sb = sl.seg_ptr[t];
      const int se = sl.seg_ptr[t + 1];
      for (int k = sb + (int)threadIdx.x; k < se; k += TLH_THREADS) {
        const int a = sl.seg_start[k];
        const int b = (k + 1 < se) ? sl.seg_start[k + 1] : len;
        double* dst = out + (int64_t)sl.seg_target[k] * W + c0;
        for (int c = 0; c < cw; ++c) {
          double sum = stage[a * cw + c];
          for (int p = a + 1; p < b; ++p) sum += stage[p * cw + c];
          dst[c] = sum;
        }
      }
    }
  }
}

// D[v] = sum of the variable's partials (csr order), mirrored to full d x d.
__global__ void __launch_bounds__(TLH_THREADS)
    diag_finish_kernel(double* __restrict__ D, const double* __restrict__ partials,
                       const int32_t* __restrict__ csr_ptr, const int32_t* __restrict__ csr_idx,
                       int64_t count, int dim) {
  const int W = dim * (dim + 1) / 2;
  const int64_t e = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x;
  if (e >= count * W) return;
  const int64_t v = e / W;
  const int k = (int)(e - v * W);
  int i, j;
  upper_ij(k, dim, i, j);
  double sum = 0.0;
  const int k1 = csr_ptr[v + 1];
  for (int p = csr_ptr[v]; p < k1; ++p) sum += partials[(int64_t)csr_idx[p] * W + k];
  double* blk = D + v * dim * dim;
  blk[i * dim + j] = sum;
  blk[j * dim + i] = sum;
}

__global__ void __launch_bounds__(TLH_THREADS)
    diag_finish_wave(double* __restrict__ D, const double* __restrict__ partials, const int32_t* __restrict__ csr_ptr,
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
    D[v * dim * dim + i * dim + j] = s;
    D[v * dim * dim + j * dim + i] = s;
  }
}

template <typename Real>
int diag_impl(const tlh_operator_t* op, double* const* out, hipStream_t stream, int64_t lds_limit) {
  for (int gi = 0; gi < op->ngroups; ++gi) {
    const tlh_group_t& g = op->groups[gi];
    if (g.count == 0 || g.ntiles == 0) continue;
    GroupArgs<Real> a{};
    a.packed = static_cast<const Real*>(g.packed);
    a.stride = g.stride;
    a.tile_start = g.tile_start;
    a.ntiles = g.ntiles;
    a.tiles_per_wg = 1;
    a.rows = g.rows;
    a.nslots = g.nslots;
    int alias[3] = {0, 0, 0};
    int wmax = 1;
    bool skipped = false;
    for (int s = 0; s < g.nslots; ++s) {
      const tlh_slot_t& sl = g.slot[s];
      if (sl.mode != TLH_REDUCE_SPARSE) return TLH_EINVAL;
      SlotArgs<Real>& o = a.slot[s];
      skipped = skipped || !out[sl.type];
      o.partials = out[sl.type] ? static_cast<Real*>(op->types[sl.type].partials) : nullptr;
      o.index = sl.index;
      o.rank = sl.rank;
      o.seg_ptr = sl.seg_ptr;
      o.seg_start = sl.seg_start;
      o.seg_target = sl.seg_target;
      o.dim = sl.dim;
      o.mode = sl.mode;
      for (int q = 0; q < s; ++q)
        if (g.slot[q].type == sl.type) alias[s] |= 1 << q;
      if (out[sl.type]) wmax = std::max(wmax, sl.dim * (sl.dim + 1) / 2);  // LDS sized for computed slots
    }
    int chunk = std::min<int64_t>(wmax, lds_limit / ((int64_t)g.tile_cap * 8));
    if (chunk < 1) return TLH_ELDS;
    const int64_t lds = (int64_t)g.tile_cap * chunk * 8;
    const bool pair = g.nslots == 2 && g.rows == 2 && alias[1] == 0;
    if (pair && ((g.slot[0].dim == 6 && g.slot[1].dim == 3) || (g.slot[0].dim == 3 && g.slot[1].dim == 6))) {
      const void* k = g.slot[0].dim == 6 ? (const void*)diag_pair_kernel<Real, 2, 6, 3>
                                         : (const void*)diag_pair_kernel<Real, 2, 3, 6>;
      TLH_RETURN_IF_ERROR(allow_lds(k, lds));
      if (g.slot[0].dim == 6)
        hipLaunchKernelGGL((diag_pair_kernel<Real, 2, 6, 3>), dim3(g.ntiles), dim3(TLH_THREADS), (size_t)lds, stream,
                           a, chunk);
      else
        hipLaunchKernelGGL((diag_pair_kernel<Real, 2, 3, 6>), dim3(g.ntiles), dim3(TLH_THREADS), (size_t)lds, stream,
                           a, chunk);
      TLH_RETURN_IF_LAUNCH_FAILED();
      continue;
    }
    if (skipped) return TLH_EINVAL;  // only the two-slot pair kernel skips a type
    TLH_RETURN_IF_ERROR(allow_lds((const void*)diag_tile_kernel<Real>, lds));
    hipLaunchKernelGGL(diag_tile_kernel<Real>, dim3(g.ntiles), dim3(TLH_THREADS), (size_t)lds, stream,
                       a, chunk, alias[0], alias[1], alias[2]);
    TLH_RETURN_IF_LAUNCH_FAILED();
  }
  for (int ti = 0; ti < op->ntypes; ++ti) {
    const tlh_type_t& t = op->types[ti];
    const int64_t total = t.count * (t.dim * (t.dim + 1) / 2);
    if (total == 0 || !out[ti]) continue;
    if (t.dense_blocks != 0) return TLH_EINVAL;
    if (csr_heavy(t.csr_entries, t.count))
      hipLaunchKernelGGL(diag_finish_wave, dim3((unsigned)ceil_div(total, TLH_THREADS / 64)), dim3(TLH_THREADS), 0,
                         stream, out[ti], static_cast<const double*>(t.partials), t.csr_ptr, t.csr_idx, t.count,
                         t.dim);
    else
      hipLaunchKernelGGL(diag_finish_kernel, dim3((unsigned)ceil_div(total, TLH_THREADS)), dim3(TLH_THREADS), 0,
                         stream, out[ti], static_cast<const double*>(t.partials), t.csr_ptr, t.csr_idx,
                         t.count, t.dim);
    TLH_RETURN_IF_LAUNCH_FAILED();
  }
  return 0;
}

}  // namespace

extern "C" int tlh_diagonal_blocks(const tlh_operator_t* op, double* const* out, int64_t lds_limit,
                                   tlh_stream_t stream) {
  if (!op || !out || op->ntypes < 1) return TLH_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (op->dtype == TLH_F32) return diag_impl<float>(op, out, s, lds_limit);
  if (op->dtype == TLH_F64) return diag_impl<double>(op, out, s, lds_limit);
  return TLH_EINVAL;
}
