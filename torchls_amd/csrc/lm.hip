// float64 reductions of the Levenberg-Marquardt loop.
//
//  * tlh_rmatvec_f64: g = J^T v from the linearization's own blocks (reference
//    layout, any float dtype) in float64, deterministic (all-SPARSE tile plan).
//    This is the outer gradient that is recomputed in float64 every LM
//    iteration even when the PCG runs on float32 blocks.
//    Replaces Linearization.gradient / BlockJacobian.rmatvec (problem.py:40-42,
//    sparse.py:43-55).
//  * tlh_squared_matvec_norm: ||J x||^2 accumulated in float64, slots summed
//    before squaring (sparse.py:57-81).
//  * tlh_trial_objective: 0.5 sum rho(||r_new||^2) and the trial reduction
//    0.5 sum [rho(||r_old||^2) - rho(||r_new||^2)] evaluated through expanded
//    differences and the Huber divided difference (problem.py:315-348,
//    losses.py:80-107), weights optional.
#include <math.h>

#include "common.hpp"
#include "reduce.hpp"

using namespace tlh;

namespace {

// ------------------------------------------------------------ rmatvec (f64)
struct RmatvecArgs {
  const void* blocks[TLH_MAX_SLOTS];  // reference layout [count, rows, dim]
  const void* v;                      // [count * rows] in original observation order
  const int64_t* order;
  const int32_t* tile_start;
  int32_t ntiles;
  int32_t rows;
  int32_t nslots;
  int32_t block_dtype;
  int32_t v_dtype;
  int32_t dims[TLH_MAX_SLOTS];
  const uint16_t* rank[TLH_MAX_SLOTS];
  const int32_t* seg_ptr[TLH_MAX_SLOTS];
  const int32_t* seg_start[TLH_MAX_SLOTS];
  const int32_t* seg_target[TLH_MAX_SLOTS];
  double* partials[TLH_MAX_SLOTS];
};

__global__ void __launch_bounds__(TLH_THREADS) rmatvec_tile_kernel(RmatvecArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* stage = reinterpret_cast<double*>(smem_raw);
  const int t = blockIdx.x;
  const int start = a.tile_start[t];
  const int len = a.tile_start[t + 1] - start;
  const int R = a.rows;
  for (int s = 0; s < a.nslots; ++s) {
    const int D = a.dims[s];
    __syncthreads();
    for (int pos = threadIdx.x; pos < len; pos += TLH_THREADS) {
      const int64_t n = start + pos;
      const int64_t o = a.order[n];
      const int rk = a.rank[s] ? (int)a.rank[s][n] : pos;
      double vr[16];
      for (int r = 0; r < R; ++r) vr[r] = load_any(a.v, o * R + r, a.v_dtype);
      for (int c = 0; c < D; ++c) {
        double acc = 0.0;
        for (int r = 0; r < R; ++r) acc += load_any(a.blocks[s], (o * R + r) * D + c, a.block_dtype) * vr[r];
        stage[rk * D + c] = acc;
      }
    }
    __syncthreads();
    const int sb = a.seg_ptr[s][t], se = a.seg_ptr[s][t + 1];
    for (int k = sb + (int)threadIdx.x; k < se; k += TLH_THREADS) {
      const int lo = a.seg_start[s][k];
      const int hi = (k + 1 < se) ? a.seg_start[s][k + 1] : len;
      double* dst = a.partials[s] + (int64_t)a.seg_target[s][k] * D;
      for (int c = 0; c < D; ++c) {
        double sum = stage[lo * D + c];
        for (int i = lo + 1; i < hi; ++i) sum += stage[i * D + c];
        dst[c] = sum;
      }
    }
  }
}

// Fixed-shape rmatvec (float64 blocks and v, two slots, R rows): each
// observation's v row and block rows are read as 16-byte vectors in one burst;
// both slots' J^T v are formed once, then staged slot by slot for the same
// segment sums as rmatvec_tile_kernel.
// BT: block storage type (float64, or float32 blocks of a mixed-precision
// linearization); v and the sums stay float64.
template <typename BT, int R, int D0, int D1>
__global__ void __launch_bounds__(TLH_THREADS) rmatvec_pair_kernel(RmatvecArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* stage = reinterpret_cast<double*>(smem_raw);
  typedef double v2d __attribute__((ext_vector_type(2)));
  typedef BT v2b __attribute__((ext_vector_type(2)));
  constexpr int K0 = R * D0, K1 = R * D1, DM = D0 > D1 ? D0 : D1;
  static_assert(K0 % 2 == 0 && K1 % 2 == 0 && R % 2 == 0, "16-byte rows");
  const int t = blockIdx.x;
  const int start = a.tile_start[t];
  const int len = a.tile_start[t + 1] - start;
  constexpr int PER = 4;  // tile positions per thread (tile_cap <= 1024)
  double g0[PER][D0], g1[PER][D1];
#pragma unroll
  for (int it = 0; it < PER; ++it) {
    const int pos = (int)threadIdx.x + it * TLH_THREADS;
    if (pos >= len) continue;
    const int64_t o = a.order[start + pos];
    double vr[R], j0[K0], j1[K1];
    const v2d* pv = reinterpret_cast<const v2d*>(static_cast<const double*>(a.v) + o * R);
#pragma unroll
    for (int k = 0; k < R / 2; ++k) {
      const v2d v = pv[k];
      vr[2 * k] = v[0];
      vr[2 * k + 1] = v[1];
    }
#pragma unroll
    for (int k = 0; k < K0; ++k) j0[k] = 0.0;
#pragma unroll
    for (int k = 0; k < K1; ++k) j1[k] = 0.0;
    if (a.blocks[0]) {
      const v2b* p0 = reinterpret_cast<const v2b*>(static_cast<const BT*>(a.blocks[0]) + o * K0);
#pragma unroll
      for (int k = 0; k < K0 / 2; ++k) {
        const v2b v = p0[k];
        j0[2 * k] = (double)v[0];
        j0[2 * k + 1] = (double)v[1];
      }
    }
    if (a.blocks[1]) {
      const v2b* p1 = reinterpret_cast<const v2b*>(static_cast<const BT*>(a.blocks[1]) + o * K1);
#pragma unroll
      for (int k = 0; k < K1 / 2; ++k) {
        const v2b v = p1[k];
        j1[2 * k] = (double)v[0];
        j1[2 * k + 1] = (double)v[1];
      }
    }
#pragma unroll
    for (int c = 0; c < D0; ++c) {
      double acc = 0.0;
#pragma unroll
      for (int r = 0; r < R; ++r) acc += j0[r * D0 + c] * vr[r];
      g0[it][c] = acc;
    }
#pragma unroll
    for (int c = 0; c < D1; ++c) {
      double acc = 0.0;
#pragma unroll
      for (int r = 0; r < R; ++r) acc += j1[r * D1 + c] * vr[r];
      g1[it][c] = acc;
    }
  }
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int D = s == 0 ? D0 : D1;
    if (!a.blocks[s]) continue;  // slot reduced elsewhere (uniform over the workgroup)
    __syncthreads();
#pragma unroll
    for (int it = 0; it < PER; ++it) {
      const int pos = (int)threadIdx.x + it * TLH_THREADS;
      if (pos >= len) continue;
      const int64_t n = start + pos;
      const int rk = a.rank[s] ? (int)a.rank[s][n] : pos;
#pragma unroll
      for (int c = 0; c < DM; ++c)
        if (c < D) stage[rk * D + c] = s == 0 ? g0[it][c < D0 ? c : 0] : g1[it][c < D1 ? c : 0];
    }
    __syncthreads();
    const int sb = a.seg_ptr[s][t], se = a.seg_ptr[s][t + 1];
    for (int k = sb + (int)threadIdx.x; k < se; k += TLH_THREADS) {
      const int lo = a.seg_start[s][k];
      const int hi = (k + 1 < se) ? a.seg_start[s][k + 1] : len;
      double* dst = a.partials[s] + (int64_t)a.seg_target[s][k] * D;
      for (int c = 0; c < D; ++c) {
        double sum = stage[lo * D + c];
        for (int i = lo + 1; i < hi; ++i) sum += stage[i * D + c];
        dst[c] = sum;
      }
    }
  }
}

__global__ void __launch_bounds__(TLH_THREADS)
    rmatvec_finish_kernel(double* __restrict__ out, const double* __restrict__ partials,
                          const int32_t* __restrict__ csr_ptr, const int32_t* __restrict__ csr_idx,
                          int64_t count, int dim) {
  const int64_t i = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x;
  if (i >= count * dim) return;
  const int64_t v = i / dim;
  const int c = (int)(i - v * dim);
  double sum = 0.0;
  for (int k = csr_ptr[v]; k < csr_ptr[v + 1]; ++k) sum += partials[(int64_t)csr_idx[k] * dim + c];
  out[i] = sum;
}

__global__ void __launch_bounds__(TLH_THREADS)
    rmatvec_finish_wave(double* __restrict__ out, const double* __restrict__ partials,
                        const int32_t* __restrict__ csr_ptr, const int32_t* __restrict__ csr_idx, int64_t count,
                        int dim) {
  const int lane = threadIdx.x & 63;
  const int64_t e = (int64_t)blockIdx.x * (TLH_THREADS / 64) + (threadIdx.x >> 6);
  if (e >= count * dim) return;
  const int64_t v = e / dim;
  const int c = (int)(e - v * dim);
  const double s = csr_wave_sum<double>(partials, csr_idx, csr_ptr[v], csr_ptr[v + 1], dim, c, lane);
  if (lane == 0) out[e] = s;
}

// ------------------------------------------------------- ||J x||^2 (float64)
struct RawGroup {
  int64_t count;
  int32_t rows;
  int32_t nslots;
  int32_t dims[TLH_MAX_SLOTS];
  int64_t offsets[TLH_MAX_SLOTS];
  const int64_t* index[TLH_MAX_SLOTS];
  const void* blocks[TLH_MAX_SLOTS];
};

__global__ void __launch_bounds__(TLH_THREADS)
    sqnorm_kernel(RawGroup g, int block_dtype, const void* x, int x_dtype, double* out, double* partials,
                  uint32_t* counter) {
  __shared__ double smem[RED_SMEM];
  double v[2] = {0.0, 0.0};
  const int R = g.rows;
  for (int64_t o = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x; o < g.count;
       o += (int64_t)gridDim.x * TLH_THREADS) {
    double q[16];
    for (int r = 0; r < R; ++r) q[r] = 0.0;
    for (int s = 0; s < g.nslots; ++s) {
      const int D = g.dims[s];
      const int64_t col = g.offsets[s] + g.index[s][o] * D;
      for (int c = 0; c < D; ++c) {
        const double xv = load_any(x, col + c, x_dtype);
        for (int r = 0; r < R; ++r) q[r] += load_any(g.blocks[s], (o * R + r) * D + c, block_dtype) * xv;
      }
    }
    for (int r = 0; r < R; ++r) v[0] += q[r] * q[r];
  }
  grid_sum2(v, partials, counter, out, smem);
}

// sqnorm_kernel for two-slot groups of a fixed shape: the same products in the
// same order, fully unrolled, so an observation's block and x loads are all in
// flight together instead of one runtime-indexed load at a time.
template <int R, int D0, int D1, typename BT, typename XT>
__global__ void __launch_bounds__(TLH_THREADS)
    sqnorm_pair_kernel(RawGroup g, const void* x, double* out, double* partials, uint32_t* counter) {
  __shared__ double smem[RED_SMEM];
  const BT* __restrict__ b0 = static_cast<const BT*>(g.blocks[0]);
  const BT* __restrict__ b1 = static_cast<const BT*>(g.blocks[1]);
  const XT* __restrict__ xv = static_cast<const XT*>(x);
  double v[2] = {0.0, 0.0};
  for (int64_t o = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x; o < g.count;
       o += (int64_t)gridDim.x * TLH_THREADS) {
    const int64_t c0 = g.offsets[0] + g.index[0][o] * D0;
    const int64_t c1 = g.offsets[1] + g.index[1][o] * D1;
    double x0[D0], x1[D1], q[R];
#pragma unroll
    for (int c = 0; c < D0; ++c) x0[c] = (double)xv[c0 + c];
#pragma unroll
    for (int c = 0; c < D1; ++c) x1[c] = (double)xv[c1 + c];
#pragma unroll
    for (int r = 0; r < R; ++r) q[r] = 0.0;
#pragma unroll
    for (int c = 0; c < D0; ++c)
#pragma unroll
      for (int r = 0; r < R; ++r) q[r] += (double)b0[(o * R + r) * D0 + c] * x0[c];
#pragma unroll
    for (int c = 0; c < D1; ++c)
#pragma unroll
      for (int r = 0; r < R; ++r) q[r] += (double)b1[(o * R + r) * D1 + c] * x1[c];
#pragma unroll
    for (int r = 0; r < R; ++r) v[0] += q[r] * q[r];
  }
  grid_sum2(v, partials, counter, out, smem);
}

template <int R, int D0, int D1>
bool launch_sqnorm_pair(const RawGroup& g, int block_dtype, const void* x, int x_dtype, double* out2,
                        double* partials, uint32_t* counter, hipStream_t s) {
  if (g.nslots != 2 || g.rows != R || g.dims[0] != D0 || g.dims[1] != D1) return false;
  void (*kernel)(RawGroup, const void*, double*, double*, uint32_t*);
  if (block_dtype == TLH_F64)
    kernel = x_dtype == TLH_F64 ? sqnorm_pair_kernel<R, D0, D1, double, double>
                                : sqnorm_pair_kernel<R, D0, D1, double, float>;
  else
    kernel = x_dtype == TLH_F64 ? sqnorm_pair_kernel<R, D0, D1, float, double>
                                : sqnorm_pair_kernel<R, D0, D1, float, float>;
  hipLaunchKernelGGL(kernel, dim3(red_grid(g.count)), dim3(TLH_THREADS), 0, s, g, x, out2, partials, counter);
  return true;
}

// y[o*R + r] = sum_s J_s[o] x_s in float64, written in the output dtype (sparse.py:31-41).
__global__ void __launch_bounds__(TLH_THREADS)
    block_matvec_kernel(RawGroup g, int block_dtype, const void* x, int x_dtype, void* y, int y_dtype) {
  const int R = g.rows;
  const int64_t o = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x;
  if (o >= g.count) return;
  double q[16];
  for (int r = 0; r < R; ++r) q[r] = 0.0;
  for (int s = 0; s < g.nslots; ++s) {
    const int D = g.dims[s];
    const int64_t col = g.offsets[s] + g.index[s][o] * D;
    for (int c = 0; c < D; ++c) {
      const double xv = load_any(x, col + c, x_dtype);
      for (int r = 0; r < R; ++r) q[r] += load_any(g.blocks[s], (o * R + r) * D + c, block_dtype) * xv;
    }
  }
  for (int r = 0; r < R; ++r) {
    if (y_dtype == TLH_F64)
      static_cast<double*>(y)[o * R + r] = q[r];
    else
      static_cast<float*>(y)[o * R + r] = (float)q[r];
  }
}

__global__ void __launch_bounds__(TLH_THREADS)
    trial_kernel(const void* before, int before_dtype, const void* after, int after_dtype, int64_t count, int rows,
                 int loss, double delta, const double* weights, int weight_stride, double* out, double* partials,
                 uint32_t* counter) {
  __shared__ double smem[RED_SMEM];
  double v[2] = {0.0, 0.0};
  for (int64_t o = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x; o < count;
       o += (int64_t)gridDim.x * TLH_THREADS) {
    double s_old = 0.0, s_new = 0.0, diff = 0.0;
    for (int r = 0; r < rows; ++r) {
      const double a = before ? load_any(before, o * rows + r, before_dtype) : 0.0;
      const double b = load_any(after, o * rows + r, after_dtype);
      s_old += a * a;
      s_new += b * b;
      diff += (a - b) * (a + b);
    }
    double value, change;
    if (loss == 1) {
      value = huber_value(s_new, delta);
      change = huber_reduction(s_old, s_new, diff, delta);
    } else {
      value = s_new;
      change = diff;
    }
    if (weights) {
      const double w = weights[o * weight_stride];
      value *= w;
      change *= w;
    }
    v[0] += value;
    v[1] += change;
  }
  grid_sum2(v, partials, counter, out, smem);
}

// |x| folded into a running maximum that turns NaN for good once a NaN is seen
// (comparisons with NaN are false), like torch's max of abs.
__device__ __forceinline__ double max_abs_nan(double m, double x) {
  const double a = fabs(x);
  return (a > m || a != a) ? a : m;
}

// Four statistics of two vectors in one pass (a or b may be NULL): out = {a.b,
// b.b, max|b|, max|a|}; sums in a fixed order (thread strides, tree, partials in
// block order by the last block), maxima NaN-propagating.  At most RED_BLOCKS/2
// blocks (vector_grid gives at most 128), so the four partials of a block fit
// the two-sum partials buffer.
__global__ void __launch_bounds__(TLH_THREADS)
    vector_stats_kernel(const void* a, int a_dtype, const void* b, int b_dtype, int64_t n, double* out,
                        double* partials, uint32_t* counter) {
  constexpr int NWAVES = TLH_THREADS / 64;
  __shared__ double smem[4][NWAVES];
  __shared__ int last;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * TLH_THREADS) {
    const double x = a ? load_any(a, i, a_dtype) : 0.0;
    const double y = b ? load_any(b, i, b_dtype) : 0.0;
    v[0] += x * y;
    v[1] += y * y;
    v[2] = max_abs_nan(v[2], y);
    v[3] = max_abs_nan(v[3], x);
  }
  // xor-shuffle tree inside each wavefront, then the wavefronts in order
  auto tree = [&](double (&w)[4]) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      w[0] += __shfl_xor(w[0], off);
      w[1] += __shfl_xor(w[1], off);
      w[2] = max_abs_nan(w[2], __shfl_xor(w[2], off));
      w[3] = max_abs_nan(w[3], __shfl_xor(w[3], off));
    }
    if ((threadIdx.x & 63) == 0)
#pragma unroll
      for (int k = 0; k < 4; ++k) smem[k][threadIdx.x >> 6] = w[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = smem[k][0];
#pragma unroll
    for (int wv = 1; wv < NWAVES; ++wv) {
      w[0] += smem[0][wv];
      w[1] += smem[1][wv];
      w[2] = max_abs_nan(w[2], smem[2][wv]);
      w[3] = max_abs_nan(w[3], smem[3][wv]);
    }
    __syncthreads();
  };
  tree(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) partials[4 * blockIdx.x + k] = v[k];
    const uint32_t t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = t == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  double w[4] = {0.0, 0.0, 0.0, 0.0};
  for (int blk = threadIdx.x; blk < (int)gridDim.x; blk += TLH_THREADS) {
    w[0] += __hip_atomic_load(&partials[4 * blk + 0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    w[1] += __hip_atomic_load(&partials[4 * blk + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    w[2] = max_abs_nan(w[2], __hip_atomic_load(&partials[4 * blk + 2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    w[3] = max_abs_nan(w[3], __hip_atomic_load(&partials[4 * blk + 3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  }
  tree(w);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = w[k];
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// dst[di(i)] = scale * src[si(i)], si(i) = sidx ? sidx[i] : soff + i (likewise di),
// converted through float64.
template <typename S, typename D>
__global__ void __launch_bounds__(TLH_THREADS)
    copy_cast_kernel(const S* __restrict__ src, const int64_t* __restrict__ sidx, int64_t soff, D* __restrict__ dst,
                     const int64_t* __restrict__ didx, int64_t doff, int64_t count, double scale) {
  const int64_t i = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x;
  if (i >= count) return;
  const int64_t si = sidx ? sidx[i] : soff + i;
  const int64_t di = didx ? didx[i] : doff + i;
  dst[di] = (D)(scale * (double)src[si]);
}

}  // namespace

extern "C" int tlh_rmatvec_f64(const tlh_operator_t* op, int group, const void* const* blocks, int block_dtype,
                               const void* v, int v_dtype, int64_t lds_limit, tlh_stream_t stream) {
  if (!op || group < 0 || group >= op->ngroups || !blocks || !v) return TLH_EINVAL;
  const tlh_group_t& g = op->groups[group];
  if (g.count == 0) return 0;
  if (g.rows > 16) return TLH_EINVAL;
  RmatvecArgs a{};
  int dmax = 1;
  bool skipped = false;
  for (int s = 0; s < g.nslots; ++s) {
    const tlh_slot_t& sl = g.slot[s];
    if (sl.mode != TLH_REDUCE_SPARSE) return TLH_EINVAL;
    a.blocks[s] = blocks[s];
    skipped = skipped || !blocks[s];
    a.dims[s] = sl.dim;
    a.rank[s] = sl.rank;
    a.seg_ptr[s] = sl.seg_ptr;
    a.seg_start[s] = sl.seg_start;
    a.seg_target[s] = sl.seg_target;
    a.partials[s] = static_cast<double*>(op->types[sl.type].partials);
    if (blocks[s]) dmax = sl.dim > dmax ? sl.dim : dmax;
  }
  a.v = v;
  a.order = g.order;
  a.tile_start = g.tile_start;
  a.ntiles = g.ntiles;
  a.rows = g.rows;
  a.nslots = g.nslots;
  a.block_dtype = block_dtype;
  a.v_dtype = v_dtype;
  const int64_t lds = (int64_t)g.tile_cap * dmax * 8;
  if (lds > lds_limit) return TLH_ELDS;
  const uintptr_t bal = block_dtype == TLH_F64 ? 15 : 7;  // 2-element vector loads of the blocks
  const bool pair = g.nslots == 2 && g.rows == 2 && (block_dtype == TLH_F64 || block_dtype == TLH_F32) &&
                    v_dtype == TLH_F64 && g.tile_cap <= 4 * TLH_THREADS && ((uintptr_t)v & 15) == 0 &&
                    ((uintptr_t)blocks[0] & bal) == 0 && ((uintptr_t)blocks[1] & bal) == 0;
  if (pair && ((a.dims[0] == 6 && a.dims[1] == 3) || (a.dims[0] == 3 && a.dims[1] == 6))) {
    const bool f64 = block_dtype == TLH_F64, c0 = a.dims[0] == 6;
    const void* kernel = f64 ? (c0 ? (const void*)rmatvec_pair_kernel<double, 2, 6, 3>
                                   : (const void*)rmatvec_pair_kernel<double, 2, 3, 6>)
                             : (c0 ? (const void*)rmatvec_pair_kernel<float, 2, 6, 3>
                                   : (const void*)rmatvec_pair_kernel<float, 2, 3, 6>);
    TLH_RETURN_IF_ERROR(allow_lds(kernel, lds));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (f64 && c0)
      hipLaunchKernelGGL((rmatvec_pair_kernel<double, 2, 6, 3>), dim3(g.ntiles), dim3(TLH_THREADS), (size_t)lds, s, a);
    else if (f64)
      hipLaunchKernelGGL((rmatvec_pair_kernel<double, 2, 3, 6>), dim3(g.ntiles), dim3(TLH_THREADS), (size_t)lds, s, a);
    else if (c0)
      hipLaunchKernelGGL((rmatvec_pair_kernel<float, 2, 6, 3>), dim3(g.ntiles), dim3(TLH_THREADS), (size_t)lds, s, a);
    else
      hipLaunchKernelGGL((rmatvec_pair_kernel<float, 2, 3, 6>), dim3(g.ntiles), dim3(TLH_THREADS), (size_t)lds, s, a);
    TLH_RETURN_IF_LAUNCH_FAILED();
    return 0;
  }
  if (skipped) return TLH_EINVAL;  // only the two-slot pair kernel skips a slot
  TLH_RETURN_IF_ERROR(allow_lds((const void*)rmatvec_tile_kernel, lds));
  hipLaunchKernelGGL(rmatvec_tile_kernel, dim3(g.ntiles), dim3(TLH_THREADS), (size_t)lds,
                     reinterpret_cast<hipStream_t>(stream), a);
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

extern "C" int tlh_rmatvec_finish_f64(const tlh_operator_t* op, double* out, tlh_stream_t stream) {
  if (!op) return TLH_EINVAL;
  return tlh_rmatvec_finish_range_f64(op, 0, op->ntypes, out, stream);
}

extern "C" int tlh_rmatvec_finish_range_f64(const tlh_operator_t* op, int type_begin, int type_end, double* out,
                                            tlh_stream_t stream) {
  if (!op || !out || type_begin < 0 || type_end > op->ntypes || type_begin > type_end) return TLH_EINVAL;
  for (int ti = type_begin; ti < type_end; ++ti) {
    const tlh_type_t& t = op->types[ti];
    const int64_t total = t.count * t.dim;
    if (total == 0) continue;
    if (csr_heavy(t.csr_entries, t.count))
      hipLaunchKernelGGL(rmatvec_finish_wave, dim3((unsigned)ceil_div(total, TLH_THREADS / 64)), dim3(TLH_THREADS),
                         0, reinterpret_cast<hipStream_t>(stream), out + t.offset,
                         static_cast<const double*>(t.partials), t.csr_ptr, t.csr_idx, t.count, t.dim);
    else
      hipLaunchKernelGGL(rmatvec_finish_kernel, dim3((unsigned)ceil_div(total, TLH_THREADS)), dim3(TLH_THREADS), 0,
                         reinterpret_cast<hipStream_t>(stream), out + t.offset,
                         static_cast<const double*>(t.partials), t.csr_ptr, t.csr_idx, t.count, t.dim);
    TLH_RETURN_IF_LAUNCH_FAILED();
  }
  return 0;
}

extern "C" int tlh_squared_matvec_norm(int64_t count, int rows, int nslots, const int32_t* dims,
                                       const int64_t* offsets, const int64_t* const* index,
                                       const void* const* blocks, int block_dtype, const void* x, int x_dtype,
                                       double* out2, double* partials, uint32_t* counter, tlh_stream_t stream) {
  if (nslots < 1 || nslots > TLH_MAX_SLOTS || rows < 1 || rows > 16 || !out2 || !partials || !counter)
    return TLH_EINVAL;
  RawGroup g{};
  g.count = count;
  g.rows = rows;
  g.nslots = nslots;
  for (int s = 0; s < nslots; ++s) {
    g.dims[s] = dims[s];
    g.offsets[s] = offsets[s];
    g.index[s] = index[s];
    g.blocks[s] = blocks[s];
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (!launch_sqnorm_pair<2, 6, 3>(g, block_dtype, x, x_dtype, out2, partials, counter, s) &&
      !launch_sqnorm_pair<2, 3, 6>(g, block_dtype, x, x_dtype, out2, partials, counter, s) &&
      !launch_sqnorm_pair<6, 6, 6>(g, block_dtype, x, x_dtype, out2, partials, counter, s))
    hipLaunchKernelGGL(sqnorm_kernel, dim3(red_grid(count)), dim3(TLH_THREADS), 0, s, g, block_dtype, x, x_dtype,
                       out2, partials, counter);
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

extern "C" int tlh_trial_objective(const void* before, int before_dtype, const void* after, int after_dtype,
                                   int64_t count, int rows, int loss, double delta, const double* weights,
                                   int weight_stride, double* out2, double* partials, uint32_t* counter,
                                   tlh_stream_t stream) {
  if (!after || !out2 || !partials || !counter || rows < 1 || loss < 0 || loss > 1) return TLH_EINVAL;
  hipLaunchKernelGGL(trial_kernel, dim3(red_grid(count)), dim3(TLH_THREADS), 0,
                     reinterpret_cast<hipStream_t>(stream), before, before_dtype, after, after_dtype, count, rows,
                     loss, delta, weights, weight_stride, out2, partials, counter);
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

extern "C" int tlh_block_matvec(int64_t count, int rows, int nslots, const int32_t* dims, const int64_t* offsets,
                                const int64_t* const* index, const void* const* blocks, int block_dtype,
                                const void* x, int x_dtype, void* y, int y_dtype, tlh_stream_t stream) {
  if (nslots < 1 || nslots > TLH_MAX_SLOTS || rows < 1 || rows > 16 || !y || !x) return TLH_EINVAL;
  if (count == 0) return 0;
  RawGroup g{};
  g.count = count;
  g.rows = rows;
  g.nslots = nslots;
  for (int s = 0; s < nslots; ++s) {
    g.dims[s] = dims[s];
    g.offsets[s] = offsets[s];
    g.index[s] = index[s];
    g.blocks[s] = blocks[s];
  }
  hipLaunchKernelGGL(block_matvec_kernel, dim3((unsigned)ceil_div(count, TLH_THREADS)), dim3(TLH_THREADS), 0,
                     reinterpret_cast<hipStream_t>(stream), g, block_dtype, x, x_dtype, y, y_dtype);
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

extern "C" int tlh_copy_cast(const void* src, int src_dtype, const int64_t* src_index, int64_t src_offset, void* dst,
                             int dst_dtype, const int64_t* dst_index, int64_t dst_offset, int64_t count, double scale,
                             tlh_stream_t stream) {
  const bool ok = (src_dtype == TLH_F32 || src_dtype == TLH_F64) && (dst_dtype == TLH_F32 || dst_dtype == TLH_F64);
  if (count < 0 || !ok) return TLH_EINVAL;
  if (count == 0) return 0;
  if (!src || !dst) return TLH_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid((unsigned)ceil_div(count, TLH_THREADS)), block(TLH_THREADS);
  if (src_dtype == TLH_F64 && dst_dtype == TLH_F64)
    hipLaunchKernelGGL((copy_cast_kernel<double, double>), grid, block, 0, s, static_cast<const double*>(src),
                       src_index, src_offset, static_cast<double*>(dst), dst_index, dst_offset, count, scale);
  else if (src_dtype == TLH_F64)
    hipLaunchKernelGGL((copy_cast_kernel<double, float>), grid, block, 0, s, static_cast<const double*>(src),
                       src_index, src_offset, static_cast<float*>(dst), dst_index, dst_offset, count, scale);
  else if (dst_dtype == TLH_F64)
    hipLaunchKernelGGL((copy_cast_kernel<float, double>), grid, block, 0, s, static_cast<const float*>(src),
                       src_index, src_offset, static_cast<double*>(dst), dst_index, dst_offset, count, scale);
  else
    hipLaunchKernelGGL((copy_cast_kernel<float, float>), grid, block, 0, s, static_cast<const float*>(src),
                       src_index, src_offset, static_cast<float*>(dst), dst_index, dst_offset, count, scale);
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

extern "C" int tlh_vector_stats(const void* a, int a_dtype, const void* b, int b_dtype, int64_t n, double* out,
                                double* partials, uint32_t* counter, tlh_stream_t stream) {
  if (n < 1 || !out || !partials || !counter || (!a && !b)) return TLH_EINVAL;
  if ((a && a_dtype != TLH_F32 && a_dtype != TLH_F64) || (b && b_dtype != TLH_F32 && b_dtype != TLH_F64))
    return TLH_EINVAL;
  const int g = vector_grid(n);  // at most 128 <= RED_BLOCKS / 2 workgroups
  hipLaunchKernelGGL(vector_stats_kernel, dim3((unsigned)g), dim3(TLH_THREADS), 0,
                     reinterpret_cast<hipStream_t>(stream), a, a_dtype, b, b_dtype, n, out, partials, counter);
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

namespace {

// x^3 with both products carried as double-double and one final rounding, as the
// correctly rounded pow(x, 3) of the host's libm gives it.
__device__ __forceinline__ double cube_rn(double x) {
  const double h = x * x, hl = fma(x, x, -h);
  const double c = h * x, cl = fma(h, x, -c) + hl * x;
  return c + cl;
}

// One thread: the trial arithmetic of the LM loop in the reference's operations and
// order (no contraction).  Every decision and the factor rule's damping equal a
// host computation bit for bit; the adaptive rule's multiplier differs from one
// through libm's pow(x, 3) (accurate to about half an ulp) by at most an ulp of
// its cube.
__global__ void lm_decide_kernel(tlh_lm_trial_t tr, int has_trial, double* state, tlh_lm_trust_t tt, double* out) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0) return;
  const double damping = state[0], rej = state[1];
  const double nan = __builtin_nan("");
  double predicted = nan, quality = nan, step_norm = nan;
  bool usable = false, accepted = false;
  if (has_trial) {
    const double cost = *tr.cost, actual = *tr.actual, sq = *tr.sq;
    const double gd = tr.stats[0], dd = tr.stats[1];
    usable = isfinite(*tr.finite);
    step_norm = dd >= 0.0 ? sqrt(dd) : nan;
    predicted = -gd - 0.5 * sq;
    if (usable) {
      quality = predicted > 0.0 ? actual / predicted : -__builtin_inf();
      accepted = isfinite(cost) && (!tt.enabled || (actual > 0.0 && quality >= tt.quality_min));
    }
  }
  double nd = damping, nr = rej;
  bool stop = false;
  if (accepted) {
    if (tt.enabled) {
      if (tt.adaptive) {
        const double mult = fmax(1.0 / 3.0, 1.0 - cube_rn(2.0 * fmin(quality, 1.0) - 1.0));
        nd = fmax(tt.lambda_min, fmin(tt.lambda_max, damping * mult));
        nr = 2.0;
      } else if (quality > 0.75) {
        nd = fmax(tt.lambda_min, damping / tt.lambda_factor);
      }
    }
  } else {
    stop = !tt.enabled || damping >= tt.lambda_max;
    if (!stop) {
      if (tt.adaptive) {
        nd = fmin(tt.lambda_max, damping * rej);
        nr = fmin(rej * 2.0, tt.lambda_max / tt.lambda_min);
      } else {
        nd = fmin(tt.lambda_max, damping * tt.lambda_factor);
      }
    }
  }
  state[0] = nd;
  state[1] = nr;
  out[0] = predicted;
  out[1] = quality;
  out[2] = accepted ? 1.0 : 0.0;
  out[3] = usable ? 1.0 : 0.0;
  out[4] = step_norm;
  out[5] = damping;
  out[6] = stop ? 1.0 : 0.0;
  out[7] = nd;
  out[8] = nr;
}

}  // namespace

extern "C" int tlh_lm_decide(const tlh_lm_trial_t* trial, double* state, const tlh_lm_trust_t* trust, double* out,
                             tlh_stream_t stream) {
  if (!state || !trust || !out) return TLH_EINVAL;
  if (trial && (!trial->cost || !trial->actual || !trial->sq || !trial->stats || !trial->finite)) return TLH_EINVAL;
  if (trust->enabled && !(trust->lambda_min > 0.0 && trust->lambda_max >= trust->lambda_min)) return TLH_EINVAL;
  const tlh_lm_trial_t tr = trial ? *trial : tlh_lm_trial_t{};
  hipLaunchKernelGGL(lm_decide_kernel, dim3(1), dim3(64), 0, reinterpret_cast<hipStream_t>(stream), tr,
                     trial ? 1 : 0, state, *trust, out);
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}
