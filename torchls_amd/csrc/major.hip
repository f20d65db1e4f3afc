// Variable-major block accumulations for a type with few variables and many
// observations (the cameras of a bundle adjustment).
//
// The tile kernels of csrc/diagonal.hip and csrc/schur.hip walk observations in
// the order of the group's most numerous variable (points), so a camera block
// reaches every tile it is observed in and each (tile, camera) segment writes a
// float64 partial: 21 values per segment, hundreds of MB at millions of
// observations, written once and read back once.  Here observations are visited
// in camera order instead, through a permutation of the raw [N, R, D] blocks
// that the linearization wrote: one 256-thread workgroup per camera sums its
// observations in registers and reduces them in a fixed order, so each block
// row is read once and nothing but the final blocks is written.
//
//   tlh_major_gram          D_v = sum_{obs of v} J^T J            (sparse.py:104-121)
//   tlh_major_rmatvec       g_v = sum_{obs of v} J^T r            (sparse.py:43-55, problem.py:40-42)
//   tlh_schur_precond_major M_c = D_c + lambda I - sum_{edges e of c} W_e C_p^-1 W_e^T,
//                           W_e = sum_{obs in e} J_c^T J_p        (schur.py:353-362)
//
// Summation order is fixed: per thread in permutation order (stride 256), then
// an xor-shuffle tree inside each wavefront, then the four wavefront sums in
// order.  Same result every run.
#include "common.hpp"

using namespace tlh;

namespace {

constexpr int MAJOR_THREADS = 256;

// Fixed-order sum of NV values over the workgroup; the totals land in
// smem[NWAVES * NV .. NWAVES * NV + NV).
template <int NV>
__device__ __forceinline__ void major_block_sum(double (&v)[NV], double* smem) {
  constexpr int NWAVES = MAJOR_THREADS / 64;
#pragma unroll
  for (int k = 0; k < NV; ++k)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off);
  const int wave = (int)threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) smem[wave * NV + k] = v[k];
  __syncthreads();
  if ((int)threadIdx.x < NV) {
    double s = smem[threadIdx.x];
#pragma unroll
    for (int w = 1; w < NWAVES; ++w) s += smem[w * NV + threadIdx.x];
    smem[NWAVES * NV + threadIdx.x] = s;
  }
  __syncthreads();
}

// Row-major upper-triangle index of (i, j), i <= j, in a D x D block.
template <int D>
__host__ __device__ constexpr int upper(int i, int j) {
  return i * D - i * (i - 1) / 2 + (j - i);
}

template <typename BT, int R, int D>
__global__ void __launch_bounds__(MAJOR_THREADS)
    major_gram_kernel(const BT* __restrict__ blocks, const int32_t* __restrict__ order,
                      const int32_t* __restrict__ ptr, double* __restrict__ out) {
  constexpr int W = D * (D + 1) / 2;
  __shared__ double smem[(MAJOR_THREADS / 64 + 1) * W];
  const int v = blockIdx.x;
  const int k0 = ptr[v], k1 = ptr[v + 1];
  double acc[W];
#pragma unroll
  for (int e = 0; e < W; ++e) acc[e] = 0.0;
#pragma unroll 4
  for (int k = k0 + (int)threadIdx.x; k < k1; k += MAJOR_THREADS) {
    const BT* row = blocks + (int64_t)order[k] * (R * D);
    double j[R][D];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int c = 0; c < D; ++c) j[r][c] = (double)row[r * D + c];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int b = a; b < D; ++b) {
        double s = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) s += j[r][a] * j[r][b];
        acc[upper<D>(a, b)] += s;
      }
  }
  major_block_sum<W>(acc, smem);
  const double* tot = smem + (MAJOR_THREADS / 64) * W;
  for (int e = threadIdx.x; e < D * D; e += MAJOR_THREADS) {
    const int i = e / D, c = e % D;
    out[(int64_t)v * D * D + e] = tot[i <= c ? upper<D>(i, c) : upper<D>(c, i)];
  }
}

template <typename BT, int R, int DC, int DP>
__global__ void __launch_bounds__(MAJOR_THREADS)
    major_precond_kernel(const BT* __restrict__ jc, const BT* __restrict__ jp, const int32_t* __restrict__ order,
                         const int32_t* __restrict__ edge_ptr, const int32_t* __restrict__ var_edge_ptr,
                         const int32_t* __restrict__ edge_point, const double* __restrict__ cinv,
                         const double* __restrict__ D, double damping, double* __restrict__ M) {
  constexpr int W = DC * (DC + 1) / 2;
  constexpr int WP = DP * (DP + 1) / 2;
  __shared__ double smem[(MAJOR_THREADS / 64 + 1) * W];
  const int v = blockIdx.x;
  const int e0 = var_edge_ptr[v], e1 = var_edge_ptr[v + 1];
  double acc[W];
#pragma unroll
  for (int e = 0; e < W; ++e) acc[e] = 0.0;
#pragma unroll 2
  for (int e = e0 + (int)threadIdx.x; e < e1; e += MAJOR_THREADS) {
    double cm[WP];
    const double* mp = cinv + (int64_t)edge_point[e] * WP;
#pragma unroll
    for (int k = 0; k < WP; ++k) cm[k] = mp[k];
    double w[DC][DP];
#pragma unroll
    for (int i = 0; i < DC; ++i)
#pragma unroll
      for (int l = 0; l < DP; ++l) w[i][l] = 0.0;
    const int m1 = edge_ptr[e + 1];
    for (int m = edge_ptr[e]; m < m1; ++m) {
      const int64_t o = order[m];
      const BT* rc = jc + o * (R * DC);
      const BT* rp = jp + o * (R * DP);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        double q[DP];
#pragma unroll
        for (int l = 0; l < DP; ++l) q[l] = (double)rp[r * DP + l];
#pragma unroll
        for (int i = 0; i < DC; ++i) {
          const double ci = (double)rc[r * DC + i];
#pragma unroll
          for (int l = 0; l < DP; ++l) w[i][l] += ci * q[l];
        }
      }
    }
    // W C_p^-1 W^T (upper triangle)
    double wm[DC][DP];
#pragma unroll
    for (int i = 0; i < DC; ++i)
#pragma unroll
      for (int l = 0; l < DP; ++l) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < DP; ++k) s += w[i][k] * cm[k <= l ? upper<DP>(k, l) : upper<DP>(l, k)];
        wm[i][l] = s;
      }
#pragma unroll
    for (int i = 0; i < DC; ++i)
#pragma unroll
      for (int j = i; j < DC; ++j) {
        double s = 0.0;
#pragma unroll
        for (int l = 0; l < DP; ++l) s += wm[i][l] * w[j][l];
        acc[upper<DC>(i, j)] += s;
      }
  }
  major_block_sum<W>(acc, smem);
  const double* tot = smem + (MAJOR_THREADS / 64) * W;
  for (int e = threadIdx.x; e < DC * DC; e += MAJOR_THREADS) {
    const int i = e / DC, c = e % DC;
    const double s = tot[i <= c ? upper<DC>(i, c) : upper<DC>(c, i)];
    const int64_t at = (int64_t)v * DC * DC + e;
    M[at] = D[at] + (i == c ? damping : 0.0) - s;
  }
}

// g_v = sum over the variable's observations of J^T v (float64), one workgroup
// per variable, the same fixed order as major_gram_kernel.
template <typename BT, typename VT, int R, int D>
__global__ void __launch_bounds__(MAJOR_THREADS)
    major_rmatvec_kernel(const BT* __restrict__ blocks, const VT* __restrict__ vec, const int32_t* __restrict__ order,
                         const int32_t* __restrict__ ptr, double* __restrict__ out) {
  __shared__ double smem[(MAJOR_THREADS / 64 + 1) * D];
  const int v = blockIdx.x;
  const int k0 = ptr[v], k1 = ptr[v + 1];
  double acc[D];
#pragma unroll
  for (int c = 0; c < D; ++c) acc[c] = 0.0;
#pragma unroll 4
  for (int k = k0 + (int)threadIdx.x; k < k1; k += MAJOR_THREADS) {
    const int64_t o = order[k];
    const BT* row = blocks + o * (R * D);
    double vr[R];
#pragma unroll
    for (int r = 0; r < R; ++r) vr[r] = (double)vec[o * R + r];
#pragma unroll
    for (int c = 0; c < D; ++c) {
      double s = 0.0;
#pragma unroll
      for (int r = 0; r < R; ++r) s += (double)row[r * D + c] * vr[r];
      acc[c] += s;
    }
  }
  major_block_sum<D>(acc, smem);
  if ((int)threadIdx.x < D) out[(int64_t)v * D + threadIdx.x] = smem[(MAJOR_THREADS / 64) * D + threadIdx.x];
}

template <int R, int D>
void launch_rmatvec(const void* blocks, int block_dtype, const void* vec, int v_dtype, const int32_t* order,
                    const int32_t* ptr, int64_t count, double* out, hipStream_t s) {
  const dim3 grid((unsigned)count), block(MAJOR_THREADS);
  if (block_dtype == TLH_F64 && v_dtype == TLH_F64)
    hipLaunchKernelGGL((major_rmatvec_kernel<double, double, R, D>), grid, block, 0, s,
                       static_cast<const double*>(blocks), static_cast<const double*>(vec), order, ptr, out);
  else if (block_dtype == TLH_F64)
    hipLaunchKernelGGL((major_rmatvec_kernel<double, float, R, D>), grid, block, 0, s,
                       static_cast<const double*>(blocks), static_cast<const float*>(vec), order, ptr, out);
  else if (v_dtype == TLH_F64)
    hipLaunchKernelGGL((major_rmatvec_kernel<float, double, R, D>), grid, block, 0, s,
                       static_cast<const float*>(blocks), static_cast<const double*>(vec), order, ptr, out);
  else
    hipLaunchKernelGGL((major_rmatvec_kernel<float, float, R, D>), grid, block, 0, s,
                       static_cast<const float*>(blocks), static_cast<const float*>(vec), order, ptr, out);
}

template <int R, int D>
bool launch_gram(const void* blocks, int block_dtype, const int32_t* order, const int32_t* ptr, int64_t count,
                 double* out, hipStream_t s) {
  if (block_dtype == TLH_F64)
    hipLaunchKernelGGL((major_gram_kernel<double, R, D>), dim3((unsigned)count), dim3(MAJOR_THREADS), 0, s,
                       static_cast<const double*>(blocks), order, ptr, out);
  else
    hipLaunchKernelGGL((major_gram_kernel<float, R, D>), dim3((unsigned)count), dim3(MAJOR_THREADS), 0, s,
                       static_cast<const float*>(blocks), order, ptr, out);
  return true;
}

}  // namespace

extern "C" int tlh_major_gram(const void* blocks, int block_dtype, int rows, int dim, const int32_t* order,
                              const int32_t* ptr, int64_t count, double* out, tlh_stream_t stream) {
  if (count < 0 || count > INT32_MAX || (block_dtype != TLH_F64 && block_dtype != TLH_F32)) return TLH_EINVAL;
  if (count == 0) return 0;
  if (!blocks || !order || !ptr || !out) return TLH_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (rows == 2 && dim == 6)
    launch_gram<2, 6>(blocks, block_dtype, order, ptr, count, out, s);
  else if (rows == 2 && dim == 3)
    launch_gram<2, 3>(blocks, block_dtype, order, ptr, count, out, s);
  else if (rows == 6 && dim == 6)
    launch_gram<6, 6>(blocks, block_dtype, order, ptr, count, out, s);
  else
    return TLH_ENOTFOUND;
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

extern "C" int tlh_schur_precond_major(const void* jc, const void* jp, int block_dtype, int rows, int dc, int dp,
                                       const int32_t* order, const int32_t* edge_ptr, const int32_t* var_edge_ptr,
                                       const int32_t* edge_point, int64_t count, const double* cinv,
                                       const double* D, double damping, double* M, tlh_stream_t stream) {
  if (count < 0 || count > INT32_MAX || (block_dtype != TLH_F64 && block_dtype != TLH_F32)) return TLH_EINVAL;
  if (count == 0) return 0;
  if (!jc || !jp || !order || !edge_ptr || !var_edge_ptr || !edge_point || !cinv || !D || !M) return TLH_EINVAL;
  if (rows != 2 || dc != 6 || dp != 3) return TLH_ENOTFOUND;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (block_dtype == TLH_F64)
    hipLaunchKernelGGL((major_precond_kernel<double, 2, 6, 3>), dim3((unsigned)count), dim3(MAJOR_THREADS), 0, s,
                       static_cast<const double*>(jc), static_cast<const double*>(jp), order, edge_ptr, var_edge_ptr,
                       edge_point, cinv, D, damping, M);
  else
    hipLaunchKernelGGL((major_precond_kernel<float, 2, 6, 3>), dim3((unsigned)count), dim3(MAJOR_THREADS), 0, s,
                       static_cast<const float*>(jc), static_cast<const float*>(jp), order, edge_ptr, var_edge_ptr,
                       edge_point, cinv, D, damping, M);
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

extern "C" int tlh_major_rmatvec(const void* blocks, int block_dtype, const void* v, int v_dtype, int rows, int dim,
                                 const int32_t* order, const int32_t* ptr, int64_t count, double* out,
                                 tlh_stream_t stream) {
  const bool dtypes_ok = (block_dtype == TLH_F64 || block_dtype == TLH_F32) && (v_dtype == TLH_F64 || v_dtype == TLH_F32);
  if (count < 0 || count > INT32_MAX || !dtypes_ok) return TLH_EINVAL;
  if (count == 0) return 0;
  if (!blocks || !v || !order || !ptr || !out) return TLH_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (rows == 2 && dim == 6)
    launch_rmatvec<2, 6>(blocks, block_dtype, v, v_dtype, order, ptr, count, out, s);
  else if (rows == 2 && dim == 3)
    launch_rmatvec<2, 3>(blocks, block_dtype, v, v_dtype, order, ptr, count, out, s);
  else if (rows == 6 && dim == 6)
    launch_rmatvec<6, 6>(blocks, block_dtype, v, v_dtype, order, ptr, count, out, s);
  else
    return TLH_ENOTFOUND;
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}
