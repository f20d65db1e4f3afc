// Fixed-order float64 grid reductions and the block Huber loss, shared by the
// LM-loop kernels (objective, trial reduction, ||J x||^2).
#pragma once

#include <math.h>

#include "common.hpp"

namespace tlh {

constexpr int RED_BLOCKS = TLH_PCG_MAX_BLOCKS;

// Workgroups of a fixed-order grid reduction.  Each workgroup publishes its
// partial with an agent-scope release and takes a ticket; on MI355X that costs
// about 20 ns per workgroup at the end of the launch (a 1.2M-entry float64 dot:
// 28 us on 1,024 workgroups, 10 us on 128-256, 3.9 us without the tickets), so
// grids are capped and each thread strides over more entries.
// 256 workgroups keep the per-observation LM kernels (linearization, trial,
// squared norms) at full rate; 128 measured slower at 5M observations.
constexpr int RED_GRID_CAP = 256;

// Fixed-order sum of NV values over the workgroup (TLH_THREADS threads): an
// xor-shuffle tree inside each wavefront (every lane ends with the same sum),
// then the wavefront sums in wavefront order.  Every thread returns the total.
template <int NV>
__device__ __forceinline__ void wave_block_sum(double (&v)[NV], double* smem) {
  constexpr int NWAVES = TLH_THREADS / 64;
#pragma unroll
  for (int k = 0; k < NV; ++k)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) smem[k * NWAVES + (threadIdx.x >> 6)] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double sum = smem[k * NWAVES];
#pragma unroll
    for (int w = 1; w < NWAVES; ++w) sum += smem[k * NWAVES + w];
    v[k] = sum;
  }
  __syncthreads();  // the next sum rewrites smem
}

// LDS doubles the two-sum reductions need (one per wavefront and sum).
constexpr int RED_SMEM = 2 * (TLH_THREADS / 64);

__device__ __forceinline__ void block_sum2(double (&v)[2], double* smem) { wave_block_sum<2>(v, smem); }

// Fixed-order grid reduction of two sums; the last block writes out[0..1].
__device__ __forceinline__ void grid_sum2(double (&v)[2], double* partials, uint32_t* counter, double* out,
                                          double* smem) {
  __shared__ int last;
  block_sum2(v, smem);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = v[0];
    partials[RED_BLOCKS + blockIdx.x] = v[1];
    const uint32_t t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = t == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  double a[2] = {0.0, 0.0};
  for (int b = threadIdx.x; b < (int)gridDim.x; b += TLH_THREADS) {
    a[0] += __hip_atomic_load(&partials[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    a[1] += __hip_atomic_load(&partials[RED_BLOCKS + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  block_sum2(a, smem);
  if (threadIdx.x == 0) {
    out[0] = a[0];
    out[1] = a[1];
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

inline int red_grid(int64_t work) {
  int64_t g = ceil_div(work, TLH_THREADS);
  return (int)(g < 1 ? 1 : (g > RED_GRID_CAP ? RED_GRID_CAP : g));
}

// Grid of a reduction over a vector of n entries with little work per entry
// (PCG vector updates, dot products, vector statistics), whose threads each
// take one of `work` items (entries, or blocks of entries; default n): 64
// workgroups up to 2^17 entries, 128 up to 2^20, then 256 -- the measured
// balance of the ticket cost against the streaming rate.
inline int vector_grid(int64_t n, int64_t work = -1) {
  const int64_t cap = n <= (int64_t(1) << 17) ? 64 : (n <= (int64_t(1) << 20) ? 128 : 256);
  const int64_t g = ceil_div(work < 0 ? n : work, TLH_THREADS);
  return (int)(g < 1 ? 1 : (g > cap ? cap : g));
}

template <typename T>
__device__ __forceinline__ double ld(const void* p, int64_t i) {
  return (double)static_cast<const T*>(p)[i];
}

__device__ __forceinline__ double load_any(const void* p, int64_t i, int dtype) {
  return dtype == TLH_F64 ? ld<double>(p, i) : ld<float>(p, i);
}

__device__ __forceinline__ double huber_value(double s, double delta) {
  const double norm = sqrt(fmax(s, 2.2250738585072014e-308));
  return s > delta * delta ? delta * (2.0 * norm - delta) : s;
}

// Divided difference of the Huber loss between two squared norms (losses.py:95-107).
__device__ __forceinline__ double huber_reduction(double before, double after, double diff, double delta) {
  const double tiny = 2.2250738585072014e-308;
  const double a = sqrt(before), b = sqrt(after);
  const double low = fmin(a, b), high = fmax(a, b);
  const double shortfall = fmax(delta - low, 0.0);
  const double crossing = (2.0 * delta - shortfall * (shortfall / fmax(high - low, tiny))) / fmax(a + b, tiny);
  double coef;
  if (high <= delta)
    coef = 1.0;
  else if (low >= delta)
    coef = 2.0 * delta / fmax(a + b, tiny);
  else
    coef = crossing;
  return diff * coef;
}

}  // namespace tlh
