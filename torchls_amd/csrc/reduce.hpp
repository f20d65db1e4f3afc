// Fixed-order float64 grid reductions and the block Huber loss, shared by the
// LM-loop kernels (objective, trial reduction, ||J x||^2).
#pragma once

#include <math.h>

#include "common.hpp"

namespace tlh {

constexpr int RED_BLOCKS = TLH_PCG_MAX_BLOCKS;

__device__ __forceinline__ void block_sum2(double (&v)[2], double* smem) {
  smem[threadIdx.x] = v[0];
  smem[TLH_THREADS + threadIdx.x] = v[1];
  __syncthreads();
  for (int w = TLH_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      smem[threadIdx.x] += smem[threadIdx.x + w];
      smem[TLH_THREADS + threadIdx.x] += smem[TLH_THREADS + threadIdx.x + w];
    }
    __syncthreads();
  }
  v[0] = smem[0];
  v[1] = smem[TLH_THREADS];
  __syncthreads();
}

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
  return (int)(g < 1 ? 1 : (g > RED_BLOCKS ? RED_BLOCKS : g));
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
