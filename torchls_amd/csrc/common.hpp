// Shared helpers for the gfx950 kernels of the block-sparse normal-equation path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "torchls_hip.h"

#define TLH_THREADS TLH_TILE_THREADS

#define TLH_RETURN_IF_LAUNCH_FAILED()                 \
  do {                                                \
    hipError_t tlh_err_ = hipGetLastError();          \
    if (tlh_err_ != hipSuccess) return (int)tlh_err_; \
  } while (0)

#define TLH_RETURN_IF_ERROR(expr)                     \
  do {                                                \
    hipError_t tlh_err_ = (expr);                     \
    if (tlh_err_ != hipSuccess) return (int)tlh_err_; \
  } while (0)

namespace tlh {

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

template <typename T>
__device__ __forceinline__ T ldg(const T* p) {
  return __builtin_nontemporal_load(p);
}

// Kernel-side view of one slot (see tlh_slot_t).
template <typename Real>
struct SlotArgs {
  const Real* x;             // x + type offset
  Real* partials;            // type partial workspace
  const int32_t* index;
  const uint16_t* rank;
  const int32_t* seg_ptr;
  const int32_t* seg_start;
  const int32_t* seg_target;
  int32_t dim;
  int32_t mode;
  int32_t dense_block;
  int32_t type_count;
};

template <typename Real>
struct GroupArgs {
  const Real* packed;
  int64_t stride;
  const int32_t* tile_start;
  int32_t ntiles;
  int32_t tiles_per_wg;
  int32_t rows;
  int32_t nslots;
  SlotArgs<Real> slot[TLH_MAX_SLOTS];
};

// Dynamic LDS above the default limit must be enabled once per kernel.
// Largest dynamic LDS allocation of one workgroup on the current device.
inline int64_t device_lds_max() {
  static int64_t cached[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 65536;
  if (!cached[dev]) {
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || lds <= 0)
      return 65536;
    cached[dev] = lds;
  }
  return cached[dev];
}

inline hipError_t allow_lds(const void* kernel, int64_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

}  // namespace tlh

namespace tlh {

// Fixed-order sum over one variable's CSR list of partials (component c of W),
// by one wavefront: lane l adds entries begin+l, begin+l+64, ..., then a fixed
// xor-shuffle tree combines the lanes; lane 0's value is the same every run.
template <typename T>
__device__ __forceinline__ T csr_wave_sum(const T* __restrict__ partials, const int32_t* __restrict__ idx, int begin,
                                         int end, int W, int c, int lane) {
  T s = T(0);
  for (int k = begin + lane; k < end; k += 64) s += partials[(int64_t)idx[k] * W + c];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  return s;
}

// (i, j) of the k-th entry of a row-major upper triangle of a d x d matrix.
__device__ __forceinline__ void upper_index(int k, int d, int& i, int& j) {
  i = 0;
  int row = d;
  while (k >= row) {
    k -= row;
    ++i;
    --row;
  }
  j = i + k;
}

// True when a type's partial lists are long enough for one wavefront per
// (variable, component); such kernels take count * W wavefronts.
inline bool csr_heavy(int64_t csr_entries, int64_t variables) {
  return variables > 0 && csr_entries >= 8 * variables;
}

}  // namespace tlh
