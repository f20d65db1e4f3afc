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

// Poll limit of the bounded cross-workgroup waits (tile Cholesky, blocks solve)
// set by tlh_set_wait_limit; 0 keeps each kernel's own limit.  Process-wide.
extern int64_t wait_limit;

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

// Cross-lane moves by DPP (data-parallel primitives: register reads of another
// lane, no LDS traffic).  A lane whose source is outside the pattern (or whose
// row is masked off) receives 0.
constexpr int DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR2 = 0x112, DPP_ROW_SHR4 = 0x114, DPP_ROW_SHR8 = 0x118;
constexpr int DPP_WAVE_SHL1 = 0x130, DPP_WAVE_SHR1 = 0x138, DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143;

template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_int(int x) {
  return __builtin_amdgcn_update_dpp(0, x, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_real(float x) {
  return __int_as_float(dpp_int<CTRL, ROWS>(__float_as_int(x)));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_real(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = dpp_int<CTRL, ROWS>((int)(b & 0xffffffffll)), hi = dpp_int<CTRL, ROWS>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// One step of the inclusive segmented scan: (f, v) of the source lane is
// combined into this lane's unless this lane starts a segment (flag set).
template <int CTRL, int ROWS, typename Real, int D>
__device__ __forceinline__ void segmented_step(int& flag, Real (&v)[D]) {
  const int fo = dpp_int<CTRL, ROWS>(flag);
  Real vo[D];
#pragma unroll
  for (int c = 0; c < D; ++c) vo[c] = dpp_real<CTRL, ROWS>(v[c]);
  if (!flag) {
#pragma unroll
    for (int c = 0; c < D; ++c) v[c] = vo[c] + v[c];
  }
  flag |= fo;
}

// Inclusive segmented scan over the 64 lanes of a wavefront, in a fixed order:
// Hillis-Steele within each row of 16 lanes (shifts 1, 2, 4, 8), then row 0's
// total into row 1 and row 2's into row 3 (broadcast of lane 15 of the row
// before), then lanes 0-31's total into rows 2 and 3 (broadcast of lane 31).
template <typename Real, int D>
__device__ __forceinline__ void wave_segmented_scan(int& flag, Real (&v)[D]) {
  segmented_step<DPP_ROW_SHR1, 0xf>(flag, v);
  segmented_step<DPP_ROW_SHR2, 0xf>(flag, v);
  segmented_step<DPP_ROW_SHR4, 0xf>(flag, v);
  segmented_step<DPP_ROW_SHR8, 0xf>(flag, v);
  segmented_step<DPP_ROW_BCAST15, 0xa>(flag, v);
  segmented_step<DPP_ROW_BCAST31, 0xc>(flag, v);
}

// True when a type's partial lists are long enough for one wavefront per
// (variable, component); such kernels take count * W wavefronts.
inline bool csr_heavy(int64_t csr_entries, int64_t variables) {
  return variables > 0 && csr_entries >= 8 * variables;
}

}  // namespace tlh
