// SE(3) right retraction T <- T Exp(delta) for a batch of poses, in HIP.
//
// Replaces the vmapped SE3Var retraction of the LM update
// (geometry.py:40-57, 130-137; AnalyzedLeastSquaresProblem.retract):
// delta = [v; w] (translation, rotation), Exp(delta) = [[R, V v], [0, 1]] with
//     R = I + a K + b K^2,  V = I + b K + c K^2,  K = [w]x,
//     a = sin t / t, b = (1 - cos t) / t^2, c = (t - sin t) / t^3,
// and their Taylor series for t^2 < 1e-4.  One thread per pose; arithmetic in
// the poses' dtype; 3x4 or 4x4 row-major storage (a 4x4 pose's bottom row is
// multiplied through like every other row).
#include "common.hpp"

using namespace tlh;

namespace {

template <typename Real>
__global__ void __launch_bounds__(TLH_THREADS)
    se3_retract_kernel(const Real* __restrict__ poses, int rows, int64_t count, const double* __restrict__ delta,
                       int64_t delta_stride, Real* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x;
  if (i >= count) return;
  Real d[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) d[k] = (Real)delta[i * delta_stride + k];
  const Real wx = d[3], wy = d[4], wz = d[5];
  const Real t2 = wx * wx + wy * wy + wz * wz;
  Real a, b, c;
  if (t2 < Real(1e-4)) {
    a = Real(1) - t2 / Real(6) + t2 * t2 / Real(120);
    b = Real(0.5) - t2 / Real(24) + t2 * t2 / Real(720);
    c = Real(1) / Real(6) - t2 / Real(120) + t2 * t2 / Real(5040);
  } else {
    const Real t = sqrt(t2);
    const Real s = sin(t), co = cos(t);
    a = s / t;
    b = (Real(1) - co) / t2;
    c = (t - s) / (t2 * t);
  }
  const Real K[3][3] = {{Real(0), -wz, wy}, {wz, Real(0), -wx}, {-wy, wx, Real(0)}};
  Real KK[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int q = 0; q < 3; ++q) KK[r][q] = K[r][0] * K[0][q] + K[r][1] * K[1][q] + K[r][2] * K[2][q];
  Real E[4][4];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    Real V[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const Real id = r == q ? Real(1) : Real(0);
      E[r][q] = id + a * K[r][q] + b * KK[r][q];
      V[q] = id + b * K[r][q] + c * KK[r][q];
    }
    E[r][3] = V[0] * d[0] + V[1] * d[1] + V[2] * d[2];
  }
  E[3][0] = E[3][1] = E[3][2] = Real(0);
  E[3][3] = Real(1);
  const Real* T = poses + i * rows * 4;
  Real* O = out + i * rows * 4;
  for (int r = 0; r < rows; ++r) {
    const Real t0 = T[4 * r], t1 = T[4 * r + 1], t2r = T[4 * r + 2], t3 = T[4 * r + 3];
#pragma unroll
    for (int q = 0; q < 4; ++q) O[4 * r + q] = t0 * E[0][q] + t1 * E[1][q] + t2r * E[2][q] + t3 * E[3][q];
  }
}

}  // namespace

extern "C" int tlh_se3_retract(int dtype, const void* poses, int rows, int64_t count, const double* delta,
                               int64_t delta_stride, void* out, tlh_stream_t stream) {
  if ((dtype != TLH_F32 && dtype != TLH_F64) || !poses || !delta || !out || count < 0 || (rows != 3 && rows != 4) ||
      delta_stride < 6)
    return TLH_EINVAL;
  if (count == 0) return 0;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid((unsigned)ceil_div(count, TLH_THREADS)), block(TLH_THREADS);
  if (dtype == TLH_F64)
    hipLaunchKernelGGL(se3_retract_kernel<double>, grid, block, 0, s, static_cast<const double*>(poses), rows, count,
                       delta, delta_stride, static_cast<double*>(out));
  else
    hipLaunchKernelGGL(se3_retract_kernel<float>, grid, block, 0, s, static_cast<const float*>(poses), rows, count,
                       delta, delta_stride, static_cast<float*>(out));
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}
