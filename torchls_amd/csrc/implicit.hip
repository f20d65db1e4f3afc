// Backward kernel of the implicitly differentiated bundle-adjustment solve.
//
// At the optimum the gradient g = sum_i rho'_i J_i^T r_i vanishes.  With the
// adjoint w = (J~^T J~ + lambda I)^-1 dL/dtheta (one PCG solve on the solver's
// own operator), dL/dobs_i = -(dg/dobs_i)^T w.  For the reprojection residual
// r_i = pi(T_c p) - obs_i of examples/named_bundle_adjustment.py:15-17 the
// residual-vs-observation Jacobian is -I, so with the block Huber loss
// (losses.py:56-90)
//     dL/dobs_i = (rho'_i I + 2 rho''_i r_i r_i^T) (J_c,i w_c + J_p,i w_p).
// The kernel recomputes r_i, J_c,i and J_p,i in float64 from the solved
// poses and points instead of storing the Jacobian (one thread per
// observation).  No counterpart exists in the reference (its solve is
// @torch.no_grad, solvers.py:288, 307-308).
#include "common.hpp"

using namespace tlh;

namespace {

__global__ void __launch_bounds__(TLH_THREADS)
    reprojection_vjp_kernel(const double* __restrict__ cams, const double* __restrict__ pts,
                            const double* __restrict__ obs, const int64_t* __restrict__ ci,
                            const int64_t* __restrict__ pi, const double* __restrict__ w, int64_t cam_offset,
                            int64_t pt_offset, double delta, int64_t count, double* __restrict__ grad) {
  const int64_t o = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x;
  if (o >= count) return;
  const double* T = cams + ci[o] * 16;
  const double* p = pts + pi[o] * 3;
  double R[3][3], l[3];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) R[i][j] = T[i * 4 + j];
    l[i] = R[i][0] * p[0] + R[i][1] * p[1] + R[i][2] * p[2] + T[i * 4 + 3];
  }
  const double iz = 1.0 / l[2];
  const double r0 = l[0] * iz - obs[2 * o], r1 = l[1] * iz - obs[2 * o + 1];
  // D = d(uv)/d(local), DR = D R
  const double D[2][3] = {{iz, 0.0, -l[0] * iz * iz}, {0.0, iz, -l[1] * iz * iz}};
  double DR[2][3];
  for (int r = 0; r < 2; ++r)
    for (int j = 0; j < 3; ++j) DR[r][j] = D[r][0] * R[0][j] + D[r][1] * R[1][j] + D[r][2] * R[2][j];
  const double* wc = w + cam_offset + ci[o] * 6;
  const double* wp = w + pt_offset + pi[o] * 3;
  // J_c = [DR, -DR [p]x];  -DR [p]x w_rot = DR (w_rot x p)
  const double cx = wc[4] * p[2] - wc[5] * p[1];
  const double cy = wc[5] * p[0] - wc[3] * p[2];
  const double cz = wc[3] * p[1] - wc[4] * p[0];
  double u[2];
  for (int r = 0; r < 2; ++r)
    u[r] = DR[r][0] * (wc[0] + cx + wp[0]) + DR[r][1] * (wc[1] + cy + wp[1]) + DR[r][2] * (wc[2] + cz + wp[2]);
  double g0 = u[0], g1 = u[1];
  if (delta > 0.0) {
    const double s = r0 * r0 + r1 * r1;
    if (s > delta * delta) {
      const double norm = sqrt(s);
      const double first = delta / norm;            // rho'
      const double second = -0.5 * first / s;       // rho''
      const double ru = r0 * u[0] + r1 * u[1];
      g0 = first * u[0] + 2.0 * second * r0 * ru;
      g1 = first * u[1] + 2.0 * second * r1 * ru;
    }
  }
  grad[2 * o] = g0;
  grad[2 * o + 1] = g1;
}

}  // namespace

extern "C" int tlh_reprojection_obs_vjp(const double* cams, const double* pts, const double* obs,
                                        const int64_t* cam_idx, const int64_t* pt_idx, const double* w,
                                        int64_t cam_offset, int64_t pt_offset, double huber_delta, int64_t count,
                                        double* grad, tlh_stream_t stream) {
  if (!cams || !pts || !obs || !cam_idx || !pt_idx || !w || !grad || count < 0) return TLH_EINVAL;
  if (count == 0) return 0;
  hipLaunchKernelGGL(reprojection_vjp_kernel, dim3((unsigned)ceil_div(count, TLH_THREADS)), dim3(TLH_THREADS), 0,
                     reinterpret_cast<hipStream_t>(stream), cams, pts, obs, cam_idx, pt_idx, w, cam_offset,
                     pt_offset, huber_delta, count, grad);
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

// ------------------------------------------------------------------ exact Hessian-vector product
// For f = 1/2 sum_i w_i rho(||r_i(delta)||^2) with r_i(delta) = pi((T_c Exp(delta_c)) (p + delta_p)) - obs_i
// (tangent [translation; rotation], right update, as the solver retracts), the exact
// Hessian-vector product at delta = 0 is, per observation and tangent coordinate j,
//   (H v)_j = d/de [ w rho'(s) sum_k r_k dr_k/ddelta_j ](delta = e v)
//           = w [ rho''(s) s' sum_k r_k J_kj + rho'(s) sum_k ((J v)_k J_kj + r_k (d2 r_k)[v, e_j]) ],
// s' = 2 r . (J v).  The residual is evaluated once in hyper-dual arithmetic: every
// scalar carries a value, an e-part (direction v), nine h_j-parts (coordinate
// directions) and nine e h_j-parts, with e^2 = h_i h_j = 0.  Because delta has no
// value part, Exp(delta) is exactly I + K + K^2/2 (rotation) and V t = t + K t / 2
// (translation) in this algebra, so no Rodrigues coefficients are needed.
namespace {

struct HD {
  double a, b;     // value, e-part
  double c[9];     // h_j parts
  double d[9];     // e h_j parts
};

__device__ __forceinline__ HD hd_const(double a) {
  HD x;
  x.a = a;
  x.b = 0.0;
#pragma unroll
  for (int j = 0; j < 9; ++j) x.c[j] = x.d[j] = 0.0;
  return x;
}

__device__ __forceinline__ HD hd_var(double a, double b, int j) {
  HD x = hd_const(a);
  x.b = b;
  x.c[j] = 1.0;
  return x;
}

__device__ __forceinline__ HD operator+(const HD& x, const HD& y) {
  HD z;
  z.a = x.a + y.a;
  z.b = x.b + y.b;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    z.c[j] = x.c[j] + y.c[j];
    z.d[j] = x.d[j] + y.d[j];
  }
  return z;
}

__device__ __forceinline__ HD operator-(const HD& x, const HD& y) {
  HD z;
  z.a = x.a - y.a;
  z.b = x.b - y.b;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    z.c[j] = x.c[j] - y.c[j];
    z.d[j] = x.d[j] - y.d[j];
  }
  return z;
}

__device__ __forceinline__ HD operator*(const HD& x, const HD& y) {
  HD z;
  z.a = x.a * y.a;
  z.b = x.a * y.b + x.b * y.a;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    z.c[j] = x.a * y.c[j] + x.c[j] * y.a;
    z.d[j] = x.a * y.d[j] + x.b * y.c[j] + x.c[j] * y.b + x.d[j] * y.a;
  }
  return z;
}

__device__ __forceinline__ HD scale(const HD& x, double s) {
  HD z;
  z.a = x.a * s;
  z.b = x.b * s;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    z.c[j] = x.c[j] * s;
    z.d[j] = x.d[j] * s;
  }
  return z;
}

// 1 / x: value 1/a, e-part -b/a^2, h-part -c/a^2, e h-part 2bc/a^3 - d/a^2
__device__ __forceinline__ HD reciprocal(const HD& x) {
  const double i = 1.0 / x.a, i2 = i * i;
  HD z;
  z.a = i;
  z.b = -x.b * i2;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    z.c[j] = -x.c[j] * i2;
    z.d[j] = 2.0 * x.b * x.c[j] * i2 * i - x.d[j] * i2;
  }
  return z;
}

__device__ __forceinline__ void cross(const HD* u, const HD* v, HD* out) {
  out[0] = u[1] * v[2] - u[2] * v[1];
  out[1] = u[2] * v[0] - u[0] * v[2];
  out[2] = u[0] * v[1] - u[1] * v[0];
}

__global__ void __launch_bounds__(TLH_THREADS)
    reprojection_hvp_kernel(const double* __restrict__ cams, const double* __restrict__ pts,
                            const double* __restrict__ obs, const int64_t* __restrict__ ci,
                            const int64_t* __restrict__ pi, const double* __restrict__ v, int64_t cam_offset,
                            int64_t pt_offset, double delta, const double* __restrict__ weights, int64_t count,
                            double* __restrict__ out) {
  const int64_t o = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x;
  if (o >= count) return;
  const double* T = cams + ci[o] * 16;
  const double* p = pts + pi[o] * 3;
  const double* vc = v + cam_offset + ci[o] * 6;
  const double* vp = v + pt_offset + pi[o] * 3;
  HD tr[3], w[3], P[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    tr[k] = hd_var(0.0, vc[k], k);
    w[k] = hd_var(0.0, vc[3 + k], 3 + k);
    P[k] = hd_var(p[k], vp[k], 6 + k);
  }
  // q = P + w x P + (w x (w x P)) / 2 + t + (w x t) / 2   (the point in the perturbed frame)
  HD wp[3], wwp[3], wt[3], q[3];
  cross(w, P, wp);
  cross(w, wp, wwp);
  cross(w, tr, wt);
#pragma unroll
  for (int k = 0; k < 3; ++k) q[k] = P[k] + wp[k] + scale(wwp[k], 0.5) + tr[k] + scale(wt[k], 0.5);
  HD l[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    l[i] = scale(q[0], T[i * 4 + 0]) + scale(q[1], T[i * 4 + 1]) + scale(q[2], T[i * 4 + 2]) + hd_const(T[i * 4 + 3]);
  const HD iz = reciprocal(l[2]);
  HD r[2];
  r[0] = l[0] * iz - hd_const(obs[2 * o]);
  r[1] = l[1] * iz - hd_const(obs[2 * o + 1]);
  const double s = r[0].a * r[0].a + r[1].a * r[1].a;
  const double ds = 2.0 * (r[0].a * r[0].b + r[1].a * r[1].b);
  double first = 1.0, second = 0.0;  // rho'(s), rho''(s)
  if (delta > 0.0 && s > delta * delta) {
    first = delta / sqrt(s);
    second = -0.5 * first / s;
  }
  const double wt_o = weights ? weights[o] : 1.0;
  double* dst = out + o * 9;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    const double rj = r[0].a * r[0].c[j] + r[1].a * r[1].c[j];
    const double tj = r[0].b * r[0].c[j] + r[1].b * r[1].c[j] + r[0].a * r[0].d[j] + r[1].a * r[1].d[j];
    dst[j] = wt_o * (second * ds * rj + first * tj);
  }
}

// out[var * dim + c] = sum over k in [ptr[var], ptr[var+1]) of values[order[k] * stride + col0 + c]:
// one wavefront per variable, lane-strided partial sums, a fixed xor tree (deterministic).
__global__ void __launch_bounds__(TLH_THREADS)
    segment_sum_kernel(const double* __restrict__ values, int64_t stride, int col0, int dim,
                       const int32_t* __restrict__ order, const int32_t* __restrict__ ptr, int64_t count,
                       double* __restrict__ out) {
  const int64_t var = ((int64_t)blockIdx.x * TLH_THREADS + threadIdx.x) / 64;
  const int lane = threadIdx.x & 63;
  if (var >= count) return;
  const int32_t k0 = ptr[var], k1 = ptr[var + 1];
  for (int c = 0; c < dim; ++c) {
    double acc = 0.0;
    for (int32_t k = k0 + lane; k < k1; k += 64) acc += values[(int64_t)order[k] * stride + col0 + c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) out[var * dim + c] = acc;
  }
}

}  // namespace

extern "C" int tlh_reprojection_hvp(const double* cams, const double* pts, const double* obs, const int64_t* cam_idx,
                                    const int64_t* pt_idx, const double* v, int64_t cam_offset, int64_t pt_offset,
                                    double huber_delta, const double* weights, int64_t count, double* out,
                                    tlh_stream_t stream) {
  if (!cams || !pts || !obs || !cam_idx || !pt_idx || !v || !out || count < 0) return TLH_EINVAL;
  if (count == 0) return 0;
  hipLaunchKernelGGL(reprojection_hvp_kernel, dim3((unsigned)ceil_div(count, TLH_THREADS)), dim3(TLH_THREADS), 0,
                     reinterpret_cast<hipStream_t>(stream), cams, pts, obs, cam_idx, pt_idx, v, cam_offset, pt_offset,
                     huber_delta, weights, count, out);
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

extern "C" int tlh_segment_sum_f64(const double* values, int64_t stride, int col0, int dim, const int32_t* order,
                                   const int32_t* ptr, int64_t count, double* out, tlh_stream_t stream) {
  if (!values || !order || !ptr || !out || count < 0 || dim < 1 || col0 < 0 || col0 + dim > stride) return TLH_EINVAL;
  if (count == 0) return 0;
  hipLaunchKernelGGL(segment_sum_kernel, dim3((unsigned)ceil_div(count, TLH_THREADS / 64)), dim3(TLH_THREADS), 0,
                     reinterpret_cast<hipStream_t>(stream), values, stride, col0, dim, order, ptr, count, out);
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}
