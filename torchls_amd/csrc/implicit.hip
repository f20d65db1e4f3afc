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
