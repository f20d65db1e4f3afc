// Reprojection residual batches of the bundle-adjustment problems, in HIP.
//
// The residual is examples/named_bundle_adjustment.py:15-17,
//     r_i = pi(R_c p + t_c) - obs_i,   pi(x, y, z) = (x / z, y / z),
// with a world-to-camera SE(3) pose (rows [R | t], 3x4 or 4x4 storage)
// updated on the right with tangent [v; w], so that
//     J_c = [D R, -D R [p]x],  J_p = D R,  D = d pi / d(x, y, z).
// tlh_reprojection_linearize replaces, for this residual, the vmapped
// jacrev/jacfwd linearization and the loss rescaling of
// linearization.py:62-128 / losses.py:56-90: one thread per observation
// gathers its pose and point, writes the raw residual, the Gauss-Newton model
// residual and both blocks scaled by sqrt(w_i rho'_i), and the float64
// objective sum w_i rho(||r_i||^2) and the count of observations with a
// nonfinite residual or block go through the fixed-order grid reduction.
// tlh_reprojection_trial fuses the trial residual with the objective and
// divided-difference reduction of problem.py:315-348 (no residual written).
// Arithmetic runs in the parameters' dtype like the torch path; the loss
// factors and sums are float64.
#include "common.hpp"
#include "reduce.hpp"

using namespace tlh;

namespace {

template <typename Real>
struct ReprojArgs {
  const Real* cams;
  const Real* pts;
  const Real* obs;
  const int64_t* ci;
  const int64_t* pi;
  const double* weights;
  int64_t cam_stride;
  int64_t count;
  double delta;  // Huber delta; <= 0: squared norm
};

template <typename Real>
struct Projection {
  Real R[3][3];
  Real l[3];
  Real r[2];
};

template <typename Real>
__device__ __forceinline__ void project(const ReprojArgs<Real>& a, int64_t o, const Real* p, Projection<Real>& q) {
  const Real* T = a.cams + a.ci[o] * a.cam_stride;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) q.R[i][j] = T[4 * i + j];
    q.l[i] = q.R[i][0] * p[0] + q.R[i][1] * p[1] + q.R[i][2] * p[2] + T[4 * i + 3];
  }
  q.r[0] = q.l[0] / q.l[2] - a.obs[2 * o];
  q.r[1] = q.l[1] / q.l[2] - a.obs[2 * o + 1];
}

// rho(s) and sqrt(rho'(s)) of the block Huber loss (or the squared norm).
__device__ __forceinline__ double loss_value(double s, double delta, double& sqrt_first) {
  sqrt_first = 1.0;
  if (delta <= 0.0) return s;
  if (s > delta * delta) sqrt_first = sqrt(delta / sqrt(fmax(s, 2.2250738585072014e-308)));
  return huber_value(s, delta);
}

// Block rows leave through LDS: a workgroup's 256 consecutive observations own
// contiguous [256, 2, 6] and [256, 2, 3] slices, written with coalesced stores
// instead of one strided row per lane.  Padded index: one extra word every 32.
__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 5); }

// BReal: storage type of the blocks (the parameters' type, or float32 blocks under
// float64 parameters: mixed precision keeps the residuals and their sums float64).
template <typename Real, typename BReal>
__global__ void __launch_bounds__(TLH_THREADS)
    linearize_kernel(ReprojArgs<Real> a, Real* __restrict__ raw, Real* __restrict__ model, BReal* __restrict__ jc,
                     BReal* __restrict__ jp, double* out, double* partials, uint32_t* counter) {
  __shared__ double smem[RED_SMEM];
  constexpr int NJ = TLH_THREADS * 18;
  __shared__ BReal sj[NJ + NJ / 32];
  double v[2] = {0.0, 0.0};
  for (int64_t base = (int64_t)blockIdx.x * TLH_THREADS; base < a.count; base += (int64_t)gridDim.x * TLH_THREADS) {
    const int rows = (int)min((int64_t)TLH_THREADS, a.count - base);
    const int64_t o = base + threadIdx.x;
    if ((int)threadIdx.x < rows) {
      const Real* p = a.pts + a.pi[o] * 3;
      Projection<Real> q;
      project(a, o, p, q);
      const Real inv = Real(1) / q.l[2];
      const Real D[2][3] = {{inv, Real(0), -q.l[0] * inv * inv}, {Real(0), inv, -q.l[1] * inv * inv}};
      Real DR[2][3];
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int j = 0; j < 3; ++j) DR[r][j] = D[r][0] * q.R[0][j] + D[r][1] * q.R[1][j] + D[r][2] * q.R[2][j];
      const double s = (double)q.r[0] * (double)q.r[0] + (double)q.r[1] * (double)q.r[1];
      double sf;
      double value = loss_value(s, a.delta, sf);
      const Real wl = (Real)sf;
      Real ws = Real(1);
      if (a.weights) {
        value *= a.weights[o];
        ws = (Real)sqrt(a.weights[o]);
      }
      v[0] += value;
      const Real m0 = q.r[0] * wl * ws, m1 = q.r[1] * wl * ws;
      Real c[2][6], e[2][3];
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        // -(D R [p]x) columns: [p]x = [[0,-p2,p1],[p2,0,-p0],[-p1,p0,0]]
        const Real k0 = DR[r][1] * p[2] - DR[r][2] * p[1];
        const Real k1 = DR[r][2] * p[0] - DR[r][0] * p[2];
        const Real k2 = DR[r][0] * p[1] - DR[r][1] * p[0];
        c[r][0] = DR[r][0] * wl * ws;
        c[r][1] = DR[r][1] * wl * ws;
        c[r][2] = DR[r][2] * wl * ws;
        c[r][3] = -k0 * wl * ws;
        c[r][4] = -k1 * wl * ws;
        c[r][5] = -k2 * wl * ws;
        e[r][0] = DR[r][0] * wl * ws;
        e[r][1] = DR[r][1] * wl * ws;
        e[r][2] = DR[r][2] * wl * ws;
      }
      // observations with a nonfinite residual or block entry, for the solver's finiteness check
      bool ok = isfinite(q.r[0]) && isfinite(q.r[1]) && isfinite(m0) && isfinite(m1);
#pragma unroll
      for (int r = 0; r < 2; ++r) {
#pragma unroll
        for (int j = 0; j < 6; ++j) ok = ok && isfinite((BReal)c[r][j]);
#pragma unroll
        for (int j = 0; j < 3; ++j) ok = ok && isfinite((BReal)e[r][j]);
      }
      v[1] += ok ? 0.0 : 1.0;
      if (raw) {
        raw[2 * o] = q.r[0];
        raw[2 * o + 1] = q.r[1];
      }
      model[2 * o] = m0;
      model[2 * o + 1] = m1;
      // stage: [0, 12*256) camera rows, [12*256, 18*256) point rows
#pragma unroll
      for (int r = 0; r < 2; ++r) {
#pragma unroll
        for (int j = 0; j < 6; ++j) sj[lds_pad(threadIdx.x * 12 + r * 6 + j)] = (BReal)c[r][j];
#pragma unroll
        for (int j = 0; j < 3; ++j) sj[lds_pad(12 * TLH_THREADS + threadIdx.x * 6 + r * 3 + j)] = (BReal)e[r][j];
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < rows * 12; i += TLH_THREADS) jc[base * 12 + i] = sj[lds_pad(i)];
    for (int i = threadIdx.x; i < rows * 6; i += TLH_THREADS) jp[base * 6 + i] = sj[lds_pad(12 * TLH_THREADS + i)];
    __syncthreads();  // the stage is rewritten by the next chunk
  }
  grid_sum2(v, partials, counter, out, smem);
}

template <typename Real>
__global__ void __launch_bounds__(TLH_THREADS)
    trial_kernel(ReprojArgs<Real> a, const void* before, int before_dtype, double* out, double* partials,
                 uint32_t* counter) {
  __shared__ double smem[RED_SMEM];
  double v[2] = {0.0, 0.0};
  for (int64_t o = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x; o < a.count;
       o += (int64_t)gridDim.x * TLH_THREADS) {
    Projection<Real> q;
    project(a, o, a.pts + a.pi[o] * 3, q);
    double s_old = 0.0, s_new = 0.0, diff = 0.0;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const double x = before ? load_any(before, 2 * o + r, before_dtype) : 0.0;
      const double y = (double)q.r[r];
      s_old += x * x;
      s_new += y * y;
      diff += (x - y) * (x + y);
    }
    double value, change;
    if (a.delta > 0.0) {
      value = huber_value(s_new, a.delta);
      change = huber_reduction(s_old, s_new, diff, a.delta);
    } else {
      value = s_new;
      change = diff;
    }
    if (a.weights) {
      value *= a.weights[o];
      change *= a.weights[o];
    }
    v[0] += value;
    v[1] += change;
  }
  grid_sum2(v, partials, counter, out, smem);
}

template <typename Real>
ReprojArgs<Real> make_args(const void* cams, int64_t cam_stride, const void* pts, const void* obs,
                           const int64_t* ci, const int64_t* pi, int64_t count, double delta, const double* w) {
  ReprojArgs<Real> a;
  a.cams = static_cast<const Real*>(cams);
  a.pts = static_cast<const Real*>(pts);
  a.obs = static_cast<const Real*>(obs);
  a.ci = ci;
  a.pi = pi;
  a.weights = w;
  a.cam_stride = cam_stride;
  a.count = count;
  a.delta = delta;
  return a;
}

bool valid(int dtype, const void* cams, int64_t cam_stride, const void* pts, const void* obs, const int64_t* ci,
           const int64_t* pi, int64_t count) {
  return (dtype == TLH_F32 || dtype == TLH_F64) && cams && pts && obs && ci && pi && count >= 0 &&
         (cam_stride == 12 || cam_stride == 16);
}

}  // namespace

extern "C" int tlh_reprojection_linearize(int dtype, int block_dtype, const void* cams, int64_t cam_stride,
                                          const void* pts, const void* obs, const int64_t* cam_idx,
                                          const int64_t* pt_idx, int64_t count, double huber_delta,
                                          const double* weights, void* raw, void* model, void* jc, void* jp,
                                          double* out2, double* partials, uint32_t* counter, tlh_stream_t stream) {
  if (!valid(dtype, cams, cam_stride, pts, obs, cam_idx, pt_idx, count) || !model || !jc || !jp || !out2 ||
      !partials || !counter)
    return TLH_EINVAL;
  // float64 parameters may store float32 blocks; float32 parameters store float32 blocks
  if (!(block_dtype == dtype || (dtype == TLH_F64 && block_dtype == TLH_F32))) return TLH_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid(red_grid(count)), block(TLH_THREADS);
  if (dtype == TLH_F64 && block_dtype == TLH_F64)
    hipLaunchKernelGGL((linearize_kernel<double, double>), grid, block, 0, s,
                       make_args<double>(cams, cam_stride, pts, obs, cam_idx, pt_idx, count, huber_delta, weights),
                       static_cast<double*>(raw), static_cast<double*>(model), static_cast<double*>(jc),
                       static_cast<double*>(jp), out2, partials, counter);
  else if (dtype == TLH_F64)
    hipLaunchKernelGGL((linearize_kernel<double, float>), grid, block, 0, s,
                       make_args<double>(cams, cam_stride, pts, obs, cam_idx, pt_idx, count, huber_delta, weights),
                       static_cast<double*>(raw), static_cast<double*>(model), static_cast<float*>(jc),
                       static_cast<float*>(jp), out2, partials, counter);
  else
    hipLaunchKernelGGL((linearize_kernel<float, float>), grid, block, 0, s,
                       make_args<float>(cams, cam_stride, pts, obs, cam_idx, pt_idx, count, huber_delta, weights),
                       static_cast<float*>(raw), static_cast<float*>(model), static_cast<float*>(jc),
                       static_cast<float*>(jp), out2, partials, counter);
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

extern "C" int tlh_reprojection_trial(int dtype, const void* cams, int64_t cam_stride, const void* pts,
                                      const void* obs, const int64_t* cam_idx, const int64_t* pt_idx, int64_t count,
                                      double huber_delta, const double* weights, const void* before,
                                      int before_dtype, double* out2, double* partials, uint32_t* counter,
                                      tlh_stream_t stream) {
  if (!valid(dtype, cams, cam_stride, pts, obs, cam_idx, pt_idx, count) || !out2 || !partials || !counter)
    return TLH_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid(red_grid(count)), block(TLH_THREADS);
  if (dtype == TLH_F64)
    hipLaunchKernelGGL(trial_kernel<double>, grid, block, 0, s,
                       make_args<double>(cams, cam_stride, pts, obs, cam_idx, pt_idx, count, huber_delta, weights),
                       before, before_dtype, out2, partials, counter);
  else
    hipLaunchKernelGGL(trial_kernel<float>, grid, block, 0, s,
                       make_args<float>(cams, cam_stride, pts, obs, cam_idx, pt_idx, count, huber_delta, weights),
                       before, before_dtype, out2, partials, counter);
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}
