// Device-resident PCG with float64 scalars and a fixed-order reduction.
//
// Replaces warp_pcg.py (_step 33-50, _direction 53-67, _finish 70-92) and the
// three torch.dot calls per iteration (warp_pcg.py:139,147-148), plus the
// block-Jacobi factorisation/apply of solvers.py:259-285 and
// preconditioning.py:12-17 (closed-form float64 Cholesky inverse per block).
//
// State lives on the device (control = {iterations, status, stop_at};
// scalars = {rz, thr^2, curvature, new_rz, ||r||^2, ...}) so one iteration can
// be captured in a hipGraph and replayed; the host reads control only between
// replay batches, like warp_pcg._run.  Vectors are float32 or float64; every
// dot product accumulates in float64, and alpha/beta are float64.
//
// Reductions: each block sums its grid-stride slice in a fixed order, writes a
// partial, and the last block to finish (ticket counter) adds the partials in
// block order -- deterministic, no float atomics.
#include <math.h>

#include "common.hpp"
#include "reduce.hpp"

using namespace tlh;

namespace {

constexpr int MAX_BLOCKS = TLH_PCG_MAX_BLOCKS;

__device__ __forceinline__ bool finite(double v) { return isfinite(v); }

__device__ __forceinline__ bool frozen(const int32_t* control) {
  return control[1] != 0 || control[0] >= control[2];
}

// Block-wide fixed-order sum of NV values (wave_block_sum); every thread gets the totals.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* smem) {
  wave_block_sum<NV>(v, smem);
}

// Publish this block's NV sums; the last block returns true with the grid totals in v.
template <int NV>
__device__ __forceinline__ bool grid_sum(double (&v)[NV], double* partials, uint32_t* counter,
                                         double* smem) {
  __shared__ int last;
  block_sum<NV>(v, smem);
  if (threadIdx.x == 0) {
    for (int k = 0; k < NV; ++k) partials[k * MAX_BLOCKS + blockIdx.x] = v[k];
    const uint32_t ticket =
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = ticket == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return false;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  double acc[NV];
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += TLH_THREADS)
    for (int k = 0; k < NV; ++k)
      acc[k] += __hip_atomic_load(&partials[k * MAX_BLOCKS + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  block_sum<NV>(acc, smem);
  for (int k = 0; k < NV; ++k) v[k] = acc[k];
  if (threadIdx.x == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// Vector kernels reduce over a few workgroups (vector_grid): their tickets cost
// more than their streaming at hundreds of workgroups.
inline int grid_for(int64_t n, int64_t work = -1) { return vector_grid(n, work); }

// ---------------------------------------------------------------- block inverse
template <int D>
__global__ void __launch_bounds__(TLH_THREADS)
    block_inverse_kernel(const double* __restrict__ blocks, int64_t count, double damping,
                         double* __restrict__ inverse, int32_t* __restrict__ status) {
  const int64_t v = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x;
  if (v >= count) return;
  const double* A = blocks + v * D * D;
  double L[D][D];
  bool ok = true;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    double s = A[j * D + j] + damping;
#pragma unroll
    for (int k = 0; k < j; ++k) s -= L[j][k] * L[j][k];
    if (!(s > 0.0) || !finite(s)) ok = false;
    const double d = sqrt(s > 0.0 ? s : 1.0);
    L[j][j] = d;
    const double inv_d = 1.0 / d;
#pragma unroll
    for (int i = j + 1; i < D; ++i) {
      double t = A[i * D + j];
#pragma unroll
      for (int k = 0; k < j; ++k) t -= L[i][k] * L[j][k];
      L[i][j] = t * inv_d;
    }
  }
  // M = L^-1 (lower), then inverse = M^T M.
  double M[D][D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
#pragma unroll
    for (int i = 0; i < D; ++i) M[i][j] = 0.0;
    M[j][j] = 1.0 / L[j][j];
#pragma unroll
    for (int i = j + 1; i < D; ++i) {
      double t = 0.0;
#pragma unroll
      for (int k = j; k < i; ++k) t -= L[i][k] * M[k][j];
      M[i][j] = t / L[i][i];
    }
  }
  double* out = inverse + v * (D * (D + 1) / 2);
  int e = 0;
#pragma unroll
  for (int i = 0; i < D; ++i) {
#pragma unroll
    for (int j = i; j < D; ++j) {
      double t = 0.0;
#pragma unroll
      for (int k = j; k < D; ++k) t += M[k][i] * M[k][j];
      out[e++] = t;
    }
  }
  if (!ok) __hip_atomic_fetch_or(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// z_block = inverse_block * r_block (float64 arithmetic), symmetric packed upper.
template <typename Real>
__device__ __forceinline__ void apply_block(const double* inv, int d, const Real* r, Real* z, double& rz,
                                            double& rr) {
  double rv[16];
  for (int i = 0; i < d; ++i) rv[i] = (double)r[i];
  for (int i = 0; i < d; ++i) {
    double t = 0.0;
    for (int j = 0; j < d; ++j) {
      const int a = i < j ? i : j, b = i < j ? j : i;
      t += inv[a * d - a * (a - 1) / 2 + (b - a)] * rv[j];
    }
    const Real zi = (Real)t;
    z[i] = zi;
    rz += rv[i] * (double)zi;
    rr += rv[i] * rv[i];
  }
}

template <typename Real, int D>
__device__ __forceinline__ void apply_block_fixed(const double* inv, const Real* r, Real* z, double& rz,
                                                  double& rr) {
  constexpr int W = D * (D + 1) / 2;
  double m[W], rv[D];
#pragma unroll
  for (int k = 0; k < W; ++k) m[k] = inv[k];
#pragma unroll
  for (int i = 0; i < D; ++i) rv[i] = (double)r[i];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    double t = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int a = i < j ? i : j, b = i < j ? j : i;
      t += m[a * D - a * (a - 1) / 2 + (b - a)] * rv[j];
    }
    const Real zi = (Real)t;
    z[i] = zi;
    rz += rv[i] * (double)zi;
    rr += rv[i] * rv[i];
  }
}

// Block-Jacobi apply with the BA block sizes unrolled (same arithmetic order).
template <typename Real>
__device__ __forceinline__ void apply_any(const double* inv, int d, const Real* r, Real* z, double& rz, double& rr) {
  if (d == 6)
    apply_block_fixed<Real, 6>(inv, r, z, rz, rr);
  else if (d == 3)
    apply_block_fixed<Real, 3>(inv, r, z, rz, rr);
  else
    apply_block<Real>(inv, d, r, z, rz, rr);
}

__device__ __forceinline__ int find_type(const tlh_block_diag_t& pc, int64_t b, int64_t& local) {
  int t = 0;
  int64_t start = 0;
  while (t + 1 < pc.ntypes && b >= start + pc.count[t]) {
    start += pc.count[t];
    ++t;
  }
  local = b - start;
  return t;
}

// ---------------------------------------------------------------- PCG kernels
template <typename Real>
__global__ void __launch_bounds__(TLH_THREADS)
    dot_kernel(const Real* __restrict__ a, const Real* __restrict__ b, int64_t n, double* scalars, int slot,
               double* partials, uint32_t* counter) {
  __shared__ double smem[TLH_THREADS / 64];
  double v[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * TLH_THREADS)
    v[0] += (double)a[i] * (double)b[i];
  if (grid_sum<1>(v, partials, counter, smem) && threadIdx.x == 0) scalars[slot] = v[0];
}

// Step (x += alpha p, r -= alpha Ap; warp_pcg._step), then z = M^-1 r and the
// float64 sums r.z and r.r (warp_pcg.py:146-148), fused.  mode:
// 0 = block-Jacobi (pc), 1 = identity (z = r), 2 = step only (caller applies M).
template <typename Real>
__global__ void __launch_bounds__(TLH_THREADS)
    step_kernel(tlh_block_diag_t pc, int mode, Real* __restrict__ x, Real* __restrict__ r, Real* __restrict__ z,
                const Real* __restrict__ p, const Real* __restrict__ ap, int64_t n, const int32_t* control,
                double* scalars, double* partials, uint32_t* counter) {
  __shared__ double smem[2 * (TLH_THREADS / 64)];
  // The entries are loaded before the scalars are awaited, one memory round trip
  // fewer: the stores are unconditional selects (a frozen state rewrites the same
  // values), so no load is sunk below the branch on the scalars.
  const double c = scalars[TLH_PCG_CURVATURE];
  const double rz = scalars[TLH_PCG_RZ];
  const bool stopped = frozen(control);
  double v[2] = {0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * TLH_THREADS;
  if (mode == 0) {
    int64_t blocks = 0;
    for (int t = 0; t < pc.ntypes; ++t) blocks += pc.count[t];
    for (int64_t bidx = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x; bidx < blocks; bidx += stride) {
      int64_t local;
      const int t = find_type(pc, bidx, local);
      const int d = pc.dim[t];
      const int64_t e0 = pc.offset[t] + local * d;
      const bool move = !stopped && finite(c) && finite(rz) && c > 0.0 && rz > 0.0;
      const double alpha = move ? rz / c : 0.0;
      if (d == 6 || d == 3) {
        Real xv[6], rv[6], pv[6], av[6];
#pragma unroll
        for (int k = 0; k < 6; ++k)
          if (k < d) {
            xv[k] = x[e0 + k];
            rv[k] = r[e0 + k];
            pv[k] = p[e0 + k];
            av[k] = ap[e0 + k];
          }
#pragma unroll
        for (int k = 0; k < 6; ++k)
          if (k < d) {
            x[e0 + k] = move ? (Real)((double)xv[k] + alpha * (double)pv[k]) : xv[k];
            r[e0 + k] = move ? (Real)((double)rv[k] - alpha * (double)av[k]) : rv[k];
          }
      } else if (move) {
        for (int k = 0; k < d; ++k) {
          x[e0 + k] = (Real)((double)x[e0 + k] + alpha * (double)p[e0 + k]);
          r[e0 + k] = (Real)((double)r[e0 + k] - alpha * (double)ap[e0 + k]);
        }
      }
      apply_any<Real>(pc.inv[t] + local * (d * (d + 1) / 2), d, r + e0, z + e0, v[0], v[1]);
    }
  } else {
    const bool move = !stopped && finite(c) && finite(rz) && c > 0.0 && rz > 0.0;
    const double alpha = move ? rz / c : 0.0;
    for (int64_t i = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x; i < n; i += stride) {
      const Real xi = x[i], ri0 = r[i], pi = p[i], ai = ap[i];
      const Real ri = move ? (Real)((double)ri0 - alpha * (double)ai) : ri0;
      x[i] = move ? (Real)((double)xi + alpha * (double)pi) : xi;
      r[i] = ri;
      if (mode == 1) {
        const double rd = (double)ri;
        z[i] = ri;
        v[0] += rd * rd;
        v[1] += rd * rd;
      }
    }
  }
  if (mode == 2) return;
  if (grid_sum<2>(v, partials, counter, smem) && threadIdx.x == 0) {
    scalars[TLH_PCG_NEW_RZ] = v[0];
    scalars[TLH_PCG_NORM2] = v[1];
  }
}

// p = z + beta p (warp_pcg._direction); the last block then applies the
// status machine of warp_pcg._finish (70-92).
template <typename Real>
__global__ void __launch_bounds__(TLH_THREADS)
    direction_kernel(Real* __restrict__ p, const Real* __restrict__ z, int64_t n, int32_t* control,
                     double* scalars, double* partials, uint32_t* counter) {
  __shared__ double smem[TLH_THREADS / 64];
  const bool stopped = frozen(control);
  const double c = scalars[TLH_PCG_CURVATURE];
  const double rz = scalars[TLH_PCG_RZ];
  const double nrz = scalars[TLH_PCG_NEW_RZ];
  const bool move = !stopped && finite(c) && finite(rz) && finite(nrz) && c > 0.0 && rz > 0.0;
  const double beta = move ? nrz / rz : 0.0;
  // loads issued before the scalars are awaited; a frozen state rewrites p unchanged
  for (int64_t i = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * TLH_THREADS) {
    const Real zi = z[i], pi = p[i];
    p[i] = move ? (Real)((double)zi + beta * (double)pi) : pi;
  }
  double v[1] = {0.0};
  if (grid_sum<1>(v, partials, counter, smem) && threadIdx.x == 0 && !stopped) {
    const double norm2 = scalars[TLH_PCG_NORM2];
    control[0] += 1;
    if (!finite(c) || !finite(rz) || !finite(norm2))
      control[1] = 2;
    else if (c <= 0.0 || rz <= 0.0)
      control[1] = 3;
    else if (norm2 <= scalars[TLH_PCG_THR2])
      control[1] = 1;
    else if (!finite(nrz))
      control[1] = 2;
    scalars[TLH_PCG_RZ] = nrz;
  }
}

// One whole PCG iteration after Ap, for systems that fit one workgroup: the
// curvature p.Ap, the step, z = M^-1 r with r.z and r.r, the direction and the
// status machine -- the work of dot_kernel + step_kernel + direction_kernel in
// one launch, with the same guards and status codes (warp_pcg.py:33-92) and
// workgroup-wide fixed-order sums.  Each thread updates p for exactly the
// entries whose z it wrote, so no cross-thread global-memory handoff is needed.
constexpr int SMALL_THREADS = 1024;
constexpr int64_t SMALL_MAX_N = 16384;

template <int NV, int NTH = SMALL_THREADS>
__device__ __forceinline__ void small_sum(double (&v)[NV], double* smem) {
  // fixed order: an xor-shuffle tree inside each wavefront (every lane ends with
  // the same sum), then the wavefront sums in wavefront order
  constexpr int NWAVES = NTH / 64;
#pragma unroll
  for (int k = 0; k < NV; ++k)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off);
  const int wave = (int)threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) smem[k * NWAVES + wave] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double sum = smem[k * NWAVES];
    for (int w = 1; w < NWAVES; ++w) sum += smem[k * NWAVES + w];
    v[k] = sum;
  }
  __syncthreads();  // the next sum rewrites smem
}

template <typename Real>
__global__ void __launch_bounds__(SMALL_THREADS)
    small_iteration_kernel(tlh_block_diag_t pc, int identity, Real* __restrict__ x, Real* __restrict__ r,
                           Real* __restrict__ z, Real* __restrict__ p, const Real* __restrict__ ap, int64_t n,
                           int32_t* control, double* scalars) {
  __shared__ double smem[2 * (SMALL_THREADS / 64)];
  double cv[1] = {0.0};
  for (int64_t i = threadIdx.x; i < n; i += SMALL_THREADS) cv[0] += (double)p[i] * (double)ap[i];
  small_sum<1>(cv, smem);
  const double c = cv[0];
  const double rz = scalars[TLH_PCG_RZ];
  const bool stopped = frozen(control);
  const bool move = !stopped && finite(c) && finite(rz) && c > 0.0 && rz > 0.0;
  const double alpha = move ? rz / c : 0.0;
  double v[2] = {0.0, 0.0};
  int64_t blocks = 0;
  if (!identity)
    for (int t = 0; t < pc.ntypes; ++t) blocks += pc.count[t];
  if (!identity) {
    for (int64_t bidx = threadIdx.x; bidx < blocks; bidx += SMALL_THREADS) {
      int64_t local;
      const int t = find_type(pc, bidx, local);
      const int d = pc.dim[t];
      const int64_t e0 = pc.offset[t] + local * d;
      if (move) {
        for (int k = 0; k < d; ++k) {
          x[e0 + k] = (Real)((double)x[e0 + k] + alpha * (double)p[e0 + k]);
          r[e0 + k] = (Real)((double)r[e0 + k] - alpha * (double)ap[e0 + k]);
        }
      }
      apply_any<Real>(pc.inv[t] + local * (d * (d + 1) / 2), d, r + e0, z + e0, v[0], v[1]);
    }
  } else {
    for (int64_t i = threadIdx.x; i < n; i += SMALL_THREADS) {
      if (move) {
        x[i] = (Real)((double)x[i] + alpha * (double)p[i]);
        r[i] = (Real)((double)r[i] - alpha * (double)ap[i]);
      }
      const double ri = (double)r[i];
      z[i] = r[i];
      v[0] += ri * ri;
      v[1] += ri * ri;
    }
  }
  small_sum<2>(v, smem);
  const double nrz = v[0], norm2 = v[1];
  const bool turn = !stopped && finite(c) && finite(rz) && finite(nrz) && c > 0.0 && rz > 0.0;
  const double beta = turn ? nrz / rz : 0.0;
  if (turn) {
    if (!identity) {
      for (int64_t bidx = threadIdx.x; bidx < blocks; bidx += SMALL_THREADS) {
        int64_t local;
        const int t = find_type(pc, bidx, local);
        const int d = pc.dim[t];
        const int64_t e0 = pc.offset[t] + local * d;
        for (int k = 0; k < d; ++k) p[e0 + k] = (Real)((double)z[e0 + k] + beta * (double)p[e0 + k]);
      }
    } else {
      for (int64_t i = threadIdx.x; i < n; i += SMALL_THREADS) p[i] = (Real)((double)z[i] + beta * (double)p[i]);
    }
  }
  if (threadIdx.x == 0) {
    scalars[TLH_PCG_CURVATURE] = c;
    scalars[TLH_PCG_NEW_RZ] = nrz;
    scalars[TLH_PCG_NORM2] = norm2;
    if (!stopped) {
      control[0] += 1;
      if (!finite(c) || !finite(rz) || !finite(norm2))
        control[1] = 2;
      else if (c <= 0.0 || rz <= 0.0)
        control[1] = 3;
      else if (norm2 <= scalars[TLH_PCG_THR2])
        control[1] = 1;
      else if (!finite(nrz))
        control[1] = 2;
      scalars[TLH_PCG_RZ] = nrz;
    }
  }
}

// The same iteration for one block-Jacobi type of at most NTH blocks of
// dimension D (the reduced camera system): thread b owns block b and keeps its
// p, Ap, x, r, z and inverse in registers, all loaded at the start, so the
// iteration makes one trip to memory instead of one per phase.
template <typename Real, int D, int NTH>
__global__ void __launch_bounds__(NTH)
    small_block_iteration_kernel(const double* __restrict__ inv, int64_t count, Real* __restrict__ x,
                                 Real* __restrict__ r, Real* __restrict__ z, Real* __restrict__ p,
                                 const Real* __restrict__ ap, int32_t* control, double* scalars) {
  constexpr int W = D * (D + 1) / 2;
  __shared__ double smem[2 * (NTH / 64)];
  const int64_t bi = threadIdx.x;
  const bool own = bi < count;
  const int64_t e0 = bi * D;
  Real pv[D], av[D], xv[D], rv[D];
  double m[W];
  if (own) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      pv[k] = p[e0 + k];
      av[k] = ap[e0 + k];
      xv[k] = x[e0 + k];
      rv[k] = r[e0 + k];
    }
#pragma unroll
    for (int k = 0; k < W; ++k) m[k] = inv[bi * W + k];
  }
  double cv[1] = {0.0};
  if (own) {
#pragma unroll
    for (int k = 0; k < D; ++k) cv[0] += (double)pv[k] * (double)av[k];
  }
  small_sum<1, NTH>(cv, smem);
  const double c = cv[0];
  const double rz = scalars[TLH_PCG_RZ];
  const bool stopped = frozen(control);
  const bool move = !stopped && finite(c) && finite(rz) && c > 0.0 && rz > 0.0;
  const double alpha = move ? rz / c : 0.0;
  double v[2] = {0.0, 0.0};
  Real zv[D];
  if (own) {
    if (move) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        xv[k] = (Real)((double)xv[k] + alpha * (double)pv[k]);
        rv[k] = (Real)((double)rv[k] - alpha * (double)av[k]);
        x[e0 + k] = xv[k];
        r[e0 + k] = rv[k];
      }
    }
    double rd[D];
#pragma unroll
    for (int k = 0; k < D; ++k) rd[k] = (double)rv[k];
#pragma unroll
    for (int i = 0; i < D; ++i) {
      double t = 0.0;
#pragma unroll
      for (int j = 0; j < D; ++j) {
        const int a = i < j ? i : j, b = i < j ? j : i;
        t += m[a * D - a * (a - 1) / 2 + (b - a)] * rd[j];
      }
      zv[i] = (Real)t;
      z[e0 + i] = zv[i];
      v[0] += rd[i] * (double)zv[i];
      v[1] += rd[i] * rd[i];
    }
  }
  small_sum<2, NTH>(v, smem);
  const double nrz = v[0], norm2 = v[1];
  const bool turn = !stopped && finite(c) && finite(rz) && finite(nrz) && c > 0.0 && rz > 0.0;
  const double beta = turn ? nrz / rz : 0.0;
  if (own && turn) {
#pragma unroll
    for (int k = 0; k < D; ++k) p[e0 + k] = (Real)((double)zv[k] + beta * (double)pv[k]);
  }
  if (threadIdx.x == 0) {
    scalars[TLH_PCG_CURVATURE] = c;
    scalars[TLH_PCG_NEW_RZ] = nrz;
    scalars[TLH_PCG_NORM2] = norm2;
    if (!stopped) {
      control[0] += 1;
      if (!finite(c) || !finite(rz) || !finite(norm2))
        control[1] = 2;
      else if (c <= 0.0 || rz <= 0.0)
        control[1] = 3;
      else if (norm2 <= scalars[TLH_PCG_THR2])
        control[1] = 1;
      else if (!finite(nrz))
        control[1] = 2;
      scalars[TLH_PCG_RZ] = nrz;
    }
  }
}

template <typename Real>
int launch_small(const tlh_block_diag_t* pc, int mode, void* x, void* r, void* z, void* p, const void* ap,
                 int64_t n, const tlh_pcg_state_t* st, hipStream_t s) {
  Real* xr = static_cast<Real*>(x);
  Real* rr = static_cast<Real*>(r);
  Real* zr = static_cast<Real*>(z);
  Real* pr = static_cast<Real*>(p);
  const Real* ar = static_cast<const Real*>(ap);
  // one block type in registers: 256 threads when they cover the blocks (four
  // wavefronts, no register cap), else 1,024 (6x6 float64 blocks would spill)
  const bool one = mode == 0 && pc->ntypes == 1 && (pc->dim[0] == 6 || pc->dim[0] == 3);
  const int64_t cnt = one ? pc->count[0] : 0;
  const int64_t o = one ? pc->offset[0] : 0;
  void (*kernel)(const double*, int64_t, Real*, Real*, Real*, Real*, const Real*, int32_t*, double*) = nullptr;
  int nth = 256;
  if (one && cnt <= 256) {
    kernel = pc->dim[0] == 6 ? small_block_iteration_kernel<Real, 6, 256> : small_block_iteration_kernel<Real, 3, 256>;
  } else if (one && cnt <= SMALL_THREADS && pc->dim[0] == 3) {
    kernel = small_block_iteration_kernel<Real, 3, SMALL_THREADS>;
    nth = SMALL_THREADS;
  } else if constexpr (sizeof(Real) == 4) {
    if (one && cnt <= SMALL_THREADS) {
      kernel = small_block_iteration_kernel<Real, 6, SMALL_THREADS>;
      nth = SMALL_THREADS;
    }
  }
  if (kernel) {
    hipLaunchKernelGGL(kernel, dim3(1), dim3(nth), 0, s, pc->inv[0], cnt, xr + o, rr + o, zr + o, pr + o, ar + o,
                       st->control, st->scalars);
  } else {
    tlh_block_diag_t empty{};
    hipLaunchKernelGGL(small_iteration_kernel<Real>, dim3(1), dim3(SMALL_THREADS), 0, s, mode == 0 ? *pc : empty,
                       mode, xr, rr, zr, pr, ar, n, st->control, st->scalars);
  }
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

// r = src; optionally x = 0; z = M^-1 r; p = z; rz = r.z (warp_pcg.py:130-132, 169-172, 210-212).
template <typename Real>
__global__ void __launch_bounds__(TLH_THREADS)
    reset_kernel(tlh_block_diag_t pc, int identity, const Real* __restrict__ src, Real* __restrict__ x,
                 Real* __restrict__ r, Real* __restrict__ z, Real* __restrict__ p, int64_t n, double* scalars,
                 double* partials, uint32_t* counter) {
  __shared__ double smem[2 * (TLH_THREADS / 64)];
  double v[2] = {0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * TLH_THREADS;
  if (!identity) {
    int64_t blocks = 0;
    for (int t = 0; t < pc.ntypes; ++t) blocks += pc.count[t];
    for (int64_t bidx = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x; bidx < blocks; bidx += stride) {
      int64_t local;
      const int t = find_type(pc, bidx, local);
      const int d = pc.dim[t];
      const int64_t e0 = pc.offset[t] + local * d;
      for (int k = 0; k < d; ++k) {
        r[e0 + k] = src[e0 + k];
        if (x) x[e0 + k] = Real(0);
      }
      apply_any<Real>(pc.inv[t] + local * (d * (d + 1) / 2), d, r + e0, z + e0, v[0], v[1]);
      for (int k = 0; k < d; ++k) p[e0 + k] = z[e0 + k];
    }
  } else {
    for (int64_t i = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x; i < n; i += stride) {
      const Real ri = src[i];
      r[i] = ri;
      z[i] = ri;
      p[i] = ri;
      if (x) x[i] = Real(0);
      v[0] += (double)ri * (double)ri;
    }
  }
  if (grid_sum<2>(v, partials, counter, smem) && threadIdx.x == 0) scalars[TLH_PCG_RZ] = v[0];
}

// true_r = rhs - Ax; ||true_r||^2 and ||true_r - r||^2 (warp_pcg.py:189-204).
template <typename Real>
__global__ void __launch_bounds__(TLH_THREADS)
    true_residual_kernel(const Real* __restrict__ rhs, const Real* __restrict__ ax, const Real* __restrict__ r,
                         Real* __restrict__ true_r, int64_t n, const int32_t* control, double* scalars,
                         double* partials, uint32_t* counter) {
  __shared__ double smem[2 * (TLH_THREADS / 64)];
  double v[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * TLH_THREADS) {
    const Real t = rhs[i] - ax[i];
    true_r[i] = t;
    const double dt = (double)t, dd = (double)t - (double)r[i];
    v[0] += dt * dt;
    v[1] += dd * dd;
  }
  if (grid_sum<2>(v, partials, counter, smem) && threadIdx.x == 0) {
    scalars[TLH_PCG_TRUE2] = v[0];
    scalars[TLH_PCG_DRIFT2] = v[1];
    // the control word beside them, so the host reads one contiguous slice
    scalars[TLH_PCG_ITERATIONS] = (double)control[0];
    scalars[TLH_PCG_STATUS] = (double)control[1];
  }
}

bool state_ok(const tlh_pcg_state_t* st) {
  return st && st->control && st->scalars && st->partials && st->counter;
}

int64_t precond_blocks(const tlh_block_diag_t* pc) {
  int64_t b = 0;
  for (int t = 0; t < pc->ntypes; ++t) b += pc->count[t];
  return b;
}

template <typename Real>
int step_impl(const tlh_block_diag_t* pc, int mode, void* x, void* r, void* z, const void* p, const void* ap,
              int64_t n, const tlh_pcg_state_t* st, hipStream_t s) {
  tlh_block_diag_t empty{};
  const int64_t work = mode == 0 ? precond_blocks(pc) : n;
  hipLaunchKernelGGL(step_kernel<Real>, dim3(grid_for(n, work)), dim3(TLH_THREADS), 0, s, mode == 0 ? *pc : empty,
                     mode, (Real*)x, (Real*)r, (Real*)z, (const Real*)p, (const Real*)ap, n, st->control,
                     st->scalars, st->partials, st->counter);
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

template <typename Real>
int reset_impl(const tlh_block_diag_t* pc, const void* src, void* x, void* r, void* z, void* p, int64_t n,
               const tlh_pcg_state_t* st, hipStream_t s) {
  tlh_block_diag_t empty{};
  const int64_t work = pc ? precond_blocks(pc) : n;
  hipLaunchKernelGGL(reset_kernel<Real>, dim3(grid_for(n, work)), dim3(TLH_THREADS), 0, s, pc ? *pc : empty,
                     pc ? 0 : 1, (const Real*)src, (Real*)x, (Real*)r, (Real*)z, (Real*)p, n, st->scalars,
                     st->partials, st->counter);
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

bool precond_ok(const tlh_block_diag_t* pc, int64_t n) {
  if (!pc) return true;
  if (pc->ntypes < 1 || pc->ntypes > TLH_MAX_TYPES) return false;
  int64_t total = 0;
  for (int t = 0; t < pc->ntypes; ++t) {
    if (pc->dim[t] < 1 || pc->dim[t] > 16 || !pc->inv[t]) return false;
    total += pc->count[t] * pc->dim[t];
  }
  return total == n;
}

}  // namespace

extern "C" int tlh_pcg_iteration_small(const tlh_block_diag_t* pc, int mode, void* x, void* r, void* z, void* p,
                                       const void* ap, int64_t n, int dtype, const tlh_pcg_state_t* st,
                                       tlh_stream_t stream) {
  if (!x || !r || !z || !p || !ap || n < 1 || n > SMALL_MAX_N || !state_ok(st)) return TLH_EINVAL;
  if (mode != 0 && mode != 1) return TLH_EINVAL;
  if (mode == 0 && (!pc || !precond_ok(pc, n))) return TLH_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (dtype == TLH_F32) return launch_small<float>(pc, mode, x, r, z, p, ap, n, st, s);
  if (dtype == TLH_F64) return launch_small<double>(pc, mode, x, r, z, p, ap, n, st, s);
  return TLH_EINVAL;
}

extern "C" int tlh_block_inverse(const double* blocks, int64_t count, int dim, double damping, double* inverse,
                                 int32_t* status, tlh_stream_t stream) {
  if (count < 0 || !status || (count && (!blocks || !inverse))) return TLH_EINVAL;
  if (count == 0) return 0;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid((unsigned)ceil_div(count, TLH_THREADS)), block(TLH_THREADS);
  switch (dim) {
#define TLH_INV_CASE(D)                                                                               \
  case D:                                                                                             \
    hipLaunchKernelGGL(block_inverse_kernel<D>, grid, block, 0, s, blocks, count, damping, inverse, status); \
    break;
    TLH_INV_CASE(1)
    TLH_INV_CASE(2)
    TLH_INV_CASE(3)
    TLH_INV_CASE(4)
    TLH_INV_CASE(5)
    TLH_INV_CASE(6)
    TLH_INV_CASE(7)
    TLH_INV_CASE(8)
    TLH_INV_CASE(9)
    TLH_INV_CASE(10)
    TLH_INV_CASE(11)
    TLH_INV_CASE(12)
    TLH_INV_CASE(13)
    TLH_INV_CASE(14)
    TLH_INV_CASE(15)
    TLH_INV_CASE(16)
#undef TLH_INV_CASE
    default:
      return TLH_ENOTFOUND;
  }
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

extern "C" int tlh_dot(const void* a, const void* b, int64_t n, int dtype, const tlh_pcg_state_t* st, int slot,
                       tlh_stream_t stream) {
  if (!state_ok(st) || slot < 0 || slot >= TLH_PCG_NSCALARS || n < 0) return TLH_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid(grid_for(n)), block(TLH_THREADS);
  if (dtype == TLH_F32)
    hipLaunchKernelGGL(dot_kernel<float>, grid, block, 0, s, (const float*)a, (const float*)b, n, st->scalars,
                       slot, st->partials, st->counter);
  else if (dtype == TLH_F64)
    hipLaunchKernelGGL(dot_kernel<double>, grid, block, 0, s, (const double*)a, (const double*)b, n, st->scalars,
                       slot, st->partials, st->counter);
  else
    return TLH_EINVAL;
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

namespace {

__global__ void begin_kernel(double* scalars, int32_t* control, double atol, double rtol, int32_t stop_at,
                             const int32_t* pd_status) {
  if (threadIdx.x != 0) return;
  const double rhs2 = scalars[TLH_PCG_AUX];
  const double norm = rhs2 >= 0.0 ? sqrt(rhs2) : (double)NAN;
  const double scaled = rtol * norm;
  const double thr = scaled > atol ? scaled : atol;  // max(atol, rtol * norm); a NaN norm gives atol
  scalars[TLH_PCG_THR2] = thr * thr;
  int32_t status = 0;
  if (pd_status && *pd_status != 0)
    status = 4;
  else if (!isfinite(norm))
    status = 2;
  else if (norm <= thr)
    status = 1;
  control[0] = 0;
  control[1] = status;
  control[2] = stop_at;
  control[3] = 0;
}

}  // namespace

extern "C" int tlh_pcg_begin(const tlh_pcg_state_t* st, double atol, double rtol, int32_t stop_at,
                             const int32_t* pd_status, tlh_stream_t stream) {
  if (!state_ok(st) || stop_at < 0) return TLH_EINVAL;
  hipLaunchKernelGGL(begin_kernel, dim3(1), dim3(64), 0, reinterpret_cast<hipStream_t>(stream), st->scalars,
                     st->control, atol, rtol, stop_at, pd_status);
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

extern "C" int tlh_pcg_reset(const tlh_block_diag_t* pc, const void* src, void* x, void* r, void* z, void* p,
                             int64_t n, int dtype, const tlh_pcg_state_t* st, tlh_stream_t stream) {
  if (!state_ok(st) || !precond_ok(pc, n) || !src || !r || !z || !p) return TLH_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (dtype == TLH_F32) return reset_impl<float>(pc, src, x, r, z, p, n, st, s);
  if (dtype == TLH_F64) return reset_impl<double>(pc, src, x, r, z, p, n, st, s);
  return TLH_EINVAL;
}

extern "C" int tlh_pcg_step(const tlh_block_diag_t* pc, int mode, void* x, void* r, void* z, const void* p,
                            const void* ap, int64_t n, int dtype, const tlh_pcg_state_t* st, tlh_stream_t stream) {
  if (!state_ok(st) || mode < 0 || mode > 2 || (mode == 0 && (!pc || !precond_ok(pc, n)))) return TLH_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (dtype == TLH_F32) return step_impl<float>(pc, mode, x, r, z, p, ap, n, st, s);
  if (dtype == TLH_F64) return step_impl<double>(pc, mode, x, r, z, p, ap, n, st, s);
  return TLH_EINVAL;
}

extern "C" int tlh_pcg_direction(void* p, const void* z, int64_t n, int dtype, const tlh_pcg_state_t* st,
                                 tlh_stream_t stream) {
  if (!state_ok(st)) return TLH_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid(grid_for(n)), block(TLH_THREADS);
  if (dtype == TLH_F32)
    hipLaunchKernelGGL(direction_kernel<float>, grid, block, 0, s, (float*)p, (const float*)z, n, st->control,
                       st->scalars, st->partials, st->counter);
  else if (dtype == TLH_F64)
    hipLaunchKernelGGL(direction_kernel<double>, grid, block, 0, s, (double*)p, (const double*)z, n, st->control,
                       st->scalars, st->partials, st->counter);
  else
    return TLH_EINVAL;
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

extern "C" int tlh_pcg_true_residual(const void* rhs, const void* ax, const void* r, void* true_r, int64_t n,
                                     int dtype, const tlh_pcg_state_t* st, tlh_stream_t stream) {
  if (!state_ok(st)) return TLH_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid(grid_for(n)), block(TLH_THREADS);
  if (dtype == TLH_F32)
    hipLaunchKernelGGL(true_residual_kernel<float>, grid, block, 0, s, (const float*)rhs, (const float*)ax,
                       (const float*)r, (float*)true_r, n, st->control, st->scalars, st->partials, st->counter);
  else if (dtype == TLH_F64)
    hipLaunchKernelGGL(true_residual_kernel<double>, grid, block, 0, s, (const double*)rhs, (const double*)ax,
                       (const double*)r, (double*)true_r, n, st->control, st->scalars, st->partials,
                       st->counter);
  else
    return TLH_EINVAL;
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

namespace {
template <typename Real>
__global__ void __launch_bounds__(TLH_THREADS)
    block_apply_kernel(const double* __restrict__ inv, int d, const int32_t* __restrict__ list, int64_t count,
                       const Real* __restrict__ x, Real* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x;
  if (i >= count) return;
  const int64_t v = list ? (int64_t)list[i] : i;
  double rz = 0.0, rr = 0.0;
  apply_any<Real>(inv + v * (d * (d + 1) / 2), d, x + v * d, y + v * d, rz, rr);
}
}  // namespace

extern "C" int tlh_block_apply(const double* inverse, int dim, const int32_t* list, int64_t count, const void* x,
                               void* y, int dtype, tlh_stream_t stream) {
  if (count < 0 || dim < 1 || dim > 16 || (count && (!inverse || !x || !y))) return TLH_EINVAL;
  if (count == 0) return 0;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid((unsigned)ceil_div(count, TLH_THREADS)), block(TLH_THREADS);
  if (dtype == TLH_F32)
    hipLaunchKernelGGL(block_apply_kernel<float>, grid, block, 0, s, inverse, dim, list, count,
                       static_cast<const float*>(x), static_cast<float*>(y));
  else if (dtype == TLH_F64)
    hipLaunchKernelGGL(block_apply_kernel<double>, grid, block, 0, s, inverse, dim, list, count,
                       static_cast<const double*>(x), static_cast<double*>(y));
  else
    return TLH_EINVAL;
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

// Native batch of PCG iterations: the launches of one iteration (operator,
// then the fused update) issued back to back from here instead of a replayed
// graph, so that a batch costs one host call.  The operator entry points below
// are this library's own (normal.hip, schur.hip).
extern "C" int tlh_pcg_iterations(const tlh_pcg_operator_t* A, const tlh_block_diag_t* pc, int mode, void* x,
                                  void* r, void* z, void* p, void* ap, int64_t n, int dtype,
                                  const tlh_pcg_state_t* st, int64_t small_max, int32_t count,
                                  tlh_stream_t stream) {
  if (!A || !A->op || !x || !r || !z || !p || !ap || n < 1 || count < 0 || !state_ok(st)) return TLH_EINVAL;
  if (mode != 0 && mode != 1) return TLH_EINVAL;
  if (mode == 0 && (!pc || !precond_ok(pc, n))) return TLH_EINVAL;
  if (dtype != TLH_F32 && dtype != TLH_F64) return TLH_EINVAL;
  if (A->kind == TLH_PCG_OP_NORMAL) {
    if (A->op->size != n) return TLH_EINVAL;
  } else if (A->kind == TLH_PCG_OP_SCHUR) {
    if (!A->cinv || A->wgroup < 0 || A->wgroup >= A->op->ngroups || A->type_end < 1 ||
        A->type_end > A->op->ntypes)
      return TLH_EINVAL;
  } else {
    return TLH_EINVAL;
  }
  const bool small = n <= small_max && n <= SMALL_MAX_N;
  for (int32_t it = 0; it < count; ++it) {
    int rc = 0;
    if (A->kind == TLH_PCG_OP_NORMAL) {
      rc = tlh_normal_matvec(A->op, p, ap, A->damping, stream);
    } else {
      rc = tlh_schur_apply(A->op, A->wgroup, A->eslot, TLH_SCHUR_MATVEC, p, nullptr, nullptr, A->cinv, stream);
      for (int32_t g = 0; rc == 0 && g < A->op->ngroups; ++g)
        if (g != A->wgroup) rc = tlh_group_tiles(A->op, g, p, stream);
      if (rc == 0) rc = tlh_finish_types(A->op, 0, A->type_end, p, A->damping, ap, stream);
    }
    if (rc != 0) return rc;
    if (small) {
      rc = tlh_pcg_iteration_small(pc, mode, x, r, z, p, ap, n, dtype, st, stream);
    } else {
      rc = tlh_dot(p, ap, n, dtype, st, TLH_PCG_CURVATURE, stream);
      if (rc == 0) rc = tlh_pcg_step(pc, mode, x, r, z, p, ap, n, dtype, st, stream);
      if (rc == 0) rc = tlh_pcg_direction(p, z, n, dtype, st, stream);
    }
    if (rc != 0) return rc;
  }
  return 0;
}
