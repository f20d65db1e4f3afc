// Deterministic fused normal matvec y = J^T (J x) + lambda x for gfx950.
//
// Replaces warp_normal.py (_factor_kernel 69-103, _block_factor_kernel 26-66,
// _finish 106-122).  Instead of one float atomic per observation column, every
// workgroup reduces a tile of observations in a fixed order:
//
//   1. each thread takes observations (tid + k*256) of the tile, gathers x for
//      every slot, forms q = sum_s J_s x_s and the slot contributions J_s^T q
//      (registers; Jacobian components are read coalesced from the packed
//      component-major layout);
//   2. per slot, contributions are staged in LDS in the tile's order by that
//      slot's variable index (precomputed rank), and each precomputed segment
//      (a run of one variable) is summed serially by one thread;
//   3. segment sums go either to a partial slot in HBM (SPARSE) or into a
//      per-workgroup dense LDS accumulator of the whole variable type (DENSE,
//      used when the type is small, e.g. cameras), which the persistent
//      workgroup writes out once;
//   4. a finish kernel adds each variable's partials in plan order plus
//      lambda * x.
//
// Every float operation therefore happens in an order fixed by the plan, so
// results are bitwise reproducible run to run and inside hipGraph replays.
#include <algorithm>

#include "common.hpp"

using namespace tlh;

namespace {

template <int D0, int D1, int D2>
struct Dims {
  static constexpr int ns = 1 + (D1 > 0) + (D2 > 0);
  static constexpr int sum = D0 + D1 + D2;
  static constexpr int max = D0 > D1 ? (D0 > D2 ? D0 : D2) : (D1 > D2 ? D1 : D2);
  static constexpr int d(int s) { return s == 0 ? D0 : (s == 1 ? D1 : D2); }
  static constexpr int off(int s) { return s == 0 ? 0 : (s == 1 ? D0 : D0 + D1); }
};

// Sum the segments of one slot of tile `tile` from the LDS stage (entries of
// `dim` values in the slot's tile order) into partials or the dense accumulator.
template <typename Real, int D>
__device__ __forceinline__ void reduce_segments(const SlotArgs<Real>& s, int tile, int len,
                                                const Real* stage, Real* acc) {
  const int sb = s.seg_ptr[tile];
  const int se = s.seg_ptr[tile + 1];
  for (int k = sb + (int)threadIdx.x; k < se; k += TLH_THREADS) {
    const int a = s.seg_start[k];
    const int b = (k + 1 < se) ? s.seg_start[k + 1] : len;
    Real sum[D];
#pragma unroll
    for (int c = 0; c < D; ++c) sum[c] = stage[a * D + c];
    for (int i = a + 1; i < b; ++i) {
#pragma unroll
      for (int c = 0; c < D; ++c) sum[c] += stage[i * D + c];
    }
    const int target = s.seg_target[k];
    if (s.mode == TLH_REDUCE_DENSE) {
#pragma unroll
      for (int c = 0; c < D; ++c) acc[target * D + c] += sum[c];
    } else {
      Real* out = s.partials + (int64_t)target * D;
#pragma unroll
      for (int c = 0; c < D; ++c) out[c] = sum[c];
    }
  }
}

// Runtime-dimension variant used by the generic kernel.
template <typename Real>
__device__ __forceinline__ void reduce_segments_rt(const SlotArgs<Real>& s, int D, int tile,
                                                   int len, const Real* stage, Real* acc) {
  const int sb = s.seg_ptr[tile];
  const int se = s.seg_ptr[tile + 1];
  for (int k = sb + (int)threadIdx.x; k < se; k += TLH_THREADS) {
    const int a = s.seg_start[k];
    const int b = (k + 1 < se) ? s.seg_start[k + 1] : len;
    const int target = s.seg_target[k];
    for (int c = 0; c < D; ++c) {
      Real sum = stage[a * D + c];
      for (int i = a + 1; i < b; ++i) sum += stage[i * D + c];
      if (s.mode == TLH_REDUCE_DENSE)
        acc[target * D + c] += sum;
      else
        s.partials[(int64_t)target * D + c] = sum;
    }
  }
}

template <typename Real>
__device__ __forceinline__ Real* dense_regions(const GroupArgs<Real>& g, Real* base, Real* acc[3]) {
  Real* p = base;
  for (int s = 0; s < TLH_MAX_SLOTS; ++s) {
    if (s < g.nslots && g.slot[s].mode == TLH_REDUCE_DENSE) {
      acc[s] = p;
      p += (int64_t)g.slot[s].type_count * g.slot[s].dim;
    } else {
      acc[s] = nullptr;
    }
  }
  return p;
}

template <typename Real>
__device__ __forceinline__ void zero_dense(const GroupArgs<Real>& g, Real* acc[3]) {
  for (int s = 0; s < g.nslots; ++s) {
    if (acc[s] == nullptr) continue;
    const int n = g.slot[s].type_count * g.slot[s].dim;
    for (int i = threadIdx.x; i < n; i += TLH_THREADS) acc[s][i] = Real(0);
  }
}

template <typename Real>
__device__ __forceinline__ void flush_dense(const GroupArgs<Real>& g, Real* acc[3]) {
  for (int s = 0; s < g.nslots; ++s) {
    if (acc[s] == nullptr) continue;
    const int64_t n = (int64_t)g.slot[s].type_count * g.slot[s].dim;
    Real* out = g.slot[s].partials + (int64_t)(g.slot[s].dense_block + blockIdx.x) * n;
    for (int64_t i = threadIdx.x; i < n; i += TLH_THREADS) out[i] = acc[s][i];
  }
}

// Specialised kernel: residual rows R and slot dims D0..D2 known at compile time.
template <typename Real, int R, int D0, int D1, int D2, int ITEMS>
__global__ void __launch_bounds__(TLH_THREADS) normal_tile_kernel(GroupArgs<Real> g) {
  using S = Dims<D0, D1, D2>;
  constexpr int NS = S::ns;
  constexpr int DSUM = S::sum;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Real* stage = reinterpret_cast<Real*>(smem_raw);
  Real* acc[3];
  dense_regions(g, stage + TLH_THREADS * ITEMS * S::max, acc);
  zero_dense(g, acc);

  const int t0 = blockIdx.x * g.tiles_per_wg;
  const int t1 = min(t0 + g.tiles_per_wg, g.ntiles);
  for (int t = t0; t < t1; ++t) {
    const int start = g.tile_start[t];
    const int len = g.tile_start[t + 1] - start;
    Real cont[ITEMS][DSUM];
    int rk[ITEMS][NS];
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
      const int pos = (int)threadIdx.x + it * TLH_THREADS;
      const int64_t n = start + min(pos, len - 1);
      Real xv[DSUM];
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const int d = S::d(s);
        const int64_t v = g.slot[s].index[n];
        const Real* xs = g.slot[s].x + v * d;
#pragma unroll
        for (int c = 0; c < d; ++c) xv[S::off(s) + c] = xs[c];
        rk[it][s] = g.slot[s].rank ? (int)g.slot[s].rank[n] : pos;
      }
      Real jv[R][DSUM];
#pragma unroll
      for (int s = 0; s < NS; ++s) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
          for (int c = 0; c < S::d(s); ++c) {
            const int comp = R * S::off(s) + r * S::d(s) + c;
            jv[r][S::off(s) + c] = g.packed[(int64_t)comp * g.stride + n];
          }
        }
      }
      Real q[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        Real a = Real(0);
#pragma unroll
        for (int c = 0; c < DSUM; ++c) a += jv[r][c] * xv[c];
        q[r] = a;
      }
#pragma unroll
      for (int c = 0; c < DSUM; ++c) {
        Real a = Real(0);
#pragma unroll
        for (int r = 0; r < R; ++r) a += jv[r][c] * q[r];
        cont[it][c] = a;
      }
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int d = S::d(s);
      __syncthreads();
#pragma unroll
      for (int it = 0; it < ITEMS; ++it) {
        const int pos = (int)threadIdx.x + it * TLH_THREADS;
        if (pos < len) {
#pragma unroll
          for (int c = 0; c < d; ++c) stage[rk[it][s] * d + c] = cont[it][S::off(s) + c];
        }
      }
      __syncthreads();
      if (s == 0) reduce_segments<Real, D0>(g.slot[0], t, len, stage, acc[0]);
      if constexpr (D1 > 0) {
        if (s == 1) reduce_segments<Real, (D1 > 0 ? D1 : 1)>(g.slot[1], t, len, stage, acc[1]);
      }
      if constexpr (D2 > 0) {
        if (s == 2) reduce_segments<Real, (D2 > 0 ? D2 : 1)>(g.slot[2], t, len, stage, acc[2]);
      }
    }
  }
  __syncthreads();
  flush_dense(g, acc);
}

// Empty asm with the values as register inputs and a memory clobber: the values
// are computed before it and no memory access moves across it (it emits nothing).
template <typename T, int A, int B>
__device__ __forceinline__ void pin_before_loads(const T (&v)[A][B]) {
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int j = 0; j < B; ++j) asm volatile("" ::"v"(v[i][j]) : "memory");
}

template <typename T, int N>
struct VecT {
  typedef T type __attribute__((ext_vector_type(N)));
};

template <typename Real>
struct WaveArgs {
  const Real* packed;
  int64_t stride;
  int64_t count;
  int32_t ntiles;
  int32_t tiles_per_wg;
  const int32_t* idx_p;
  const int32_t* idx_s;
  const uint16_t* rank_s;
  const Real* x_p;
  const Real* x_s;
  Real* y_p;
  const uint8_t* direct_p;
  Real* part_p;
  int64_t pos_base_p;
  Real* part_s;
  int64_t pos_base_s;
  int32_t mode_s;
  int32_t dense_block_s;
  int32_t count_s;
  Real damping;
  const int32_t* rounds;  // occurrence rounds per tile (dense secondary) or NULL
  int32_t out_cap;        // staged primary outputs per wavefront (FAST kernel)
  int32_t stage_cap;      // overflow positions of one tile (FAST kernel)
};

// out[k] = base[start + b + k]; a short tile (len < T) clamps to its last position.
// STREAM marks data read once per launch from a working set larger than the
// last-level cache: the loads carry the non-temporal hint.
template <typename T, int ITEMS, bool STREAM = false>
__device__ __forceinline__ void load_items(const T* __restrict__ base, int64_t start, int b, int len, bool full,
                                           T (&out)[ITEMS]) {
  if constexpr (ITEMS == 1) {
    out[0] = base[start + min(b, len - 1)];
  } else if (full) {
    typedef typename VecT<T, ITEMS>::type V;
    const V v = STREAM ? __builtin_nontemporal_load(reinterpret_cast<const V*>(base + start + b))
                       : *reinterpret_cast<const V*>(base + start + b);
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) out[k] = v[k];
  } else {
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) out[k] = base[start + min(b + k, len - 1)];
  }
}

// Wave-kernel loads: a uniform tile base plus a 32-bit byte offset per lane, so
// that each load addresses through a scalar base and one shared offset register
// instead of a 64-bit address computed per load.
template <typename T, int ITEMS, bool STREAM = false>
__device__ __forceinline__ void load_lane(const T* __restrict__ tile, uint32_t off, T (&out)[ITEMS]) {
  const char* p = reinterpret_cast<const char*>(tile) + off;
  if constexpr (ITEMS == 1) {
    out[0] = STREAM ? __builtin_nontemporal_load(reinterpret_cast<const T*>(p)) : *reinterpret_cast<const T*>(p);
  } else {
    typedef typename VecT<T, ITEMS>::type V;
    const V v = STREAM ? __builtin_nontemporal_load(reinterpret_cast<const V*>(p)) : *reinterpret_cast<const V*>(p);
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) out[k] = v[k];
  }
}

// Loads of the pipeline are issued on every path: a load skipped on one path, or
// a generic (flat) load, makes the compiler's later wait counts conservative,
// down to waiting for every load in flight.  An absent array is therefore
// replaced, through a select, by another kernel argument array (global memory,
// always valid) and the loaded value is discarded.

// Raw per-thread data of one tile: ids, ranks, block components and the
// gathered primary x (and secondary x when it is not held in LDS).
template <typename Real, int R, int DP, int DS, int ITEMS, bool XL>
struct WaveTile {
  static constexpr int RW = (ITEMS + 1) / 2;
  int ip[ITEMS], is[ITEMS];
  uint32_t rk[RW];
  int rounds;
  Real jp[R][DP][ITEMS], js[R][DS][ITEMS];
  Real xp[ITEMS][DP], xs[ITEMS][XL ? 1 : DS];
  uint8_t direct[ITEMS];  // raw flag bytes (a widened copy would force a wait)
  __device__ __forceinline__ int rank(int k) const {
    return ITEMS == 1 ? (int)rk[0] : (int)((k & 1) ? (rk[k >> 1] >> 16) : (rk[k >> 1] & 0xffffu));
  }
};

// Whole-tile loads: the packed stride and the id/rank arrays of a wave group are
// padded to ntiles * T positions (ids with a valid variable), so no load needs a
// bounds branch; positions past the group's count are masked by the caller.
// Ranks are read as single 16-bit loads (zero-extended, no unpacking) so that
// nothing forces a wait between the id loads and the block loads.
template <typename Real, int ITEMS, int NT, bool FAST, typename Tile>
__device__ __forceinline__ void load_wave_ids(const WaveArgs<Real>& a, int t, int b, Tile& d) {
  constexpr int T = NT * ITEMS;
  const int64_t start = (int64_t)t * T;
  const uint32_t off = (uint32_t)b * 4u;
  load_lane<int32_t, ITEMS>(a.idx_p + start, off, d.ip);
  load_lane<int32_t, ITEMS>(a.idx_s + start, off, d.is);
  if constexpr (FAST) {
    // occurrence ranks are always present; the round counts are read from LDS
    const char* rp = reinterpret_cast<const char*>(a.rank_s + start) + (uint32_t)b * 2u;
    if constexpr (ITEMS == 1) {
      d.rk[0] = *reinterpret_cast<const uint16_t*>(rp);
    } else {
#pragma unroll
      for (int w = 0; w < ITEMS / 2; ++w) d.rk[w] = reinterpret_cast<const uint32_t*>(rp)[w];
    }
    d.rounds = 0;
    return;
  }
  // without ranks (the tile is in secondary order) the loads read the id array
  // (valid memory of at least twice the size) and the identity ranks are selected
  const bool ranked = a.rank_s != nullptr;
  const uint16_t* rank = ranked ? a.rank_s : reinterpret_cast<const uint16_t*>(a.idx_s);
  if constexpr (ITEMS == 1) {
    const uint32_t v = rank[start + b];
    d.rk[0] = ranked ? v : (uint32_t)b;
  } else {
    const uint32_t* r32 = reinterpret_cast<const uint32_t*>(rank + start + b);
#pragma unroll
    for (int w = 0; w < ITEMS / 2; ++w) {
      const uint32_t v = r32[w];
      d.rk[w] = ranked ? v : ((uint32_t)(b + 2 * w) | ((uint32_t)(b + 2 * w + 1) << 16));
    }
  }
  const int32_t nr = *(a.rounds ? a.rounds + t : a.idx_p);
  d.rounds = a.rounds ? nr : 0;
}

// Block loads of tile t: no dependence on anything loaded before them, so they
// are issued a whole tile ahead of their use.
template <typename Real, int R, int DP, int DS, int OP, int OS, int ITEMS, int NT, bool STREAM, bool XL>
__device__ __forceinline__ void load_wave_blocks(const WaveArgs<Real>& a, int t, int64_t stride, int b,
                                                 WaveTile<Real, R, DP, DS, ITEMS, XL>& d) {
  constexpr int T = NT * ITEMS;
  const Real* tile = a.packed + (int64_t)t * T;
  const uint32_t off = (uint32_t)b * (uint32_t)sizeof(Real);
#pragma unroll
  for (int r = 0; r < R; ++r) {
#pragma unroll
    for (int c = 0; c < DP; ++c)
      load_lane<Real, ITEMS, STREAM>(tile + (int64_t)(OP + r * DP + c) * stride, off, d.jp[r][c]);
#pragma unroll
    for (int c = 0; c < DS; ++c)
      load_lane<Real, ITEMS, STREAM>(tile + (int64_t)(OS + r * DS + c) * stride, off, d.js[r][c]);
  }
}

// The x gathers addressed by a tile's ids (loaded into the tile earlier: no id
// is copied between registers, so no copy waits for a load in flight).
template <typename Real, int R, int DP, int DS, int ITEMS, bool XL>
__device__ __forceinline__ void load_wave_gathers(const WaveArgs<Real>& a, WaveTile<Real, R, DP, DS, ITEMS, XL>& d) {
  const WaveTile<Real, R, DP, DS, ITEMS, XL>& ids = d;
  const bool has_direct = a.direct_p != nullptr;
  const uint8_t* direct = has_direct ? a.direct_p : reinterpret_cast<const uint8_t*>(a.idx_p);
  // 32-bit byte offsets: the launcher checks that each x slice is below 4 GiB
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const char* xp = reinterpret_cast<const char*>(a.x_p) + (uint32_t)ids.ip[k] * (uint32_t)(DP * sizeof(Real));
#pragma unroll
    for (int c = 0; c < DP; ++c) d.xp[k][c] = reinterpret_cast<const Real*>(xp)[c];
    if constexpr (!XL) {
      const char* xs = reinterpret_cast<const char*>(a.x_s) + (uint32_t)ids.is[k] * (uint32_t)(DS * sizeof(Real));
#pragma unroll
      for (int c = 0; c < DS; ++c) d.xs[k][c] = reinterpret_cast<const Real*>(xs)[c];
    }
    // raw byte (garbage without the array): tested against direct_p where used, so
    // that no select right after the load forces a wait for it
    d.direct[k] = direct[has_direct ? (uint32_t)ids.ip[k] : 0u];
  }
}

// Empty asm that takes every block component of a tile as a read-write register
// operand, with a memory clobber: memory accesses before it (the next tile's
// block loads) are issued first, and arithmetic on the components cannot move
// above it.
template <typename Real, int R, int DP, int DS, int ITEMS, bool XL>
__device__ __forceinline__ void pin_blocks(WaveTile<Real, R, DP, DS, ITEMS, XL>& d) {
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
#pragma unroll
      for (int c = 0; c < DP; ++c) asm volatile("" : "+v"(d.jp[r][c][k])::"memory");
#pragma unroll
      for (int c = 0; c < DS; ++c) asm volatile("" : "+v"(d.js[r][c][k])::"memory");
    }
}

// Wave kernel for two-slot groups (bundle adjustment, pose-graph edges).
// Tiles are aligned: tile t covers sorted positions [t*T, t*T + T), T = NT *
// ITEMS, and thread i of the workgroup owns the ITEMS consecutive positions
// ITEMS*i .. ITEMS*i + ITEMS-1, so every block component and id array is read
// with one vector load per thread.  Two register buffers alternate so that the
// next tile's loads are in flight while the current tile is reduced.
// Primary slot (the tile is sorted by its variable): each thread first sums
// runs of one variable over its own items, then an inclusive segmented scan
// over the 64 lanes (DPP moves, fixed order) carries every run across lanes; a
// run is cut at the end of each wavefront's 64*ITEMS positions, and its last
// item writes y directly (variables observed only there) or a positional
// partial.
// Secondary slot, dense accumulator with occurrence rounds (when no tile needs
// more than 8): round r of a tile adds the contributions of the positions
// holding the r-th occurrence of their variable in the tile; rounds
// are separated by barriers, so every accumulator entry has one writer per
// round and a fixed summation order.  Otherwise contributions and variable ids
// are staged in LDS in the tile's order by that variable and each segment head
// sums its run serially into the dense LDS accumulator or a positional partial.
// Positions past the group's count (the last tile may be short) are masked.
// XL: a dense secondary type whose x slice also fits in LDS (2,000 cameras in
// float32 do) has that slice copied into LDS once per workgroup, so the
// per-observation secondary gathers are LDS reads rather than scattered L2
// requests.
// FAST (dense secondary with occurrence rounds or the overflow encoding,
// primary outputs that fit in LDS): the workgroup's round (or overflow) counts
// are read into LDS once, and every piece end is staged in its wavefront's LDS
// region (value and target) and written to y or the partials after the last
// tile.  The tile loop then issues no
// global store, so the compiler's load waits keep the next tile's block loads
// in flight (a pending store makes them wait for every load).  out_cap, the
// most piece ends of one wavefront over the workgroup's tiles, comes from the
// plan; a wavefront that would exceed it writes that tile's ends directly.
template <int DP, typename Real>
__host__ __device__ constexpr int64_t out_region_bytes(int cap) {
  return ((int64_t)cap * (DP * (int64_t)sizeof(Real) + 4) + 15) / 16 * 16;
}

template <typename Real, int R, int D0, int D1, int P, int ITEMS, int NT, bool STREAM, bool XL, bool FAST, bool OVF>
__global__ void __launch_bounds__(NT) normal_wave_kernel(WaveArgs<Real> a) {
  constexpr int DP = P == 0 ? D0 : D1;
  constexpr int DS = P == 0 ? D1 : D0;
  constexpr int OP = P == 0 ? 0 : R * D0;
  constexpr int OS = P == 0 ? R * D0 : 0;
  constexpr int T = NT * ITEMS;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // LDS: [stage: T contributions + T ids (staged reduction only)] [dense accumulator] [XL: x slice]
  //      FAST: [dense accumulator] [XL: x slice] [round counts] [per-wavefront output regions]
  Real* cstage = reinterpret_cast<Real*>(smem_raw);
  int32_t* cid = reinterpret_cast<int32_t*>(cstage + T * DS);
  Real* acc = (FAST || a.rounds) ? reinterpret_cast<Real*>(smem_raw) : reinterpret_cast<Real*>(cid + T);
  Real* xl = acc + (int64_t)a.count_s * DS;
  const bool dense = FAST || a.mode_s == TLH_REDUCE_DENSE;
  if (dense) {
    for (int i = threadIdx.x; i < a.count_s * DS; i += NT) acc[i] = Real(0);
    if constexpr (!FAST) __syncthreads();  // the first occurrence round may add to any entry
  }
  const int lane = threadIdx.x & 63;
  const int b = ITEMS * (int)threadIdx.x;  // first tile-local position of this thread
  const int t0 = blockIdx.x * a.tiles_per_wg;
  const int t1 = min(t0 + a.tiles_per_wg, a.ntiles);
  int32_t* nrounds = reinterpret_cast<int32_t*>(
      smem_raw + (((int64_t)a.count_s * DS * (XL ? 2 : 1) * (int64_t)sizeof(Real) + 15) / 16) * 16);
  unsigned char* out_base =
      reinterpret_cast<unsigned char*>(nrounds) + ((int64_t)a.tiles_per_wg * 4 + 15) / 16 * 16;
  Real* out_val = reinterpret_cast<Real*>(out_base + (threadIdx.x >> 6) * out_region_bytes<DP, Real>(a.out_cap));
  int32_t* out_tgt = reinterpret_cast<int32_t*>(out_val + (int64_t)a.out_cap * DP);
  // FAST overflow stage (overflow encoding only): contributions of the second
  // and later occurrences of a variable in the tile, in (variable, position)
  // order, and their variables
  Real* ov_val = reinterpret_cast<Real*>(out_base + (NT / 64) * out_region_bytes<DP, Real>(a.out_cap));
  int32_t* ov_var = reinterpret_cast<int32_t*>(ov_val + (int64_t)a.stage_cap * DS);
  int out_n = 0;  // piece ends staged by this wavefront (uniform over its lanes)
  if constexpr (FAST) {
    for (int i = threadIdx.x; i < t1 - t0; i += NT) nrounds[i] = a.rounds[t0 + i];
  }

  // One tile: the next tile's block loads into the other register buffer, then
  // this tile's contributions from `cur`, the next tile's gathers and the ids of
  // the tile after it, then the primary scan and the secondary LDS reduction.
  // Two tiles of block loads are thus in flight while a tile's data is awaited.
  auto process = [&](WaveTile<Real, R, DP, DS, ITEMS, XL>& cur, WaveTile<Real, R, DP, DS, ITEMS, XL>& nxt,
                     int t) {
    // The loads are unconditional so that the compiler's wait counts stay exact
    // (a skipped load on one path makes every later wait conservative): after
    // the workgroup's last tile they read the group's first tile, which every
    // workgroup reads there (so it stays in the L2 caches), and discard it.
    const int tn = t + 1 < t1 ? t + 1 : 0;
    load_wave_ids<Real, ITEMS, NT, FAST>(a, tn, b, nxt);
    load_wave_blocks<Real, R, DP, DS, OP, OS, ITEMS, NT, STREAM, XL>(a, tn, a.stride, b, nxt);
    pin_blocks(cur);
    const int64_t start = (int64_t)t * T;
    // t == t1 is the padding half of the last pair (an odd tile count): it
    // computes on stand-in data, writes nothing and takes no rounds
    const int len = t < t1 ? (int)min((int64_t)T, a.count - start) : 0;
    const int left = len - b;  // items of this thread inside the tile (may be <= 0)
    int ip[ITEMS], is[ITEMS];
    uint8_t direct[ITEMS];
    Real xp[ITEMS][DP], xs[ITEMS][DS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      ip[k] = cur.ip[k];
      is[k] = cur.is[k];
      direct[k] = cur.direct[k];
#pragma unroll
      for (int c = 0; c < DP; ++c) xp[k][c] = cur.xp[k][c];
#pragma unroll
      for (int c = 0; c < DS; ++c) xs[k][c] = XL ? xl[is[k] * DS + c] : cur.xs[k][XL ? 0 : c];
    }
    // contributions J_p^T q and J_s^T q with q = J_p x_p + J_s x_s
    Real cp[ITEMS][DP], cs[ITEMS][DS];
    int key[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      Real q[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        Real v = Real(0);
#pragma unroll
        for (int c = 0; c < DP; ++c) v += cur.jp[r][c][k] * xp[k][c];
#pragma unroll
        for (int c = 0; c < DS; ++c) v += cur.js[r][c][k] * xs[k][c];
        q[r] = v;
      }
      const bool valid = k < left;
      key[k] = valid ? ip[k] : -1;
#pragma unroll
      for (int c = 0; c < DP; ++c) {
        Real v = Real(0);
#pragma unroll
        for (int r = 0; r < R; ++r) v += cur.jp[r][c][k] * q[r];
        cp[k][c] = valid ? v : Real(0);
      }
#pragma unroll
      for (int c = 0; c < DS; ++c) {
        Real v = Real(0);
#pragma unroll
        for (int r = 0; r < R; ++r) v += cur.js[r][c][k] * q[r];
        cs[k][c] = v;
      }
    }
    // The next tile's gathers wait for its ids, which were issued after this
    // tile's blocks (vmcnt waits are in issue order): they follow every use of
    // this tile's blocks, and the ids of tile t+2 follow them.
    pin_before_loads(cp);
    pin_before_loads(cs);
    load_wave_gathers<Real, R, DP, DS, ITEMS, XL>(a, nxt);
    // primary slot: runs inside the thread, then a segmented scan over lanes
#pragma unroll
    for (int k = 1; k < ITEMS; ++k)
      if (key[k] == key[k - 1]) {
#pragma unroll
        for (int c = 0; c < DP; ++c) cp[k][c] += cp[k - 1][c];
      }
    bool same = true;
#pragma unroll
    for (int k = 1; k < ITEMS; ++k) same = same && key[k] == key[0];
    const int prev_last = dpp_int<DPP_WAVE_SHR1, 0xf>(key[ITEMS - 1]);
    const int next_first = dpp_int<DPP_WAVE_SHL1, 0xf>(key[0]);
    const bool joins = lane > 0 && key[0] == prev_last;  // first run continues the previous lane's
    int flag = (joins && same) ? 0 : 1;
    Real v[DP];
#pragma unroll
    for (int c = 0; c < DP; ++c) v[c] = cp[ITEMS - 1][c];
    wave_segmented_scan<Real, DP>(flag, v);
    // carry into this lane's first run: the previous lane's scanned trailing run
    Real carry[DP];
#pragma unroll
    for (int c = 0; c < DP; ++c) {
      const Real up = dpp_real<DPP_WAVE_SHR1, 0xf>(v[c]);
      carry[c] = joins ? up : Real(0);
    }
    bool first_run = true;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      if (k > 0 && key[k] != key[k - 1]) first_run = false;
      const bool tail = k < left && (k + 1 < ITEMS ? key[k + 1] != key[k] : (lane == 63 || next_first != key[k]));
      const bool dk = a.direct_p != nullptr && direct[k] != 0;  // read only at piece ends
      if constexpr (FAST) {
        const uint64_t ends = __ballot(tail);
        const int count = __popcll(ends);
        if (out_n + count <= a.out_cap) {
          if (tail) {
            const int slot = out_n + __popcll(ends & ((1ull << lane) - 1ull));
#pragma unroll
            for (int c = 0; c < DP; ++c) {
              const Real sc = first_run ? cp[k][c] + carry[c] : cp[k][c];
              out_val[slot * DP + c] = dk ? sc + a.damping * xp[k][c] : sc;
            }
            out_tgt[slot] = dk ? -(ip[k] + 1) : (int32_t)(a.pos_base_p + start + b + k);
          }
          out_n += count;
        } else {
          // beyond the planned capacity (a plan that does not match the group):
          // direct writes, then a wait for every memory operation so that no
          // store stays pending in the tile loop
          if (tail) {
            Real* out = dk ? a.y_p + (int64_t)ip[k] * DP : a.part_p + (a.pos_base_p + start + b + k) * DP;
#pragma unroll
            for (int c = 0; c < DP; ++c) {
              const Real sc = first_run ? cp[k][c] + carry[c] : cp[k][c];
              out[c] = dk ? sc + a.damping * xp[k][c] : sc;
            }
          }
          __builtin_amdgcn_s_waitcnt(0);
        }
      } else if (tail) {
        Real s[DP];
#pragma unroll
        for (int c = 0; c < DP; ++c) s[c] = first_run ? cp[k][c] + carry[c] : cp[k][c];
        const int64_t vv = ip[k];
        if (dk) {
#pragma unroll
          for (int c = 0; c < DP; ++c) a.y_p[vv * DP + c] = s[c] + a.damping * xp[k][c];
        } else {
          Real* out = a.part_p + (a.pos_base_p + start + b + k) * DP;
#pragma unroll
          for (int c = 0; c < DP; ++c) out[c] = s[c];
        }
      }
    }
    if constexpr (FAST && OVF) {
      // secondary slot, overflow encoding: the first occurrence of each variable
      // in the tile adds into the accumulator (no two share a variable); every
      // later occurrence is staged at its slot (rank - 1) of the tile's
      // (variable, position) order.  After a barrier the head of each staged run
      // sums the run in order and adds the sum.  Each entry thus receives, per
      // tile, its first contribution and then the sum of the others: a fixed
      // order, two barriers (used where occurrence rounds would be many).
#pragma unroll
      for (int k = 0; k < ITEMS; ++k) {
        if (k < left) {
          const int rk = cur.rank(k);
          if (rk == 0) {
#pragma unroll
            for (int c = 0; c < DS; ++c) acc[is[k] * DS + c] += cs[k][c];
          } else if (rk <= a.stage_cap) {
#pragma unroll
            for (int c = 0; c < DS; ++c) ov_val[(rk - 1) * DS + c] = cs[k][c];
            ov_var[rk - 1] = is[k];
          }
        }
      }
      __syncthreads();
      const int nov = t < t1 ? min(nrounds[t - t0], a.stage_cap) : 0;
      for (int p = threadIdx.x; p < nov; p += NT) {
        const int v = ov_var[p];
        if (p > 0 && ov_var[p - 1] == v) continue;
        Real sum[DS];
#pragma unroll
        for (int c = 0; c < DS; ++c) sum[c] = ov_val[p * DS + c];
        for (int q = p + 1; q < nov && ov_var[q] == v; ++q) {
#pragma unroll
          for (int c = 0; c < DS; ++c) sum[c] += ov_val[q * DS + c];
        }
#pragma unroll
        for (int c = 0; c < DS; ++c) acc[v * DS + c] += sum[c];
      }
      __syncthreads();
      return;
    }
    if (FAST || a.rounds) {
      // secondary slot, occurrence rounds: round r adds every r-th occurrence of a
      // variable in this tile, so each accumulator entry has one writer per round
      // (a plain read-modify-write; LDS float atomics measured twice as slow) and
      // receives the tile's contributions in position order
      const int nr = t < t1 ? (FAST ? nrounds[t - t0] : cur.rounds) : 0;
      for (int r = 0; r < nr; ++r) {
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
          if (k < left && cur.rank(k) == r) {
#pragma unroll
            for (int c = 0; c < DS; ++c) acc[is[k] * DS + c] += cs[k][c];
          }
        }
        __syncthreads();
      }
      return;
    }
    // secondary slot: LDS stage in variable order, on-the-fly segment heads
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      if (k < left) {
        const int rk = cur.rank(k);
#pragma unroll
        for (int c = 0; c < DS; ++c) cstage[rk * DS + c] = cs[k][c];
        cid[rk] = is[k];
      }
    }
    __syncthreads();
    for (int p = threadIdx.x; p < len; p += NT) {
      const int vsec = cid[p];
      if (p > 0 && cid[p - 1] == vsec) continue;
      Real sum[DS];
#pragma unroll
      for (int c = 0; c < DS; ++c) sum[c] = cstage[p * DS + c];
      for (int q = p + 1; q < len && cid[q] == vsec; ++q) {
#pragma unroll
        for (int c = 0; c < DS; ++c) sum[c] += cstage[q * DS + c];
      }
      if (dense) {
#pragma unroll
        for (int c = 0; c < DS; ++c) acc[vsec * DS + c] += sum[c];
      } else {
        Real* out = a.part_s + (a.pos_base_s + start + p) * DS;
#pragma unroll
        for (int c = 0; c < DS; ++c) out[c] = sum[c];
      }
    }
    __syncthreads();
  };
  // Two register buffers used alternately (no copies between them): the block
  // loads of tile t+1 are issued before tile t's data is awaited.
  WaveTile<Real, R, DP, DS, ITEMS, XL> ta, tb;
  if (t0 < t1) {
    load_wave_ids<Real, ITEMS, NT, FAST>(a, t0, b, ta);
    load_wave_blocks<Real, R, DP, DS, OP, OS, ITEMS, NT, STREAM, XL>(a, t0, a.stride, b, ta);
    load_wave_gathers<Real, R, DP, DS, ITEMS, XL>(a, ta);
  }
  if constexpr (XL) {
    for (int i = threadIdx.x; i < a.count_s * DS; i += NT) xl[i] = a.x_s[i];
  }
  if constexpr (XL || FAST) __syncthreads();
  // pairs of tiles with no branch between the halves (a conditional half would
  // move loop-carried registers and force waits for loads in flight)
  for (int t = t0; t < t1; t += 2) {
    process(ta, tb, t);
    process(tb, ta, t + 1);
  }
  if constexpr (FAST) {
    // the staged piece ends of this wavefront (the rounds' barriers have made
    // every lane's LDS writes visible)
    for (int e = lane; e < out_n; e += 64) {
      const int32_t tg = out_tgt[e];
      Real* out = tg < 0 ? a.y_p + (int64_t)(-(tg + 1)) * DP : a.part_p + (int64_t)tg * DP;
#pragma unroll
      for (int c = 0; c < DP; ++c) out[c] = out_val[e * DP + c];
    }
  }
  if (dense) {
    const int64_t nn = (int64_t)a.count_s * DS;
    Real* out = a.part_s + (int64_t)(a.dense_block_s + blockIdx.x) * nn;
    for (int64_t i = threadIdx.x; i < nn; i += NT) out[i] = acc[i];
  }
}

// Generic kernel: any rows (<= 16) and slot dims (<= 16); re-reads J per slot.
constexpr int GEN_RMAX = 16;
constexpr int GEN_DSUM = 48;

template <typename Real>
__global__ void __launch_bounds__(TLH_THREADS)
    normal_tile_generic(GroupArgs<Real> g, int dmax, int cap) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Real* stage = reinterpret_cast<Real*>(smem_raw);
  Real* acc[3];
  // The stage holds one slot of a full tile (cap entries of dmax values).
  dense_regions(g, stage + (int64_t)cap * dmax, acc);
  zero_dense(g, acc);
  const int R = g.rows;
  int off[TLH_MAX_SLOTS + 1];
  off[0] = 0;
  for (int s = 0; s < TLH_MAX_SLOTS; ++s) off[s + 1] = off[s] + (s < g.nslots ? g.slot[s].dim : 0);

  const int t0 = blockIdx.x * g.tiles_per_wg;
  const int t1 = min(t0 + g.tiles_per_wg, g.ntiles);
  for (int t = t0; t < t1; ++t) {
    const int start = g.tile_start[t];
    const int len = g.tile_start[t + 1] - start;
    for (int s = 0; s < g.nslots; ++s) {
      const int d = g.slot[s].dim;
      __syncthreads();
      for (int pos = threadIdx.x; pos < len; pos += TLH_THREADS) {
        const int64_t n = start + pos;
        Real q[GEN_RMAX];
        for (int r = 0; r < R; ++r) q[r] = Real(0);
        for (int u = 0; u < g.nslots; ++u) {
          const int du = g.slot[u].dim;
          const Real* xs = g.slot[u].x + (int64_t)g.slot[u].index[n] * du;
          for (int r = 0; r < R; ++r)
            for (int c = 0; c < du; ++c)
              q[r] += g.packed[(int64_t)(R * off[u] + r * du + c) * g.stride + n] * xs[c];
        }
        const int rk = g.slot[s].rank ? (int)g.slot[s].rank[n] : pos;
        for (int c = 0; c < d; ++c) {
          Real a = Real(0);
          for (int r = 0; r < R; ++r) a += g.packed[(int64_t)(R * off[s] + r * d + c) * g.stride + n] * q[r];
          stage[rk * d + c] = a;
        }
      }
      __syncthreads();
      reduce_segments_rt(g.slot[s], d, t, len, stage, acc[s]);
    }
  }
  __syncthreads();
  flush_dense(g, acc);
}

// Finishes: y[v] = sum of the variable's partials (dense blocks, then its CSR
// list) + damping * x[v].  Every type of one finish call is a job of a single
// launch (the jobs write disjoint slices of y), each job taking one of three
// shapes below; a workgroup finds its job from the jobs' block counts.
constexpr int FIN_COLS = 32;
constexpr int FIN_COLS_NARROW = 8;
enum FinishRole { FIN_LIST = 0, FIN_WAVE = 1, FIN_DENSE = 2, FIN_DENSE_NARROW = 3, FIN_DENSE_VEC = 4 };

template <typename Real>
struct FinishJob {
  Real* y;
  const Real* x;
  const Real* partials;
  const int32_t* csr_ptr;
  const int32_t* csr_idx;
  const int32_t* list;
  int64_t count;
  int64_t nlist;
  int32_t dim;
  int32_t dense_blocks;
  int32_t role;
  int32_t blocks;  // workgroups of this job
};

template <typename Real>
struct FinishJobs {
  FinishJob<Real> job[TLH_MAX_TYPES];
  int32_t njobs;
  Real damping;
};

// One thread per (listed variable, component).
template <typename Real>
__device__ __forceinline__ void finish_list(const FinishJob<Real>& j, Real damping, int64_t bid) {
  const int64_t e = bid * TLH_THREADS + threadIdx.x;
  if (e >= j.nlist * j.dim) return;
  const int64_t le = e / j.dim;
  const int c = (int)(e - le * j.dim);
  const int64_t v = j.list ? (int64_t)j.list[le] : le;
  const int64_t i = v * j.dim + c;
  Real sum = Real(0);
  for (int b = 0; b < j.dense_blocks; ++b) sum += j.partials[((int64_t)b * j.count + v) * j.dim + c];
  const int k1 = j.csr_ptr[v + 1];
  for (int k = j.csr_ptr[v]; k < k1; ++k) sum += j.partials[(int64_t)j.csr_idx[k] * j.dim + c];
  j.y[i] = sum + damping * j.x[i];
}

// Types with many dense blocks: a workgroup owns COLS consecutive columns; its
// 256 / COLS column lanes each sum every (256 / COLS)-th dense block (fixed
// stride), then the lane sums are added in lane order.  Few columns (a small
// camera set) take COLS = 8 so the grid still spreads over the chip.
template <typename Real, int COLS>
__device__ __forceinline__ void finish_dense(const FinishJob<Real>& j, Real damping, int64_t bid, Real* lane_sum) {
  constexpr int LANES = TLH_THREADS / COLS;
  const int col = threadIdx.x % COLS;
  const int lane = threadIdx.x / COLS;
  const int64_t total = j.count * j.dim;
  const int64_t i = bid * COLS + col;
  Real sum = Real(0);
  if (i < total) {
#pragma unroll 8
    for (int b = lane; b < j.dense_blocks; b += LANES) sum += j.partials[(int64_t)b * total + i];
  }
  lane_sum[lane * COLS + col] = sum;
  __syncthreads();
  if (lane == 0 && i < total) {
    Real acc = lane_sum[col];
#pragma unroll
    for (int l = 1; l < LANES; ++l) acc += lane_sum[l * COLS + col];
    const int64_t v = i / j.dim;
    const int c = (int)(i - v * j.dim);
    const int k1 = j.csr_ptr[v + 1];
    for (int k = j.csr_ptr[v]; k < k1; ++k) acc += j.partials[(int64_t)j.csr_idx[k] * j.dim + c];
    j.y[i] = acc + damping * j.x[i];
  }
}

// Dense blocks of a wide type, vectorised: a workgroup owns FIN_VCOLS
// consecutive columns (FIN_QUADS vectors of 16 bytes); its FIN_RANGES block
// ranges each sum a contiguous run of dense blocks with every load of the run in
// flight (each wavefront reads FIN_RANGES / 4 rows of 256 contiguous bytes per
// instruction), then the range sums are added in range order.  Requires the
// type's column count to be a multiple of the vector width.
constexpr int FIN_QUADS = 16;
constexpr int FIN_RANGES = TLH_THREADS / FIN_QUADS;

template <typename Real>
constexpr int fin_vec() {
  return 16 / (int)sizeof(Real);
}

template <typename Real>
__device__ __forceinline__ void finish_dense_vec(const FinishJob<Real>& j, Real damping, int64_t bid, Real* lane_sum) {
  constexpr int V = fin_vec<Real>();
  constexpr int COLS = FIN_QUADS * V;
  typedef typename VecT<Real, V>::type Vec;
  const int quad = threadIdx.x % FIN_QUADS;
  const int range = threadIdx.x / FIN_QUADS;
  const int64_t total = j.count * j.dim;
  const int64_t i0 = bid * COLS + (int64_t)quad * V;  // first of this thread's V columns
  const int per = (j.dense_blocks + FIN_RANGES - 1) / FIN_RANGES;
  const int b0 = range * per, b1 = min(j.dense_blocks, b0 + per);
  Vec sum = (Vec)Real(0);
  if (i0 < total) {
    const Real* col = j.partials + i0;
#pragma unroll 8
    for (int b = b0; b < b1; ++b) sum += *reinterpret_cast<const Vec*>(col + (int64_t)b * total);
  }
#pragma unroll
  for (int c = 0; c < V; ++c) lane_sum[range * COLS + quad * V + c] = sum[c];
  __syncthreads();
  if (threadIdx.x < COLS) {
    const int64_t i = bid * COLS + threadIdx.x;
    if (i < total) {
      Real acc = lane_sum[threadIdx.x];
#pragma unroll
      for (int r = 1; r < FIN_RANGES; ++r) acc += lane_sum[r * COLS + threadIdx.x];
      const int64_t v = i / j.dim;
      const int c = (int)(i - v * j.dim);
      const int k1 = j.csr_ptr[v + 1];
      for (int k = j.csr_ptr[v]; k < k1; ++k) acc += j.partials[(int64_t)j.csr_idx[k] * j.dim + c];
      j.y[i] = acc + damping * j.x[i];
    }
  }
}

// One wavefront per (variable, component) for types with long partial lists.
template <typename Real>
__device__ __forceinline__ void finish_wave(const FinishJob<Real>& j, Real damping, int64_t bid) {
  const int lane = threadIdx.x & 63;
  const int64_t e = bid * (TLH_THREADS / 64) + (threadIdx.x >> 6);
  if (e >= j.nlist * j.dim) return;
  const int64_t le = e / j.dim;
  const int c = (int)(e - le * j.dim);
  const int64_t v = j.list ? (int64_t)j.list[le] : le;
  const Real sum = csr_wave_sum<Real>(j.partials, j.csr_idx, j.csr_ptr[v], j.csr_ptr[v + 1], j.dim, c, lane);
  if (lane == 0) {
    Real d = Real(0);
    for (int b = 0; b < j.dense_blocks; ++b) d += j.partials[((int64_t)b * j.count + v) * j.dim + c];
    j.y[v * j.dim + c] = (d + sum) + damping * j.x[v * j.dim + c];
  }
}

template <typename Real>
__global__ void __launch_bounds__(TLH_THREADS) finish_jobs_kernel(FinishJobs<Real> jobs) {
  __shared__ Real lane_sum[TLH_THREADS * fin_vec<Real>()];  // FIN_DENSE_VEC: [FIN_RANGES][FIN_QUADS * vector]
  int64_t bid = blockIdx.x;
  int k = 0;
  while (k + 1 < jobs.njobs && bid >= jobs.job[k].blocks) {
    bid -= jobs.job[k].blocks;
    ++k;
  }
  const FinishJob<Real>& j = jobs.job[k];
  switch (j.role) {  // uniform over the workgroup
    case FIN_DENSE: finish_dense<Real, FIN_COLS>(j, jobs.damping, bid, lane_sum); break;
    case FIN_DENSE_NARROW: finish_dense<Real, FIN_COLS_NARROW>(j, jobs.damping, bid, lane_sum); break;
    case FIN_DENSE_VEC: finish_dense_vec<Real>(j, jobs.damping, bid, lane_sum); break;
    case FIN_WAVE: finish_wave<Real>(j, jobs.damping, bid); break;
    default: finish_list<Real>(j, jobs.damping, bid); break;
  }
}

// Optional column scale of the packed blocks (J diag(s), sparse.py:86-102): entry
// (slot, row, col) of sorted position n is multiplied by scale[off + index[n] * dim
// + col], off = the slot type's first column, in float64 before the conversion.
struct PackScale {
  const double* scale;             // [op->size] or nullptr
  const int32_t* index[TLH_MAX_SLOTS];
  int64_t offset[TLH_MAX_SLOTS];
};

template <typename In, typename Out>
__global__ void __launch_bounds__(TLH_THREADS)
    pack_kernel(Out* __restrict__ packed, int64_t stride, const int64_t* __restrict__ order,
                int64_t count, int rows, int nslots, const In* b0, const In* b1, const In* b2,
                int d0, int d1, int d2, PackScale sc) {
  const int64_t n = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x;
  if (n >= count) return;
  const int64_t o = order[n];
  const In* src[3] = {b0, b1, b2};
  const int dims[3] = {d0, d1, d2};
  int comp = 0;
  for (int s = 0; s < nslots; ++s) {
    const int d = dims[s];
    const int k = rows * d;
    const In* p = src[s] + o * k;
    if (sc.scale) {
      const double* cs = sc.scale + sc.offset[s] + (int64_t)sc.index[s][n] * d;
      for (int e = 0; e < k; ++e) packed[(int64_t)(comp + e) * stride + n] = (Out)((double)p[e] * cs[e % d]);
    } else {
      for (int e = 0; e < k; ++e) packed[(int64_t)(comp + e) * stride + n] = (Out)p[e];
    }
    comp += k;
  }
}

// Two-slot pack with each observation's block rows read as 8/16-byte vectors
// in one burst (observation order is usually a random permutation, so a
// row-at-a-time gather touches every cache line once instead of once per
// element).  K0, K1 = rows * dims of the two slots (even); D0, D1 = the dims.
template <typename In, typename Out, int K0, int K1, int D0, int D1, bool SCALED>
__global__ void __launch_bounds__(TLH_THREADS)
    pack_pair_kernel(Out* __restrict__ packed, int64_t stride, const int64_t* __restrict__ order, int64_t count,
                     const In* __restrict__ b0, const In* __restrict__ b1, PackScale sc) {
  static_assert(K0 % 2 == 0 && K1 % 2 == 0, "rows of even length");
  typedef typename VecT<In, 2>::type V2;
  const int64_t n = (int64_t)blockIdx.x * TLH_THREADS + threadIdx.x;
  if (n >= count) return;
  const int64_t o = order[n];
  In r0[K0], r1[K1];
  const V2* p0 = reinterpret_cast<const V2*>(b0 + o * K0);
  const V2* p1 = reinterpret_cast<const V2*>(b1 + o * K1);
#pragma unroll
  for (int k = 0; k < K0 / 2; ++k) {
    const V2 v = p0[k];
    r0[2 * k] = v[0];
    r0[2 * k + 1] = v[1];
  }
#pragma unroll
  for (int k = 0; k < K1 / 2; ++k) {
    const V2 v = p1[k];
    r1[2 * k] = v[0];
    r1[2 * k + 1] = v[1];
  }
  if constexpr (SCALED) {
    double s0[D0], s1[D1];
    const double* c0 = sc.scale + sc.offset[0] + (int64_t)sc.index[0][n] * D0;
    const double* c1 = sc.scale + sc.offset[1] + (int64_t)sc.index[1][n] * D1;
#pragma unroll
    for (int c = 0; c < D0; ++c) s0[c] = c0[c];
#pragma unroll
    for (int c = 0; c < D1; ++c) s1[c] = c1[c];
#pragma unroll
    for (int e = 0; e < K0; ++e) packed[(int64_t)e * stride + n] = (Out)((double)r0[e] * s0[e % D0]);
#pragma unroll
    for (int e = 0; e < K1; ++e) packed[(int64_t)(K0 + e) * stride + n] = (Out)((double)r1[e] * s1[e % D1]);
  } else {
#pragma unroll
    for (int e = 0; e < K0; ++e) packed[(int64_t)e * stride + n] = (Out)r0[e];
#pragma unroll
    for (int e = 0; e < K1; ++e) packed[(int64_t)(K0 + e) * stride + n] = (Out)r1[e];
  }
}

// ------------------------------------------------------------------- host side
template <typename Real>
GroupArgs<Real> make_args(const tlh_operator_t* op, const tlh_group_t& g, const void* x) {
  GroupArgs<Real> a{};
  a.packed = static_cast<const Real*>(g.packed);
  a.stride = g.stride;
  a.tile_start = g.tile_start;
  a.ntiles = g.ntiles;
  a.tiles_per_wg = g.tiles_per_wg;
  a.rows = g.rows;
  a.nslots = g.nslots;
  for (int s = 0; s < g.nslots; ++s) {
    const tlh_slot_t& sl = g.slot[s];
    const tlh_type_t& ty = op->types[sl.type];
    SlotArgs<Real>& o = a.slot[s];
    o.x = static_cast<const Real*>(x) + ty.offset;
    o.partials = static_cast<Real*>(ty.partials);
    o.index = sl.index;
    o.rank = sl.rank;
    o.seg_ptr = sl.seg_ptr;
    o.seg_start = sl.seg_start;
    o.seg_target = sl.seg_target;
    o.dim = sl.dim;
    o.mode = sl.mode;
    o.dense_block = sl.dense_block;
    o.type_count = (int32_t)ty.count;
  }
  return a;
}

template <typename Real>
int64_t dense_bytes(const tlh_operator_t* op, const tlh_group_t& g) {
  int64_t b = 0;
  for (int s = 0; s < g.nslots; ++s)
    if (g.slot[s].mode == TLH_REDUCE_DENSE)
      b += (int64_t)op->types[g.slot[s].type].count * g.slot[s].dim * (int64_t)sizeof(Real);
  return b;
}

template <typename Real>
constexpr int items_for() {
  return sizeof(Real) == 4 ? 4 : 2;
}

template <typename Real, int R, int D0, int D1, int D2, int ITEMS>
int launch_special_items(const tlh_operator_t* op, const tlh_group_t& g, const void* x,
                         hipStream_t stream) {
  if (g.tile_cap > TLH_THREADS * ITEMS) return TLH_EINVAL;
  const int64_t lds = (int64_t)TLH_THREADS * ITEMS * Dims<D0, D1, D2>::max * sizeof(Real) +
                      dense_bytes<Real>(op, g);
  auto kernel = normal_tile_kernel<Real, R, D0, D1, D2, ITEMS>;
  TLH_RETURN_IF_ERROR(allow_lds((const void*)kernel, lds));
  hipLaunchKernelGGL(kernel, dim3(g.nwg), dim3(TLH_THREADS), (size_t)lds, stream,
                     make_args<Real>(op, g, x));
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

// Tiles of at most one observation per thread (plans of groups too small to
// fill the chip with full tiles) take the one-item kernel: the items of a
// thread are processed one after another, so a short tile on the multi-item
// kernel would still pay for every item.
template <typename Real, int R, int D0, int D1, int D2>
int launch_special(const tlh_operator_t* op, const tlh_group_t& g, const void* x,
                   hipStream_t stream) {
  if (g.tile_cap <= TLH_THREADS) return launch_special_items<Real, R, D0, D1, D2, 1>(op, g, x, stream);
  return launch_special_items<Real, R, D0, D1, D2, items_for<Real>()>(op, g, x, stream);
}

template <typename Real>
int64_t generic_lds(const tlh_operator_t* op, const tlh_group_t& g) {
  int dmax = 0;
  for (int s = 0; s < g.nslots; ++s) dmax = std::max(dmax, (int)g.slot[s].dim);
  return (int64_t)g.tile_cap * dmax * sizeof(Real) + dense_bytes<Real>(op, g);
}

template <typename Real>
int launch_generic(const tlh_operator_t* op, const tlh_group_t& g, const void* x,
                   hipStream_t stream) {
  int dmax = 0, dsum = 0;
  for (int s = 0; s < g.nslots; ++s) {
    dmax = std::max(dmax, (int)g.slot[s].dim);
    dsum += g.slot[s].dim;
  }
  if (g.rows > GEN_RMAX || dsum > GEN_DSUM) return TLH_ENOTFOUND;
  const int64_t lds = generic_lds<Real>(op, g);
  auto kernel = normal_tile_generic<Real>;
  TLH_RETURN_IF_ERROR(allow_lds((const void*)kernel, lds));
  hipLaunchKernelGGL(kernel, dim3(g.nwg), dim3(TLH_THREADS), (size_t)lds, stream,
                     make_args<Real>(op, g, x), dmax, (int)g.tile_cap);
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

// LDS of a wave group.  Staged reduction: stage (contribution + id per
// position), the dense secondary accumulator and, when all of that fits, the
// secondary x slice.  Occurrence rounds: the accumulator and the x slice when
// it fits.  FAST (out_entries > 0): accumulator, x slice when it fits, the
// workgroup's round (or overflow) counts, one output region per wavefront and,
// with the overflow encoding, the overflow stage; a FAST plan that does not
// fit is an error (the overflow encoding admits no other kernel).
struct WaveLds {
  int64_t bytes;
  bool xl;
  bool fast;
};

inline int64_t align16(int64_t b) { return (b + 15) / 16 * 16; }

template <typename Real>
WaveLds wave_lds(const tlh_group_t& g, const tlh_type_t& ts, int64_t lds_max) {
  const tlh_slot_t& ss = g.slot[1 - g.primary];
  const int ds = ss.dim;
  const int dp = g.slot[g.primary].dim;
  const bool dense = ss.mode == TLH_REDUCE_DENSE;
  const bool rounds = dense && ss.rounds;
  const int64_t slice = dense ? ts.count * ds * (int64_t)sizeof(Real) : 0;
  if (g.out_entries > 0) {
    // positional partial ids of staged outputs are int32
    const bool ids_fit = g.slot[g.primary].pos_base + (int64_t)g.ntiles * g.tile_cap <= INT32_MAX;
    if (!rounds || !ss.rank || !ids_fit || g.stage_entries < 0) return {-1, false, true};
    const int64_t counts_b = align16((int64_t)g.tiles_per_wg * 4);
    const int64_t out_b = (TLH_WAVE_THREADS / 64) * align16((int64_t)g.out_entries * (dp * (int64_t)sizeof(Real) + 4));
    const int64_t stage_b = align16((int64_t)g.stage_entries * (ds * (int64_t)sizeof(Real) + 4));
    for (int x = 1; x >= 0; --x) {
      const int64_t b = align16(slice * (x ? 2 : 1)) + counts_b + out_b + stage_b;
      if (b <= lds_max) return {b, x == 1, true};
    }
    return {-1, false, true};
  }
  int64_t b = rounds ? 0 : (int64_t)g.tile_cap * (ds * (int64_t)sizeof(Real) + 4);
  bool xl = false;
  if (dense) {
    b += slice;
    xl = b + slice <= lds_max;
    if (xl) b += slice;
  }
  return {b, xl, false};
}

// mode 0: general (direct stores), 1: FAST with occurrence rounds, 2: FAST with
// the overflow encoding (separate kernels: the other path's code in the tile
// loop measurably slows the rounds kernel)
template <typename Real, int R, int D0, int D1, int P, int ITEMS, int NT, bool STREAM, bool XL>
auto wave_kernel_for(int mode) {
  return mode == 2   ? normal_wave_kernel<Real, R, D0, D1, P, ITEMS, NT, STREAM, XL, true, true>
         : mode == 1 ? normal_wave_kernel<Real, R, D0, D1, P, ITEMS, NT, STREAM, XL, true, false>
                     : normal_wave_kernel<Real, R, D0, D1, P, ITEMS, NT, STREAM, XL, false, false>;
}

template <typename Real, int R, int D0, int D1, int P>
int launch_wave_t(const tlh_operator_t* op, const tlh_group_t& g, const void* x, void* y, double damping,
                  hipStream_t stream) {
  constexpr int NT = TLH_WAVE_THREADS;
  constexpr int ITEMS = TLH_WAVE_TILE / NT / (int)(sizeof(Real) / 4);
  const tlh_slot_t& sp = g.slot[P];
  const tlh_slot_t& ss = g.slot[1 - P];
  const tlh_type_t& tp = op->types[sp.type];
  const tlh_type_t& ts = op->types[ss.type];
  WaveArgs<Real> a{};
  // aligned tiles of exactly NT * ITEMS sorted positions; whole-tile loads need the
  // packed stride (and the caller's id/rank arrays) to cover ntiles * tile_cap positions
  if (g.tile_cap != NT * ITEMS || g.ntiles != ceil_div(g.count, NT * ITEMS) || g.stride % ITEMS != 0 ||
      g.stride < (int64_t)g.ntiles * g.tile_cap || g.primary != P)
    return TLH_EINVAL;
  // the x gathers address each slot's slice with 32-bit byte offsets
  if (tp.count * tp.dim * (int64_t)sizeof(Real) > UINT32_MAX || ts.count * ts.dim * (int64_t)sizeof(Real) > UINT32_MAX)
    return TLH_EINVAL;
  a.packed = static_cast<const Real*>(g.packed);
  a.stride = g.stride;
  a.count = g.count;
  a.ntiles = g.ntiles;
  a.tiles_per_wg = g.tiles_per_wg;
  a.idx_p = sp.index;
  a.idx_s = ss.index;
  a.rank_s = ss.rank;
  a.x_p = static_cast<const Real*>(x) + tp.offset;
  a.x_s = static_cast<const Real*>(x) + ts.offset;
  a.y_p = static_cast<Real*>(y) + tp.offset;
  a.direct_p = tp.direct;
  a.part_p = static_cast<Real*>(tp.partials);
  a.pos_base_p = sp.pos_base;
  a.part_s = static_cast<Real*>(ts.partials);
  a.pos_base_s = ss.pos_base;
  a.mode_s = ss.mode;
  a.dense_block_s = ss.dense_block;
  a.count_s = (int32_t)ts.count;
  a.damping = (Real)damping;
  a.rounds = ss.mode == TLH_REDUCE_DENSE ? ss.rounds : nullptr;
  const WaveLds lds = wave_lds<Real>(g, ts, device_lds_max());
  if (lds.bytes < 0) return TLH_EINVAL;
  a.out_cap = lds.fast ? g.out_entries : 0;
  a.stage_cap = lds.fast ? g.stage_entries : 0;
  // blocks that cannot stay in the 256 MB last-level cache between matvecs are streamed
  const bool stream_blocks = (int64_t)g.stride * g.rows * (D0 + D1) * (int64_t)sizeof(Real) > (192ll << 20);
  const int mode = !lds.fast ? 0 : (g.stage_entries > 0 ? 2 : 1);
  auto kernel = stream_blocks ? (lds.xl ? wave_kernel_for<Real, R, D0, D1, P, ITEMS, NT, true, true>(mode)
                                        : wave_kernel_for<Real, R, D0, D1, P, ITEMS, NT, true, false>(mode))
                              : (lds.xl ? wave_kernel_for<Real, R, D0, D1, P, ITEMS, NT, false, true>(mode)
                                        : wave_kernel_for<Real, R, D0, D1, P, ITEMS, NT, false, false>(mode));
  TLH_RETURN_IF_ERROR(allow_lds((const void*)kernel, lds.bytes));
  hipLaunchKernelGGL(kernel, dim3(g.nwg), dim3(NT), (size_t)lds.bytes, stream, a);
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

template <typename Real>
int launch_wave(const tlh_operator_t* op, const tlh_group_t& g, const void* x, void* y, double damping,
                hipStream_t stream) {
  if (g.nslots != 2 || !y) return TLH_EINVAL;
  const int R = g.rows, d0 = g.slot[0].dim, d1 = g.slot[1].dim;
  const int P = g.primary;
  if (P < 0 || P > 1 || g.slot[P].rank != nullptr) return TLH_EINVAL;
  if (R == 2 && d0 == 6 && d1 == 3)
    return P == 0 ? launch_wave_t<Real, 2, 6, 3, 0>(op, g, x, y, damping, stream)
                  : launch_wave_t<Real, 2, 6, 3, 1>(op, g, x, y, damping, stream);
  if (R == 2 && d0 == 3 && d1 == 6)
    return P == 0 ? launch_wave_t<Real, 2, 3, 6, 0>(op, g, x, y, damping, stream)
                  : launch_wave_t<Real, 2, 3, 6, 1>(op, g, x, y, damping, stream);
  return TLH_ENOTFOUND;
}

// Shapes with a specialised kernel; everything else takes the generic kernel.
template <typename Real>
int launch_group(const tlh_operator_t* op, const tlh_group_t& g, const void* x, void* y, double damping,
                 hipStream_t stream) {
  if (g.kernel == 1) return launch_wave<Real>(op, g, x, y, damping, stream);
  const int R = g.rows;
  const int d0 = g.slot[0].dim;
  const int d1 = g.nslots > 1 ? g.slot[1].dim : 0;
  const int d2 = g.nslots > 2 ? g.slot[2].dim : 0;
  if (d2 == 0 && g.tile_cap <= TLH_THREADS * items_for<Real>()) {
    if (R == 2 && d0 == 6 && d1 == 3) return launch_special<Real, 2, 6, 3, 0>(op, g, x, stream);
    if (R == 2 && d0 == 3 && d1 == 6) return launch_special<Real, 2, 3, 6, 0>(op, g, x, stream);
    if (R == 6 && d0 == 6 && d1 == 6) return launch_special<Real, 6, 6, 6, 0>(op, g, x, stream);
    if (R == 3 && d0 == 3 && d1 == 0) return launch_special<Real, 3, 3, 0, 0>(op, g, x, stream);
    if (R == 6 && d0 == 6 && d1 == 0) return launch_special<Real, 6, 6, 0, 0>(op, g, x, stream);
  }
  return launch_generic<Real>(op, g, x, stream);
}

template <typename Real>
int finish_impl(const tlh_operator_t* op, int t0, int t1, const void* x, double damping, void* y,
                hipStream_t stream);

template <typename Real>
int normal_matvec_impl(const tlh_operator_t* op, const void* x, void* y, double damping,
                       hipStream_t stream) {
  for (int gi = 0; gi < op->ngroups; ++gi) {
    const tlh_group_t& g = op->groups[gi];
    if (g.count == 0 || g.ntiles == 0) continue;
    const int rc = launch_group<Real>(op, g, x, y, damping, stream);
    if (rc) return rc;
  }
  return finish_impl<Real>(op, 0, op->ntypes, x, damping, y, stream);
}

template <typename Real>
int finish_impl(const tlh_operator_t* op, int t0, int t1, const void* x, double damping, void* y,
                hipStream_t stream) {
  FinishJobs<Real> jobs{};
  jobs.damping = (Real)damping;
  int64_t blocks = 0;
  for (int ti = t0; ti < t1; ++ti) {
    const tlh_type_t& t = op->types[ti];
    const int64_t total = t.count * t.dim;
    if (total == 0) continue;
    const int64_t nlist = t.finish_list ? t.finish_count : t.count;
    if (nlist == 0) continue;
    FinishJob<Real>& j = jobs.job[jobs.njobs++];
    j.y = static_cast<Real*>(y) + t.offset;
    j.x = static_cast<const Real*>(x) + t.offset;
    j.partials = static_cast<const Real*>(t.partials);
    j.csr_ptr = t.csr_ptr;
    j.csr_idx = t.csr_idx;
    j.list = t.finish_list;
    j.count = t.count;
    j.nlist = nlist;
    j.dim = t.dim;
    j.dense_blocks = t.dense_blocks;
    int64_t nb;
    if (t.dense_blocks >= 2 * (TLH_THREADS / FIN_COLS) && !t.finish_list) {
      const bool narrow = ceil_div(total, FIN_COLS) < 256;
      constexpr int vcols = FIN_QUADS * fin_vec<Real>();
      // vector loads need 16-byte aligned rows (total * sizeof(Real) % 16 == 0);
      // enough column groups to spread over the chip
      const bool vec = total % fin_vec<Real>() == 0 && ceil_div(total, vcols) >= 128;
      j.role = vec ? FIN_DENSE_VEC : (narrow ? FIN_DENSE_NARROW : FIN_DENSE);
      nb = ceil_div(total, vec ? vcols : (narrow ? FIN_COLS_NARROW : FIN_COLS));
    } else if (csr_heavy(t.csr_entries, nlist)) {
      j.role = FIN_WAVE;
      nb = ceil_div(nlist * t.dim, TLH_THREADS / 64);
    } else {
      j.role = FIN_LIST;
      nb = ceil_div(nlist * t.dim, TLH_THREADS);
    }
    if (nb > INT32_MAX) return TLH_EINVAL;
    j.blocks = (int32_t)nb;
    blocks += nb;
  }
  if (jobs.njobs == 0) return 0;
  if (blocks > INT32_MAX) return TLH_EINVAL;
  hipLaunchKernelGGL(finish_jobs_kernel<Real>, dim3((unsigned)blocks), dim3(TLH_THREADS), 0, stream, jobs);
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

int check_op(const tlh_operator_t* op);

template <typename In, typename Out>
int pack_impl(const tlh_operator_t* op, const tlh_group_t& g, const void* const* blocks, const double* scale,
              hipStream_t stream) {
  const In* b[3] = {nullptr, nullptr, nullptr};
  int d[3] = {0, 0, 0};
  PackScale sc{};
  sc.scale = scale;
  for (int s = 0; s < g.nslots; ++s) {
    b[s] = static_cast<const In*>(blocks[s]);
    d[s] = g.slot[s].dim;
    sc.index[s] = g.slot[s].index;
    sc.offset[s] = op->types[g.slot[s].type].offset;
  }
  const dim3 grid((unsigned)ceil_div(g.count, TLH_THREADS)), block(TLH_THREADS);
  Out* out = static_cast<Out*>(g.packed);
  const int k0 = g.rows * d[0], k1 = g.rows * d[1];
  const uintptr_t vec = 2 * sizeof(In) - 1;  // the pair kernel reads 2-element vectors
  const bool aligned = g.nslots == 2 && ((uintptr_t)b[0] & vec) == 0 && ((uintptr_t)b[1] & vec) == 0;
  const bool ba = g.rows == 2 && ((d[0] == 6 && d[1] == 3) || (d[0] == 3 && d[1] == 6));
  if (aligned && ba && k0 == 12) {
    if (scale)
      hipLaunchKernelGGL((pack_pair_kernel<In, Out, 12, 6, 6, 3, true>), grid, block, 0, stream, out, g.stride,
                         g.order, g.count, b[0], b[1], sc);
    else
      hipLaunchKernelGGL((pack_pair_kernel<In, Out, 12, 6, 6, 3, false>), grid, block, 0, stream, out, g.stride,
                         g.order, g.count, b[0], b[1], sc);
  } else if (aligned && ba && k0 == 6) {
    if (scale)
      hipLaunchKernelGGL((pack_pair_kernel<In, Out, 6, 12, 3, 6, true>), grid, block, 0, stream, out, g.stride,
                         g.order, g.count, b[0], b[1], sc);
    else
      hipLaunchKernelGGL((pack_pair_kernel<In, Out, 6, 12, 3, 6, false>), grid, block, 0, stream, out, g.stride,
                         g.order, g.count, b[0], b[1], sc);
  } else {
    hipLaunchKernelGGL((pack_kernel<In, Out>), grid, block, 0, stream, out, g.stride, g.order, g.count, g.rows,
                       g.nslots, b[0], b[1], b[2], d[0], d[1], d[2], sc);
  }
  TLH_RETURN_IF_LAUNCH_FAILED();
  return 0;
}

int pack_dispatch(const tlh_operator_t* op, int group, const void* const* blocks, int in_dtype, const double* scale,
                  tlh_stream_t stream) {
  if (int rc = check_op(op)) return rc;
  if (group < 0 || group >= op->ngroups || !blocks) return TLH_EINVAL;
  const tlh_group_t& g = op->groups[group];
  if (g.count == 0) return 0;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (in_dtype == TLH_F32 && op->dtype == TLH_F32) return pack_impl<float, float>(op, g, blocks, scale, s);
  if (in_dtype == TLH_F64 && op->dtype == TLH_F32) return pack_impl<double, float>(op, g, blocks, scale, s);
  if (in_dtype == TLH_F64 && op->dtype == TLH_F64) return pack_impl<double, double>(op, g, blocks, scale, s);
  if (in_dtype == TLH_F32 && op->dtype == TLH_F64) return pack_impl<float, double>(op, g, blocks, scale, s);
  return TLH_EINVAL;
}

int check_op(const tlh_operator_t* op) {
  if (!op || (op->dtype != TLH_F32 && op->dtype != TLH_F64)) return TLH_EINVAL;
  if (op->ngroups < 0 || op->ntypes < 1 || !op->types) return TLH_EINVAL;
  for (int gi = 0; gi < op->ngroups; ++gi) {
    const tlh_group_t& g = op->groups[gi];
    if (g.nslots < 1 || g.nslots > TLH_MAX_SLOTS) return TLH_EINVAL;
    for (int s = 0; s < g.nslots; ++s)
      if (g.slot[s].type < 0 || g.slot[s].type >= op->ntypes) return TLH_EINVAL;
  }
  return 0;
}

}  // namespace

extern "C" int tlh_device_info(int device, tlh_device_info_t* out) {
  if (!out) return TLH_EINVAL;
  hipDeviceProp_t prop;
  TLH_RETURN_IF_ERROR(hipGetDeviceProperties(&prop, device));
  out->cu_count = prop.multiProcessorCount;
  int lds = 0;
  TLH_RETURN_IF_ERROR(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
  out->lds_per_block = lds;
  out->warp_size = prop.warpSize;
  int arch = 0;
  for (const char* c = prop.gcnArchName; *c && *c != ':'; ++c)
    if (*c >= '0' && *c <= '9') arch = arch * 10 + (*c - '0');
  out->gcn_arch = arch;
  return 0;
}

extern "C" int64_t tlh_group_lds_bytes(const tlh_operator_t* op, int group) {
  if (check_op(op) || group < 0 || group >= op->ngroups) return -1;
  const tlh_group_t& g = op->groups[group];
  if (g.kernel == 1) {
    const tlh_type_t& ts = op->types[g.slot[1 - g.primary].type];
    return op->dtype == TLH_F32 ? wave_lds<float>(g, ts, device_lds_max()).bytes
                                : wave_lds<double>(g, ts, device_lds_max()).bytes;
  }
  int dmax = 0;
  for (int s = 0; s < g.nslots; ++s) dmax = std::max(dmax, (int)g.slot[s].dim);
  if (op->dtype == TLH_F32) {
    const int64_t special = (int64_t)TLH_THREADS * items_for<float>() * dmax * 4 + dense_bytes<float>(op, g);
    return std::max(special, generic_lds<float>(op, g));
  }
  const int64_t special = (int64_t)TLH_THREADS * items_for<double>() * dmax * 8 + dense_bytes<double>(op, g);
  return std::max(special, generic_lds<double>(op, g));
}

extern "C" int tlh_pack_blocks(const tlh_operator_t* op, int group, const void* const* blocks,
                               int in_dtype, tlh_stream_t stream) {
  return pack_dispatch(op, group, blocks, in_dtype, nullptr, stream);
}

extern "C" int tlh_pack_blocks_scaled(const tlh_operator_t* op, int group, const void* const* blocks,
                                      int in_dtype, const double* col_scale, tlh_stream_t stream) {
  if (!col_scale) return TLH_EINVAL;
  return pack_dispatch(op, group, blocks, in_dtype, col_scale, stream);
}

extern "C" int tlh_normal_matvec(const tlh_operator_t* op, const void* x, void* y, double damping,
                                 tlh_stream_t stream) {
  if (int rc = check_op(op)) return rc;
  if (!x || !y) return TLH_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (op->dtype == TLH_F32) return normal_matvec_impl<float>(op, x, y, damping, s);
  return normal_matvec_impl<double>(op, x, y, damping, s);
}

extern "C" int tlh_finish_types(const tlh_operator_t* op, int type_begin, int type_end, const void* base,
                                double base_scale, void* y, tlh_stream_t stream) {
  if (int rc = check_op(op)) return rc;
  if (type_begin < 0 || type_end > op->ntypes || type_begin > type_end || !base || !y) return TLH_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (op->dtype == TLH_F32) return finish_impl<float>(op, type_begin, type_end, base, base_scale, y, s);
  return finish_impl<double>(op, type_begin, type_end, base, base_scale, y, s);
}

extern "C" int tlh_group_tiles(const tlh_operator_t* op, int group, const void* x, tlh_stream_t stream) {
  if (int rc = check_op(op)) return rc;
  if (group < 0 || group >= op->ngroups || !x) return TLH_EINVAL;
  const tlh_group_t& g = op->groups[group];
  if (g.count == 0 || g.ntiles == 0) return 0;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (g.kernel != 0) return TLH_EINVAL;
  if (op->dtype == TLH_F32) return launch_group<float>(op, g, x, nullptr, 0.0, s);
  return launch_group<double>(op, g, x, nullptr, 0.0, s);
}
