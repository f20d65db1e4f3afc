/*
 * torchls_hip.h -- C ABI of the MI355X (gfx950) block-sparse normal-equation path.
 *
 * Every entry point takes plain pointers and sizes, never torch types.  All
 * buffers are device memory owned by the caller (the Python host side keeps
 * them in PyTorch's caching allocator, as the reference keeps its plan arrays
 * in Torch's pool: warp_normal.py:222-229).  No entry point allocates device
 * memory, synchronizes the stream, or reads device memory on the host, so
 * every compute entry point may be recorded into a hipGraph.  Work is enqueued
 * on the stream passed in (the caller's current stream).
 *
 * Return value: 0 on success, otherwise a hipError_t code (launch failure) or
 * one of the TLH_E* codes below (invalid arguments).
 *
 * Symbolic structure ("plan") and values are separated exactly as in the
 * reference, where "only the symbolic column map is reused between nonlinear
 * linearizations" (warp_normal.py:5-6, warp_diagonal.py:3-4):
 *   - tlh_group_t / tlh_type_t describe the fixed incidence structure built
 *     once per solve (reference: WarpNormalPlan.__init__, warp_normal.py:132-185);
 *   - tlh_pack_blocks copies each linearization's Jacobian blocks into the
 *     plan's tiled component-major layout (reference: WarpNormalLinearization,
 *     warp_normal.py:205-246).
 */
#ifndef TORCHLS_HIP_H
#define TORCHLS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* tlh_stream_t; /* == hipStream_t */

#define TLH_F32 0
#define TLH_F64 1

#define TLH_MAX_SLOTS 3
#define TLH_TILE_THREADS 256

#define TLH_REDUCE_SPARSE 0 /* one partial per (tile, variable) segment */
#define TLH_REDUCE_DENSE 1  /* per-workgroup dense LDS accumulator of the whole type */

#define TLH_EINVAL 1000     /* invalid argument (shape, dtype, unsupported block shape) */
#define TLH_ELDS 1001       /* LDS request exceeds the device limit */
#define TLH_ENOTFOUND 1002  /* no kernel instantiated for this block shape */

/* One slot (variable argument) of a residual group, in the plan's sorted order. */
typedef struct tlh_slot {
  int32_t type;              /* layout entry (variable type) index */
  int32_t dim;               /* tangent dimension of that type */
  int32_t mode;              /* TLH_REDUCE_SPARSE / TLH_REDUCE_DENSE */
  int32_t dense_block;       /* DENSE: first dense partial block of this slot */
  const int32_t* index;      /* [count] variable index per sorted observation */
  const uint16_t* rank;      /* [count] position inside its tile when the tile is
                                ordered by this slot's index; NULL = identity */
  const int32_t* seg_ptr;    /* [ntiles+1] segment range of each tile */
  const int32_t* seg_start;  /* [nseg] first position of the segment (tile-local) */
  const int32_t* seg_target; /* [nseg] SPARSE: partial id; DENSE: variable index */
} tlh_slot_t;

/* One residual group: N observations of r rows, 1..3 slots (JacobianGroup,
 * sparse.py:11-22).  Observations are stored sorted by the primary slot's
 * variable and cut into tiles of at most tile_cap observations. */
typedef struct tlh_group {
  int64_t count;             /* observations */
  int32_t rows;              /* residual rows r */
  int32_t nslots;
  int32_t ntiles;
  int32_t tile_cap;          /* max observations per tile (<= 256 * items) */
  int32_t nwg;               /* workgroups launched (DENSE slots: persistent) */
  int32_t tiles_per_wg;
  int64_t stride;            /* component stride of the packed blocks (>= count) */
  const int32_t* tile_start; /* [ntiles+1] */
  const int64_t* order;      /* [count] original observation of each sorted position */
  void* packed;              /* [rows * sum(dim)][stride], component (slot, row, col) major */
  tlh_slot_t slot[TLH_MAX_SLOTS];
} tlh_group_t;

/* One variable type of the tangent layout (VariableLayout, variables.py:100-117)
 * and the deterministic reduction of its partial sums. */
typedef struct tlh_type {
  int64_t count;             /* variables of this type */
  int32_t dim;               /* tangent dimension */
  int32_t dense_blocks;      /* dense partial blocks (ids b*count + v) */
  int64_t offset;            /* first global tangent column */
  int64_t sparse_partials;   /* sparse partial ids follow the dense region */
  const int32_t* csr_ptr;    /* [count+1] sparse partials of each variable */
  const int32_t* csr_idx;    /* partial ids in fixed (group, tile, slot) order */
  void* partials;            /* workspace [(dense_blocks*count + sparse_partials) * dim] */
} tlh_type_t;

typedef struct tlh_operator {
  int32_t dtype;             /* TLH_F32 / TLH_F64: packed blocks, vectors and partials */
  int32_t ngroups;
  int32_t ntypes;
  int32_t items;             /* observations per thread per tile (tile = 256*items) */
  int64_t size;              /* tangent dimension n */
  const tlh_group_t* groups; /* host array [ngroups] */
  const tlh_type_t* types;   /* host array [ntypes] */
} tlh_operator_t;

typedef struct tlh_device_info {
  int32_t cu_count;
  int32_t lds_per_block;     /* bytes */
  int32_t warp_size;
  int32_t gcn_arch;          /* e.g. 950 */
} tlh_device_info_t;

/* Device properties used by the planner to size tiles and LDS accumulators. */
int tlh_device_info(int device, tlh_device_info_t* out);

/* LDS bytes the normal-matvec kernel of a group needs (planner check). */
int64_t tlh_group_lds_bytes(const tlh_operator_t* op, int group);

/* Copy one group's blocks into its packed tiled layout, converting dtype.
 * blocks[s] points to the reference layout [count, rows, dim_s] (sparse.py:15)
 * in dtype in_dtype; the packed copy is in op->dtype.
 * Replaces: WarpNormalPlan.bind / WarpNormalLinearization (warp_normal.py:187-246). */
int tlh_pack_blocks(const tlh_operator_t* op, int group, const void* const* blocks,
                    int in_dtype, tlh_stream_t stream);

/* y = J^T (J x) + damping * x, deterministic (fixed-order) reductions.
 * x and y are [size] vectors in op->dtype.
 * Replaces: WarpNormalLinearization.normal_matvec (warp_normal.py:248-294),
 * BlockJacobian.normal_matvec (sparse.py:83-84). */
int tlh_normal_matvec(const tlh_operator_t* op, const void* x, void* y, double damping,
                      tlh_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* TORCHLS_HIP_H */
