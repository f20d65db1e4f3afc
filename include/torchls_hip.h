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
/* Wave-kernel groups (tlh_group_t.kernel == 1): workgroups of TLH_WAVE_THREADS
 * threads over aligned tiles of TLH_WAVE_TILE (float32) or TLH_WAVE_TILE / 2
 * (float64) sorted positions; tile t starts at t * tile_cap.  The packed stride
 * and every slot's index (and rank) array must cover ntiles * tile_cap positions,
 * the padding indices naming valid variables. */
#define TLH_WAVE_THREADS 512
#define TLH_WAVE_TILE 2048

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
  int64_t pos_base;          /* wave kernel: positional partial ids start here */
  const int32_t* rounds;     /* wave kernel, DENSE secondary slot: [ntiles] or NULL.  When set,
                                rank holds each position's occurrence index (earlier positions
                                of its tile with the same variable) and tile t adds its
                                contributions to the accumulator in rounds[t] rounds, round k
                                adding the k-th occurrences; NULL: staged segment sums */
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
  int32_t kernel;            /* 0: segment-array tile kernel; 1: wave kernel (2 slots): the
                                primary slot is reduced by a wavefront segmented scan in
                                64-position pieces (piece end n -> partial pos_base + n, or a
                                direct write of y when the type's `direct` flag is set), the
                                other slot by on-the-fly segment heads (DENSE accumulator or
                                partial pos_base + tile_start + head) */
  int32_t primary;           /* slot the observations are sorted by */
  int32_t out_entries;       /* wave kernel with a DENSE secondary slot that has rank and rounds:
                                0, or the most primary-slot piece ends of one wavefront over the
                                tiles of one workgroup (the y / partial writes are then staged in
                                LDS, and a plan that does not fit is an error) */
  int32_t stage_entries;     /* out_entries > 0: 0 (occurrence rounds), or the overflow encoding:
                                the secondary rank holds 0 for the first occurrence of its
                                variable in the tile, else 1 + its slot in the tile's (variable,
                                position) order of later occurrences, rounds holds the number of
                                later occurrences per tile, and stage_entries their maximum */
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
  const uint8_t* direct;     /* [count] or NULL: 1 = written directly by a wave kernel */
  const int32_t* finish_list;/* [finish_count] variables the finish writes; NULL = all */
  int64_t finish_count;
  int64_t csr_entries;       /* csr_ptr[count]: sparse partials the finish reads */
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

/* tlh_pack_blocks of J diag(col_scale): entry (row, col) of a slot's block is
 * multiplied by col_scale[offset_t + index * dim_t + col] (float64 [op->size], in
 * op's layout, offset_t = op->types[slot type].offset) before the conversion, so
 * every operator reading the packed blocks (normal matvec, diagonal and Schur
 * kernels) works on the column-scaled Jacobian without a scaled copy of J.
 * Replaces: BlockJacobian.scale_columns (sparse.py:86-102) as used by
 * jacobi_scaling / damping_mode="diagonal" (solvers.py:430-472). */
int tlh_pack_blocks_scaled(const tlh_operator_t* op, int group, const void* const* blocks,
                           int in_dtype, const double* col_scale, tlh_stream_t stream);

/* y = J^T (J x) + damping * x, deterministic (fixed-order) reductions.
 * x and y are [size] vectors in op->dtype.
 * Replaces: WarpNormalLinearization.normal_matvec (warp_normal.py:248-294),
 * BlockJacobian.normal_matvec (sparse.py:83-84). */
int tlh_normal_matvec(const tlh_operator_t* op, const void* x, void* y, double damping,
                      tlh_stream_t stream);

/* Variable diagonal blocks in float64: out[t] is [count_t, dim_t, dim_t] (full,
 * symmetric).  op must be an all-SPARSE plan sharing the normal plan's order and
 * tiles (same packed blocks); each type's `partials` is a float64 workspace of
 * [sparse_partials, dim*(dim+1)/2].  lds_limit: bytes of LDS per workgroup.
 * out[t] == NULL skips type t (allowed only for types whose groups are
 * two-slot 2x6 / 2x3 pairs); tlh_major_gram then supplies its blocks.
 * Replaces: WarpDiagonalPlan.diagonal_blocks (warp_diagonal.py:123-179),
 * BlockJacobian.diagonal_blocks (sparse.py:104-121). */
int tlh_diagonal_blocks(const tlh_operator_t* op, double* const* out, int64_t lds_limit,
                        tlh_stream_t stream);

/* ---------------------------------------------------------------- block Jacobi */
#define TLH_MAX_TYPES 8

/* Block-Jacobi preconditioner: per type, float64 inverse blocks of (D + lambda I)
 * stored as the row-major upper triangle [count][dim*(dim+1)/2]. */
typedef struct tlh_block_diag {
  int32_t ntypes;
  int32_t dim[TLH_MAX_TYPES];
  int64_t count[TLH_MAX_TYPES];
  int64_t offset[TLH_MAX_TYPES];
  const double* inv[TLH_MAX_TYPES];
} tlh_block_diag_t;

/* inverse = (blocks + damping I)^-1 through a float64 Cholesky per block
 * (dim <= 16; joint preconditioner groups reach 9-16); *status |= 1 when a block is not positive definite (the analogue
 * of torch.linalg.cholesky's LinAlgError, solvers.py:267-272, 595-596).
 * Replaces: block_jacobi's factorisation (solvers.py:259-285). */
int tlh_block_inverse(const double* blocks, int64_t count, int dim, double damping,
                      double* inverse, int32_t* status, tlh_stream_t stream);

/* y[v] = inverse[v] x[v] for the variables v in list (int32, count entries; NULL =
 * v < count) of one type: dim-vectors at v * dim, float64 packed upper-triangle
 * inverses as written by tlh_block_inverse, x and y in dtype, float64 arithmetic.
 * Replaces: the per-block solves C^-1 b of DampedSchurSystem.backsubstitute
 * (schur.py:292-301) for eliminated variables no coupling residual observes. */
int tlh_block_apply(const double* inverse, int dim, const int32_t* list, int64_t count, const void* x, void* y,
                    int dtype, tlh_stream_t stream);

/* ---------------------------------------------------------------- PCG */
#define TLH_PCG_MAX_BLOCKS 1024
#define TLH_PCG_RZ 0        /* r.z of the current iterate (state[0] of warp_pcg) */
#define TLH_PCG_THR2 1      /* threshold^2 (state[1]) */
#define TLH_PCG_CURVATURE 2 /* p.Ap */
#define TLH_PCG_NEW_RZ 3
#define TLH_PCG_NORM2 4     /* ||r||^2 after the step */
#define TLH_PCG_TRUE2 5     /* ||rhs - A x||^2 */
#define TLH_PCG_DRIFT2 6    /* ||(rhs - A x) - r||^2 */
#define TLH_PCG_AUX 7       /* free slot for tlh_dot */
#define TLH_PCG_ITERATIONS 8 /* control[0] as float64, written by tlh_pcg_true_residual */
#define TLH_PCG_STATUS 9     /* control[1] as float64, written by tlh_pcg_true_residual */
#define TLH_PCG_NSCALARS 10

/* Device-resident PCG state (warp_pcg.py:131-132): control = {iterations,
 * status, stop_at, spare}; status 0 active, 1 recurrence tolerance, 2
 * nonfinite, 3 nonpositive curvature (warp_pcg.py:78-79), 4 the damped
 * preconditioner blocks are not positive definite (tlh_pcg_begin).  partials holds
 * 2 * TLH_PCG_MAX_BLOCKS doubles; counter must start at zero (kernels reset it). */
typedef struct tlh_pcg_state {
  int32_t* control;
  double* scalars;
  double* partials;
  uint32_t* counter;
} tlh_pcg_state_t;

/* SE(3) between factors r_f = Log(Z_f^-1 T_i^-1 T_j) (pose graph; the logarithm of
 * torchls_amd/geometry.py, quaternion branch) with their exact 6x6 blocks for right
 * updates T <- T Exp(delta), tangent [translation; rotation], by forward-mode dual
 * numbers; poses and measurements row-major 4x4 (stride 16) or 3x4 (12).  Writes
 * residual [count][6], jac_i / jac_j [count][6][6] in dtype, out2[0] = sum ||r||^2 and
 * out2[1] = count of nonfinite entries (float64).  partials: 2*TLH_PCG_MAX_BLOCKS doubles,
 * counter: one zeroed uint32.
 * Replaces: the vmapped jacrev linearization (linearization.py:62-128) of this residual. */
int tlh_between_linearize(int dtype, const void* poses, int64_t pose_stride, const int64_t* idx_i,
                          const int64_t* idx_j, const void* meas, int64_t meas_stride, int64_t count,
                          void* residual, void* jac_i, void* jac_j, double* out2, double* partials,
                          uint32_t* counter, tlh_stream_t stream);

/* scalars[slot] = a . b accumulated in float64 (deterministic).
 * Replaces: torch.dot in warp_pcg.py:139,147-148 and solvers.py:198-246. */
int tlh_dot(const void* a, const void* b, int64_t n, int dtype, const tlh_pcg_state_t* st,
            int slot, tlh_stream_t stream);

/* Starts a solve on the device, with no host read: from scalars[AUX] = rhs.rhs,
 * threshold = max(atol, rtol * ||rhs||) (scalars[THR2] = threshold^2) and
 * control = {0, status, stop_at, 0}, status 4 if *pd_status != 0 (pd_status may be
 * NULL), else 2 if ||rhs|| is not finite, else 1 if ||rhs|| <= threshold (the early
 * exits of warp_pcg.py:119-128, which the host reports after its first read), else 0.
 * Replaces: the threshold computation and first host read of warp_pcg.py:117-129. */
int tlh_pcg_begin(const tlh_pcg_state_t* st, double atol, double rtol, int32_t stop_at,
                  const int32_t* pd_status, tlh_stream_t stream);

/* r = src, x = 0 (when x != NULL), z = M^-1 r, p = z, scalars[RZ] = r.z.
 * pc == NULL means no preconditioner.  Replaces warp_pcg.py:130-132, 169-172, 210-212. */
int tlh_pcg_reset(const tlh_block_diag_t* pc, const void* src, void* x, void* r, void* z, void* p,
                  int64_t n, int dtype, const tlh_pcg_state_t* st, tlh_stream_t stream);

/* Guarded step x += alpha p, r -= alpha Ap (alpha = rz / p.Ap in float64), then
 * mode 0: z = M^-1 r with pc; 1: z = r; 2: step only (caller applies M and dots);
 * modes 0/1 also set scalars[NEW_RZ] = r.z and scalars[NORM2] = r.r.
 * Replaces warp_pcg._step (33-50) fused with the preconditioner and two dots. */
int tlh_pcg_step(const tlh_block_diag_t* pc, int mode, void* x, void* r, void* z, const void* p,
                 const void* ap, int64_t n, int dtype, const tlh_pcg_state_t* st,
                 tlh_stream_t stream);

/* Guarded p = z + (new_rz / rz) p, then the status update of warp_pcg._finish.
 * Replaces warp_pcg._direction (53-67) and _finish (70-92). */
int tlh_pcg_direction(void* p, const void* z, int64_t n, int dtype, const tlh_pcg_state_t* st,
                      tlh_stream_t stream);

/* One whole iteration after ap = A p for n <= 16384 (one workgroup): p.Ap into
 * scalars[CURVATURE], the guarded step, mode 0: z = M^-1 r with pc / 1: z = r,
 * scalars[NEW_RZ] and [NORM2], the guarded direction and the status update --
 * the work of tlh_dot(p, ap, CURVATURE) + tlh_pcg_step + tlh_pcg_direction in
 * one launch.  Returns TLH_EINVAL for larger n.
 * Replaces warp_pcg._step/_direction/_finish (33-92) and the dots of
 * warp_pcg.py:139, 147-148 for small (e.g. reduced Schur) systems. */
int tlh_pcg_iteration_small(const tlh_block_diag_t* pc, int mode, void* x, void* r, void* z, void* p,
                            const void* ap, int64_t n, int dtype, const tlh_pcg_state_t* st,
                            tlh_stream_t stream);

/* true_r = rhs - ax; scalars[TRUE2] = ||true_r||^2, scalars[DRIFT2] = ||true_r - r||^2,
 * and scalars[ITERATIONS], scalars[STATUS] = control[0], control[1] (as float64), so
 * the host reads a check's values as one slice [TRUE2, STATUS].
 * Replaces the true-residual check of warp_pcg.py:189-204. */
int tlh_pcg_true_residual(const void* rhs, const void* ax, const void* r, void* true_r, int64_t n,
                          int dtype, const tlh_pcg_state_t* st, tlh_stream_t stream);

/* Operator of a native PCG batch (tlh_pcg_iterations).
 * TLH_PCG_OP_NORMAL: ap = J^T J p + damping p (tlh_normal_matvec of op).
 * TLH_PCG_OP_SCHUR: the implicit Schur operator of a Schur plan: tlh_schur_apply
 * (MATVEC) over W-group wgroup with eliminated slot eslot and inverses cinv,
 * tlh_group_tiles of every other group of op, then tlh_finish_types of the
 * retained types [0, type_end) with base p scaled by damping. */
#define TLH_PCG_OP_NORMAL 0
#define TLH_PCG_OP_SCHUR 1
typedef struct tlh_pcg_operator {
  int32_t kind;
  int32_t wgroup;
  int32_t eslot;
  int32_t type_end;
  const tlh_operator_t* op;
  const double* cinv;
  double damping;
} tlh_pcg_operator_t;

/* count PCG iterations issued back to back on stream, no host read or
 * synchronisation: ap = A p, then the iteration update -- for n <= small_max
 * (and n <= 16384) tlh_pcg_iteration_small, else tlh_dot(p, ap, CURVATURE) +
 * tlh_pcg_step + tlh_pcg_direction; mode 0: block-Jacobi pc applied in the step,
 * 1: no preconditioner.  Iterations after the control word stopped are frozen
 * by the device (as replayed graph iterations are).  ap is an [n] workspace.
 * Replaces: the graph replays of one batch, warp_pcg.py:184-185, with the
 * iteration's launches issued from native code. */
int tlh_pcg_iterations(const tlh_pcg_operator_t* A, const tlh_block_diag_t* pc, int mode, void* x, void* r,
                       void* z, void* p, void* ap, int64_t n, int dtype, const tlh_pcg_state_t* st,
                       int64_t small_max, int32_t count, tlh_stream_t stream);

/* ---------------------------------------------------------------- LM float64 reductions */

/* Partial sums of g = J^T v for one group of an all-SPARSE plan, in float64, read
 * from the reference-layout blocks (blocks[s]: [count, rows, dim_s], block_dtype)
 * and v ([count*rows] in original observation order, v_dtype).  Each type's
 * `partials` must be a float64 [sparse_partials, dim] workspace.  Call once per
 * group, then tlh_rmatvec_finish_f64.  blocks[s] == NULL skips slot s (two-slot
 * 2x6 / 2x3 float64 groups only); tlh_major_rmatvec then supplies its type.
 * Replaces: BlockJacobian.rmatvec (sparse.py:43-55), Linearization.gradient
 * (problem.py:40-42). */
int tlh_rmatvec_f64(const tlh_operator_t* op, int group, const void* const* blocks, int block_dtype,
                    const void* v, int v_dtype, int64_t lds_limit, tlh_stream_t stream);

/* out[offset_t + ...] = sum of each variable's rmatvec partials (float64 [size]). */
int tlh_rmatvec_finish_f64(const tlh_operator_t* op, double* out, tlh_stream_t stream);

/* tlh_rmatvec_finish_f64 for the types [type_begin, type_end) only (the others
 * are left untouched, e.g. types summed by tlh_major_rmatvec). */
int tlh_rmatvec_finish_range_f64(const tlh_operator_t* op, int type_begin, int type_end, double* out,
                                 tlh_stream_t stream);

/* out2[0] = ||J_g x||^2 of one residual group in float64 (slots summed before
 * squaring); index[s] are the group's int64 variable indices (original order),
 * offsets[s] the first tangent column of each slot's type.  partials holds
 * 2*TLH_PCG_MAX_BLOCKS doubles, counter one zeroed uint32.
 * Replaces: BlockJacobian.squared_matvec_norm (sparse.py:57-81). */
int tlh_squared_matvec_norm(int64_t count, int rows, int nslots, const int32_t* dims,
                            const int64_t* offsets, const int64_t* const* index,
                            const void* const* blocks, int block_dtype, const void* x, int x_dtype,
                            double* out2, double* partials, uint32_t* counter, tlh_stream_t stream);

/* y = J_g x for one group, y in original observation order [count*rows] (float64
 * accumulation, y_dtype output).  Replaces: BlockJacobian.matvec (sparse.py:31-41). */
int tlh_block_matvec(int64_t count, int rows, int nslots, const int32_t* dims, const int64_t* offsets,
                     const int64_t* const* index, const void* const* blocks, int block_dtype,
                     const void* x, int x_dtype, void* y, int y_dtype, tlh_stream_t stream);

/* out2 = { sum w*rho(||after_i||^2), sum w*[rho(||before_i||^2) - rho(||after_i||^2)] }
 * over `count` residual blocks of `rows` rows (before may be NULL: zeros).
 * loss 0 = squared norm, 1 = Huber(delta) with the divided-difference reduction;
 * weights (float64, stride weight_stride, may be NULL).
 * Replaces: AnalyzedLeastSquaresProblem.cost/_evaluate_trial (problem.py:301-348),
 * HuberLoss.value/_reduction (losses.py:80-107). */
int tlh_trial_objective(const void* before, int before_dtype, const void* after, int after_dtype,
                        int64_t count, int rows, int loss, double delta, const double* weights,
                        int weight_stride, double* out2, double* partials, uint32_t* counter,
                        tlh_stream_t stream);

/* ---------------------------------------------------------------- Schur complement */
#define TLH_SCHUR_MATVEC 0  /* camera partials of J_c^T (J_c x - J_p C^-1 W^T x) */
#define TLH_SCHUR_RHS 1     /* camera partials of -W C^-1 b_p */
#define TLH_SCHUR_BACKSUB 2 /* outp = C^-1 (b_p - W^T x) */

/* One pass of the implicit Schur operator over the W-group `wgroup` of a Schur
 * plan (2 slots, slot `eslot` eliminated and primary: sorted, every eliminated
 * variable inside one tile).  xc: retained vector (its type offset is taken
 * from the operator); bp/outp: eliminated vectors starting at that type;
 * cinv: float64 upper triangles of (D_e + lambda I)^-1.  MATVEC/RHS leave
 * retained-type partials for tlh_normal_matvec-style finishing.
 * Replaces: DampedSchurSystem.matvec / reduced_rhs / backsubstitute
 * (schur.py:285-328) and the CSR W, W^T products (schur.py:270-283, 290, 296, 328). */
int tlh_schur_apply(const tlh_operator_t* op, int wgroup, int eslot, int mode, const void* xc,
                    const void* bp, void* outp, const double* cinv, tlh_stream_t stream);

/* Retained preconditioner blocks M = D + lambda I - sum_edges W_e C^-1 W_e^T (full
 * float64 [count, dim, dim]) over an all-SPARSE Schur plan (float64 partials
 * [sparse_partials, dim*(dim+1)/2] in the retained type's `partials`).
 * Replaces: DampedSchurSystem.preconditioner (schur.py:330-378). */
int tlh_schur_precond(const tlh_operator_t* op, int wgroup, int eslot, const double* cinv,
                      const double* D, double damping, double* M, int64_t lds_limit,
                      tlh_stream_t stream);

/* Variable-major diagonal blocks of one slot: out[v] (float64 [count, dim, dim],
 * full, symmetric) = sum over k in [ptr[v], ptr[v+1]) of B_o^T B_o with
 * o = order[k] and B_o the o-th [rows, dim] block of the raw row-major blocks
 * ([N, rows, dim], float32 or float64, as the linearization wrote them).
 * order: int32 observation indices sorted by the slot's variable; ptr: int32
 * [count + 1].  One workgroup per variable, fixed summation order.  For types
 * with few variables and many observations each (cameras): no partials.
 * Shapes: (rows, dim) in {(2, 6), (2, 3), (6, 6)}, else TLH_ENOTFOUND.
 * Replaces: BlockJacobian.diagonal_blocks (sparse.py:104-121) for that slot. */
int tlh_major_gram(const void* blocks, int block_dtype, int rows, int dim, const int32_t* order,
                   const int32_t* ptr, int64_t count, double* out, tlh_stream_t stream);

/* Variable-major float64 J^T v of one slot: out[var * dim + c] (float64
 * [count, dim]) = sum over k in [ptr[var], ptr[var+1]) of sum_r B_o[r][c] v[o *
 * rows + r], o = order[k]; blocks as in tlh_major_gram, v in original row order
 * (float32 or float64).  Shapes as tlh_major_gram.
 * Replaces: BlockJacobian.rmatvec (sparse.py:43-55) for that slot, used by
 * Linearization.gradient (problem.py:40-42). */
int tlh_major_rmatvec(const void* blocks, int block_dtype, const void* v, int v_dtype, int rows, int dim,
                      const int32_t* order, const int32_t* ptr, int64_t count, double* out, tlh_stream_t stream);

/* Retained preconditioner blocks in camera order: M[c] = D[c] + lambda I -
 * sum over edges e of camera c of W_e C_p^-1 W_e^T, W_e = sum over the edge's
 * observations of J_c^T J_p (float64).  jc, jp: raw row-major blocks [N, rows,
 * dc] and [N, rows, dp]; order: int32 observations sorted by (camera, point);
 * edge_ptr: int32 [E + 1] runs of one (camera, point) pair in that order;
 * var_edge_ptr: int32 [count + 1] edges of each camera; edge_point: int32 [E]
 * point of each edge; cinv: float64 upper triangles of C_p^-1; D, M: float64
 * [count, dc, dc].  Shapes: rows 2, dc 6, dp 3, else TLH_ENOTFOUND.
 * Replaces: DampedSchurSystem.preconditioner (schur.py:330-378), like
 * tlh_schur_precond, without per-tile camera partials. */
int tlh_schur_precond_major(const void* jc, const void* jp, int block_dtype, int rows, int dc, int dp,
                            const int32_t* order, const int32_t* edge_ptr, const int32_t* var_edge_ptr,
                            const int32_t* edge_point, int64_t count, const double* cinv, const double* D,
                            double damping, double* M, tlh_stream_t stream);

/* Subtract W C^-1 W^T from a dense float64 reduced matrix M (leading dimension
 * ld, holding blockdiag(D + lambda I)): one thread per camera pair key
 * (key_ci <= key_cj) summing its observation pairs pair_i/pair_j
 * [pair_ptr[k], pair_ptr[k+1]) in order (deterministic, no atomics).  Camera ci's
 * rows start at the retained type's offset in op (op->types[..].offset) plus
 * ci * dim, so other retained types may precede it in M.  work
 * (float64 [count][2 * DC * DP], may be NULL) holds W_o and W_o C_p^-1 per
 * observation for the 2x6 / 2x3 bundle-adjustment blocks, computed once instead
 * of once per camera pair.
 * Replaces: DampedSchurSystem.dense_matrix (schur.py:380-397). */
int tlh_schur_dense(const tlh_operator_t* op, int wgroup, int eslot, const double* cinv,
                    const int32_t* pair_ptr, const int32_t* pair_i, const int32_t* pair_j,
                    const int32_t* key_ci, const int32_t* key_cj, int64_t npairs, double* M,
                    int64_t ld, double* work, tlh_stream_t stream);

/* Finish for a subset of types: y[t] = sum of partials + base_scale * base[t] for
 * t in [type_begin, type_end) (retained types of a Schur plan).  x-like vectors
 * are indexed by each type's offset.  Used after tlh_schur_apply and the normal
 * tile kernels of retained-only groups. */
int tlh_finish_types(const tlh_operator_t* op, int type_begin, int type_end, const void* base,
                     double base_scale, void* y, tlh_stream_t stream);

/* Normal-matvec tile pass of a single group (no finish): adds that group's
 * J^T J x contributions to its types' partials. */
int tlh_group_tiles(const tlh_operator_t* op, int group, const void* x, tlh_stream_t stream);

/* out[0..3] = {a.b, b.b, max|b|, max|a|} of two [n] vectors (float32 or float64 each;
 * a or b may be NULL, its terms are then 0), float64 sums in a fixed order, maxima
 * NaN-propagating (a NaN entry makes the maximum NaN, an infinite one infinite), in
 * one launch.  partials: 2 * TLH_PCG_MAX_BLOCKS doubles; counter: one zeroed uint32
 * (left at zero).
 * Replaces: torch.dot(gradient, delta), the step norm, the isfinite(delta).all()
 * step check and max |g| of the LM loop (solvers.py:386-397, 541, 591-593, 612-613). */
int tlh_vector_stats(const void* a, int a_dtype, const void* b, int b_dtype, int64_t n, double* out,
                     double* partials, uint32_t* counter, tlh_stream_t stream);

/* Trust region of tlh_lm_decide (TrustRegionConfig): enabled = 0 without one (every
 * finite trial is accepted, the damping stays), adaptive = 1 for the "adaptive"
 * strategy, 0 for the factor rule. */
typedef struct tlh_lm_trust {
  int32_t enabled;
  int32_t adaptive;
  double lambda_min;
  double lambda_max;
  double lambda_factor;
  double quality_min;
} tlh_lm_trust_t;

/* Where a trial's float64 values are on the device: the trial objective, the actual
 * reduction, ||J delta||^2, the step statistics {g.delta, delta.delta, ...} of
 * tlh_vector_stats, and a value finite exactly when every step entry is. */
typedef struct tlh_lm_trial {
  const double* cost;
  const double* actual;
  const double* sq;
  const double* stats;
  const double* finite;
} tlh_lm_trial_t;

/* Gain ratio, acceptance and damping update of one LM trial, on the device (one
 * thread).  trial == NULL: a trial without a usable step (rejected).  state[0..1] =
 * {damping, rejection factor} (float64, device) is read and updated in place.
 * out[0..8] (float64, device) = {predicted reduction -g.delta - ||J delta||^2 / 2,
 * gain ratio, accepted, usable, ||delta||, the trial's damping, stop (rejected at
 * lambda_max, or rejected without a trust region), damping and rejection factor
 * after the update}; flags are 0 or 1.
 * Replaces: the predicted reduction, gain ratio, acceptance and lambda update of
 * solvers.py:607-617, 633-655, 683-693. */
int tlh_lm_decide(const tlh_lm_trial_t* trial, double* state, const tlh_lm_trust_t* trust, double* out,
                  tlh_stream_t stream);

/* dst[di(i)] = scale * src[si(i)] for i < count, si(i) = src_index ? src_index[i]
 * : src_offset + i (likewise di), converted through float64 (float32/float64 on
 * either side): the gathers and scatters of a Schur step in one launch each.
 * Replaces: the slicing, index gathers/scatters and casts of
 * DampedSchurSystem.reduced_rhs / backsubstitute (schur.py:285-301). */
int tlh_copy_cast(const void* src, int src_dtype, const int64_t* src_index, int64_t src_offset, void* dst,
                  int dst_dtype, const int64_t* dst_index, int64_t dst_offset, int64_t count, double scale,
                  tlh_stream_t stream);

/* ---------------------------------------------------------------- dense Cholesky */
/* In-place lower Cholesky A = L L^T (row-major, leading dimension ld, float64):
 * blocked right-looking, 64-wide panels, MFMA (v_mfma_f64_16x16x4_f64) trailing
 * updates, one step of look-ahead (the next panel's columns are updated first,
 * then that panel runs beside the rest of the trailing update).  Only the lower triangle is read and written meaningfully; the upper
 * triangle of the diagonal blocks is left undefined.  *status |= 1 on a
 * non-positive pivot.  Replaces: torch.linalg.cholesky (solvers.py:517). */
int tlh_cholesky_f64(double* A, int64_t n, int64_t ld, int32_t* status, tlh_stream_t stream);

/* In-place lower Cholesky A = L L^T (row-major, leading dimension ld, float64,
 * n <= 8192) as one persistent cooperative launch over 64x64 tiles: a
 * workgroup factors the diagonal tiles (16x16 blocks, MFMA updates) and solves
 * the subdiagonal ones, the others own the remaining tiles, apply their MFMA
 * updates as soon as the L tiles they need are published (agent-scope
 * release/acquire of per-tile states) and solve with the published inverse
 * 16x16 diagonal blocks.  Only the lower triangle is written.  work:
 * tlh_cholesky_tiles_workspace(n) bytes (tile states and inverse blocks; the
 * launch clears what it needs).  *status |= 1 on a non-positive or non-finite
 * pivot; |= 4 when a bounded wait gave up (the factor is then unusable: factor a
 * fresh copy with tlh_cholesky_f64).  Returns TLH_ELDS when the device cannot
 * keep two workgroups resident.  Replaces: torch.linalg.cholesky (solvers.py:517). */
int64_t tlh_cholesky_tiles_workspace(int64_t n);
int tlh_cholesky_tiles_f64(double* A, int64_t n, int64_t ld, void* work, int64_t work_bytes, int32_t* status,
                           tlh_stream_t stream);

/* Poll limit of the bounded cross-workgroup waits of tlh_cholesky_tiles_f64 and
 * tlh_cholesky_solve_blocks_f64 (and its no-allocation variant): polls >= 1 caps
 * them at that many polls, 0 restores each kernel's own limit.  Process-wide, not
 * synchronised with launches in flight.  No counterpart in the reference: it lets
 * tests force the give-up paths (status bits 4 and 2) and the callers' fallbacks. */
int tlh_set_wait_limit(int64_t polls);

/* b <- (L L^T)^-1 b with the factor from tlh_cholesky_f64 (float64, in place).
 * Replaces: torch.cholesky_solve (solvers.py:518). */
int tlh_cholesky_solve_f64(const double* L, int64_t n, int64_t ld, double* b, tlh_stream_t stream);

/* Same result as tlh_cholesky_solve_f64, spread over one workgroup per 64-row
 * block: forward then backward substitution, each block waiting on the solved
 * values of the blocks it depends on.  work: float64 workspace of 2 * n entries,
 * filled by the call itself (a fill kernel, graph-capturable).  *status |= 2 if
 * a wait gave up (the result is then not defined).
 * Replaces: torch.cholesky_solve (solvers.py:518). */
int tlh_cholesky_solve_blocks_f64(const double* L, int64_t n, int64_t ld, double* b, double* work, int32_t* status,
                                  tlh_stream_t stream);

/* tlh_cholesky_solve_blocks_f64 without the work argument: the 2 * n float64
 * workspace is a library-owned device array of 2 * 8192 entries (128 KiB per
 * device, held until the process ends), resolved once per device on the first
 * call made with that device current.  No device memory is allocated by the call,
 * so it suits callers that cannot allocate (code under hipGraph capture, serving
 * loops, harnesses that forbid hipMalloc); the first call on a device should be
 * made outside capture.  n > 8192 returns TLH_EINVAL.  Every call on a device
 * shares that device's array (the call's fill kernel resets it): calls that
 * overlap on one device (concurrent streams) race.
 * No counterpart in the reference. */
int tlh_cholesky_solve_blocks_f64_noalloc(const double* L, int64_t n, int64_t ld, double* b, int32_t* status,
                                          tlh_stream_t stream);

/* ---------------------------------------------------------------- reprojection residual */
/* One reprojection residual batch r_i = pi(R_c p + t_c) - obs_i (pi(x,y,z) = (x/z, y/z);
 * poses world-to-camera, rows [R | t], cam_stride 12 (3x4) or 16 (4x4) elements per
 * pose; tangent [v; w] applied on the right) linearized in `dtype` (TLH_F32/TLH_F64;
 * cams, pts, obs, raw and model in that dtype; jc and jp in block_dtype, which is dtype
 * or, for float64 parameters, TLH_F32: float32 Jacobian storage with float64 residuals):
 *   raw[i]  = r_i (may be NULL),  model[i] = sqrt(w_i rho'_i) r_i,
 *   jc[i]   = sqrt(w_i rho'_i) [D R, -D R [p]x]  ([count,2,6]),
 *   jp[i]   = sqrt(w_i rho'_i) D R               ([count,2,3]),
 *   out2[0] = sum_i w_i rho(||r_i||^2) in float64 (fixed order), out2[1] = the number
 *   of observations whose residual or blocks are not finite.
 * rho is the block Huber loss for huber_delta > 0, else the squared norm; weights
 * (float64 [count]) may be NULL.  partials: 2*TLH_PCG_MAX_BLOCKS doubles, counter: one
 * zeroed uint32 (left zeroed).
 * Replaces, for this residual: AnalyzedLeastSquaresProblem.linearize's vmapped
 * jacrev/jacfwd and loss rescaling (linearization.py:62-128, losses.py:56-90). */
int tlh_reprojection_linearize(int dtype, int block_dtype, const void* cams, int64_t cam_stride, const void* pts,
                               const void* obs, const int64_t* cam_idx, const int64_t* pt_idx,
                               int64_t count, double huber_delta, const double* weights, void* raw,
                               void* model, void* jc, void* jp, double* out2, double* partials,
                               uint32_t* counter, tlh_stream_t stream);

/* Trial objective of the same batch at new poses/points without materializing the
 * residual: out2 = { sum w rho(||r_new||^2), sum w [rho(||before||^2) - rho(||r_new||^2)] }
 * (divided-difference Huber reduction), before = the linearization's raw residual
 * ([count*2], before_dtype, may be NULL).
 * Replaces: AnalyzedLeastSquaresProblem._evaluate_trial for this residual
 * (problem.py:315-348). */
int tlh_reprojection_trial(int dtype, const void* cams, int64_t cam_stride, const void* pts,
                           const void* obs, const int64_t* cam_idx, const int64_t* pt_idx, int64_t count,
                           double huber_delta, const double* weights, const void* before,
                           int before_dtype, double* out2, double* partials, uint32_t* counter,
                           tlh_stream_t stream);

/* out[i] = poses[i] Exp(delta[i]) for `count` SE(3) poses (row-major, rows 3 or 4 of 4
 * columns, `dtype` storage; delta float64 [count, delta_stride], first 6 entries =
 * [translation, rotation]); arithmetic in `dtype` like the torch retraction.
 * Replaces: the vmapped SE3Var retraction of the LM update (geometry.py:40-57,
 * 130-137). */
int tlh_se3_retract(int dtype, const void* poses, int rows, int64_t count, const double* delta,
                    int64_t delta_stride, void* out, tlh_stream_t stream);

/* ---------------------------------------------------------------- implicit gradient */
/* grad[i] = (rho'_i I + 2 rho''_i r_i r_i^T)(J_c,i w_c + J_p,i w_p) for every reprojection
 * residual r_i = pi(T_c p) - obs_i (float64, Jacobian recomputed from cams [n_c,4,4] and
 * pts [n_p,3]); w is the adjoint tangent vector, cam_offset/pt_offset the first tangent
 * column of each type; huber_delta <= 0 means no loss.  This is dL/dobs for
 * w = (J~^T J~ + lambda I)^-1 dL/dtheta.  No counterpart in the reference
 * (solvers.py:288 solves under torch.no_grad). */
int tlh_reprojection_obs_vjp(const double* cams, const double* pts, const double* obs,
                             const int64_t* cam_idx, const int64_t* pt_idx, const double* w,
                             int64_t cam_offset, int64_t pt_offset, double huber_delta, int64_t count,
                             double* grad, tlh_stream_t stream);

/* Per-observation exact Hessian-vector products of f = 1/2 sum_i w_i rho(||r_i||^2) for the
 * reprojection residual at delta = 0 of the solver's retraction (poses T Exp(delta_c),
 * points p + delta_p): out[i][0..5] (camera tangent) and out[i][6..8] (point tangent) =
 * d/de grad f_i(e v) for the tangent vector v (float64, cam_offset / pt_offset the first
 * column of each type), i.e. J_i^T (rho' I + 2 rho'' r r^T) J_i v plus the second-order
 * residual and retraction terms rho' sum_k r_k d2 r_k[v, .].  Evaluated in hyper-dual
 * arithmetic from cams [n_c,4,4], pts [n_p,3], obs [N,2] (float64); weights may be NULL;
 * huber_delta <= 0 means no loss.  Sum the rows per variable with tlh_segment_sum_f64.
 * No counterpart in the reference (its solve is not differentiated: solvers.py:288);
 * it replaces the torch.func Hessian-vector products of an exact-Hessian adjoint solve. */
int tlh_reprojection_hvp(const double* cams, const double* pts, const double* obs, const int64_t* cam_idx,
                         const int64_t* pt_idx, const double* v, int64_t cam_offset, int64_t pt_offset,
                         double huber_delta, const double* weights, int64_t count, double* out, tlh_stream_t stream);

/* out[var * dim + c] = sum over k in [ptr[var], ptr[var+1]) of values[order[k] * stride + col0 + c]
 * (float64, fixed summation order: one wavefront per variable, lane-strided sums and a
 * fixed xor tree).  order: int32 rows grouped by variable; ptr: int32 [count + 1].
 * Reduces per-observation rows (e.g. of tlh_reprojection_hvp) into per-variable vectors. */
int tlh_segment_sum_f64(const double* values, int64_t stride, int col0, int dim, const int32_t* order,
                        const int32_t* ptr, int64_t count, double* out, tlh_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* TORCHLS_HIP_H */
