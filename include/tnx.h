/*
 * tnx.h -- C ABI of libtnx, the MI355X (gfx950) block-sparse contraction engine.
 *
 * Plain pointers and sizes only.  Device buffers are owned by the caller
 * (torch / any HIP allocator); the library owns plan objects and the device
 * work lists / intermediate buffers inside them.  All device work is enqueued
 * on the caller's hipStream_t (passed as void*).  Every entry point returns a
 * status code (TNX_OK = 0, negative on failure); tnx_last_error() returns the
 * message of the last failure on the calling thread.
 *
 * Packed block-sparse storage (one tensor = one contiguous float64 buffer):
 * blocks in lexicographically sorted key order, each block row-major in its
 * own leg order; offsets[i] is the element offset of block i.  This is the
 * reference's canonical pytree order (src/tenax/core/tensor.py:572-578).
 */
#ifndef TNX_H
#define TNX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define TNX_OK 0
#define TNX_ERR_BAD_STRUCTURE (-1)
#define TNX_ERR_CONSERVATION (-2)
#define TNX_ERR_LABEL_COUNT (-3)
#define TNX_ERR_DTYPE (-4)
#define TNX_ERR_HIP (-5)
#define TNX_ERR_OOM (-6)
#define TNX_ERR_ARG (-7)
#define TNX_ERR_CAPACITY (-8)

/* arithmetic used by the block GEMMs; storage is always float64 */
#define TNX_COMPUTE_F64 0       /* FP64 MFMA, FP64 accumulation            */
#define TNX_COMPUTE_F32 1       /* FP32 MFMA, FP32 accumulation            */
#define TNX_COMPUTE_SPLIT_F32 2 /* hi/lo FP32 split, 3 FP32 MFMA products,
                                   FP32 per 16-wide K chunk, FP64 across   */

/* planner modes */
#define TNX_MODE_PAIRWISE_FILTERED 0 /* contractor.py:291-484 semantics        */
#define TNX_MODE_MULTI_REACHABLE 1   /* dmrg.py:1048-1151 semantics (no filter) */

/* One leg as the planner sees it: sorted distinct charges with their
 * multiplicities (number of basis states carrying the charge).  Replaces the
 * per-leg TensorIndex charge arrays (src/tenax/core/index.py:43-77). */
typedef struct {
  int32_t n_charges;
  const int32_t *charges; /* host, ascending */
  const int32_t *mult;    /* host, > 0 */
  int32_t flow;           /* +1 IN, -1 OUT (index.py:27-40) */
} tnx_leg;

/* Block table of one packed tensor (host arrays). */
typedef struct {
  int32_t rank;
  int32_t n_blocks;
  const int32_t *keys;    /* n_blocks*rank, sorted lexicographically */
  const int32_t *shapes;  /* n_blocks*rank */
  const int64_t *offsets; /* n_blocks, element offsets */
} tnx_blocks;

typedef struct tnx_plan_s tnx_plan;

int tnx_version(void);
int tnx_last_error(char *buf, size_t n);
int tnx_device_sync(void);

/* Valid-sector enumeration on the device.
 * Replaces tenax.core.tensor._compute_valid_blocks (src/tenax/core/tensor.py:53-165)
 * and _block_slices shapes (tensor.py:168-188).  sym_n = 0 for U(1), n for Z_n.
 * Keys come back sorted lexicographically (the order of tensor.py:576/651).
 * If the count exceeds `capacity`, returns TNX_ERR_CAPACITY with *n_out set. */
int tnx_valid_blocks(const tnx_leg *legs, int32_t rank, int32_t sym_n,
                     int32_t has_target, int32_t target, int32_t capacity,
                     int32_t *keys_out, int32_t *shapes_out, int32_t *n_out,
                     void *stream);

/* Pairwise block-sparse contraction plan, built on the device.
 * Replaces the matching/filter/output-key stage of
 * tenax.contraction.contractor._contract_symmetric (contractor.py:344-441)
 * for two operands, and the two-operand case of dmrg._blockwise_contract
 * (dmrg.py:1079-1136) in TNX_MODE_MULTI_REACHABLE.
 * Operand layouts: A legs = [free..., contracted...] (a_kfirst=0) or
 * [contracted..., free...] (a_kfirst=1); B legs must be
 * [contracted..., free...] (b_kfirst=1; other values are rejected -- callers
 * transpose B first); the n_contracted contracted legs correspond in order.
 * Output legs = [free A..., free B...] described by out_legs. */
int tnx_plan_pairwise(const tnx_blocks *a, const tnx_blocks *b, int32_t a_kfirst,
                      int32_t b_kfirst, int32_t n_contracted,
                      const tnx_leg *out_legs, int32_t sym_n, int32_t mode,
                      int32_t has_target, int32_t target, void *stream,
                      tnx_plan **plan);

/* Pairwise plan whose output legs are stored in another order: as
 * tnx_plan_pairwise (out_legs still describe [free A..., free B...]), and
 * output leg j of the result is natural leg out_perm[j].  The output blocks
 * (tnx_plan_output_copy) are keyed, sorted and laid out in that order and the
 * grouped GEMM (and its split-K reduction) writes every element straight to
 * its place, where the reference contracts and then transposes
 * (contractor.py:474-484, tensor.py:832-861).  Bosonic charges only; such a
 * plan has no per-block path and no adjoint.  TNX_ERR_ARG when a side (the A
 * or B free legs) needs more than 3 strides after merging the legs that stay
 * adjacent (callers then transpose the natural result instead). */
int tnx_plan_pairwise_permuted(const tnx_blocks *a, const tnx_blocks *b, int32_t a_kfirst,
                               int32_t b_kfirst, int32_t n_contracted, const tnx_leg *out_legs,
                               int32_t sym_n, int32_t mode, int32_t has_target, int32_t target,
                               const int32_t *out_perm, void *stream, tnx_plan **plan);

/* Pairwise plan for graded (fermionic) charges: as tnx_plan_pairwise, and
 * each block pair enters with the exchange sign of the reference
 * (contractor.py:206-285 inversion pairs, applied per combination at
 * 462-470): sign_pairs holds n_sign_pairs pairs (x, y) of leg positions in
 * the concatenated [A legs..., B legs...] of the operands as passed; a pair
 * whose two charges are both odd (parity = charge & 1, FermionParity and
 * FermionicU1) flips the sign of the block product.  The caller derives the
 * pairs from the contraction's subscripts. */
int tnx_plan_pairwise_graded(const tnx_blocks *a, const tnx_blocks *b, int32_t a_kfirst,
                             int32_t b_kfirst, int32_t n_contracted, const tnx_leg *out_legs,
                             int32_t sym_n, int32_t mode, int32_t has_target, int32_t target,
                             int32_t n_sign_pairs, const int32_t *sign_pairs, void *stream,
                             tnx_plan **plan);

/* Output structure of a plan (pairwise / transpose plans). */
int tnx_plan_output_info(const tnx_plan *plan, int32_t *n_blocks, int32_t *rank,
                         int64_t *total_elems);
int tnx_plan_output_copy(const tnx_plan *plan, int32_t *keys, int32_t *shapes,
                         int64_t *offsets);
/* Algorithmic work of one execute: flops (2*M*N*K summed over block GEMMs)
 * and bytes (operands read once + output written once, float64). */
int tnx_plan_stats(const tnx_plan *plan, double *flops, double *bytes,
                   int32_t *n_pairs, int32_t *n_tasks);
int tnx_plan_free(tnx_plan *plan);

/* Tiling of one grouped-GEMM launch of a plan (no reference counterpart: the
 * reference dispatches one einsum per block pair).  stage: 0 for a pairwise
 * plan or a batch of them (tnx_batch_pairwise); 1 (L.theta) or 3 (.R) for a
 * matvec / environment plan.
 *   shape  0 = 64x64, 1 = 128x64, 2 = 64x128, 3 = 128x128 output tiles per workgroup
 *   kchunk K tiles (of 16) per task: 0 = whole K, -2 = balanced two-wave split
 *          (tiles whole up to the mean work of one workgroup slot, longer ones
 *          cut into a head of that length, dispatched with the whole tiles,
 *          and short tail pieces dispatched after them), other values < 0 =
 *          equal chunks chosen from the work per workgroup slot; outputs split
 *          in K are summed by a second launch in a fixed order (deterministic)
 *   xcd    1 = place the tiles of one output (or tile row) on one XCD, bins balanced
 * The result stays deterministic for a given tiling; different tilings differ
 * by rounding.  tnx_plan_get_tiling returns the active tiling and task count. */
int tnx_plan_set_tiling(tnx_plan *plan, int32_t stage, int32_t shape, int32_t kchunk,
                        int32_t xcd, void *stream);
int tnx_plan_get_tiling(const tnx_plan *plan, int32_t stage, int32_t *shape, int32_t *kchunk,
                        int32_t *xcd, int32_t *n_tasks);

/* Grouped execution: every block pair of the plan in ONE launch (output tiles
 * of uneven blocks, K-segments of all contributing pairs accumulated in a
 * fixed order, no atomics).  Replaces the per-combination opt_einsum loop of
 * contractor.py:444-476. */
int tnx_contract_execute(const tnx_plan *plan, const double *a, const double *b,
                         double *c, int32_t compute, void *stream);

/* Adjoint (vector-Jacobian product) of tnx_contract_execute: given g_c laid
 * out like the plan's output, writes g_a = dL/dA laid out like A and
 * g_b = dL/dB laid out like B (blocks no pair touches are zero); either may
 * be null when not wanted.  The adjoint work lists are derived once from the
 * forward plan's pair list and kept in the plan: dB_j = sum A_i^T G_o in one
 * grouped launch (A read in the other layout), dA_i = sum G_o B_j^T in one
 * grouped launch after a block-wise transpose of B (and, for a K-first A, a
 * transpose of the result into place).  Replaces the reverse pass of
 * jax.custom_vjp over the block contraction (the differentiated per-block
 * einsums of contractor.py:444-476; the custom_vjp of ad_utils.py:694-764
 * calls it through the CTM step).  The plan owns the scratch (transposed B). */
int tnx_contract_vjp(const tnx_plan *plan, const double *g_c, const double *a, const double *b,
                     double *g_a, double *g_b, int32_t compute, void *stream);

/* Plain per-block GPU path: one GEMM launch per block pair (the structure of
 * contractor.py:444-476 on a GPU).  Same arithmetic; used as the comparison
 * baseline. */
int tnx_contract_execute_perblock(const tnx_plan *plan, const double *a,
                                  const double *b, double *c, int32_t compute,
                                  void *stream);

/* Leg permutation of a packed tensor (SymmetricTensor.transpose,
 * src/tenax/core/tensor.py:832-861): re-sorted keys on the device, one
 * block-transpose launch. */
int tnx_plan_transpose(const tnx_blocks *a, const tnx_leg *legs, const int32_t *perm,
                       void *stream, tnx_plan **plan);
int tnx_transpose_execute(const tnx_plan *plan, const double *a, double *out,
                          void *stream);
/* Leg permutation of a graded (fermionic) tensor, graded != 0: each block is
 * multiplied by its Koszul sign (tensor.py:32-50, 853-856: -1 per inversion
 * of the permutation between two legs with odd charges, parity = charge & 1). */
int tnx_plan_transpose_graded(const tnx_blocks *a, const tnx_leg *legs, const int32_t *perm,
                              int32_t graded, void *stream, tnx_plan **plan);

/* Many independent instances in one launch (the many-temperature HOTRG step,
 * src/tenax/algorithms/hotrg.py:109-215, run for every instance at once; the
 * reference contracts each instance with its own per-block einsums).
 * tnx_batch_pairwise binds pairwise plans[i] to its operands a[i], b[i] and
 * output c[i] (device buffers laid out as the plan's operands and output) and
 * merges their block pairs into one grouped-GEMM program (default 64 x 64
 * tiles, longest K first over all instances); tnx_batch_transpose does the
 * same for transpose plans (in[i] -> out[i]).  The outputs must not overlap
 * each other; all plans of one pairwise batch must share the A layout
 * (a_kfirst).  tnx_batch_execute runs a batch (compute as in
 * tnx_contract_execute; ignored for transposes).  A batch is a plan object:
 * it keeps the buffer addresses it was built with and is freed with
 * tnx_plan_free; the instance plans may be freed after the batch is built. */
int tnx_batch_pairwise(const tnx_plan *const *plans, int32_t n, const double *const *a,
                       const double *const *b, double *const *c, void *stream, tnx_plan **batch);
int tnx_batch_transpose(const tnx_plan *const *plans, int32_t n, const double *const *in,
                        double *const *out, void *stream, tnx_plan **batch);
int tnx_batch_execute(const tnx_plan *batch, int32_t compute, void *stream);

/* Batched 2-D strided copy: desc[i] copies rows x cols float64 elements
 * from src (row pitch src_ld) to dst (row pitch dst_ld), each multiplied by
 * scale, all copies in one launch.  Assembles the sector matrices of the
 * block-sparse SVD from the stored blocks and unpacks the kept singular
 * vectors into the U / Vh blocks, in place of the reference's per-block
 * slice updates (src/tenax/linalg.py:163-187 and 268-309).  desc is DEVICE
 * memory (n records); copy i is done by workgroups task0 .. task0 +
 * ceil(rows * cols / TNX_COPY2D_CHUNK) - 1, task0 being the sum of those
 * counts over copies 0 .. i-1, and n_wg the sum over all copies.  Ranges may
 * not overlap between destinations; src and dst must not overlap. */
#define TNX_COPY2D_CHUNK 4096
typedef struct {
  const double *src;
  double *dst;
  int64_t src_ld, dst_ld;
  int64_t task0;
  int32_t rows, cols;
  double scale;
} tnx_copy2d;
int tnx_copy2d_execute(const tnx_copy2d *desc, int32_t n, int64_t n_wg, void *stream);

/* Work-list validation (no reference counterpart; a debug and test aid).
 * tnx_plan_validate checks every grouped-GEMM work list of a plan (pairwise,
 * matvec or environment; default and tiled task lists) and the matvec's mix
 * lists against the block layouts of the operands and intermediates: every
 * range a launch can address lies inside one block of its operand, every task
 * is a tile of its output with a non-empty K range, every split-K slot lies in
 * the workspace.  TNX_ERR_BAD_STRUCTURE with the first violation otherwise.
 * With TNX_VALIDATE=1 in the environment every plan is validated when it is
 * built and when its tiling changes. */
int tnx_plan_validate(const tnx_plan *plan, void *stream);
/* The same check on host arrays (no device work): pairs / outs / tasks are
 * n records of the plan's internal layout (32 bytes each):
 *   pair {int64 a_off, b_off; int32 K, lda, ldb, neg}
 *   out  {int64 c_off; int32 M, N, ldc, pair_begin, pair_end, pad}
 *   task {int32 out, m0, n0, kt0, kt1, slot, pad0, pad1}
 * a_layout 0: A blocks M x K (MK), 1: K x M (KM); B is K x N.  *_blocks are
 * n pairs {offset, size} (elements) of each operand's blocks; bm x bn is the
 * task tile and ws_tiles the split-K workspace size in tiles. */
int tnx_check_gemm_worklist(int32_t a_layout, const void *pairs, int32_t n_pairs, const void *outs,
                            int32_t n_outs, const void *tasks, int32_t n_tasks, int32_t bm, int32_t bn,
                            int64_t ws_tiles, const int64_t *a_blocks, int32_t n_a, const int64_t *b_blocks,
                            int32_t n_b, const int64_t *c_blocks, int32_t n_c);

#ifdef __cplusplus
}
#endif
#endif /* TNX_H */
