/*
 * tnx_dmrg.h -- DMRG contractions of libtnx (C ABI).
 *
 * The two-site effective-Hamiltonian matvec, the environment updates and the
 * Lanczos eigensolver of the reference's symmetric DMRG backend, as device
 * plans.  Conventions as in tnx.h (caller-owned float64 device buffers,
 * caller stream, status codes).  Block tables use the packed layout of tnx.h.
 *
 * Factored evaluation (same sums as the reference's 5-/4-block products):
 *   matvec:  X[(b,c),d] = sum_a L[a,(b,c)] th[a,d]            (grouped GEMM)
 *            Y[c,(d,f)] = sum_b C[b,f] X[(b,c),d],
 *                         C[b,f] = sum_e Wl[b,p,s,e] Wr[e,q,t,f]  (mix)
 *            out[c,g]  += sum_(d,f) Y[c,(d,f)] R[(d,f),g]     (grouped GEMM)
 *   env:     the same three stages with one MPO tensor.
 * Requires physical legs of multiplicity one (distinct physical charges).
 */
#ifndef TNX_DMRG_H
#define TNX_DMRG_H

#include "tnx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Plan of out[c,s,t,g] = L[a,b,c] th[a,p,q,d] Wl[b,p,s,e] Wr[e,q,t,f] R[d,f,g]
 * ("abc,apqd,bpse,eqtf,dfg->cstg", src/tenax/algorithms/dmrg.py:1282-1289,
 * evaluated by _blockwise_contract, dmrg.py:1048-1151: equality matching of
 * shared charges, no conservation filter).  `theta` is the input AND output
 * block table (the full sector layout of the two-site tensor); output blocks
 * that no block combination reaches are written as zeros. */
int tnx_matvec_plan(const tnx_blocks *L, const tnx_blocks *theta, const tnx_blocks *Wl,
                    const tnx_blocks *Wr, const tnx_blocks *R, void *stream, tnx_plan **plan);

/* Reachability of the matvec plan: n_terms (theta block -> output block) edges. */
int tnx_matvec_reach(const tnx_plan *plan, int32_t *n_edges, int32_t *src, int32_t *dst,
                     int32_t capacity);

/* y = H_eff x for the planned structure. */
int tnx_matvec_execute(const tnx_plan *plan, const double *L, const double *Wl, const double *Wr,
                       const double *R, const double *x, double *y, int32_t compute, void *stream);

/* One stage of the fused matvec alone (profiling/roofline): stage 0 = MPO
 * coefficient kernel, 1 = L.theta grouped GEMM, 2 = coefficient mix,
 * 3 = .R grouped GEMM.  Stages read the intermediates the previous stage left
 * in the plan, so stage k is meaningful after stages 0..k-1 ran once.
 * stage | TNX_STAGE_GEMM_ONLY runs the GEMM kernel of stage 1 or 3 without
 * its split-K reduction (its output is then incomplete), and
 * stage | TNX_STAGE_REDUCE_ONLY that reduction alone (nothing when the stage's
 * tiling has no split tiles): the two in sequence equal the whole stage, so
 * each kernel can be timed in place inside a matvec. */
#define TNX_STAGE_GEMM_ONLY 16
#define TNX_STAGE_REDUCE_ONLY 32
int tnx_matvec_execute_stage(const tnx_plan *plan, int32_t stage, const double *L,
                             const double *Wl, const double *Wr, const double *R,
                             const double *x, double *y, int32_t compute, void *stream);

/* Algorithmic flops of each stage of a fused plan (array of 4: coef, GEMM1, mix, GEMM2). */
int tnx_fused_stage_flops(const tnx_plan *plan, double *flops4);

/* Plain per-block GPU path of the same matvec: one GEMM launch per block
 * product of both stages (the reference's per-combination structure on a GPU). */
int tnx_matvec_execute_perblock(const tnx_plan *plan, const double *L, const double *Wl,
                                const double *Wr, const double *R, const double *x, double *y,
                                int32_t compute, void *stream);

/* Environment update plan:
 *   new[d,e,f] = E[a,b,c] A[a,p,d] W[b,p,x,e] A[c,x,f]
 *   ("abc,apd,bpxe,cxf->def", the left update of dmrg.py:1165-1198).
 * The right update of dmrg.py:1201-1232 ("abc,dpa,epxb,fxc->def") is the same
 * sum with the site transposed to (a,p,d) and the MPO tensor to (b,p,x,e); the
 * caller passes those transposes.  `side` must be 0.  The MPS site A is used
 * for ket and bra (real data, so A.conj() == A).  out_legs: the three output
 * legs (candidate-key space); output structure (reachable keys, sorted) via
 * tnx_plan_output_*. */
int tnx_env_plan(int32_t side, const tnx_blocks *E, const tnx_blocks *A, const tnx_blocks *W,
                 const tnx_leg *out_legs, void *stream, tnx_plan **plan);
int tnx_env_execute(const tnx_plan *plan, const double *E, const double *A, const double *W,
                    double *out, int32_t compute, void *stream);
/* Plain per-block GPU path of the same update: one GEMM launch per block
 * product of each stage (the per-combination structure of dmrg.py:1165-1232
 * on a GPU); the comparison baseline of the sweep. */
int tnx_env_execute_perblock(const tnx_plan *plan, const double *E, const double *A, const double *W,
                             double *out, int32_t compute, void *stream);

/* Lanczos ground state of the planned matvec (dmrg.py:978-1045), entirely on
 * the device: FP64 alpha/beta and vector updates, optional one-pass classical
 * Gram-Schmidt re-orthogonalisation against every previous Krylov vector,
 * stop when beta < tol (later steps become no-ops on the device), lowest
 * eigenpair of the tridiagonal matrix by Sturm bisection + inverse iteration
 * on the device, Ritz vector sum_k c_k v_k normalised.  theta is overwritten
 * with the eigenvector; *energy and *n_steps are returned to the host. */
int tnx_lanczos(const tnx_plan *matvec_plan, const double *L, const double *Wl, const double *Wr,
                const double *R, double *theta, int32_t max_iter, double tol, int32_t reorth,
                int32_t compute, double *energy, int32_t *n_steps, void *stream);

/* Allocation-free variant of tnx_lanczos.  No counterpart in the reference;
 * it is for callers that cannot allocate device memory (code under graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): it calls no
 * hipMalloc / hipFree and no allocating library routine.  Same arguments,
 * kernels and results (bitwise) as tnx_lanczos.  The Krylov basis, scalars
 * and solver state live in a static device array of
 * TNX_LANCZOS_NOALLOC_SCRATCH_BYTES that every GPU this entry point runs on
 * holds until the process ends; a solve needs
 *   256 + 8 * ((max_iter + 1) * n + 1033 * (max_iter + 2))  bytes
 * (n = the plan's theta size) and a larger one returns TNX_ERR_CAPACITY
 * before any device work.  All calls on one GPU share that array: calls that
 * overlap on one GPU (two host threads, or two streams) race and must be
 * serialised by the caller.  The plan itself must have been created (and its
 * tiling set) beforehand, since plan creation allocates. */
#define TNX_LANCZOS_NOALLOC_SCRATCH_BYTES (64u << 20)
int tnx_lanczos_noalloc(const tnx_plan *matvec_plan, const double *L, const double *Wl, const double *Wr,
                        const double *R, double *theta, int32_t max_iter, double tol, int32_t reorth,
                        int32_t compute, double *energy, int32_t *n_steps, void *stream);

/* Captured Lanczos solve (hipGraph).  No counterpart in the reference, whose
 * Lanczos is a Python loop under JAX's asynchronous dispatch
 * (src/tenax/algorithms/dmrg.py:978-1045): the whole device solve of
 * tnx_lanczos -- every launch of its max_iter steps, the tridiagonal solve and
 * the Ritz vector -- recorded once into an instantiated graph that owns its
 * Krylov workspace, then replayed with one launch on any stream.  Results are
 * bitwise those of tnx_lanczos with the same arguments.
 *   create   records the solve for these operand and theta pointers;
 *   update   re-records it for new pointers (or after the plan's tiling
 *            changed) and updates the executable graph in place;
 *   rebind   re-records it for another matvec plan (the next two-site
 *            update of a sweep) and its buffers, growing the workspace when
 *            the new problem is larger; the executable graph is updated in
 *            place when the new recording has the same launch topology, and
 *            instantiated anew otherwise;
 *   launch   replays it (TNX_ERR_ARG if tnx_plan_set_tiling changed the plan
 *            since the last create/update);
 *   result   synchronises the stream and returns energy and step count
 *            (theta then holds the eigenvector);
 *   free     releases the graph and its workspace.
 * The operand buffers must outlive the graph, and so must the plan:
 * tnx_plan_free refuses (TNX_ERR_ARG) a plan with graphs not yet freed.  The
 * graph has its own Krylov workspace but shares the plan's intermediates
 * (matvec stages, split-K workspace) with eager tnx_matvec_execute /
 * tnx_lanczos calls on the same plan: replaying a graph while an eager call
 * on the same plan runs on another stream races, so order them on one stream
 * (or synchronise between them). */
typedef struct tnx_graph_s tnx_graph;
int tnx_lanczos_graph_create(const tnx_plan *matvec_plan, const double *L, const double *Wl,
                             const double *Wr, const double *R, double *theta, int32_t max_iter,
                             double tol, int32_t reorth, int32_t compute, void *stream,
                             tnx_graph **graph);
int tnx_lanczos_graph_update(tnx_graph *graph, const double *L, const double *Wl, const double *Wr,
                             const double *R, double *theta, void *stream);
int tnx_lanczos_graph_rebind(tnx_graph *graph, const tnx_plan *matvec_plan, const double *L,
                             const double *Wl, const double *Wr, const double *R, double *theta,
                             void *stream);
int tnx_graph_launch(tnx_graph *graph, void *stream);
int tnx_lanczos_graph_result(tnx_graph *graph, void *stream, double *energy, int32_t *n_steps);
int tnx_graph_free(tnx_graph *graph);

#ifdef __cplusplus
}
#endif
#endif /* TNX_DMRG_H */
