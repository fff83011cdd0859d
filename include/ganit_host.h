/* ganit_host.h -- C ABI of libganit.so, the host side of the MI355X PDHG engine: problem I/O, scaling,
 * workload generators and the GPU solve.  This is what a ctypes / cgo / N-API binding calls; the C++
 * interface it wraps mirrors the reference's public API (include/ganit/pdhg.hpp:10-47 and
 * include/ganit/lp.hpp:19-50).
 *
 * Errors: functions returning int give 0 on success; functions returning a handle give NULL on failure.
 * ganit_last_error() then holds the calling thread's message.
 */
#ifndef GANIT_HOST_H
#define GANIT_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ganit_lp ganit_lp; /* an LP/QP in original (unscaled) space: the reference's ganit::LP */

const char* ganit_last_error(void);

/* read_mps (reference src/mps.cpp:69-272): free MPS / QPS, .gz through gzip. */
ganit_lp* ganit_lp_read_mps(const char* path);
/* csr_from_triplets (reference src/lp.cpp:9-53) + the LP fields (reference include/ganit/lp.hpp:34-47).
 * Q triplets may be absent (qnnz = 0). */
ganit_lp* ganit_lp_from_triplets(int m, int n, int64_t nnz, const int* rows, const int* cols, const double* vals,
                                 const double* c, const double* l, const double* u, const double* lo,
                                 const double* hi, double obj_const, int maximize, int64_t qnnz, const int* qrows,
                                 const int* qcols, const double* qvals);
/* Seeded workloads (no reference counterpart; the reference ships Python generators instead):
 *   kind 0 random LP        p = [m, n, m_eq, nnz_min, nnz_max, tighten_frac]
 *   kind 1 blending LP      p = [crudes, products, specs, periods]
 *   kind 2 portfolio QP     p = [assets, factors, gamma]
 *   kind 3 differentiable QP p = [n, m_eq, m_le, density, qdensity]
 *   kind 4 batched-LP instance p = [m_lo, m_hi, n_lo, n_hi, nnz_min, nnz_max, tighten_frac]: m, n
 *                           log-uniform from the seed, half the rows equalities, then kind 0 with that seed */
ganit_lp* ganit_lp_generate(int kind, const double* p, uint64_t seed);
void ganit_lp_free(ganit_lp* lp);
/* dims[0..4] = m, n, nnz(A), nnz(Q), maximize */
void ganit_lp_dims(const ganit_lp* lp, int64_t* dims);
/* Copy out the original data; any pointer may be NULL.  Sizes from ganit_lp_dims. */
void ganit_lp_export(const ganit_lp* lp, int* aptr, int* aidx, double* aval, int* qptr, int* qidx, double* qval,
                     double* c, double* l, double* u, double* lo, double* hi, double* obj_const);
/* The planted point of a generated instance (random LP: x0, y0; differentiable QP: x0).  Returns nonzero
 * when the instance has none. */
int ganit_lp_planted(const ganit_lp* lp, double* x0, double* y0);
int ganit_lp_write_mps(const ganit_lp* lp, const char* path);

/* scale_lp + estimate_norm (reference src/scaling.cpp:36-155).
 * scal[0..8] = bscale, cscale, normQ, obj_const, bnorm, cnorm, normA, nnz(AT), maximize; any array pointer
 * may be NULL (CSR of A, AT, Q; then c, l, u, lo, hi, R, C, Rres, Cres). */
int ganit_lp_scaled(const ganit_lp* lp, int ruiz_iters, int pock_chambolle, double* scal, int* aptr, int* aidx,
                    double* aval, int* tptr, int* tidx, double* tval, int* qptr, int* qidx, double* qval, double* c,
                    double* l, double* u, double* lo, double* hi, double* R, double* C, double* Rres, double* Cres);
/* The same with the matrix half on GPU `device` (Ruiz / Pock-Chambolle passes, transpose, the power iteration's
 * SpMVs; no reference counterpart -- the reference's set-up is host code).  Bitwise the same outputs. */
int ganit_lp_scaled_gpu(const ganit_lp* lp, int ruiz_iters, int pock_chambolle, int device, double* scal, int* aptr,
                        int* aidx, double* aval, int* tptr, int* tidx, double* tval, int* qptr, int* qidx,
                        double* qval, double* c, double* l, double* u, double* lo, double* hi, double* R, double* C,
                        double* Rres, double* Cres);

/* solve_pdhg_gpu (reference include/ganit/pdhg.hpp:45-47, src/backend_cuda.cu:382-393).
 * opts[0..14] = tol, max_iter, time_limit, check_every, ruiz_iters, pock_chambolle, verbose, infeas_tol,
 *               precision (0 double, 1 mixed, 2 dynamic), stall_checks, stall_checks_l2,
 *               values (0 FP32, 1 BF16 matrix values in L1/L2, 2 BF16 then FP32: the L1 phase runs on BF16
 *               values until stall_checks checks without progress, then on FP32 values until stall_checks_l2),
 *               device,
 *               eval_fp64 (1: L1/L2 checks evaluated in FP64 with the FP64 matrix),
 *               flags: bit 0 kahan (L1's FP32 checks sum FP32-rounded terms with Kahan compensation),
 *                      bit 1 detect_unbounded (LPs: test the primal ray x - xr as a dual-infeasibility
 *                      certificate; ignored by ganit_batch_*),
 *                      bits 2-3 where the set-up's matrix work runs: 0 automatic (the GPU from a million nonzeros,
 *                      when no row or column has more than 4096 entries), 1 host, 2 GPU -- bitwise the same
 *                      scaled problem and norm estimate either way (ignored by ganit_batch_*),
 *                      bit 4 polish (LPs: project an optimal point onto its active face on the GPU and keep it
 *                      when its FP64 KKT errors are all <= 1e-9; ignored by ganit_batch_*),
 *                      bit 5 scale_q (QPs: the column scaling also equilibrates Q, the block matrix [Q A'; A 0]),
 *                      bit 6 (ganit_batch_* only): no lanes-per-row tail -- a batch whose running instances are
 *                      few and small (at most 8, 2^18 nonzeros) otherwise iterates them 8-32 lanes per row (row
 *                      sums by a lane tree) instead of on row blocks (each row in row order, the CPU's bits),
 *                      bit 7 (ganit_batch_* only): the reduced-precision levels end once the running instances are
 *                      such a tail, which then runs in FP64 (FP32 saves nothing when the launches bound it; a batch
 *                      policy, not the reference's ladder rule)
 * out[0..17]  = status (0 optimal, 1 primal infeasible, 2 iteration limit, 3 time limit, 4 numerical,
 *                       5 dual infeasible -- only with detect_unbounded),
 *               pobj, dobj, rel_p, rel_d, rel_gap, iterations, fp32_iterations, l2_iterations, restarts,
 *               setup_s, solve_s, normA, upload_s, L1_s, L2_s, L3_s, checks
 * x (n) and y (m) receive the original-space solution; either may be NULL. */
int ganit_solve_gpu(const ganit_lp* lp, const double* opts, double* out, double* x, double* y);

/* The same solve split into stages (no reference counterpart: the reference always re-uploads per solve).
 * create = HIP context + scaling + norm estimate (opts as above); upload = device-resident problem in every
 * precision the ladder needs; run = one full solve from the initial point with the problem already in HBM
 * (out / x / y as ganit_solve_gpu); kernel_times = device ms of the fused iteration kernels at level 0 (L1),
 * 1 (L2) or 2 (L3), six entries as ganit_hip_pdhg_time_kernels: ms[0] primal and ms[1] dual kernel with
 * nothing of theirs in L2 (as in a solve), ms[2] one whole iteration (both kernels in solve order), ms[3] primal
 * and ms[4] dual kernel launched back to back, ms[5] the 64 MiB streaming read the cold figures are taken
 * against. */
typedef struct ganit_solver ganit_solver;
ganit_solver* ganit_solver_create(const ganit_lp* lp, const double* opts);
int ganit_solver_upload(ganit_solver* s);
int ganit_solver_run(ganit_solver* s, double* out, double* x, double* y);
int ganit_solver_kernel_times(ganit_solver* s, int level, int reps, double* ms);
/* bytes per column index the fused kernels read for A (0), A^T (1), Q (2): 2 (compact) or 4 */
int ganit_solver_index_bytes(ganit_solver* s, int which, int* bytes);
/* the kernel variant kept for A (0), A^T (1), Q (2): 0 = LDS-staged row blocks, else lanes per row (1..64) */
int ganit_solver_variant(ganit_solver* s, int which, int* lanes);
/* Differentiation through the last ganit_solver_run (no reference counterpart): gx = dL/dx* (n);
 * aopts[0..5] = precision (0 FP64, 1 FP32 matrix + vectors, which needs a mixed/dynamic solver), tol,
 * max operator applications, method (0 GMRES on (I - J^T) lambda = g, tol on the relative residual;
 * 1 plain fixed-point iteration, tol on the relative change), restart length (GMRES) or check interval
 * (fixed point) (0: 200 / 64), bound-activity tolerance (< 0: automatic).  Outputs (any may be NULL):
 * dc = dL/dc (n), db = dL/d(active bound of each row) (m; dL/db of equality and one-sided rows, 0 for inactive
 * rows), dq = dL/dQ_jj (n); info[0..4] = iterations, final relative change, free variables, active rows,
 * seconds. */
int ganit_solver_adjoint(ganit_solver* s, const double* gx, const double* aopts, double* dc, double* db, double* dq,
                         double* info);
/* Make (x (n), y (m)), an original-space solution found elsewhere (ganit_batch_run), the solver's solution as if
 * ganit_solver_run had ended there, so that ganit_solver_adjoint differentiates at it; fp32_iterate = 1 when
 * the point came out of an FP32-vector level (needs a mixed or dynamic solver).  No reference counterpart. */
int ganit_solver_set_solution(ganit_solver* s, const double* x, const double* y, int fp32_iterate);
/* info[0..3] = set-up seconds (scaling + norm estimate), ||A'|| estimate, 1 if the set-up's matrix work ran on the
 * GPU, 1 if the last run kept a polished point (no reference counterpart). */
int ganit_solver_info(const ganit_solver* s, double* info);
void ganit_solver_free(ganit_solver* s);

/* Many LPs and convex QPs solved together by the block-diagonal batched kernels (no reference counterpart:
 * the reference solves one problem per call; this serves branch-and-bound node relaxations, strong-branching
 * probes and batches of QP layers).
 * Each instance gets the single solve's scaling, initial point, step sizes, restart rules and FP64
 * certification; opts as ganit_solve_gpu (check_every must be even; values 2 runs as
 * FP32).
 * run: out[18*k..] = instance k's results in ganit_solve_gpu's layout (solve_s is the whole batch's);
 * x, y = the instances' original-space solutions concatenated in order (either may be NULL).
 * stats[0..23] = rows, cols, nnz, row blocks of A, row blocks of A^T, upload_s, solve_s, certify_s,
 *                L1_s, L2_s, L3_s, instances in L1, L2, L3, max iterations, total iterations,
 *                autotuned partition (-1 none), ms per batched iteration of the 4 candidates, setup_s,
 *                batched iterations run lanes-per-row (opts bit 6), 1 reserved (0). */
typedef struct ganit_batch ganit_batch;
ganit_batch* ganit_batch_create(const ganit_lp* const* lps, int ninst, const double* opts);
int ganit_batch_upload(ganit_batch* b);
int ganit_batch_run(ganit_batch* b, double* out, double* x, double* y);
int ganit_batch_stats(const ganit_batch* b, double* stats);
/* Branch and bound on a resident batch (no reference counterpart; the reference's branch-and-cut re-solves a
 * node with its dual simplex after set_col_bounds, warm-started from the parent's basis: src/mip.cpp:306-323,
 * src/simplex.hpp:64-73, tests/test_warm.cpp:8-38).
 * set_bounds: replace instance k's variable bounds (original space, n_k entries each) in place; the scaled
 *   and device copies follow, nothing else is rebuilt.
 * set_start: start instance k's next runs from the original-space point x (n_k), y (m_k) -- projected onto its
 *   bounds -- with primal weight w (w <= 0: its own initial weight); x = NULL returns it to the cold start.
 * set_options: new tolerances, limits, stall rules, check interval or verbosity (opts as ganit_solve_gpu);
 *   precision, values, scaling and device must stay as created.
 * weights: the primal weight each instance ended the last run with (ninst entries). */
int ganit_batch_set_bounds(ganit_batch* b, int k, const double* l, const double* u);
int ganit_batch_set_start(ganit_batch* b, int k, const double* x, const double* y, double w);
int ganit_batch_set_options(ganit_batch* b, const double* opts);
int ganit_batch_weights(const ganit_batch* b, double* w);
/* Differentiation through the last ganit_batch_run, every instance at once (no reference counterpart): the
 * GMRES adjoint of ganit_solver_adjoint for all instances in lockstep on the block-diagonal matrices.
 * gx = dL/dx* of the instances concatenated (sum of n_k); aopts as ganit_solver_adjoint (method must be 0);
 * dc, dq (sum of n_k), db (sum of m_k) concatenated in instance order (any may be NULL); info[5k..5k+4] as
 * ganit_solver_adjoint's info for instance k (seconds: the whole call). */
int ganit_batch_adjoint(ganit_batch* b, const double* gx, const double* aopts, double* dc, double* db, double* dq,
                        double* info);
void ganit_batch_free(ganit_batch* b);

#ifdef __cplusplus
}
#endif
#endif /* GANIT_HOST_H */
