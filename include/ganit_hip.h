/* ganit_hip.h -- C ABI of libganit_hip.so, the MI355X (gfx950) restarted-PDHG engine.
 *
 * This is the drop-in boundary for the reference's GPU backend.  The reference drives a compile-time
 * backend concept from its driver (src/pdhg_impl.hpp:67-68, 201-202, 243-244); its CUDA implementation is
 * src/backend_cuda.cu:286-378.  Every entry point below either replaces one member of that concept (the
 * citation says which) or is a coarse, fused operation used by this library's own driver.
 *
 * Conventions
 *   - Every function returns 0 on success and a nonzero ganit_status otherwise; no exception crosses the
 *     ABI.  ganit_hip_last_error() returns the calling thread's last message (the reference throws
 *     std::runtime_error from GANIT_CUDA_CHECK instead, src/backend_cuda.cu:22-29).
 *   - Device vectors are plain device pointers; their element type is given by a ganit_dtype argument.
 *     Host arrays are caller-owned; the library never frees caller memory.
 *   - All work of a context is ordered on the context's own HIP stream.  Functions that return host
 *     scalars synchronise that stream; the others are asynchronous.
 *   - Reductions are deterministic: the same inputs give the same bits on every call.
 */
#ifndef GANIT_HIP_H
#define GANIT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Element types.  Matrix values may be F64, F32 or BF16; vectors F64 or F32. */
enum ganit_dtype { GANIT_F64 = 0, GANIT_F32 = 1, GANIT_BF16 = 2 };

enum ganit_status {
    GANIT_OK = 0,
    GANIT_EINVAL = 1,    /* bad argument (null pointer, size mismatch, unsupported dtype) */
    GANIT_EHIP = 2,      /* a HIP runtime call failed */
    GANIT_ENOMEM = 3,    /* device allocation failed */
    GANIT_EINTERNAL = 4  /* anything else */
};

typedef struct ganit_hip_ctx ganit_hip_ctx;   /* device + stream + reduction scratch */
typedef struct ganit_hip_csr ganit_hip_csr;   /* a CSR matrix resident on the device */
typedef struct ganit_hip_pdhg ganit_hip_pdhg; /* a scaled LP/QP resident on the device + PDHG state */

/* ------------------------------------------------------------------ errors / version */
const char* ganit_hip_last_error(void);
int ganit_hip_version(void);

/* ------------------------------------------------------------------ context
 * Replaces the reference's implicit CUDA device/default stream (src/backend_cuda.cu:299-302 reads the
 * SM count for its reduction grid; here the context owns the stream and scratch instead). */
int ganit_hip_ctx_create(int device, ganit_hip_ctx** out);
int ganit_hip_ctx_destroy(ganit_hip_ctx* ctx);
int ganit_hip_ctx_sync(ganit_hip_ctx* ctx);
int ganit_hip_ctx_stream(ganit_hip_ctx* ctx, void** hip_stream);
/* Device name and compute-unit count (the reference's solve_pdhg_gpu reads cudaDeviceProp,
 * src/backend_cuda.cu:383-388). */
int ganit_hip_ctx_device_info(ganit_hip_ctx* ctx, char* name, int name_len, int* compute_units);

/* ------------------------------------------------------------------ device vectors
 * Replace Vec = DevBuf<V> and its helpers (src/backend_cuda.cu:44-74, 305-322, 345-350). */
int ganit_hip_alloc(ganit_hip_ctx* ctx, size_t bytes, void** dptr);
int ganit_hip_free(ganit_hip_ctx* ctx, void* dptr);
/* from_host: narrow a host double array to dtype vdt on the device (src/backend_cuda.cu:306 upload_as). */
int ganit_hip_upload(ganit_hip_ctx* ctx, int vdt, const double* host, int64_t n, void* dst);
/* to_host: widen to double and copy back; synchronises (src/backend_cuda.cu:307-315). */
int ganit_hip_download(ganit_hip_ctx* ctx, int vdt, const void* src, int64_t n, double* host);
/* fill (src/backend_cuda.cu:320-322), copy (:316-319), sub (:345-347), scale_into (:348-350). */
int ganit_hip_vec_fill(ganit_hip_ctx* ctx, int vdt, void* dst, int64_t n, double a);
int ganit_hip_vec_copy(ganit_hip_ctx* ctx, int vdt, const void* src, void* dst, int64_t n);
int ganit_hip_vec_sub(ganit_hip_ctx* ctx, int vdt, const void* a, const void* b, void* d, int64_t n);
int ganit_hip_vec_scale_into(ganit_hip_ctx* ctx, int vdt, const void* src, double a, void* dst, int64_t n);
/* Element-type conversion on the device (src/backend_cuda.cu:154-163 k_convert / k_to_double). */
int ganit_hip_vec_convert(ganit_hip_ctx* ctx, int src_dt, const void* src, int dst_dt, void* dst, int64_t n);

/* ------------------------------------------------------------------ CSR matrices
 * Replaces DevCsr / to_device (src/backend_cuda.cu:252-284).  Indices are uploaded once; values are kept
 * resident in every dtype whose bit (1 << dtype) is set in dtype_mask, so precision levels share one
 * upload.  The row-length histogram picks the lanes-per-row variant (the reference's rule is a single
 * avg-row-length > 12 switch between thread- and 32-lane-warp-per-row, :281-282). */
int ganit_hip_csr_upload(ganit_hip_ctx* ctx, int rows, int cols, const int* ptr, const int* idx,
                         const double* val, unsigned dtype_mask, ganit_hip_csr** out);
int ganit_hip_csr_free(ganit_hip_csr* a);
/* lanes_per_row: the variant in use (0 = row blocks, else 1..64); hist[k] (k = 0..7) counts rows with length in
 * [0,1], (1,2], (2,4], ... , (64,128], and hist[8] rows longer than 128. */
int ganit_hip_csr_plan(const ganit_hip_csr* a, int* lanes_per_row, int64_t* hist9);
/* Override the variant: a power of two in 1..64 selects row groups of that many lanes per row; 0 selects
 * LDS-staged row blocks (up to 256 rows / 2048 nonzeros per workgroup, products staged in LDS and each row
 * summed sequentially in row order, so those sums are bit-identical to a scalar CPU row loop). */
int ganit_hip_csr_set_lanes(ganit_hip_csr* a, int lanes_per_row);
/* Time every variant with `reps` stand-alone SpMVs in (mdt, vdt) and keep the fastest. */
int ganit_hip_csr_autotune(ganit_hip_ctx* ctx, ganit_hip_csr* a, int mdt, int vdt, int reps);

/* ------------------------------------------------------------------ backend-concept kernels
 * spmv_A / spmv_AT / spmv_Q (src/backend_cuda.cu:324-335): y = a * x, accumulated in vdt. */
int ganit_hip_spmv(ganit_hip_ctx* ctx, const ganit_hip_csr* a, int mdt, int vdt, const void* x, void* y);
/* primal_step (src/backend_cuda.cu:337-340, kernel :102-117); qx may be NULL (LP). */
int ganit_hip_primal_step(ganit_hip_ctx* ctx, int vdt, int n, const void* x, const void* aty, const void* qx,
                          const void* c, const void* l, const void* u, double tau, void* xn, void* xbar,
                          void* sumx);
/* dual_step (src/backend_cuda.cu:341-344, kernel :119-133). */
int ganit_hip_dual_step(ganit_hip_ctx* ctx, int vdt, int m, const void* y, const void* axbar, const void* lo,
                        const void* hi, double sigma, void* yn, void* sumy);
/* FP64-accumulated reductions (src/backend_cuda.cu:363-377 with functors :199-249); synchronise. */
int ganit_hip_primal_res2(ganit_hip_ctx* ctx, int vdt, int m, const void* ax, const void* lo, const void* hi,
                          const void* rres, double* out);
int ganit_hip_dual_res(ganit_hip_ctx* ctx, int vdt, int n, const void* aty, const void* qx, const void* c,
                       const void* l, const void* u, const void* cres, int with_c, double* rd2, double* vobj);
int ganit_hip_row_dual_obj(ganit_hip_ctx* ctx, int vdt, int m, const void* y, const void* lo, const void* hi,
                           double* out);
int ganit_hip_dot(ganit_hip_ctx* ctx, int vdt, int64_t n, const void* a, const void* b, double* out);
int ganit_hip_diff2(ganit_hip_ctx* ctx, int vdt, int64_t n, const void* a, const void* b, double* out);
/* ganit_hip_dot without a context and without any allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The block partials go to a fixed 16 KiB device array inside the library (one
 * copy per GPU, held until the process ends; the reduction grid never exceeds 1024 workgroups, so every n
 * fits); the result, bitwise equal to ganit_hip_dot's, is written ASYNCHRONOUSLY on `hip_stream` (a
 * hipStream_t of the current device, NULL = its null stream) to `out`, a DEVICE pointer to one double.
 * Nothing synchronises.  Calls that overlap on one GPU (other streams, other host threads) share that
 * array and race: order them on one stream.  No hipMalloc, hipFree or allocating library call is made;
 * the first call on each device looks up the array's address there (hipGetSymbolAddress). */
int ganit_hip_dot_noalloc(void* hip_stream, int vdt, int64_t n, const void* a, const void* b, double* out);

/* ------------------------------------------------------------------ device-resident PDHG
 * The fused path behind this library's own driver.  One ganit_hip_pdhg holds the scaled problem
 * (ScaledLP, src/scaling.hpp:8-23) uploaded ONCE, with matrix values in every dtype of mdt_mask and
 * problem vectors in every dtype of vdt_mask, plus the iterate working set per vector dtype.  It replaces
 * the per-phase backend rebuild and per-evaluation re-upload of the reference (src/pdhg_impl.hpp:73, 207). */
typedef struct {
    int m, n;
    const int* a_ptr; const int* a_idx; const double* a_val;     /* scaled A, CSR m x n      */
    const int* at_ptr; const int* at_idx; const double* at_val;  /* its stored transpose     */
    const int* q_ptr; const int* q_idx; const double* q_val;     /* scaled Q (n x n) or NULL */
    const double *c, *l, *u, *lo, *hi, *rres, *cres;            /* scaled vectors           */
} ganit_hip_lp_desc;

int ganit_hip_pdhg_create(ganit_hip_ctx* ctx, const ganit_hip_lp_desc* lp, unsigned mdt_mask, unsigned vdt_mask,
                          ganit_hip_pdhg** out);
int ganit_hip_pdhg_destroy(ganit_hip_pdhg* pd);
/* Bytes per column index the fused kernels read for matrix `which` (0 A, 1 A^T, 2 Q): 2 when the matrix runs
 * on LDS-staged row blocks with compact indices (each block's columns = a base column + 16-bit offsets), else 4.
 * No reference counterpart (the reference always reads 32-bit indices). */
int ganit_hip_pdhg_index_bytes(const ganit_hip_pdhg* pd, int which, int* bytes);
/* Whether matrix `which` (0 A, 1 A^T, 2 Q) stages its LDS windows swizzled (entry w at slot w ^ (w >> 4)):
 * chosen at upload from the bank pattern of its gathers (GANIT_HIP_WINDOW_SWIZZLE=0/1 forces it); the row sums
 * are the same bits either way.  No reference counterpart (the reference has no LDS windows). */
int ganit_hip_pdhg_window_layout(const ganit_hip_pdhg* pd, int which, int* swizzled);
/* The matrices' variant plans (A, AT, Q); the variant may be overridden per matrix (which: 0 = A, 1 = AT,
 * 2 = Q), with the lanes_per_row values of ganit_hip_csr_set_lanes.  ganit_hip_pdhg_create picks each
 * matrix's variant by timing them (ganit_hip_csr_autotune) unless GANIT_HIP_AUTOTUNE=0. */
int ganit_hip_pdhg_plan(const ganit_hip_pdhg* pd, int which, int* lanes_per_row, int64_t* hist9);
int ganit_hip_pdhg_set_lanes(ganit_hip_pdhg* pd, int which, int lanes_per_row);

/* Load an iterate (x, y) and restart point (xr, yr) from the host into the vdt working set and clear the
 * running sums (the vector set-up of pdhg_phase, src/pdhg_impl.hpp:78-84).  xr / yr may be NULL (= x, y). */
int ganit_hip_pdhg_load(ganit_hip_pdhg* pd, int vdt, const double* x, const double* y, const double* xr,
                        const double* yr);
/* Device-side hand-over between precision levels: copy (x, y, xr, yr) of working set from_vdt into
 * working set to_vdt with conversion, and clear to_vdt's running sums (replaces to_host / from_host of
 * PdhgState across phases, src/pdhg_impl.hpp:191-197, 78-82). */
int ganit_hip_pdhg_handover(ganit_hip_pdhg* pd, int from_vdt, int to_vdt);
/* Run `iters` fused PDHG iterations (src/pdhg_impl.hpp:130-141) in matrix dtype mdt and vector dtype vdt
 * with step sizes tau, sigma.  Each iteration is two kernels: [AT y (+ Q x) -> primal prox] and
 * [A xbar -> dual prox].  Blocks of the same length are replayed from a captured HIP graph. */
int ganit_hip_pdhg_iterate(ganit_hip_pdhg* pd, int mdt, int vdt, int iters, double tau, double sigma);

/* Check flags and the layout of the sums ganit_hip_pdhg_check returns (all FP64-accumulated). */
#define GANIT_CHECK_AVG 1  /* also form (xa, ya) = (sumx, sumy) * inv_k and evaluate them          */
#define GANIT_CHECK_RAY 2  /* also the infeasibility ray y - yr (src/pdhg_impl.hpp:110-118)        */
#define GANIT_CHECK_FP64 4 /* with vdt = GANIT_F32: evaluate the FP32 iterate (and its FP32 average)
                              in FP64 -- each entry widened as the check reads it, against the FP64
                              problem data with matrix values mdt (GANIT_F64 for a full FP64
                              evaluation).  Requires the FP64 working set (vdt_mask bit 0). */
#define GANIT_CHECK_KAHAN 8 /* sums formed from FP32-rounded terms with Kahan-compensated FP32 accumulation per
                             * thread, combined in FP64 (extension; without it every sum is FP64) */
#define GANIT_CHECK_DRAY 16 /* also the primal ray d = x - xr of an LP: c'd and d's excess over the recession cones
                             * of the row and variable bounds (extension: the dual-infeasibility test that the
                             * reference lists as missing, README.md:197; its src/pdhg_impl.hpp:110-118,163
                             * tests only primal infeasibility, on the ray y - yr) */
#define GANIT_CHECK_NOUT 22
/* out[ 0.. 3]  current point : rd2, vobj, c'x, x'Qx      (dual_res, dot_c, dot with Qx)
 * out[ 4.. 7]  average point : rd2, vobj, c'x, x'Qx
 * out[ 8.. 9]  ray           : rd2, vobj with cw = 0 and no Qx
 * out[10..11]  ||x - xr||^2, ||xa - xr||^2
 * out[12..13]  current point : primal_res2, row_dual_obj
 * out[14..15]  average point : primal_res2, row_dual_obj
 * out[16]      ray           : row_dual_obj(y - yr)
 * out[17..18]  ||y - yr||^2, ||ya - yr||^2
 * out[19]      primal ray    : sum_i ((A d)_i's distance to its row's recession cone / row scale)^2
 * out[20..21]  primal ray    : c'd, sum_j (d_j's distance to its bounds' recession cone * column scale)^2
 * Entries whose flag is off are returned as 0.  Replaces the 8-12 separate reductions and host syncs of
 * one check (src/pdhg_impl.hpp:92-118, 144-147, 179) with two kernels (three with split rows, plus an averaging
 * kernel when a matrix runs a lanes-per-row variant) and one synchronisation. */
int ganit_hip_pdhg_check(ganit_hip_pdhg* pd, int mdt, int vdt, int flags, double inv_k, double* out);
/* Restart (src/pdhg_impl.hpp:177-188): optionally adopt the average, then (xr, yr) = (x, y) and clear the
 * running sums. */
int ganit_hip_pdhg_restart(ganit_hip_pdhg* pd, int vdt, int adopt_average);
/* Make the average the current iterate without restarting (the best_is_avg hand-back at the end of a
 * phase, src/pdhg_impl.hpp:191-192). */
int ganit_hip_pdhg_adopt_average(ganit_hip_pdhg* pd, int vdt);
/* Copy the current iterate (x, y) and restart point (xr, yr) of working set vdt to the host as doubles;
 * any pointer may be NULL.  Synchronises. */
int ganit_hip_pdhg_store(ganit_hip_pdhg* pd, int vdt, double* x, double* y, double* xr, double* yr);

/* ------------------------------------------------------------------ problem set-up on the device
 * The matrix half of the reference's scale_lp (src/scaling.cpp:36-70: ruiz_iters inf-norm Ruiz passes, then the
 * Pock-Chambolle 1-norm pass when pock_chambolle != 0), the stable transpose of the pattern (src/lp.cpp:55-72)
 * and the SpMVs of estimate_norm's power iteration (src/scaling.cpp:131-155), on A (CSR, FP64, unscaled)
 * uploaded once.  Every result is bitwise the reference's host arithmetic (maxima exact; row and column 1-norms
 * summed by one thread in the reference's order; the same factor products; sequential row sums without FMA). */
typedef struct ganit_hip_setup ganit_hip_setup;
int ganit_hip_setup_create(ganit_hip_ctx* ctx, int m, int n, const int* a_ptr, const int* a_idx, const double* a_val,
                           int ruiz_iters, int pock_chambolle, ganit_hip_setup** out);
/* R (m) and C (n): the accumulated row / column factors; a_val (nnz): the scaled values of A in A's order;
 * at_ptr (n + 1), at_idx, at_val (nnz): the scaled transpose.  Any pointer may be NULL.  Synchronises. */
int ganit_hip_setup_result(const ganit_hip_setup* s, double* R, double* C, double* a_val, int* at_ptr, int* at_idx,
                           double* at_val);
/* estimate_norm of the scaled matrix (src/scaling.cpp:131-155): `iters` power-iteration steps on A'^T A' from the
 * normalised all-ones vector, the SpMVs on the device and each step's ||A'^T A' v|| summed on the host in the
 * reference's serial order; *norm = sqrt(lambda) * 1.01 (1 for an empty matrix).  Synchronises. */
int ganit_hip_setup_norm(ganit_hip_setup* s, int iters, double* norm);
int ganit_hip_setup_destroy(ganit_hip_setup* s);

/* ------------------------------------------------------------------ batched PDHG (many uneven LPs / QPs)
 * No reference counterpart: the reference solves one LP per call and its branch-and-cut uses only the
 * dual simplex (src/mip.cpp:116, 312-323).  The instances (scaled LPs or convex QPs) are packed
 * block-diagonally (A, A^T and, when any instance has one, Q); each keeps its own restart state, step sizes
 * (with the QP cap of src/pdhg_impl.hpp:87-91), primal weight and status on the device, and the check logic
 * of src/pdhg_impl.hpp:144-189 is applied per instance by a device kernel. */
typedef struct ganit_hip_batch ganit_hip_batch;
/* inst_scalars[k] = {eta, ||A||, obj_scale, obj_const, bnorm, cnorm, w0} of instance k, followed by ||Q'||
 * when instance k has a Q (lps[k]->q_ptr non-NULL). */
int ganit_hip_batch_create(ganit_hip_ctx* ctx, int ninst, const ganit_hip_lp_desc* const* lps,
                           const double* const* inst_scalars, unsigned mdt_mask, unsigned vdt_mask,
                           ganit_hip_batch** out);
int ganit_hip_batch_destroy(ganit_hip_batch* bt);
/* dims = {instances, rows, columns, nnz, row blocks of A, row blocks of A^T} */
int ganit_hip_batch_dims(const ganit_hip_batch* bt, int64_t* dims);
/* Initial iterate (concatenated x, y) and primal weights; every instance starts running. */
int ganit_hip_batch_load(ganit_hip_batch* bt, int vdt, const double* x, const double* y, const double* w);
/* Replace instance k's variable bounds in place, in scaled space (l, u: that instance's n entries), in every
 * resident vector dtype; the next load and phase run on them.  Nothing is rebuilt: a branch-and-bound child
 * differs from its parent only here.  No reference counterpart (its branch-and-cut changes bounds in the dual
 * simplex, src/simplex.hpp:64-73 set_col_bounds and warm start, src/mip.cpp:306-307). */
int ganit_hip_batch_set_bounds(ganit_hip_batch* bt, int k, const double* l, const double* u);
/* Convert every instance's (x, y, xr, yr) between working sets and clear the target's running sums. */
int ganit_hip_batch_handover(ganit_hip_batch* bt, int from_vdt, int to_vdt);
/* The instances with sel[k] != 0 run in the next phase; their iteration count, restarts, weight and
 * restart error carry over (the PdhgState hand-over, src/pdhg_impl.hpp:55-63). */
int ganit_hip_batch_select(ganit_hip_batch* bt, const int* sel);
/* Run one precision phase for the selected instances until each converges, stalls (stall_exit),
 * is certified infeasible or hits the iteration limit; params = {tol, infeas_tol, max_iter,
 * check_every (even), stall_checks, stall_exit, time_limit_s, eval_fp64}: eval_fp64 != 0 evaluates a
 * reduced-precision phase's checks in FP64 (FP64 matrix and problem data, the iterate widened as it is read), as
 * ganit_hip_pdhg_check's GANIT_CHECK_FP64 does for one problem. */
int ganit_hip_batch_phase(ganit_hip_batch* bt, int mdt, int vdt, const double* params, int* running_left);
/* Autotuner of the batched kernels: times `reps` iterations at (mdt, vdt) for each row-block partition
 * candidate (rows, nonzeros per workgroup) and keeps the fastest; ms[c] = ms per iteration of candidate c
 * (up to 8, may be NULL), *chosen = its index.  Advances the loaded iterate: load again afterwards. */
int ganit_hip_batch_autotune(ganit_hip_batch* bt, int mdt, int vdt, int reps, double* ms, int* chosen);
/* Launch-bound tail (no reference counterpart): while at most 8 instances run, with at most 2^18 nonzeros, the
 * batched iteration kernels run 8, 16 or 32 lanes per row (a fixed lane tree per row) instead of the row blocks
 * (each row summed in row order, the CPU's bits).  mode bit 0 (the default, mode 1) allows it -- 0 keeps the row
 * blocks to the end; bit 1 also ends a reduced-precision phase (stall exit on) once its instances are such a
 * tail: they leave as stalled at the next check, so the ladder carries them on in FP64 (a batch policy, not the
 * reference's ladder rule); -1 leaves the mode.  tail_iters (may be NULL) = iterations run lanes-per-row so far. */
int ganit_hip_batch_tail(ganit_hip_batch* bt, int mode, int64_t* tail_iters);
/* FP64 KKT of every instance's FP64 iterate with matrix values mdt (the certification). */
int ganit_hip_batch_certify(ganit_hip_batch* bt, int mdt);
/* ints[5k..] = status (0 running, 1 converged, 2 infeasible, 3 iteration limit, 4 numerical, 5 stalled,
 * 6 time limit), iterations, restarts, best_is_avg, active; dbl[13k..] = phase-result KKT (pobj, dobj,
 * rel_p, rel_d, rel_gap, err), certified FP64 KKT (same six), primal weight. */
int ganit_hip_batch_results(ganit_hip_batch* bt, int* ints, double* dbl);
int ganit_hip_batch_store(ganit_hip_batch* bt, int vdt, double* x, double* y);
/* The adjoint of ganit_hip_pdhg_adjoint (GMRES form, below) for every instance at once: restarted GMRES on
 * (I - J_k^T) lambda_k = (gx_k, 0) for all k in lockstep, J_k^T applied for all instances by the block-diagonal
 * matrices in one set of launches, dot products and Hessenberg systems per instance.  Linearised at the FP64
 * working set's iterate; gx (N, scaled space), per-instance steps tau[k], sigma[k] and bound-activity tolerance
 * mask_tol[k]; params = {max operator applications, tol, restart length (1..1024)}.  p_out (N) = p and mu_out
 * (M) = F o lambda_y at the solution, as ganit_hip_pdhg_adjoint returns them per instance; info[4k..4k+3] =
 * operator applications, relative residual, free variables and active rows of instance k.  No reference
 * counterpart (the reference has no differentiation). */
int ganit_hip_batch_adjoint(ganit_hip_batch* bt, int mdt, int vdt, const double* gx, const double* tau,
                            const double* sigma, const double* mask_tol, const double* params, double* p_out,
                            double* mu_out, double* info);

/* Adjoint of the solution map at the FP64 working set's iterate (a converged forward solve; no reference
 * counterpart).  gx = dL/dx in the scaled space (n).  Solves lambda = J^T lambda + (gx, 0), J the PDHG
 * iteration's Jacobian with the projections frozen at the iterate (D_j = 0 when x_j lies within
 * mask_tol (1 + |bound|) of l_j or u_j; F_i = 1 for equality rows and rows with y_i != 0), in matrix values mdt
 * and working precision vdt, with steps (tau, sigma).  params = {method, max_iter, tol, every, mask_tol}:
 * method 0 iterates lambda <- J^T lambda + (gx, 0) and stops when ||lambda' - lambda|| / ||lambda'|| <= tol
 * (checked every `every` iterations); method 1 runs restarted GMRES(every) on (I - J^T) lambda = (gx, 0)
 * (the same operator, Krylov-accelerated) until ||residual|| / ||gx|| <= tol, or until a restart cycle lowers the
 * residual by less than 1% (the working precision's floor).  Outputs
 * p = D o (lx - 2 sigma A^T (F o ly)) (n) and mu = F o ly (m), from which dL/dc = -tau p,
 * dL/dQ_jj = -tau p_j x_j and dL/d(active bound of row i) = sigma mu_i (scaled space).
 * info[0..3] = iterations (operator applications), final relative change / residual, free variables
 * (sum D), active rows (sum F). */
int ganit_hip_pdhg_adjoint(ganit_hip_pdhg* pd, int mdt, int vdt, const double* gx, double tau, double sigma,
                           const double* params, double* p, double* mu, double* info);

/* Polish an LP's FP64 iterate (no reference counterpart; the reference lists crossover / polish as planned,
 * README.md:195).  The active set is read off the iterate with tolerance params[0] (a variable within it of a
 * bound is fixed there; a row is active when it is an equality, its dual is not negligible, or its activity is
 * that close to a bound), and two least-squares projections are solved by conjugate gradients (relative residual
 * params[1], at most params[2] products each): x onto the face {A_T x = b_T, fixed variables at their bounds},
 * y onto {free variables' reduced costs = 0, y = 0 off the active rows}.  The result replaces the FP64 iterate
 * (and restart point); judge it with a check (its FP64 KKT errors) and reload the old point to reject it.
 * info[0..5] = free variables, active rows, primal CG products, primal CG relative residual, dual CG products,
 * dual CG relative residual.  LPs only. */
int ganit_hip_pdhg_polish(ganit_hip_pdhg* pd, const double* params, double* info);

/* Device time in milliseconds of the fused iteration kernels, measured with HIP events on the context stream from
 * the current iterate, each figure averaged over `reps` launches (a launch's duration plus the dispatch gap to
 * the next one).  Six entries:
 *   ms[0] primal kernel (A^T y [+ Q x] + primal prox) and ms[1] dual kernel (A xbar + dual prox), each launched
 *         after a 64 MiB streaming read that leaves none of its data in L2 -- as in a solve, where the two
 *         kernels alternate -- less the read's own time (ms[5]);
 *   ms[2] one whole iteration, both kernels in the order a solve launches them;
 *   ms[3] primal and ms[4] dual kernel launched back to back, where part of a kernel's own data stays in L2
 *         from one launch to the next;
 *   ms[5] the 64 MiB streaming read alone.
 * Allocates the 64 MiB for the duration of the call.  Advances the iterate; no reference counterpart (the
 * reference has no per-kernel timers). */
int ganit_hip_pdhg_time_kernels(ganit_hip_pdhg* pd, int mdt, int vdt, int reps, double tau, double sigma,
                                double* ms);

#ifdef __cplusplus
}
#endif
#endif /* GANIT_HIP_H */
