// ganit restarted-PDHG LP/QP solver on MI355X -- public C++ API.
// Mirrors the reference's include/ganit/pdhg.hpp:10-47 (Precision, Options, Status, Result,
// solve_pdhg_gpu); additions are marked "extension".
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "lp.hpp"

namespace ganit {

// Double = L3 only; Mixed = L1 -> L3; Dynamic = L1 -> L2 -> L3 (reference pdhg.hpp:10, pdhg_impl.hpp:13-22).
enum class Precision { Double, Mixed, Dynamic };
// extension: element type of the matrix values in the reduced-precision levels L1 and L2.
// Matrix values of the reduced-precision levels: FP32, BF16, or BF16 first and FP32 once BF16 stalls
// (extension: the L1 phase of a mixed or dynamic ladder runs as a BF16 phase, left after stall_checks checks
// without progress, then an FP32 phase, left after stall_checks_l2; both on the FP32 vectors).
enum class Values { F32, BF16, BF16_F32 };
// extension: where the matrix half of the set-up runs (scale_lp's Ruiz / Pock-Chambolle passes and the transpose,
// estimate_norm's SpMVs): on the host, on the GPU (bitwise the same results), or Auto = the GPU for problems of at
// least a million nonzeros whose rows and columns all have at most 4096 entries (each is swept by one thread).
enum class Setup { Auto, Host, Device };

struct Options {
    double tol = 1e-4;           // relative KKT tolerance (primal, dual, gap)
    long max_iter = 10000000;
    double time_limit = 3600.0;  // seconds
    int check_every = 64;        // evaluate KKT every N iterations
    int ruiz_iters = 10;
    bool pock_chambolle = true;
    int verbose = 1;
    double log_interval = 1.0;   // seconds between progress lines
    double infeas_tol = 1e-8;    // relative tolerance for infeasibility certificates
    Precision precision = Precision::Double;
    int stall_checks = 12;       // checks without 20% progress before leaving L1
    int stall_checks_l2 = 24;    // same for L2
    // extensions
    Values values = Values::F32;  // matrix values in L1/L2
    bool eval_fp64 = false;       // L1/L2 checks (restart candidates, stall and convergence tests) evaluated
                                  // in FP64 with the FP64 matrix instead of at the level's own precision
    int device = 0;               // HIP device ordinal
    bool kahan = false;           // L1 checks at FP32: FP32-rounded terms, Kahan-compensated FP32 sums per
                                  // thread, FP64 across threads (instead of FP64 sums)
    bool detect_unbounded = false;  // LPs: also test the primal ray x - xr as a dual-infeasibility
                                    // (unboundedness) certificate, reported as Status::DualInfeasible
    Setup setup = Setup::Auto;      // where the set-up's matrix work runs
    bool polish = false;            // LPs: after an optimal solve, project the point onto the face its active set
                                    // names (crossover-style polish, on the GPU) and keep it when its FP64 KKT
                                    // errors are all <= polish_tol
    double polish_tol = 1e-9;
    bool scale_q = false;           // QPs: column scaling also equilibrates Q (scale_lp's q_columns); changes the
                                    // trajectory, so it is off unless asked for
    bool tail_lanes = true;         // batches: once the running instances are few and small (launch-bound), iterate
                                    // them lanes-per-row (row sums by a lane tree) instead of on row blocks (row
                                    // order, the CPU's bits); false keeps the row blocks to the end
    bool tail_fp64 = false;         // batches: the reduced-precision levels end once the running instances are such
                                    // a tail, which the ladder then carries on in FP64 (a batch policy: FP32 saves
                                    // nothing on launch-bound iterations; not the reference's ladder rule)
};

// DualInfeasible (an extension, only with Options::detect_unbounded) comes last so the reference's codes keep
// their values.
enum class Status { Optimal, PrimalInfeasible, IterationLimit, TimeLimit, NumericalError, DualInfeasible };
const char* to_string(Status s);

struct Result {
    Status status = Status::NumericalError;
    std::vector<double> x;  // primal solution (original space)
    std::vector<double> y;  // row duals (original space, reduced cost = c - A^T y)
    double pobj = 0, dobj = 0;
    double rel_primal = 0, rel_dual = 0, rel_gap = 0;
    long iterations = 0;
    long fp32_iterations = 0;
    long l2_iterations = 0;
    int restarts = 0;
    double setup_seconds = 0, solve_seconds = 0;
    std::string device;
    // extensions
    double norm_a = 0;
    double upload_seconds = 0;        // device upload of the scaled problem (inside solve_seconds)
    bool setup_on_device = false;     // the set-up's matrix work ran on the GPU
    bool polished = false;            // Options::polish replaced the PDHG point by its polished vertex
    double polish_seconds = 0;        // inside solve_seconds
    double iterate_seconds[3] = {0};  // wall time in L1 / L2 / L3 iteration blocks + checks
    long checks = 0;
};

// Runs restarted PDHG on the MI355X through libganit_hip.so; throws std::runtime_error on HIP errors.
// Timing follows the reference: setup_seconds = scaling + norm estimate; solve_seconds includes the device
// upload and the FP64 certification (pdhg_impl.hpp:273-274, 305).
Result solve_pdhg_gpu(const LP& lp, const Options& opt);

struct ScaledLP;

// extension: options and result of the adjoint (differentiation through a solve).
struct AdjointOptions {
    bool fp32 = false;     // FP32 matrix values and vectors (needs a mixed/dynamic solver), else FP64
    bool gmres = true;     // GMRES on (I - J^T) lambda = g; false: the plain fixed-point iteration
    double tol = 1e-8;     // GMRES: relative residual; fixed point: relative change of lambda
    int max_iter = 20000;  // operator applications
    int restart = 200;     // GMRES restart length
    int check_every = 64;  // fixed point: iterations between convergence checks
    double mask_tol = -1;  // bound-activity tolerance; < 0: 1e-12, or 1e-6 after an FP32-vector level
};
struct AdjointResult {
    std::vector<double> dc;  // dL/dc (n)
    std::vector<double> db;  // dL/d(active bound of each row) (m): dL/db for equality and one-sided rows
    std::vector<double> dq;  // dL/dQ_jj (n)
    long iterations = 0;
    double rel_change = 0;   // relative change of lambda at the last check
    int free_vars = 0, active_rows = 0;
    double seconds = 0;
};

// extension: the same solve split into set-up (scaling, norm), device upload and repeated runs from the
// initial point, so a caller can time solves with the problem already resident in HBM.
class GpuSolver {
   public:
    GpuSolver(const LP& lp, const Options& opt);  // HIP context + scaling + norm estimate
    ~GpuSolver();
    GpuSolver(const GpuSolver&) = delete;
    GpuSolver& operator=(const GpuSolver&) = delete;
    void upload();   // device-resident problem in every precision the ladder needs; idempotent
    Result run();    // one full solve from the initial point (uploads first if needed)
    // Device milliseconds of the fused iteration kernels at a precision level (0 = L1, 1 = L2, 2 = L3): ms[0]
    // primal kernel, ms[1] dual kernel (each launched after a 64 MiB streaming read, which leaves none of its data
    // in L2, as in a solve), ms[2] one whole iteration in solve order, ms[3] / ms[4] primal / dual launched back to
    // back, ms[5] the 64 MiB streaming read alone.
    void kernel_times(int level, int reps, double* ms);
    // Bytes per column index the fused kernels read for A (0), A^T (1) or Q (2): 2 with compact indices, else 4.
    int index_bytes(int which);
    // The kernel variant the autotuner kept for A (0), A^T (1) or Q (2): 0 = LDS-staged row blocks, else lanes per
    // row (1..64).
    int variant(int which);
    // Gradient of L(x*) with respect to c, the active row bounds and diag Q, given gx = dL/dx* at the last
    // run()'s solution: the adjoint PDHG iteration of libganit_hip.so at that solution.
    AdjointResult adjoint(const std::vector<double>& gx, const AdjointOptions& opt);
    // Make (x, y) -- an original-space solution found elsewhere, e.g. by the batched solver -- this solver's
    // solution, as if run() had ended there: adjoint() then differentiates at it.  fp32_iterate: the point
    // came out of an FP32-vector level (its bound activity is judged to FP32 rounding).
    void set_solution(const std::vector<double>& x, const std::vector<double>& y, bool fp32_iterate);
    const ScaledLP& scaled() const;
    double norm_a() const;
    double setup_seconds() const;  // scaling + norm estimate, as Result::setup_seconds reports it
    bool setup_on_device() const;  // the set-up's matrix work ran on the GPU
    bool last_polished() const;    // the last run kept a polished point (Options::polish)

   private:
    struct Impl;
    std::unique_ptr<Impl> p_;
};

// extension: many LPs and convex QPs solved together by the block-diagonal batched kernels.  Each instance gets the single
// solve's scaling, initial point, step sizes and restart rules; a precision level runs for all selected
// instances at once, every result is certified in FP64, and only the instances the ladder sends on take
// part in the next level.
struct BatchStats {
    int64_t rows = 0, cols = 0, nnz = 0, blocks_a = 0, blocks_at = 0;
    double upload_seconds = 0, solve_seconds = 0, certify_seconds = 0;
    double level_seconds[3] = {0, 0, 0};  // wall time of the L1 / L2 / L3 batched phases
    int level_instances[3] = {0, 0, 0};   // instances taking part in each level
    long max_iterations = 0, total_iterations = 0;
    int autotune_pick = -1;               // chosen row-block partition candidate (-1: not tuned)
    double autotune_ms[8] = {0};          // ms per batched iteration of each candidate
    long tail_iterations = 0;             // batched iterations run lanes-per-row (Options::tail_lanes), last run
};

class BatchSolver {
   public:
    BatchSolver(const std::vector<const LP*>& lps, const Options& opt);  // context + per-instance set-up
    ~BatchSolver();
    BatchSolver(const BatchSolver&) = delete;
    BatchSolver& operator=(const BatchSolver&) = delete;
    void upload();              // pack + upload every instance, autotune the partition; idempotent
    std::vector<Result> run();  // one solve of every instance from its initial point
    // Branch and bound on a resident batch (extension; the reference re-solves nodes with its dual simplex,
    // src/mip.cpp:306-323, warm-started from the parent's basis, tests/test_warm.cpp:8-38):
    // replace instance k's variable bounds (original space, n_k entries each) in place -- the scaled copy and
    // the device copy follow, nothing is rebuilt or re-uploaded besides those two vectors ...
    void set_bounds(int k, const std::vector<double>& l, const std::vector<double>& u);
    // ... start instance k's next runs from an original-space point (x: n_k, y: m_k; e.g. its parent's solution,
    // projected onto the instance's bounds) with primal weight w (w <= 0: the instance's own initial weight) ...
    void set_start(int k, const std::vector<double>& x, const std::vector<double>& y, double w);
    void clear_start(int k);  // ... or from the cold start again (x = proj(0), y = 0)
    // Change tolerances, limits, stall rules, check interval or verbosity between runs; precision, matrix
    // values, scaling and device stay as constructed (extension).
    void set_options(const Options& opt);
    // The primal weight each instance ended the last run with (what a child's warm start takes over).
    std::vector<double> final_weights() const;
    // Differentiation through the last run (extension): every instance's adjoint at its solution at once --
    // GpuSolver::adjoint's GMRES form for all instances in lockstep on the block-diagonal matrices.  gx = dL/dx*
    // of every instance concatenated (sum of n_k entries); one result per instance.  GMRES only.
    std::vector<AdjointResult> adjoint(const std::vector<double>& gx, const AdjointOptions& ao);
    int size() const;
    double setup_seconds() const;
    const ScaledLP& scaled(int k) const;
    const BatchStats& stats() const;  // of the last run

   private:
    struct Impl;
    std::unique_ptr<Impl> p_;
};

}  // namespace ganit
