// Batched restarted PDHG: many independent LPs (branch-and-bound node relaxations, strong-branching probes)
// solved together by the block-diagonal kernels of libganit_hip.so (ganit_hip_batch_*).
//
// Every instance follows the single-instance driver's rules (reference src/pdhg_impl.hpp:67-328): its own
// scaling, norm estimate, initial point and primal weight; the same step sizes, checks, adaptive restarts,
// stall exit and precision ladder.  The difference is the schedule: a precision level runs for all selected
// instances at once (the device decides per instance after every block of check_every iterations), then
// every instance's result is certified in FP64 in one pass, and only the instances that the ladder's rule
// sends on (pdhg_impl.hpp:276-299) take part in the next level.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include "../../include/ganit_hip.h"
#include "pdhg.hpp"
#include "scaling.hpp"

namespace ganit {

namespace {
void chk(int rc) {
    if (rc != GANIT_OK) throw std::runtime_error(std::string("HIP backend: ") + ganit_hip_last_error());
}
using Clock = std::chrono::steady_clock;
double since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

// device status codes of ganit_hip_batch_results
enum { kRunning = 0, kConverged = 1, kInfeasible = 2, kIterLimit = 3, kNumerical = 4, kStalled = 5, kTimeLimit = 6 };

Status status_of(int code) {
    switch (code) {
        case kConverged: return Status::Optimal;
        case kInfeasible: return Status::PrimalInfeasible;
        case kTimeLimit: return Status::TimeLimit;
        case kNumerical: return Status::NumericalError;
        default: return Status::IterationLimit;  // iteration limit, stalled (pdhg_impl.hpp:232-240)
    }
}
bool converged6(const double* k, double tol) { return k[2] <= tol && k[3] <= tol && k[4] <= tol; }
}  // namespace

struct BatchSolver::Impl {
    Options opt;
    ganit_hip_ctx* ctx = nullptr;
    ganit_hip_batch* bt = nullptr;
    std::string device;
    std::vector<ScaledLP> s;
    std::vector<double> norm_a, w0;
    std::vector<int> row0, col0;
    int rows = 0, cols = 0;
    double setup_seconds = 0, autotune_ms[8] = {0};
    int autotune_pick = -1;
    bool mixed = false;
    int low_mdt = GANIT_F32;
    BatchStats stats;
    // per-instance start of the next runs in scaled space (warm[k] = 0: the cold start proj(0), y = 0)
    std::vector<char> warm;
    std::vector<std::vector<double>> xw, yw;
    std::vector<double> ww;
    std::vector<double> w_final;  // primal weights at the end of the last run
    std::vector<char> final_fp32;  // per instance: its last run ended in the FP32-vector level
    bool solved = false;           // a run has completed and nothing has changed the instances since
    explicit Impl(const Options& o) : opt(o) {}
    ~Impl() {
        if (bt) ganit_hip_batch_destroy(bt);
        if (ctx) ganit_hip_ctx_destroy(ctx);
    }
};

BatchSolver::BatchSolver(const std::vector<const LP*>& lps, const Options& opt) : p_(new Impl(opt)) {
    Impl& I = *p_;
    if (lps.empty()) throw std::runtime_error("empty batch");
    if (opt.check_every < 2 || opt.check_every % 2) throw std::runtime_error("batched check_every must be even");
    for (const LP* lp : lps)
        if (!lp) throw std::runtime_error("null batch instance");
    chk(ganit_hip_ctx_create(opt.device, &I.ctx));
    char name[256] = {0};
    chk(ganit_hip_ctx_device_info(I.ctx, name, sizeof(name), nullptr));
    I.device = std::string("gpu:") + name;
    const auto t0 = Clock::now();
    const int K = (int)lps.size();
    I.s.resize(K);
    I.norm_a.resize(K);
    I.w0.resize(K);
    // independent set-ups, one instance per thread (the per-instance arithmetic is the single solve's)
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < K; ++k) {
        I.s[k] = scale_lp(*lps[k], opt.ruiz_iters, opt.pock_chambolle, true, opt.scale_q);
        I.norm_a[k] = estimate_norm(I.s[k]);
        const double cn = norm2(I.s[k].c), bn = finite_bound_norm(I.s[k].lo, I.s[k].hi);
        I.w0[k] = (cn > 1e-10 && bn > 1e-10) ? cn / bn : 1.0;
    }
    I.row0.resize(K + 1);
    I.col0.resize(K + 1);
    I.row0[0] = I.col0[0] = 0;
    for (int k = 0; k < K; ++k) {
        I.row0[k + 1] = I.row0[k] + I.s[k].m;
        I.col0[k + 1] = I.col0[k] + I.s[k].n;
    }
    I.rows = I.row0[K];
    I.cols = I.col0[K];
    I.warm.assign(K, 0);
    I.xw.resize(K);
    I.yw.resize(K);
    I.ww.assign(K, 0.0);
    I.mixed = opt.precision != Precision::Double;
    I.low_mdt = opt.values == Values::BF16 ? GANIT_BF16 : GANIT_F32;
    I.setup_seconds = since(t0);
}

BatchSolver::~BatchSolver() = default;

int BatchSolver::size() const { return (int)p_->s.size(); }
double BatchSolver::setup_seconds() const { return p_->setup_seconds; }
const ScaledLP& BatchSolver::scaled(int k) const { return p_->s.at(k); }
const BatchStats& BatchSolver::stats() const { return p_->stats; }

void BatchSolver::upload() {
    Impl& I = *p_;
    if (I.bt) return;
    const int K = size();
    std::vector<ganit_hip_lp_desc> d(K);
    std::vector<const ganit_hip_lp_desc*> dp(K);
    std::vector<std::array<double, 8>> sc(K);
    std::vector<const double*> scp(K);
    for (int k = 0; k < K; ++k) {
        const ScaledLP& s = I.s[k];
        ganit_hip_lp_desc& e = d[k];
        e = ganit_hip_lp_desc{};
        e.m = s.m; e.n = s.n;
        e.a_ptr = s.A.ptr.data(); e.a_idx = s.A.idx.data(); e.a_val = s.A.val.data();
        e.at_ptr = s.AT.ptr.data(); e.at_idx = s.AT.idx.data(); e.at_val = s.AT.val.data();
        e.c = s.c.data(); e.l = s.l.data(); e.u = s.u.data(); e.lo = s.lo.data(); e.hi = s.hi.data();
        e.rres = s.Rres.data(); e.cres = s.Cres.data();
        if (s.has_q()) { e.q_ptr = s.Q.ptr.data(); e.q_idx = s.Q.idx.data(); e.q_val = s.Q.val.data(); }
        dp[k] = &d[k];
        sc[k] = {0.998 / I.norm_a[k], I.norm_a[k], s.cscale * s.bscale, s.obj_const, s.bnorm, s.cnorm, I.w0[k],
                 s.normQ};
        scp[k] = sc[k].data();
    }
    const unsigned mmask = (1u << GANIT_F64) | (I.mixed ? (1u << I.low_mdt) : 0u);
    const unsigned vmask = (1u << GANIT_F64) | (I.mixed ? (1u << GANIT_F32) : 0u);
    chk(ganit_hip_batch_create(I.ctx, K, dp.data(), scp.data(), mmask, vmask, &I.bt));
    // pick the row-block partition for the level that runs first (the batch's bulk of iterations), timed on the
    // row blocks whatever the tail mode
    chk(ganit_hip_batch_tail(I.bt, 0, nullptr));
    const char* at = std::getenv("GANIT_HIP_AUTOTUNE");
    if (!(at && at[0] == '0')) {
        const int vdt = I.mixed ? GANIT_F32 : GANIT_F64, mdt = I.mixed ? I.low_mdt : GANIT_F64;
        std::vector<double> x0(I.cols), y0(I.rows, 0.0);
        for (int k = 0; k < K; ++k)
            for (int j = 0; j < I.s[k].n; ++j) x0[I.col0[k] + j] = std::min(std::max(0.0, I.s[k].l[j]), I.s[k].u[j]);
        chk(ganit_hip_batch_load(I.bt, vdt, x0.data(), y0.data(), I.w0.data()));
        chk(ganit_hip_batch_autotune(I.bt, mdt, vdt, 20, I.autotune_ms, &I.autotune_pick));
    }
}

void BatchSolver::set_bounds(int k, const std::vector<double>& l, const std::vector<double>& u) {
    Impl& I = *p_;
    if (k < 0 || k >= size()) throw std::runtime_error("batch instance index out of range");
    ScaledLP& s = I.s[k];
    if ((int)l.size() != s.n || (int)u.size() != s.n) throw std::runtime_error("bounds must have n entries");
    for (int j = 0; j < s.n; ++j)
        if (std::isnan(l[j]) || std::isnan(u[j]) || l[j] > u[j])
            throw std::runtime_error("variable bounds must satisfy l <= u and not be NaN");
    // scale_lp's arithmetic for the variable bounds (divide by the column factor, then by bscale); nothing else in
    // the scaled problem depends on them (bscale and the initial weight come from the row bounds)
    for (int j = 0; j < s.n; ++j) {
        s.l[j] = l[j] / s.C[j] / s.bscale;
        s.u[j] = u[j] / s.C[j] / s.bscale;
    }
    if (I.bt) chk(ganit_hip_batch_set_bounds(I.bt, k, s.l.data(), s.u.data()));
    I.solved = false;  // the last run's solution is no longer that of these instances
}

void BatchSolver::set_start(int k, const std::vector<double>& x, const std::vector<double>& y, double w) {
    Impl& I = *p_;
    if (k < 0 || k >= size()) throw std::runtime_error("batch instance index out of range");
    const ScaledLP& s = I.s[k];
    if ((int)x.size() != s.n || (int)y.size() != s.m) throw std::runtime_error("start point has the wrong size");
    const double sign = s.maximize ? -1.0 : 1.0;
    // the inverse of the solution's unscaling x = x' C bscale, y = sign y' R cscale (pdhg_impl.hpp:313-314)
    I.xw[k].resize(s.n);
    I.yw[k].resize(s.m);
    for (int j = 0; j < s.n; ++j) I.xw[k][j] = x[j] / s.C[j] / s.bscale;
    for (int i = 0; i < s.m; ++i) I.yw[k][i] = sign * y[i] / s.R[i] / s.cscale;
    I.ww[k] = w;
    I.warm[k] = 1;
}

void BatchSolver::clear_start(int k) {
    Impl& I = *p_;
    if (k < 0 || k >= size()) throw std::runtime_error("batch instance index out of range");
    I.warm[k] = 0;
}

void BatchSolver::set_options(const Options& o) {
    Impl& I = *p_;
    const Options& c = I.opt;
    if (o.precision != c.precision || o.values != c.values || o.device != c.device || o.ruiz_iters != c.ruiz_iters ||
        o.pock_chambolle != c.pock_chambolle)
        throw std::runtime_error("a batch keeps the precision, matrix values, scaling and device it was built with");
    if (o.check_every < 2 || o.check_every % 2) throw std::runtime_error("batched check_every must be even");
    I.opt = o;
}

std::vector<double> BatchSolver::final_weights() const { return p_->w_final; }

std::vector<AdjointResult> BatchSolver::adjoint(const std::vector<double>& gx, const AdjointOptions& ao) {
    Impl& I = *p_;
    const int K = size();
    if (!I.solved) throw std::runtime_error("adjoint needs a completed run()");
    if ((int)gx.size() != I.cols) throw std::runtime_error("adjoint: dL/dx has the wrong length");
    if (!ao.gmres) throw std::runtime_error("the batched adjoint runs GMRES only");
    if (ao.fp32 && !I.mixed) throw std::runtime_error("FP32 adjoint needs a mixed or dynamic precision batch");
    const auto t0 = Clock::now();
    // per instance, as GpuSolver::adjoint: the steps of its last phase (pdhg_impl.hpp:87-91), dL/dx' = dL/dx C bscale,
    // and the bound-activity tolerance of where its iterate came from
    std::vector<double> g(I.cols), tau(K), sigma(K), mtol(K);
    for (int k = 0; k < K; ++k) {
        const ScaledLP& s = I.s[k];
        const double eta = 0.998 / I.norm_a[k];
        const double w = I.w_final[k] > 0 && std::isfinite(I.w_final[k]) ? I.w_final[k] : I.w0[k];
        sigma[k] = eta * w;
        tau[k] = eta / w;
        if (s.has_q()) tau[k] = std::min(tau[k], 0.998 / (0.5 * s.normQ + sigma[k] * I.norm_a[k] * I.norm_a[k]));
        for (int j = 0; j < s.n; ++j) g[I.col0[k] + j] = gx[I.col0[k] + j] * s.C[j] * s.bscale;
        mtol[k] = ao.mask_tol >= 0 ? ao.mask_tol : (I.final_fp32[k] ? 1e-6 : 1e-12);
    }
    std::vector<double> p(I.cols), mu(I.rows), info(4 * (size_t)K), xs(I.cols);
    const int mdt = ao.fp32 ? I.low_mdt : GANIT_F64, vdt = ao.fp32 ? GANIT_F32 : GANIT_F64;
    const double params[3] = {(double)ao.max_iter, ao.tol, (double)ao.restart};
    chk(ganit_hip_batch_adjoint(I.bt, mdt, vdt, g.data(), tau.data(), sigma.data(), mtol.data(), params, p.data(),
                                mu.data(), info.data()));
    chk(ganit_hip_batch_store(I.bt, GANIT_F64, xs.data(), nullptr));  // the scaled solution (dQ_jj needs x*)
    const double secs = since(t0);
    std::vector<AdjointResult> out(K);
    for (int k = 0; k < K; ++k) {
        const ScaledLP& s = I.s[k];
        AdjointResult& r = out[k];
        const double sign = s.maximize ? -1.0 : 1.0;
        const int c0 = I.col0[k], r0 = I.row0[k];
        r.dc.resize(s.n);
        r.dq.resize(s.n);
        r.db.resize(s.m);
        // back to the original data, as GpuSolver::adjoint (scaling.cpp:36-129)
        for (int j = 0; j < s.n; ++j) {
            const double dcs = -tau[k] * p[c0 + j];
            r.dc[j] = sign * dcs * s.C[j] / s.cscale;
            r.dq[j] = sign * dcs * xs[c0 + j] * s.C[j] * s.C[j] * s.bscale / s.cscale;
        }
        for (int i = 0; i < s.m; ++i) r.db[i] = sigma[k] * mu[r0 + i] * s.R[i] / s.bscale;
        r.iterations = (long)info[4 * k];
        r.rel_change = info[4 * k + 1];
        r.free_vars = (int)info[4 * k + 2];
        r.active_rows = (int)info[4 * k + 3];
        r.seconds = secs;
    }
    return out;
}

std::vector<Result> BatchSolver::run() {
    Impl& I = *p_;
    const Options& opt = I.opt;
    const int K = size();
    const auto t_solve = Clock::now();
    upload();
    ganit_hip_batch* bt = I.bt;
    BatchStats& S = I.stats;
    S = BatchStats{};
    S.autotune_pick = I.autotune_pick;
    std::copy(I.autotune_ms, I.autotune_ms + 8, S.autotune_ms);
    int64_t dims[6];
    chk(ganit_hip_batch_dims(bt, dims));
    S.rows = dims[1]; S.cols = dims[2]; S.nnz = dims[3]; S.blocks_a = dims[4]; S.blocks_at = dims[5];

    // each instance's start: proj(0), y = 0 and its initial weight (pdhg_impl.hpp:253-262), or its warm start
    // projected onto its current bounds
    std::vector<double> x0(I.cols), y0(I.rows, 0.0), w(I.w0);
    for (int k = 0; k < K; ++k) {
        const ScaledLP& s = I.s[k];
        for (int j = 0; j < s.n; ++j)
            x0[I.col0[k] + j] = std::min(std::max(I.warm[k] ? I.xw[k][j] : 0.0, s.l[j]), s.u[j]);
        if (I.warm[k]) {
            std::copy(I.yw[k].begin(), I.yw[k].end(), y0.begin() + I.row0[k]);
            if (I.ww[k] > 0 && std::isfinite(I.ww[k])) w[k] = I.ww[k];
        }
    }
    const int first_vdt = I.mixed ? GANIT_F32 : GANIT_F64;
    chk(ganit_hip_batch_load(bt, first_vdt, x0.data(), y0.data(), w.data()));
    int64_t tail0 = 0;
    chk(ganit_hip_batch_tail(bt, (opt.tail_lanes ? 1 : 0) | (opt.tail_fp64 ? 2 : 0), &tail0));
    S.upload_seconds = since(t_solve);

    std::vector<int> ints(5 * K);
    std::vector<double> dbl(13 * K);
    auto fetch = [&]() { chk(ganit_hip_batch_results(bt, ints.data(), dbl.data())); };
    std::vector<int> done(K, 0), sel(K, 1);
    std::vector<long> l1_it(K, 0), l2_it(K, 0);
    int last_vdt = first_vdt;

    auto phase = [&](int slot, int mdt, int vdt, bool stall_exit, int stall_checks) {
        const auto t0 = Clock::now();
        const double left = std::max(0.0, opt.time_limit - since(t_solve));
        const double params[8] = {opt.tol, opt.infeas_tol, (double)opt.max_iter, (double)opt.check_every,
                                  (double)stall_checks, stall_exit ? 1.0 : 0.0, left, opt.eval_fp64 ? 1.0 : 0.0};
        if (last_vdt != vdt) chk(ganit_hip_batch_handover(bt, last_vdt, vdt));
        else chk(ganit_hip_batch_handover(bt, vdt, vdt));  // fresh running sums
        last_vdt = vdt;
        int running = 0;
        chk(ganit_hip_batch_phase(bt, mdt, vdt, params, &running));
        S.level_seconds[slot] += since(t0);
        S.level_instances[slot] += (int)std::count(sel.begin(), sel.end(), 1);
    };
    // FP64 certification of every instance's current result (pdhg_evaluate<Bd>)
    auto certify = [&]() {
        const auto t0 = Clock::now();
        if (last_vdt != GANIT_F64) {
            chk(ganit_hip_batch_handover(bt, last_vdt, GANIT_F64));
            last_vdt = GANIT_F64;
        }
        chk(ganit_hip_batch_certify(bt, GANIT_F64));
        fetch();
        S.certify_seconds += since(t0);
    };
    // the ladder's rule for leaving after a reduced-precision level (pdhg_impl.hpp:283-289)
    auto settle = [&]() {
        int left = 0;
        for (int k = 0; k < K; ++k) {
            if (done[k]) { sel[k] = 0; continue; }
            const int e = ints[5 * k];
            const double* k64 = &dbl[13 * k + 6];
            done[k] = (e == kConverged && converged6(k64, opt.tol)) || e == kTimeLimit || e == kIterLimit ||
                      e == kInfeasible;
            sel[k] = done[k] ? 0 : 1;
            left += sel[k];
        }
        return left;
    };

    if (I.mixed) {
        phase(0, I.low_mdt, GANIT_F32, true, opt.stall_checks);
        certify();
        for (int k = 0; k < K; ++k) l1_it[k] = ints[5 * k + 1];
        int left = settle();
        if (left > 0 && opt.precision == Precision::Dynamic) {
            chk(ganit_hip_batch_select(bt, sel.data()));
            phase(1, I.low_mdt, GANIT_F64, true, opt.stall_checks_l2);
            certify();
            for (int k = 0; k < K; ++k) l2_it[k] = sel[k] ? ints[5 * k + 1] - l1_it[k] : 0;
            left = settle();
        }
        if (left > 0) {
            chk(ganit_hip_batch_select(bt, sel.data()));
            phase(2, GANIT_F64, GANIT_F64, false, 0);
        }
    } else {
        phase(2, GANIT_F64, GANIT_F64, false, 0);
    }
    certify();  // final FP64 certification (pdhg_impl.hpp:300-303)

    std::vector<double> xs(I.cols), ys(I.rows);
    chk(ganit_hip_batch_store(bt, GANIT_F64, xs.data(), ys.data()));
    const double solve_s = since(t_solve);
    S.solve_seconds = solve_s;
    int64_t tail1 = 0;
    chk(ganit_hip_batch_tail(bt, -1, &tail1));
    S.tail_iterations = (long)(tail1 - tail0);
    std::vector<Result> out(K);
    I.w_final.resize(K);
    for (int k = 0; k < K; ++k) I.w_final[k] = dbl[13 * k + 12];
    I.final_fp32.assign(K, 0);
    for (int k = 0; k < K; ++k) I.final_fp32[k] = I.mixed && l1_it[k] == ints[5 * k + 1];
    I.solved = true;
    for (int k = 0; k < K; ++k) {
        const ScaledLP& s = I.s[k];
        Result& r = out[k];
        r.device = I.device;
        r.norm_a = I.norm_a[k];
        r.setup_seconds = I.setup_seconds;
        r.solve_seconds = solve_s;
        r.upload_seconds = S.upload_seconds;
        for (int q = 0; q < 3; ++q) r.iterate_seconds[q] = S.level_seconds[q];
        const double* k64 = &dbl[13 * k + 6];
        Status st = status_of(ints[5 * k]);
        if (st == Status::Optimal && !converged6(k64, opt.tol * 1.0000001)) st = Status::NumericalError;
        r.status = st;
        r.iterations = ints[5 * k + 1];
        r.fp32_iterations = I.mixed ? l1_it[k] : 0;
        r.l2_iterations = l2_it[k];
        r.restarts = ints[5 * k + 2];
        r.checks = r.iterations / opt.check_every;
        const double sign = s.maximize ? -1.0 : 1.0;
        r.x.resize(s.n);
        r.y.resize(s.m);
        for (int j = 0; j < s.n; ++j) r.x[j] = xs[I.col0[k] + j] * s.C[j] * s.bscale;  // pdhg_impl.hpp:313-314
        for (int i = 0; i < s.m; ++i) r.y[i] = sign * ys[I.row0[k] + i] * s.R[i] * s.cscale;
        r.pobj = sign * k64[0];
        r.dobj = sign * k64[1];
        r.rel_primal = k64[2];
        r.rel_dual = k64[3];
        r.rel_gap = k64[4];
        S.max_iterations = std::max(S.max_iterations, r.iterations);
        S.total_iterations += r.iterations;
    }
    if (opt.verbose) {
        int opt_n = 0;
        for (const Result& r : out) opt_n += r.status == Status::Optimal;
        std::printf("batch of %d LPs (%lld rows, %lld cols, %lld nnz): %d optimal, max %ld iterations, "
                    "solve %.3fs (L1 %.3fs, L2 %.3fs, L3 %.3fs, certification %.3fs)\n",
                    K, (long long)S.rows, (long long)S.cols, (long long)S.nnz, opt_n, S.max_iterations, solve_s,
                    S.level_seconds[0], S.level_seconds[1], S.level_seconds[2], S.certify_seconds);
    }
    return out;
}

}  // namespace ganit
