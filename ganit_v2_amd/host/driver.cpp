// Restarted-average PDHG driver over the device-resident C ABI of libganit_hip.so.
//
// The control logic is the reference's (src/pdhg_impl.hpp:67-328): the same step sizes, KKT formulas,
// convergence / limit / infeasibility / stall tests in the same order, the same adaptive-restart rule and
// primal-weight update, the same precision ladder and the same FP64 certification.  What changes is where
// the work happens: the scaled problem is uploaded once with every precision it will need; an iteration is
// two fused kernels replayed from a HIP graph; a check is three kernels and one host synchronisation that
// return every sum the decisions need; and precision hand-over and FP64 certification convert the iterate
// on the device instead of round-tripping it through host doubles.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <limits>
#include <stdexcept>
#include <string>

#include "../../include/ganit_hip.h"
#include "pdhg.hpp"
#include "scaling.hpp"

namespace ganit {

const char* to_string(Status s) {
    switch (s) {
        case Status::Optimal: return "OPTIMAL";
        case Status::PrimalInfeasible: return "PRIMAL_INFEASIBLE";
        case Status::IterationLimit: return "ITERATION_LIMIT";
        case Status::TimeLimit: return "TIME_LIMIT";
        case Status::DualInfeasible: return "DUAL_INFEASIBLE";
        default: return "NUMERICAL_ERROR";
    }
}

namespace {

void chk(int rc) {
    if (rc != GANIT_OK) throw std::runtime_error(std::string("HIP backend: ") + ganit_hip_last_error());
}

// Setup::Auto: the GPU set-up pays for its upload from about a million nonzeros on; its kernels give each row
// and each column to one thread, so a matrix with a row or column longer than kSetupRow stays on the host.
constexpr int64_t kSetupNnz = 1000000;
constexpr int kSetupRow = 4096;

bool place_setup_on_device(const LP& lp, Setup where) {
    if (where != Setup::Auto) return where == Setup::Device;
    if (lp.A.nnz() < kSetupNnz) return false;
    for (int i = 0; i < lp.m; ++i)
        if (lp.A.ptr[i + 1] - lp.A.ptr[i] > kSetupRow) return false;
    std::vector<int> cols(lp.n, 0);
    for (const int j : lp.A.idx)
        if (++cols[j] > kSetupRow) return false;
    return true;
}

struct KKT {  // reference pdhg_impl.hpp:36-42
    double pobj = 0, dobj = 0, rel_p = 0, rel_d = 0, rel_gap = 0, err = 0;
    bool finite() const {
        return std::isfinite(pobj) && std::isfinite(dobj) && std::isfinite(rel_p) && std::isfinite(rel_d);
    }
    bool converged(double tol) const { return rel_p <= tol && rel_d <= tol && rel_gap <= tol; }
};

using Clock = std::chrono::steady_clock;

struct Shared {
    const ScaledLP& s;
    const Options& opt;
    double eta, normA, obj_scale;
    Clock::time_point t_solve;
    double last_log = -1e9;
    double elapsed() const { return std::chrono::duration<double>(Clock::now() - t_solve).count(); }
};

struct State {
    double restart_err = -1.0;
    double w = 1.0;
    long it = 0;
    int restarts = 0;
    long checks = 0;
    KKT kkt;
};

enum class Exit { Converged, Stalled, IterationLimit, TimeLimit, Infeasible, DualInfeasible, Numerical };

struct Level {
    int mdt, vdt;
    const char* tag;
};

// KKT from the raw sums of one point (the eval lambda, reference pdhg_impl.hpp:92-109).
KKT kkt_from(const Shared& sh, double rd2, double vobj, double cx, double xqx, double rp2, double robj) {
    const ScaledLP& s = sh.s;
    KKT k;
    k.pobj = (cx + 0.5 * xqx) * sh.obj_scale + s.obj_const;
    k.dobj = (robj + vobj - 0.5 * xqx) * sh.obj_scale + s.obj_const;
    k.rel_p = std::sqrt(rp2) / (1.0 + s.bnorm);
    k.rel_d = std::sqrt(rd2) / (1.0 + s.cnorm);
    k.rel_gap = std::fabs(k.pobj - k.dobj) / (1.0 + std::fabs(k.pobj) + std::fabs(k.dobj));
    k.err = std::sqrt(k.rel_p * k.rel_p + k.rel_d * k.rel_d + k.rel_gap * k.rel_gap);
    return k;
}
KKT kkt_current(const Shared& sh, const double* o) { return kkt_from(sh, o[0], o[1], o[2], o[3], o[12], o[13]); }
KKT kkt_average(const Shared& sh, const double* o) { return kkt_from(sh, o[4], o[5], o[6], o[7], o[14], o[15]); }

KKT evaluate(Shared& sh, ganit_hip_pdhg* pd, int mdt, int vdt, int flags = 0);

// Options::polish (extension; the reference lists crossover / polish as planned, README.md:195): project the FP64
// iterate of an LP onto the face named by its active set (ganit_hip_pdhg_polish), for a few activity tolerances
// from loose to tight, and keep the first result whose FP64 KKT errors are all <= polish_tol and whose variables
// lie within their bounds to that tolerance; otherwise the PDHG point stays.  (xs, ys, k) are the point in and
// out; returns whether a polished point was kept.
bool polish(Shared& sh, ganit_hip_pdhg* pd, std::vector<double>& xs, std::vector<double>& ys, KKT& k) {
    const ScaledLP& s = sh.s;
    const double ptol = sh.opt.polish_tol;
    std::vector<double> xp(s.n), yp(s.m);
    for (const double act_tol : {1e-6, 1e-5, 1e-7, 1e-4, 1e-8}) {
        // up to three rounds: each re-reads the active set from the previous round's point while that improves
        double last_err = k.err;
        for (int round = 0; round < 3; ++round) {
            const double params[3] = {act_tol, 1e-14, 20000};
            double info[6];
            chk(ganit_hip_pdhg_polish(pd, params, info));
            const KKT kp = evaluate(sh, pd, GANIT_F64, GANIT_F64, 0);
            chk(ganit_hip_pdhg_store(pd, GANIT_F64, xp.data(), yp.data(), nullptr, nullptr));
            double viol = 0;  // bound violation of the polished x, relative like the activity test
            for (int j = 0; j < s.n; ++j) {
                if (std::isfinite(s.l[j])) viol = std::max(viol, (s.l[j] - xp[j]) / (1.0 + std::fabs(s.l[j])));
                if (std::isfinite(s.u[j])) viol = std::max(viol, (xp[j] - s.u[j]) / (1.0 + std::fabs(s.u[j])));
            }
            if (sh.opt.verbose)
                std::printf("  polish at activity tolerance %.0e, round %d: %d free, %d active rows, LSQR %d / %d "
                            "products, rel p %.2e d %.2e gap %.2e, bound violation %.2e\n",
                            act_tol, round, (int)info[0], (int)info[1], (int)info[2], (int)info[4], kp.rel_p,
                            kp.rel_d, kp.rel_gap, viol);
            if (kp.finite() && kp.converged(ptol) && viol <= ptol) {
                xs = xp;
                ys = yp;
                k = kp;
                return true;
            }
            if (!kp.finite() || viol > ptol || !(kp.err < last_err)) break;
            last_err = kp.err;
        }
        chk(ganit_hip_pdhg_load(pd, GANIT_F64, xs.data(), ys.data(), nullptr, nullptr));  // back to the PDHG point
    }
    return false;
}

// KKT of the current iterate of working set `vdt` evaluated with matrix values `mdt` (extra check flags,
// e.g. GANIT_CHECK_FP64, in `flags`).
KKT evaluate(Shared& sh, ganit_hip_pdhg* pd, int mdt, int vdt, int flags) {
    double o[GANIT_CHECK_NOUT];
    chk(ganit_hip_pdhg_check(pd, mdt, vdt, flags, 0.0, o));
    return kkt_current(sh, o);
}

// One precision phase (reference pdhg_phase, pdhg_impl.hpp:67-199).  On entry working set lv.vdt holds the
// iterate and restart point; on exit it holds the phase's best point (average adopted if that was best).
Exit phase(Shared& sh, ganit_hip_pdhg* pd, const Level& lv, State& st, double tol, bool stall_exit,
           int stall_checks) {
    const Options& opt = sh.opt;
    const ScaledLP& s = sh.s;
    const bool hasq = s.has_q();
    const int ce = opt.check_every;
    double w = st.w;
    chk(ganit_hip_pdhg_handover(pd, lv.vdt, lv.vdt));  // fresh running sums
    // checks of a reduced-precision level: at the level's precision (the reference's rule) or in FP64
    const bool e64 = opt.eval_fp64 && (lv.mdt != GANIT_F64 || lv.vdt != GANIT_F64);
    // Kahan-compensated FP32 check reductions apply to L1's own-precision checks (as the oracle's kahan option)
    const bool kahan = opt.kahan && !e64 && lv.vdt == GANIT_F32;
    const int ev_mdt = e64 ? GANIT_F64 : lv.mdt;
    const int ev_flags = (e64 ? GANIT_CHECK_FP64 : 0) | (kahan ? GANIT_CHECK_KAHAN : 0);

    KKT k_restart = evaluate(sh, pd, ev_mdt, lv.vdt, ev_flags);
    if (st.restart_err > 0) k_restart.err = std::max(k_restart.err, st.restart_err);
    double k_prev_cand = std::numeric_limits<double>::infinity();
    long it_since = 0;
    double best_err = k_restart.err;
    int no_progress = 0;
    Exit why = Exit::IterationLimit;
    KKT best;
    bool best_is_avg = false;
    double o[GANIT_CHECK_NOUT];

    while (true) {
        double sigma = sh.eta * w, tau = sh.eta / w;  // step sizes (pdhg_impl.hpp:87-91)
        if (hasq) tau = std::min(tau, 0.998 / (0.5 * s.normQ + sigma * sh.normA * sh.normA));
        // Iterations are only inspected at multiples of check_every (pdhg_impl.hpp:142), so a block of
        // check_every iterations runs without host involvement.
        const long to_check = ce - (st.it % ce);
        chk(ganit_hip_pdhg_iterate(pd, lv.mdt, lv.vdt, (int)to_check, tau, sigma));
        st.it += to_check;
        it_since += to_check;

        const bool rays = it_since >= 4L * ce;
        const bool dray_test = rays && opt.detect_unbounded && !hasq;
        const int flags = GANIT_CHECK_AVG | (rays ? GANIT_CHECK_RAY : 0) | (dray_test ? GANIT_CHECK_DRAY : 0);
        chk(ganit_hip_pdhg_check(pd, ev_mdt, lv.vdt, flags | ev_flags, 1.0 / static_cast<double>(it_since), o));
        ++st.checks;
        const KKT kc = kkt_current(sh, o);
        const KKT ka = kkt_average(sh, o);
        const double t = sh.elapsed();

        if (!kc.finite() && !ka.finite()) { why = Exit::Numerical; best = kc; break; }
        const bool avg_better = ka.finite() && (!kc.finite() || ka.err < kc.err);
        const KKT kcand = avg_better ? ka : kc;
        if (opt.verbose && (t - sh.last_log >= opt.log_interval)) {
            std::printf("%9ld %8.2fs %4s %15.8e %15.8e %9.2e %9.2e %9.2e %9.2e\n", st.it, t, lv.tag, kcand.pobj,
                        kcand.dobj, kcand.rel_p, kcand.rel_d, kcand.rel_gap, w);
            sh.last_log = t;
        }
        if (kc.converged(tol)) { why = Exit::Converged; best = kc; best_is_avg = false; break; }
        if (ka.converged(tol)) { why = Exit::Converged; best = ka; best_is_avg = true; break; }
        if (st.it >= opt.max_iter) { why = Exit::IterationLimit; best = kcand; best_is_avg = avg_better; break; }
        if (t >= opt.time_limit) { why = Exit::TimeLimit; best = kcand; best_is_avg = avg_better; break; }
        if (flags & GANIT_CHECK_RAY) {  // Farkas-type certificate from y - yr (pdhg_impl.hpp:110-118)
            const double dray = (o[16] + o[9]) * sh.obj_scale;
            if (dray > 0 && std::sqrt(o[8]) * (1.0 + s.bnorm) / dray <= opt.infeas_tol) {
                why = Exit::Infeasible; best = kcand; best_is_avg = avg_better; break;
            }
        }
        if (dray_test) {
            // The mirror test on the primal ray d = x - xr (an extension: the reference has no dual-infeasibility
            // test): c'd < 0 while A d and d violate the recession cones of the row and variable bounds by a
            // small multiple of -c'd.  o[19] is in row units; o[20] and o[21] carry a factor 1/obj_scale.
            const double cd = o[20] * sh.obj_scale;
            const double viol = std::sqrt(o[19] + o[21] * sh.obj_scale * sh.obj_scale);
            if (cd < 0 && viol * (1.0 + s.cnorm) / -cd <= opt.infeas_tol) {
                why = Exit::DualInfeasible; best = kcand; best_is_avg = avg_better; break;
            }
        }
        if (kcand.err < 0.8 * best_err) { best_err = kcand.err; no_progress = 0; }
        else if (++no_progress >= stall_checks && stall_exit) {
            why = Exit::Stalled; best = kcand; best_is_avg = avg_better; break;
        }
        // adaptive restart (pdhg_impl.hpp:172-189)
        const bool sufficient = kcand.err <= 0.2 * k_restart.err;
        const bool necessary = kcand.err <= 0.8 * k_restart.err && kcand.err > k_prev_cand;
        const bool artificial = it_since >= 0.36 * static_cast<double>(st.it);
        k_prev_cand = kcand.err;
        if (sufficient || necessary || artificial) {
            const double dx = std::sqrt(avg_better ? o[11] : o[10]);
            const double dy = std::sqrt(avg_better ? o[18] : o[17]);
            if (dx > 1e-10 && dy > 1e-10) w = std::exp(0.5 * std::log(dy / dx) + 0.5 * std::log(w));
            chk(ganit_hip_pdhg_restart(pd, lv.vdt, avg_better ? 1 : 0));
            it_since = 0;
            k_restart = kcand;
            k_prev_cand = std::numeric_limits<double>::infinity();
            ++st.restarts;
        }
    }
    if (best_is_avg) chk(ganit_hip_pdhg_adopt_average(pd, lv.vdt));
    st.restart_err = k_restart.err;
    st.w = w;
    st.kkt = best;
    return why;
}

Status status_from(Exit e) {
    switch (e) {
        case Exit::Converged: return Status::Optimal;
        case Exit::Infeasible: return Status::PrimalInfeasible;
        case Exit::DualInfeasible: return Status::DualInfeasible;
        case Exit::TimeLimit: return Status::TimeLimit;
        case Exit::Numerical: return Status::NumericalError;
        default: return Status::IterationLimit;
    }
}

}  // namespace

struct GpuSolver::Impl {
    const Options opt;
    ganit_hip_ctx* ctx = nullptr;
    ganit_hip_pdhg* pd = nullptr;
    std::string device;
    ScaledLP s;
    double normA = 0, setup_seconds = 0;
    bool setup_on_device = false;
    bool polished = false;
    bool mixed = false;
    int low_mdt = GANIT_F32;
    std::vector<double> x0, y0;
    double w0 = 1.0;
    // the last run's final primal weight and scaled solution (the adjoint linearises there)
    bool solved = false;
    bool final_fp32 = false;  // the last run's solution came out of an FP32-vector level
    double w_final = 1.0;
    std::vector<double> xs_final;

    explicit Impl(const Options& o) : opt(o) {}
    ~Impl() {
        if (pd) ganit_hip_pdhg_destroy(pd);
        if (ctx) ganit_hip_ctx_destroy(ctx);
    }
};

GpuSolver::GpuSolver(const LP& lp, const Options& opt) : p_(new Impl(opt)) {
    Impl& I = *p_;
    if (opt.check_every < 1) throw std::runtime_error("check_every must be positive");
    chk(ganit_hip_ctx_create(opt.device, &I.ctx));
    char name[256] = {0};
    chk(ganit_hip_ctx_device_info(I.ctx, name, sizeof(name), nullptr));
    I.device = std::string("gpu:") + name;
    const auto t0 = Clock::now();
    I.setup_on_device = !(opt.scale_q && lp.has_q()) && place_setup_on_device(lp, opt.setup);
    if (I.setup_on_device) {
        I.s = scale_lp_device(lp, opt.ruiz_iters, opt.pock_chambolle, I.ctx, &I.normA);
    } else {
        I.s = scale_lp(lp, opt.ruiz_iters, opt.pock_chambolle, true, opt.scale_q);
        I.normA = estimate_norm(I.s);
    }
    I.x0.resize(I.s.n);
    for (int j = 0; j < I.s.n; ++j) I.x0[j] = std::min(std::max(0.0, I.s.l[j]), I.s.u[j]);
    I.y0.assign(I.s.m, 0.0);
    const double cn = norm2(I.s.c), bn = finite_bound_norm(I.s.lo, I.s.hi);
    I.w0 = (cn > 1e-10 && bn > 1e-10) ? cn / bn : 1.0;
    I.mixed = opt.precision != Precision::Double;
    I.low_mdt = opt.values == Values::BF16 ? GANIT_BF16 : GANIT_F32;
    I.setup_seconds = std::chrono::duration<double>(Clock::now() - t0).count();
}

GpuSolver::~GpuSolver() = default;

const ScaledLP& GpuSolver::scaled() const { return p_->s; }
double GpuSolver::norm_a() const { return p_->normA; }
double GpuSolver::setup_seconds() const { return p_->setup_seconds; }
bool GpuSolver::setup_on_device() const { return p_->setup_on_device; }
bool GpuSolver::last_polished() const { return p_->polished; }

void GpuSolver::upload() {
    Impl& I = *p_;
    if (I.pd) return;
    const ScaledLP& s = I.s;
    ganit_hip_lp_desc d{};
    d.m = s.m; d.n = s.n;
    d.a_ptr = s.A.ptr.data(); d.a_idx = s.A.idx.data(); d.a_val = s.A.val.data();
    d.at_ptr = s.AT.ptr.data(); d.at_idx = s.AT.idx.data(); d.at_val = s.AT.val.data();
    if (s.has_q()) { d.q_ptr = s.Q.ptr.data(); d.q_idx = s.Q.idx.data(); d.q_val = s.Q.val.data(); }
    d.c = s.c.data(); d.l = s.l.data(); d.u = s.u.data(); d.lo = s.lo.data(); d.hi = s.hi.data();
    d.rres = s.Rres.data(); d.cres = s.Cres.data();
    // every precision the ladder will use is resident at once, plus FP64 for certification
    const unsigned mmask = (1u << GANIT_F64) | (I.mixed ? (1u << I.low_mdt) : 0u) |
                           (I.mixed && I.opt.values == Values::BF16_F32 ? (1u << GANIT_BF16) : 0u);
    const unsigned vmask = (1u << GANIT_F64) | (I.mixed ? (1u << GANIT_F32) : 0u);
    chk(ganit_hip_pdhg_create(I.ctx, &d, mmask, vmask, &I.pd));
}

void GpuSolver::kernel_times(int level, int reps, double* ms) {
    Impl& I = *p_;
    upload();
    const int mdt = level == 2 ? GANIT_F64 : I.low_mdt, vdt = level == 0 ? GANIT_F32 : GANIT_F64;
    if (level != 2 && !I.mixed) throw std::runtime_error("reduced-precision level not resident (precision double)");
    const double sigma = (0.998 / I.normA) * I.w0, tau = (0.998 / I.normA) / I.w0;
    chk(ganit_hip_pdhg_load(I.pd, vdt, I.x0.data(), I.y0.data(), nullptr, nullptr));
    chk(ganit_hip_pdhg_time_kernels(I.pd, mdt, vdt, reps, tau, sigma, ms));
}

int GpuSolver::index_bytes(int which) {
    upload();
    int b = 4;
    chk(ganit_hip_pdhg_index_bytes(p_->pd, which, &b));
    return b;
}

int GpuSolver::variant(int which) {
    upload();
    int lanes = 0;
    chk(ganit_hip_pdhg_plan(p_->pd, which, &lanes, nullptr));
    return lanes;
}

Result GpuSolver::run() {
    Impl& I = *p_;
    const Options& opt = I.opt;
    const ScaledLP& s = I.s;
    Result res;
    res.device = I.device;
    res.norm_a = I.normA;
    res.setup_seconds = I.setup_seconds;
    res.setup_on_device = I.setup_on_device;
    Shared sh{s, opt, 0.998 / I.normA, I.normA, s.cscale * s.bscale, Clock::now()};
    State st;
    st.w = I.w0;
    const bool mixed = I.mixed;
    const int low_mdt = I.low_mdt;
    const char* pname = opt.precision == Precision::Dynamic ? "dynamic L1->L2->L3"
                      : opt.precision == Precision::Mixed ? "mixed L1->L3" : "FP64";
    if (opt.verbose) {
        std::printf("GANIT PDHG | device %s | %s | precision %s%s | rows %d cols %d nnz %lld | ||A||~%.3e | tol %.1e\n",
                    res.device.c_str(), s.has_q() ? "QP" : "LP", pname,
                    mixed ? (opt.values != Values::F32
                                 ? (opt.values == Values::BF16 ? (opt.eval_fp64 ? " (BF16 values, FP64 checks)" : " (BF16 values)")
                                                               : (opt.eval_fp64 ? " (BF16 then FP32 values, FP64 checks)"
                                                                                : " (BF16 then FP32 values)"))
                                                        : (opt.eval_fp64 ? " (FP64 checks)" : ""))
                          : "",
                    s.m, s.n, (long long)s.A.nnz(),
                    I.normA, opt.tol);
        if (s.has_q()) std::printf("Hessian nnz %lld | ||Q||~%.3e\n", (long long)s.Q.nnz(), s.normQ);
        std::printf("%9s %9s %4s %15s %15s %9s %9s %9s %9s\n", "iter", "time", "prec", "primal obj", "dual obj",
                    "rel_p", "rel_d", "rel_gap", "weight");
    }
    sh.t_solve = Clock::now();
    upload();  // no-op when the caller uploaded beforehand
    ganit_hip_pdhg* pd = I.pd;
    const int first_vdt = mixed ? GANIT_F32 : GANIT_F64;
    chk(ganit_hip_pdhg_load(pd, first_vdt, I.x0.data(), I.y0.data(), nullptr, nullptr));
    res.upload_seconds = sh.elapsed();

    const Level L1{low_mdt, GANIT_F32, "L1"}, L2{low_mdt, GANIT_F64, "L2"}, L3{GANIT_F64, GANIT_F64, "L3"};
    int last_vdt = first_vdt;
    int solution_vdt = first_vdt;  // working set of the level that produced the current iterate
    // FP64 KKT of the phase's result (pdhg_evaluate<Bd>): hand the iterate to the FP64 set if needed.
    auto fp64_eval = [&]() {
        if (last_vdt != GANIT_F64) {
            chk(ganit_hip_pdhg_handover(pd, last_vdt, GANIT_F64));
            last_vdt = GANIT_F64;
        }
        return evaluate(sh, pd, GANIT_F64, GANIT_F64);
    };
    auto finished = [&](Exit e, const char* level, const char* next) {
        const KKT k64 = fp64_eval();
        const bool done = (e == Exit::Converged && k64.converged(opt.tol)) || e == Exit::TimeLimit ||
                          e == Exit::IterationLimit || e == Exit::Infeasible || e == Exit::DualInfeasible;
        if (opt.verbose)
            std::printf("  %s phase: %ld iters total, %.3fs, FP64-checked error %.2e -> %s\n", level, st.it,
                        sh.elapsed(), k64.err, done ? "done" : next);
        return done;
    };
    auto timed_phase = [&](int slot, const Level& lv, bool stall_exit, int stall_checks) {
        const double t0 = sh.elapsed();
        const Exit e = phase(sh, pd, lv, st, opt.tol, stall_exit, stall_checks);
        res.iterate_seconds[slot] += sh.elapsed() - t0;
        last_vdt = lv.vdt;
        solution_vdt = lv.vdt;
        return e;
    };

    Exit ex;
    long fp32_iters = 0;
    if (mixed && opt.values == Values::BF16_F32) {
        // BF16 values until they stall, then FP32 values on the same FP32 vectors, then FP64
        const Level L0{GANIT_BF16, GANIT_F32, "L0"};
        ex = timed_phase(0, L0, true, opt.stall_checks);
        bool done = finished(ex, "L0 (BF16 matrix)", "switching to L1");
        if (!done) {
            ex = timed_phase(0, L1, true, opt.stall_checks_l2);
            done = finished(ex, "L1 (FP32)", "switching to L3");
        }
        fp32_iters = st.it;
        if (!done) ex = timed_phase(2, L3, false, 0);
    } else if (mixed) {
        ex = timed_phase(0, L1, true, opt.stall_checks);
        fp32_iters = st.it;
        bool done = finished(ex, "L1 (FP32)",
                             opt.precision == Precision::Dynamic ? "switching to L2" : "switching to L3");
        if (!done && opt.precision == Precision::Dynamic) {
            ex = timed_phase(1, L2, true, opt.stall_checks_l2);
            res.l2_iterations = st.it - fp32_iters;
            done = finished(ex, "L2 (FP32 matrix, FP64 vectors)", "switching to L3");
        }
        if (!done) ex = timed_phase(2, L3, false, 0);
    } else {
        ex = timed_phase(2, L3, false, 0);
    }
    KKT fin = fp64_eval();  // final FP64 certification (pdhg_impl.hpp:300-303)
    Status status = status_from(ex);
    if (status == Status::Optimal && !fin.converged(opt.tol * 1.0000001)) status = Status::NumericalError;

    std::vector<double> xs(s.n), ys(s.m);
    chk(ganit_hip_pdhg_store(pd, GANIT_F64, xs.data(), ys.data(), nullptr, nullptr));
    if (opt.polish && status == Status::Optimal && !s.has_q()) {
        const double t0 = sh.elapsed();
        res.polished = polish(sh, pd, xs, ys, fin);
        res.polish_seconds = sh.elapsed() - t0;
        if (opt.verbose)
            std::printf("  polish: %s, FP64-checked error %.2e (%.3fs)\n", res.polished ? "vertex kept" : "no vertex",
                        fin.err, res.polish_seconds);
    }
    res.solve_seconds = sh.elapsed();
    I.solved = true;
    I.polished = res.polished;
    I.final_fp32 = solution_vdt == GANIT_F32;
    I.w_final = st.w;
    I.xs_final = xs;
    res.status = status;
    res.iterations = st.it;
    res.fp32_iterations = fp32_iters;
    res.restarts = st.restarts;
    res.checks = st.checks;
    const double sign = s.maximize ? -1.0 : 1.0;
    res.x.resize(s.n);
    res.y.resize(s.m);
    for (int j = 0; j < s.n; ++j) res.x[j] = xs[j] * s.C[j] * s.bscale;  // pdhg_impl.hpp:313-314
    for (int i = 0; i < s.m; ++i) res.y[i] = sign * ys[i] * s.R[i] * s.cscale;
    res.pobj = sign * fin.pobj;
    res.dobj = sign * fin.dobj;
    res.rel_primal = fin.rel_p;
    res.rel_dual = fin.rel_d;
    res.rel_gap = fin.rel_gap;
    if (opt.verbose) {
        std::printf("Status %s | iters %ld (L1 %ld, L2 %ld) | restarts %d | solve %.3fs (setup %.3fs)\n",
                    to_string(status), st.it, fp32_iters, res.l2_iterations, res.restarts, res.solve_seconds,
                    res.setup_seconds);
        std::printf("Primal obj %.10e | Dual obj %.10e | FP64-certified rel errors p %.1e d %.1e gap %.1e\n",
                    res.pobj, res.dobj, fin.rel_p, fin.rel_d, fin.rel_gap);
    }
    return res;
}

AdjointResult GpuSolver::adjoint(const std::vector<double>& gx, const AdjointOptions& ao) {
    Impl& I = *p_;
    const ScaledLP& s = I.s;
    if (!I.solved) throw std::runtime_error("adjoint needs a completed run()");
    if ((int)gx.size() != s.n) throw std::runtime_error("adjoint: dL/dx has the wrong length");
    const bool fp32 = ao.fp32;
    if (fp32 && !I.mixed) throw std::runtime_error("FP32 adjoint needs a mixed or dynamic precision solver");
    const auto t0 = Clock::now();
    // steps of the forward solve's last phase (reference pdhg_impl.hpp:87-91)
    const double eta = 0.998 / I.normA;
    const double sigma = eta * I.w_final;
    double tau = eta / I.w_final;
    if (s.has_q()) tau = std::min(tau, 0.998 / (0.5 * s.normQ + sigma * I.normA * I.normA));
    // x = C x' bscale: dL/dx' = dL/dx * C * bscale
    std::vector<double> g(s.n), p(s.n), mu(s.m);
    for (int j = 0; j < s.n; ++j) g[j] = gx[j] * s.C[j] * s.bscale;
    double info[4] = {0, 0, 0, 0};
    const int mdt = fp32 ? I.low_mdt : GANIT_F64, vdt = fp32 ? GANIT_F32 : GANIT_F64;
    // an iterate that ended in FP32 sits at its bounds only up to FP32 rounding of the bound
    const double mask_tol = ao.mask_tol >= 0 ? ao.mask_tol : (I.final_fp32 ? 1e-6 : 1e-12);
    const double params[5] = {ao.gmres ? 1.0 : 0.0, (double)ao.max_iter, ao.tol,
                              (double)(ao.gmres ? ao.restart : ao.check_every), mask_tol};
    chk(ganit_hip_pdhg_adjoint(I.pd, mdt, vdt, g.data(), tau, sigma, params, p.data(), mu.data(), info));
    AdjointResult r;
    const double sign = s.maximize ? -1.0 : 1.0;
    r.dc.resize(s.n);
    r.dq.resize(s.n);
    r.db.resize(s.m);
    // back to the original data: c' = sign c C / cscale, Q'_jj = sign Q_jj C_j^2 bscale / cscale,
    // row bounds' = bounds R / bscale (scaling.cpp:36-129)
    for (int j = 0; j < s.n; ++j) {
        const double dcs = -tau * p[j];
        r.dc[j] = sign * dcs * s.C[j] / s.cscale;
        r.dq[j] = sign * dcs * I.xs_final[j] * s.C[j] * s.C[j] * s.bscale / s.cscale;
    }
    for (int i = 0; i < s.m; ++i) r.db[i] = sigma * mu[i] * s.R[i] / s.bscale;
    r.iterations = (long)info[0];
    r.rel_change = info[1];
    r.free_vars = (int)info[2];
    r.active_rows = (int)info[3];
    r.seconds = std::chrono::duration<double>(Clock::now() - t0).count();
    return r;
}

void GpuSolver::set_solution(const std::vector<double>& x, const std::vector<double>& y, bool fp32_iterate) {
    Impl& I = *p_;
    const ScaledLP& s = I.s;
    if ((int)x.size() != s.n || (int)y.size() != s.m) throw std::runtime_error("set_solution: wrong lengths");
    if (fp32_iterate && !I.mixed) throw std::runtime_error("an FP32 iterate needs a mixed or dynamic solver");
    upload();
    // the inverse of run()'s unscaling: x = x' C bscale, y = sign y' R cscale
    const double sign = s.maximize ? -1.0 : 1.0;
    std::vector<double> xs(s.n), ys(s.m);
    for (int j = 0; j < s.n; ++j) xs[j] = x[j] / (s.C[j] * s.bscale);
    for (int i = 0; i < s.m; ++i) ys[i] = sign * y[i] / (s.R[i] * s.cscale);
    chk(ganit_hip_pdhg_load(I.pd, GANIT_F64, xs.data(), ys.data(), nullptr, nullptr));
    if (I.mixed) chk(ganit_hip_pdhg_load(I.pd, GANIT_F32, xs.data(), ys.data(), nullptr, nullptr));
    I.solved = true;
    I.final_fp32 = fp32_iterate;
    I.w_final = I.w0;  // the gradient at a fixed point does not depend on the step sizes; any weight serves
    I.xs_final = xs;
}

Result solve_pdhg_gpu(const LP& lp, const Options& opt) {
    GpuSolver solver(lp, opt);
    return solver.run();
}

}  // namespace ganit
