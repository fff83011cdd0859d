// Per-instance decisions of restarted PDHG on the device: one thread applies, after every block of
// check_every iterations, the reference's check logic (src/pdhg_impl.hpp:92-189) to the 19 check sums of its
// instance.  Shared by the batched path (one thread per instance) and the single-instance device loop.
#pragma once
#include <cmath>

#include "internal.hpp"

namespace ganit_hip {

constexpr int kBOut = 19;  // per-instance sums of the batch: the single check's first 19 outputs

enum BatchStatus { kRunning = 0, kConverged = 1, kInfeasible = 2, kIterLimit = 3, kNumerical = 4, kStalled = 5,
                   kTimeLimit = 6 };

struct BatchInst {
    int m, n, row0, col0;
    double eta, norm_a, obj_scale, obj_const, bnorm, cnorm, w0;
    double norm_q;  // ||Q'|| estimate; 0 for an LP
    int hasq;
};

struct BatchState {
    double w, tau, sigma, inv_k;
    double restart_err, prev_cand, best_err;
    double best[6];  // pobj, dobj, rel_p, rel_d, rel_gap, err of the phase's result
    double kkt64[6];
    double cand[6];  // candidate of the latest check (the result if the phase is stopped by the clock)
    int it, it_since, restarts, no_progress, status, best_avg, adopt, active, cand_avg;
};

struct BatchParams {
    double tol, infeas_tol;
    int max_iter, check_every, stall_checks, stall_exit;
    int end_phase;  // 1 (with stall_exit): every running instance leaves the phase at this check, as stalled
};

struct Kkt {
    double pobj, dobj, rel_p, rel_d, rel_gap, err;
    __device__ bool finite() const { return isfinite(pobj) && isfinite(dobj) && isfinite(rel_p) && isfinite(rel_d); }
    __device__ bool converged(double tol) const { return rel_p <= tol && rel_d <= tol && rel_gap <= tol; }
};

__device__ inline Kkt bt_kkt(const BatchInst& in, double rd2, double vobj, double cx, double xqx, double rp2, double robj) {
    Kkt k;  // src/pdhg_impl.hpp:102-107
    k.pobj = (cx + 0.5 * xqx) * in.obj_scale + in.obj_const;
    k.dobj = (robj + vobj - 0.5 * xqx) * in.obj_scale + in.obj_const;
    k.rel_p = sqrt(rp2) / (1.0 + in.bnorm);
    k.rel_d = sqrt(rd2) / (1.0 + in.cnorm);
    k.rel_gap = fabs(k.pobj - k.dobj) / (1.0 + fabs(k.pobj) + fabs(k.dobj));
    k.err = sqrt(k.rel_p * k.rel_p + k.rel_d * k.rel_d + k.rel_gap * k.rel_gap);
    return k;
}

__device__ inline void bt_store(double* dst, const Kkt& k) {
    dst[0] = k.pobj; dst[1] = k.dobj; dst[2] = k.rel_p; dst[3] = k.rel_d; dst[4] = k.rel_gap; dst[5] = k.err;
}

enum DecideMode { kInit = 0, kCheck = 1, kFinal = 2 };

// Step sizes from the primal weight (src/pdhg_impl.hpp:87-91), with the QP cap.
__device__ __forceinline__ void set_steps(const BatchInst& in, BatchState& s) {
    s.sigma = in.eta * s.w;
    s.tau = in.eta / s.w;
    if (in.hasq) s.tau = fmin(s.tau, 0.998 / (0.5 * in.norm_q + s.sigma * in.norm_a * in.norm_a));
}

// One instance's decision on its check sums o[0..18].  kInit: phase start (restart error from the iterate);
// kCheck: the check after check_every more iterations -- stop (status, result, adopt = 2 if the average is
// the result), restart (adopt = 0 / 1, new weight and steps) or continue; kFinal: store the FP64 KKT of the
// iterate in kkt64.  Returns whether the instance is still running.
__device__ inline bool decide_one(const BatchInst& in, BatchState& s, const double* o, const BatchParams& P,
                                  int mode) {
    if (mode == kFinal) {
        if (s.active) bt_store(s.kkt64, bt_kkt(in, o[0], o[1], o[2], o[3], o[12], o[13]));
        return false;
    }
    if (!s.active) return false;
    const int ce = P.check_every;
    if (mode == kInit) {  // phase start (src/pdhg_impl.hpp:120-125)
        double e = bt_kkt(in, o[0], o[1], o[2], o[3], o[12], o[13]).err;
        if (s.restart_err > 0) e = fmax(e, s.restart_err);
        s.restart_err = e;
        s.best_err = e;
        s.prev_cand = INFINITY;
        s.it_since = 0;
        s.no_progress = 0;
        s.adopt = -1;
        set_steps(in, s);
        s.inv_k = 1.0 / ce;
        return true;
    }
    // a check after ce more iterations (src/pdhg_impl.hpp:144-189)
    s.it += ce;
    s.it_since += ce;
    s.adopt = -1;
    const Kkt kc = bt_kkt(in, o[0], o[1], o[2], o[3], o[12], o[13]);
    const Kkt ka = bt_kkt(in, o[4], o[5], o[6], o[7], o[14], o[15]);
    int why = kRunning;
    Kkt best = kc;
    int best_avg = 0;
    const bool avg_better = ka.finite() && (!kc.finite() || ka.err < kc.err);
    const Kkt kcand = avg_better ? ka : kc;
    if (!kc.finite() && !ka.finite()) { why = kNumerical; best = kc; }
    else if (kc.converged(P.tol)) { why = kConverged; best = kc; best_avg = 0; }
    else if (ka.converged(P.tol)) { why = kConverged; best = ka; best_avg = 1; }
    else if (s.it >= P.max_iter) { why = kIterLimit; best = kcand; best_avg = avg_better; }
    else {
        if (s.it_since >= 4 * ce) {  // src/pdhg_impl.hpp:110-118
            const double dray = (o[16] + o[9]) * in.obj_scale;
            if (dray > 0 && sqrt(o[8]) * (1.0 + in.bnorm) / dray <= P.infeas_tol) {
                why = kInfeasible; best = kcand; best_avg = avg_better;
            }
        }
        if (why == kRunning) {
            if (kcand.err < 0.8 * s.best_err) { s.best_err = kcand.err; s.no_progress = 0; }
            else if (++s.no_progress >= P.stall_checks && P.stall_exit) {
                why = kStalled; best = kcand; best_avg = avg_better;
            }
        }
        if (why == kRunning && P.end_phase && P.stall_exit) {  // the batch's policy, not a stall test
            why = kStalled; best = kcand; best_avg = avg_better;
        }
        if (why == kRunning) {
            const bool sufficient = kcand.err <= 0.2 * s.restart_err;
            const bool necessary = kcand.err <= 0.8 * s.restart_err && kcand.err > s.prev_cand;
            const bool artificial = s.it_since >= 0.36 * (double)s.it;
            s.prev_cand = kcand.err;
            if (sufficient || necessary || artificial) {
                const double dx = sqrt(avg_better ? o[11] : o[10]), dy = sqrt(avg_better ? o[18] : o[17]);
                if (dx > 1e-10 && dy > 1e-10) s.w = exp(0.5 * log(dy / dx) + 0.5 * log(s.w));
                s.adopt = avg_better ? 1 : 0;
                s.it_since = 0;
                s.restart_err = kcand.err;
                s.prev_cand = INFINITY;
                ++s.restarts;
                set_steps(in, s);
            }
        }
    }
    bt_store(s.cand, kcand);
    s.cand_avg = avg_better ? 1 : 0;
    if (why != kRunning) {
        s.status = why;
        s.best_avg = best_avg;
        bt_store(s.best, best);
        s.adopt = best_avg ? 2 : -1;  // hand the average back as the phase's result
    }
    s.inv_k = 1.0 / (double)(s.it_since + ce);
    return why == kRunning;
}

}  // namespace ganit_hip
