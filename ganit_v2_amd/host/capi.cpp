// C ABI of libganit.so (declared in include/ganit_host.h).
#include <algorithm>
#include <stdexcept>
#include <string>

#include "../../include/ganit_hip.h"
#include "../../include/ganit_host.h"
#include "generators.hpp"
#include "pdhg.hpp"
#include "scaling.hpp"

namespace {
thread_local std::string g_err;
struct Handle {
    ganit::LP lp;
    std::vector<double> x0, y0;
};
template <class F>
int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}
template <class T, class S>
void put(T* dst, const S& src) {
    if (dst) std::copy(src.begin(), src.end(), dst);
}
}  // namespace

extern "C" {

const char* ganit_last_error(void) { return g_err.c_str(); }

ganit_lp* ganit_lp_read_mps(const char* path) {
    Handle* h = nullptr;
    if (guard([&] {
            h = new Handle;
            h->lp = ganit::read_mps(path);
        })) {
        delete h;
        return nullptr;
    }
    return reinterpret_cast<ganit_lp*>(h);
}

ganit_lp* ganit_lp_from_triplets(int m, int n, int64_t nnz, const int* rows, const int* cols, const double* vals,
                                 const double* c, const double* l, const double* u, const double* lo,
                                 const double* hi, double obj_const, int maximize, int64_t qnnz, const int* qrows,
                                 const int* qcols, const double* qvals) {
    Handle* h = nullptr;
    if (guard([&] {
            if (m < 0 || n < 0 || nnz < 0 || qnnz < 0) throw std::runtime_error("negative size");
            h = new Handle;
            ganit::LP& lp = h->lp;
            lp.m = m;
            lp.n = n;
            lp.A = ganit::csr_from_triplets(m, n, std::vector<int>(rows, rows + nnz), std::vector<int>(cols, cols + nnz),
                                            std::vector<double>(vals, vals + nnz));
            lp.Q = qnnz > 0 ? ganit::csr_from_triplets(n, n, std::vector<int>(qrows, qrows + qnnz),
                                                       std::vector<int>(qcols, qcols + qnnz),
                                                       std::vector<double>(qvals, qvals + qnnz))
                            : ganit::empty_csr(n, n);
            lp.c.assign(c, c + n);
            lp.l.assign(l, l + n);
            lp.u.assign(u, u + n);
            lp.lo.assign(lo, lo + m);
            lp.hi.assign(hi, hi + m);
            lp.obj_const = obj_const;
            lp.maximize = maximize != 0;
            lp.is_int.assign(n, 0);
        })) {
        delete h;
        return nullptr;
    }
    return reinterpret_cast<ganit_lp*>(h);
}

ganit_lp* ganit_lp_generate(int kind, const double* p, uint64_t seed) {
    Handle* h = nullptr;
    if (guard([&] {
            h = new Handle;
            if (kind == 0) {
                ganit::RandomLpSpec sp;
                sp.m = (int)p[0]; sp.n = (int)p[1]; sp.m_eq = (int)p[2];
                sp.nnz_min = (int)p[3]; sp.nnz_max = (int)p[4]; sp.tighten_frac = p[5];
                h->lp = ganit::gen_random_lp(sp, seed, &h->x0, &h->y0);
            } else if (kind == 1) {
                h->lp = ganit::gen_blending((int)p[0], (int)p[1], (int)p[2], (int)p[3], seed);
            } else if (kind == 2) {
                h->lp = ganit::gen_portfolio((int)p[0], (int)p[1], p[2], seed);
            } else if (kind == 3) {
                h->lp = ganit::gen_diff_qp((int)p[0], (int)p[1], (int)p[2], p[3], p[4], seed, &h->x0);
            } else if (kind == 4) {
                ganit::BatchLpSpec sp;
                sp.m_lo = (int)p[0]; sp.m_hi = (int)p[1]; sp.n_lo = (int)p[2]; sp.n_hi = (int)p[3];
                sp.nnz_min = (int)p[4]; sp.nnz_max = (int)p[5]; sp.tighten_frac = p[6];
                h->lp = ganit::gen_batch_instance(sp, seed, &h->x0, &h->y0);
            } else {
                throw std::runtime_error("unknown generator kind");
            }
        })) {
        delete h;
        return nullptr;
    }
    return reinterpret_cast<ganit_lp*>(h);
}

void ganit_lp_free(ganit_lp* h) { delete reinterpret_cast<Handle*>(h); }

void ganit_lp_dims(const ganit_lp* hh, int64_t* d) {
    const ganit::LP& lp = reinterpret_cast<const Handle*>(hh)->lp;
    d[0] = lp.m; d[1] = lp.n; d[2] = lp.A.nnz(); d[3] = lp.Q.nnz(); d[4] = lp.maximize ? 1 : 0;
}

void ganit_lp_export(const ganit_lp* hh, int* aptr, int* aidx, double* aval, int* qptr, int* qidx, double* qval,
                     double* c, double* l, double* u, double* lo, double* hi, double* obj_const) {
    const ganit::LP& lp = reinterpret_cast<const Handle*>(hh)->lp;
    put(aptr, lp.A.ptr); put(aidx, lp.A.idx); put(aval, lp.A.val);
    put(qptr, lp.Q.ptr); put(qidx, lp.Q.idx); put(qval, lp.Q.val);
    put(c, lp.c); put(l, lp.l); put(u, lp.u); put(lo, lp.lo); put(hi, lp.hi);
    if (obj_const) *obj_const = lp.obj_const;
}

int ganit_lp_planted(const ganit_lp* hh, double* x0, double* y0) {
    const Handle* h = reinterpret_cast<const Handle*>(hh);
    if (h->x0.empty()) { g_err = "instance has no planted point"; return 1; }
    put(x0, h->x0);
    put(y0, h->y0);
    return 0;
}

int ganit_lp_write_mps(const ganit_lp* hh, const char* path) {
    return guard([&] { ganit::write_mps(reinterpret_cast<const Handle*>(hh)->lp, path); });
}

namespace {
void scaled_out(const ganit::ScaledLP& s, double norm_a, double* scal, int* aptr, int* aidx, double* aval, int* tptr,
                int* tidx, double* tval, int* qptr, int* qidx, double* qval, double* c, double* l, double* u,
                double* lo, double* hi, double* R, double* C, double* Rres, double* Cres) {
    if (scal) {
        const double v[9] = {s.bscale, s.cscale, s.normQ, s.obj_const, s.bnorm, s.cnorm, norm_a, (double)s.AT.nnz(),
                             s.maximize ? 1.0 : 0.0};
        std::copy(v, v + 9, scal);
    }
    put(aptr, s.A.ptr); put(aidx, s.A.idx); put(aval, s.A.val);
    put(tptr, s.AT.ptr); put(tidx, s.AT.idx); put(tval, s.AT.val);
    put(qptr, s.Q.ptr); put(qidx, s.Q.idx); put(qval, s.Q.val);
    put(c, s.c); put(l, s.l); put(u, s.u); put(lo, s.lo); put(hi, s.hi);
    put(R, s.R); put(C, s.C); put(Rres, s.Rres); put(Cres, s.Cres);
}
}  // namespace

int ganit_lp_scaled(const ganit_lp* hh, int ruiz, int pc, double* scal, int* aptr, int* aidx, double* aval, int* tptr,
                    int* tidx, double* tval, int* qptr, int* qidx, double* qval, double* c, double* l, double* u,
                    double* lo, double* hi, double* R, double* C, double* Rres, double* Cres) {
    return guard([&] {
        const ganit::ScaledLP s = ganit::scale_lp(reinterpret_cast<const Handle*>(hh)->lp, ruiz, pc != 0);
        scaled_out(s, scal ? ganit::estimate_norm(s) : 0.0, scal, aptr, aidx, aval, tptr, tidx, tval, qptr, qidx,
                   qval, c, l, u, lo, hi, R, C, Rres, Cres);
    });
}

int ganit_lp_scaled_gpu(const ganit_lp* hh, int ruiz, int pc, int device, double* scal, int* aptr, int* aidx,
                        double* aval, int* tptr, int* tidx, double* tval, int* qptr, int* qidx, double* qval, double* c,
                        double* l, double* u, double* lo, double* hi, double* R, double* C, double* Rres,
                        double* Cres) {
    return guard([&] {
        ganit_hip_ctx* ctx = nullptr;
        if (ganit_hip_ctx_create(device, &ctx) != GANIT_OK)
            throw std::runtime_error(std::string("HIP backend: ") + ganit_hip_last_error());
        try {
            double norm_a = 0;
            const ganit::ScaledLP s =
                ganit::scale_lp_device(reinterpret_cast<const Handle*>(hh)->lp, ruiz, pc != 0, ctx, &norm_a);
            scaled_out(s, norm_a, scal, aptr, aidx, aval, tptr, tidx, tval, qptr, qidx, qval, c, l, u, lo, hi, R, C,
                       Rres, Cres);
        } catch (...) {
            ganit_hip_ctx_destroy(ctx);
            throw;
        }
        ganit_hip_ctx_destroy(ctx);
    });
}

namespace {
ganit::Options options_from(const double* o) {
    ganit::Options opt;
    opt.tol = o[0]; opt.max_iter = (long)o[1]; opt.time_limit = o[2]; opt.check_every = (int)o[3];
    opt.ruiz_iters = (int)o[4]; opt.pock_chambolle = o[5] != 0; opt.verbose = (int)o[6];
    opt.infeas_tol = o[7];
    const int pr = (int)o[8];
    opt.precision = pr == 1 ? ganit::Precision::Mixed : pr == 2 ? ganit::Precision::Dynamic : ganit::Precision::Double;
    opt.stall_checks = (int)o[9]; opt.stall_checks_l2 = (int)o[10];
    const int vals = (int)o[11];
    opt.values = vals == 1 ? ganit::Values::BF16 : vals == 2 ? ganit::Values::BF16_F32 : ganit::Values::F32;
    opt.device = (int)o[12];
    opt.eval_fp64 = o[13] != 0;
    // bit 0: kahan, bit 1: detect_unbounded, bits 2-3: set-up (0 auto, 1 host, 2 GPU), bit 4: polish,
    // bit 5: scale_q, bit 6: no lanes-per-row tail in batches, bit 7: batches end reduced levels at the tail
    const int bits = (int)o[14];
    opt.kahan = (bits & 1) != 0;
    opt.detect_unbounded = (bits & 2) != 0;
    opt.polish = (bits & 16) != 0;
    opt.scale_q = (bits & 32) != 0;
    opt.tail_lanes = (bits & 64) == 0;
    opt.tail_fp64 = (bits & 128) != 0;
    const int where = (bits >> 2) & 3;
    if (where == 3) throw std::runtime_error("set-up placement must be 0 (auto), 1 (host) or 2 (GPU)");
    opt.setup = where == 1 ? ganit::Setup::Host : where == 2 ? ganit::Setup::Device : ganit::Setup::Auto;
    return opt;
}
void result_out(const ganit::Result& r, double* out, double* x, double* y) {
    const double v[18] = {(double)(int)r.status, r.pobj, r.dobj, r.rel_primal, r.rel_dual, r.rel_gap,
                          (double)r.iterations, (double)r.fp32_iterations, (double)r.l2_iterations,
                          (double)r.restarts, r.setup_seconds, r.solve_seconds, r.norm_a, r.upload_seconds,
                          r.iterate_seconds[0], r.iterate_seconds[1], r.iterate_seconds[2], (double)r.checks};
    if (out) std::copy(v, v + 18, out);
    put(x, r.x);
    put(y, r.y);
}
}  // namespace

int ganit_solve_gpu(const ganit_lp* hh, const double* o, double* out, double* x, double* y) {
    return guard([&] {
        const ganit::Result r = ganit::solve_pdhg_gpu(reinterpret_cast<const Handle*>(hh)->lp, options_from(o));
        result_out(r, out, x, y);
    });
}

ganit_solver* ganit_solver_create(const ganit_lp* hh, const double* o) {
    ganit::GpuSolver* s = nullptr;
    if (guard([&] { s = new ganit::GpuSolver(reinterpret_cast<const Handle*>(hh)->lp, options_from(o)); })) {
        delete s;
        return nullptr;
    }
    return reinterpret_cast<ganit_solver*>(s);
}

int ganit_solver_upload(ganit_solver* h) {
    return guard([&] { reinterpret_cast<ganit::GpuSolver*>(h)->upload(); });
}

int ganit_solver_run(ganit_solver* h, double* out, double* x, double* y) {
    return guard([&] { result_out(reinterpret_cast<ganit::GpuSolver*>(h)->run(), out, x, y); });
}

int ganit_solver_kernel_times(ganit_solver* h, int level, int reps, double* ms) {
    return guard([&] { reinterpret_cast<ganit::GpuSolver*>(h)->kernel_times(level, reps, ms); });
}

int ganit_solver_set_solution(ganit_solver* h, const double* x, const double* y, int fp32_iterate) {
    return guard([&] {
        auto* sv = reinterpret_cast<ganit::GpuSolver*>(h);
        if (!sv || !x || !y) throw std::runtime_error("null argument");
        const int n = sv->scaled().n, m = sv->scaled().m;
        sv->set_solution(std::vector<double>(x, x + n), std::vector<double>(y, y + m), fp32_iterate != 0);
    });
}

int ganit_solver_adjoint(ganit_solver* h, const double* gx, const double* ao, double* dc, double* db, double* dq,
                         double* info) {
    return guard([&] {
        auto* sv = reinterpret_cast<ganit::GpuSolver*>(h);
        if (!gx || !ao) throw std::runtime_error("null argument");
        ganit::AdjointOptions o;
        o.fp32 = ao[0] != 0;
        o.tol = ao[1];
        o.max_iter = (int)ao[2];
        o.gmres = ao[3] == 0;
        if (ao[4] > 0) {
            if (o.gmres) o.restart = (int)ao[4];
            else o.check_every = (int)ao[4];
        }
        o.mask_tol = ao[5];
        const int n = sv->scaled().n;
        const ganit::AdjointResult r = sv->adjoint(std::vector<double>(gx, gx + n), o);
        put(dc, r.dc);
        put(db, r.db);
        put(dq, r.dq);
        if (info) {
            const double v[5] = {(double)r.iterations, r.rel_change, (double)r.free_vars, (double)r.active_rows,
                                 r.seconds};
            std::copy(v, v + 5, info);
        }
    });
}

int ganit_solver_index_bytes(ganit_solver* h, int which, int* bytes) {
    return guard([&] { *bytes = reinterpret_cast<ganit::GpuSolver*>(h)->index_bytes(which); });
}

int ganit_solver_variant(ganit_solver* h, int which, int* lanes) {
    return guard([&] {
        if (!h || !lanes) throw std::runtime_error("null argument");
        *lanes = reinterpret_cast<ganit::GpuSolver*>(h)->variant(which);
    });
}

int ganit_solver_info(const ganit_solver* h, double* info) {
    return guard([&] {
        if (!h || !info) throw std::runtime_error("null argument");
        auto* s = reinterpret_cast<const ganit::GpuSolver*>(h);
        info[0] = s->setup_seconds();
        info[1] = s->norm_a();
        info[2] = s->setup_on_device() ? 1.0 : 0.0;
        info[3] = s->last_polished() ? 1.0 : 0.0;
    });
}

void ganit_solver_free(ganit_solver* h) { delete reinterpret_cast<ganit::GpuSolver*>(h); }

ganit_batch* ganit_batch_create(const ganit_lp* const* lps, int ninst, const double* o) {
    ganit::BatchSolver* b = nullptr;
    if (guard([&] {
            if (ninst <= 0 || !lps) throw std::runtime_error("empty batch");
            std::vector<const ganit::LP*> v(ninst);
            for (int k = 0; k < ninst; ++k) {
                if (!lps[k]) throw std::runtime_error("null instance");
                v[k] = &reinterpret_cast<const Handle*>(lps[k])->lp;
            }
            b = new ganit::BatchSolver(v, options_from(o));
        })) {
        delete b;
        return nullptr;
    }
    return reinterpret_cast<ganit_batch*>(b);
}

int ganit_batch_upload(ganit_batch* h) {
    return guard([&] { reinterpret_cast<ganit::BatchSolver*>(h)->upload(); });
}

int ganit_batch_run(ganit_batch* h, double* out, double* x, double* y) {
    return guard([&] {
        auto* b = reinterpret_cast<ganit::BatchSolver*>(h);
        const std::vector<ganit::Result> rs = b->run();
        size_t xo = 0, yo = 0;
        for (size_t k = 0; k < rs.size(); ++k) {
            result_out(rs[k], out ? out + 18 * k : nullptr, x ? x + xo : nullptr, y ? y + yo : nullptr);
            xo += rs[k].x.size();
            yo += rs[k].y.size();
        }
    });
}

int ganit_batch_stats(const ganit_batch* h, double* st) {
    return guard([&] {
        const ganit::BatchStats& s = reinterpret_cast<const ganit::BatchSolver*>(h)->stats();
        const double v[24] = {(double)s.rows, (double)s.cols, (double)s.nnz, (double)s.blocks_a,
                              (double)s.blocks_at, s.upload_seconds, s.solve_seconds, s.certify_seconds,
                              s.level_seconds[0], s.level_seconds[1], s.level_seconds[2],
                              (double)s.level_instances[0], (double)s.level_instances[1],
                              (double)s.level_instances[2], (double)s.max_iterations, (double)s.total_iterations,
                              (double)s.autotune_pick, s.autotune_ms[0], s.autotune_ms[1], s.autotune_ms[2],
                              s.autotune_ms[3],
                              reinterpret_cast<const ganit::BatchSolver*>(h)->setup_seconds(),
                              (double)s.tail_iterations, 0.0};
        std::copy(v, v + 24, st);
    });
}

int ganit_batch_set_bounds(ganit_batch* h, int k, const double* l, const double* u) {
    return guard([&] {
        auto* b = reinterpret_cast<ganit::BatchSolver*>(h);
        if (!b || !l || !u) throw std::runtime_error("null argument");
        if (k < 0 || k >= b->size()) throw std::runtime_error("batch instance index out of range");
        const int n = b->scaled(k).n;
        b->set_bounds(k, std::vector<double>(l, l + n), std::vector<double>(u, u + n));
    });
}

int ganit_batch_set_start(ganit_batch* h, int k, const double* x, const double* y, double w) {
    return guard([&] {
        auto* b = reinterpret_cast<ganit::BatchSolver*>(h);
        if (!b) throw std::runtime_error("null batch");
        if (k < 0 || k >= b->size()) throw std::runtime_error("batch instance index out of range");
        if (!x) {
            b->clear_start(k);
            return;
        }
        if (!y) throw std::runtime_error("null dual start");
        const ganit::ScaledLP& s = b->scaled(k);
        b->set_start(k, std::vector<double>(x, x + s.n), std::vector<double>(y, y + s.m), w);
    });
}

int ganit_batch_set_options(ganit_batch* h, const double* o) {
    return guard([&] {
        if (!h || !o) throw std::runtime_error("null argument");
        reinterpret_cast<ganit::BatchSolver*>(h)->set_options(options_from(o));
    });
}

int ganit_batch_weights(const ganit_batch* h, double* w) {
    return guard([&] {
        if (!h || !w) throw std::runtime_error("null argument");
        const std::vector<double> v = reinterpret_cast<const ganit::BatchSolver*>(h)->final_weights();
        if (v.empty()) throw std::runtime_error("the batch has not run yet");
        std::copy(v.begin(), v.end(), w);
    });
}

int ganit_batch_adjoint(ganit_batch* h, const double* gx, const double* ao, double* dc, double* db, double* dq,
                        double* info) {
    return guard([&] {
        auto* b = reinterpret_cast<ganit::BatchSolver*>(h);
        if (!b || !gx || !ao) throw std::runtime_error("null argument");
        if (ao[3] != 0) throw std::runtime_error("the batched adjoint runs GMRES only (method 0)");
        ganit::AdjointOptions o;
        o.fp32 = ao[0] != 0;
        o.tol = ao[1];
        o.max_iter = (int)ao[2];
        if (ao[4] > 0) o.restart = (int)ao[4];
        o.mask_tol = ao[5];
        size_t cols = 0;
        for (int k = 0; k < b->size(); ++k) cols += b->scaled(k).n;
        const std::vector<ganit::AdjointResult> r = b->adjoint(std::vector<double>(gx, gx + cols), o);
        size_t c0 = 0, r0 = 0;
        for (size_t k = 0; k < r.size(); ++k) {
            if (dc) std::copy(r[k].dc.begin(), r[k].dc.end(), dc + c0);
            if (dq) std::copy(r[k].dq.begin(), r[k].dq.end(), dq + c0);
            if (db) std::copy(r[k].db.begin(), r[k].db.end(), db + r0);
            if (info) {
                const double v[5] = {(double)r[k].iterations, r[k].rel_change, (double)r[k].free_vars,
                                     (double)r[k].active_rows, r[k].seconds};
                std::copy(v, v + 5, info + 5 * k);
            }
            c0 += r[k].dc.size();
            r0 += r[k].db.size();
        }
    });
}

void ganit_batch_free(ganit_batch* h) { delete reinterpret_cast<ganit::BatchSolver*>(h); }

}  // extern "C"
