// Ruiz + Pock-Chambolle scaling, bound/objective rescaling and norm estimates (reference src/scaling.cpp).
// Every floating-point operation happens in the reference's order; parallel loops only split work whose
// per-element arithmetic is sequential (a row's or a column's own sum in ascending index order, or a max).
#include "scaling.hpp"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

#include "../../include/ganit_hip.h"

namespace ganit {

double norm2(const std::vector<double>& v) {
    double s = 0.0;
    for (double x : v) s += x * x;
    return std::sqrt(s);
}

double finite_bound_norm(const std::vector<double>& lo, const std::vector<double>& hi) {
    double s = 0.0;
    for (size_t i = 0; i < lo.size(); ++i) {
        const double a = std::isfinite(lo[i]) ? std::fabs(lo[i]) : 0.0;
        const double b = std::isfinite(hi[i]) ? std::fabs(hi[i]) : 0.0;
        const double v = std::max(a, b);
        s += v * v;
    }
    return std::sqrt(s);
}

namespace {

void inv_sqrt_or_one(std::vector<double>& v) {
#pragma omp parallel for schedule(static)
    for (long k = 0; k < (long)v.size(); ++k) v[k] = v[k] > 0 ? 1.0 / std::sqrt(v[k]) : 1.0;
}

// A <- diag(rs) A diag(cs); R *= rs; C *= cs   (reference scaling.cpp:27-31)
void apply(Csr& a, const std::vector<double>& rs, const std::vector<double>& cs, std::vector<double>& R,
           std::vector<double>& C) {
#pragma omp parallel for schedule(dynamic, 1024)
    for (int i = 0; i < a.rows; ++i)
        for (int p = a.ptr[i]; p < a.ptr[i + 1]; ++p) a.val[p] *= rs[i] * cs[a.idx[p]];
    for (int i = 0; i < a.rows; ++i) R[i] *= rs[i];
    for (int j = 0; j < a.cols; ++j) C[j] *= cs[j];
}

// Column view of a's pattern: entries of column j are a's entries perm[cptr[j] .. cptr[j+1]), rows ascending,
// and trow[q] is the row of entry perm[q] -- the order of the reference's stable transpose (lp.cpp:55-72),
// built as a parallel counting sort: each thread counts the columns of its own contiguous range of rows, the
// per-column offsets are laid out thread after thread, and each thread then places its rows' entries in order.
void column_view(const Csr& a, std::vector<int>& cptr, std::vector<int>& perm, std::vector<int>& trow) {
    const int m = a.rows, n = a.cols;
    const int64_t nnz = (int64_t)a.val.size();
    cptr.assign((size_t)n + 1, 0);
    perm.resize((size_t)nnz);
    trow.resize((size_t)nnz);
    // threads, bounded so that the per-thread column counters stay within about 64 MB; one inside an enclosing
    // parallel region (the batched set-up scales its instances concurrently), where nested regions run serially
    const int64_t cap = std::max<int64_t>(1, (int64_t)16 * 1024 * 1024 / std::max(n, 1));
    const int64_t avail = omp_in_parallel() ? 1 : (int64_t)omp_get_max_threads();
    const int nt = (int)std::max<int64_t>(1, std::min<int64_t>({avail, cap, (int64_t)m}));
    // row ranges of about nnz / nt entries each
    std::vector<int> rb((size_t)nt + 1, m);
    rb[0] = 0;
    for (int t = 1; t < nt; ++t) {
        const int64_t target = nnz * t / nt;
        rb[t] = (int)(std::lower_bound(a.ptr.begin(), a.ptr.begin() + m + 1, (int)target) - a.ptr.begin());
        rb[t] = std::max(rb[t], rb[t - 1]);
    }
    std::vector<int> cnt((size_t)nt * n, 0);
#pragma omp parallel for num_threads(nt) schedule(static, 1)
    for (int t = 0; t < nt; ++t) {
        int* c = cnt.data() + (size_t)t * n;
        for (int p = a.ptr[rb[t]]; p < a.ptr[rb[t + 1]]; ++p) ++c[a.idx[p]];
    }
    for (int j = 0; j < n; ++j) {  // column starts, and each thread's offset within each column
        int pos = cptr[j];
        for (int t = 0; t < nt; ++t) {
            int& c = cnt[(size_t)t * n + j];
            const int k = c;
            c = pos;
            pos += k;
        }
        cptr[j + 1] = pos;
    }
#pragma omp parallel for num_threads(nt) schedule(static, 1)
    for (int t = 0; t < nt; ++t) {
        int* at = cnt.data() + (size_t)t * n;
        for (int i = rb[t]; i < rb[t + 1]; ++i)
            for (int p = a.ptr[i]; p < a.ptr[i + 1]; ++p) {
                const int q = at[a.idx[p]]++;
                perm[q] = p;
                trow[q] = i;
            }
    }
}

// Fold the columns of the current C Q C into the column norms cs: the max (sum = false) or the sum (sum = true)
// of |C_j Q_jk C_k| over row j of the symmetric Q, combined with A's column norm by max, resp. addition.
void q_column_norms(const Csr& Q, const std::vector<double>& C, bool sum, std::vector<double>& cs) {
#pragma omp parallel for schedule(dynamic, 1024)
    for (int j = 0; j < Q.rows; ++j) {
        double t = 0.0;
        for (int p = Q.ptr[j]; p < Q.ptr[j + 1]; ++p) {
            const double v = std::fabs(C[j] * Q.val[p] * C[Q.idx[p]]);
            t = sum ? t + v : std::max(t, v);
        }
        cs[j] = sum ? cs[j] + t : std::max(cs[j], t);
    }
}

// CSR SpMV with each row summed sequentially (the reference's host loops).
void spmv_seq_rows(const Csr& a, const std::vector<double>& x, std::vector<double>& y) {
#pragma omp parallel for schedule(dynamic, 1024)
    for (int i = 0; i < a.rows; ++i) {
        double t = 0.0;
        for (int p = a.ptr[i]; p < a.ptr[i + 1]; ++p) t += a.val[p] * x[a.idx[p]];
        y[i] = t;
    }
}

}  // namespace

ScaledLP scale_lp(const LP& lp, int ruiz_iters, bool pock_chambolle, bool bound_obj_rescale, bool q_columns) {
    ScaledLP s;
    s.m = lp.m;
    s.n = lp.n;
    s.A = lp.A;
    s.R.assign(lp.m, 1.0);
    s.C.assign(lp.n, 1.0);
    s.maximize = lp.maximize;

    // Column view of A's pattern: entry q of column j is A entry perm[q] (row trow[q]); rows ascending within
    // a column.  The scaled A^T is read off it at the end.
    std::vector<int> cptr, perm, trow;
    column_view(s.A, cptr, perm, trow);

    std::vector<double> rs(lp.m), cs(lp.n);
    for (int it = 0; it < ruiz_iters; ++it) {  // inf-norm Ruiz (reference scaling.cpp:45-57)
#pragma omp parallel for schedule(dynamic, 1024)
        for (int i = 0; i < lp.m; ++i) {
            double mx = 0.0;
            for (int p = s.A.ptr[i]; p < s.A.ptr[i + 1]; ++p) mx = std::max(mx, std::fabs(s.A.val[p]));
            rs[i] = mx;
        }
#pragma omp parallel for schedule(dynamic, 1024)
        for (int j = 0; j < lp.n; ++j) {
            double mx = 0.0;
            for (int q = cptr[j]; q < cptr[j + 1]; ++q) mx = std::max(mx, std::fabs(s.A.val[perm[q]]));
            cs[j] = mx;
        }
        if (q_columns && lp.has_q()) q_column_norms(lp.Q, s.C, false, cs);
        inv_sqrt_or_one(rs);
        inv_sqrt_or_one(cs);
        apply(s.A, rs, cs, s.R, s.C);
    }
    if (pock_chambolle) {  // alpha = 1: row and column 1-norms (reference scaling.cpp:58-70)
#pragma omp parallel for schedule(dynamic, 1024)
        for (int i = 0; i < lp.m; ++i) {
            double t = 0.0;
            for (int p = s.A.ptr[i]; p < s.A.ptr[i + 1]; ++p) t += std::fabs(s.A.val[p]);
            rs[i] = t;
        }
#pragma omp parallel for schedule(dynamic, 1024)
        for (int j = 0; j < lp.n; ++j) {  // ascending row order, as the reference's row-major accumulation
            double t = 0.0;
            for (int q = cptr[j]; q < cptr[j + 1]; ++q) t += std::fabs(s.A.val[perm[q]]);
            cs[j] = t;
        }
        if (q_columns && lp.has_q()) q_column_norms(lp.Q, s.C, true, cs);
        inv_sqrt_or_one(rs);
        inv_sqrt_or_one(cs);
        apply(s.A, rs, cs, s.R, s.C);
    }
    // A^T: the column view with the scaled values (the same arrays as transpose(s.A), built in parallel)
    s.AT = empty_csr(lp.n, lp.m);
    s.AT.ptr = cptr;
    s.AT.idx = std::move(trow);
    s.AT.val.resize(perm.size());
#pragma omp parallel for schedule(static)
    for (long q = 0; q < (long)perm.size(); ++q) s.AT.val[q] = s.A.val[perm[q]];

    scale_vectors(lp, bound_obj_rescale, s);
    return s;
}

// The part of scale_lp after the matrix (reference src/scaling.cpp:72-128): objective, bounds, the uniform
// bound / objective rescaling, the residual weights, Q and its norm, the norms of the data as given.
void scale_vectors(const LP& lp, bool bound_obj_rescale, ScaledLP& s) {
    const double sign = lp.maximize ? -1.0 : 1.0;
    s.c.resize(lp.n);
    s.l.resize(lp.n);
    s.u.resize(lp.n);
    for (int j = 0; j < lp.n; ++j) {
        s.c[j] = sign * lp.c[j] * s.C[j];
        s.l[j] = lp.l[j] / s.C[j];
        s.u[j] = lp.u[j] / s.C[j];
    }
    s.lo.resize(lp.m);
    s.hi.resize(lp.m);
    for (int i = 0; i < lp.m; ++i) {
        s.lo[i] = lp.lo[i] * s.R[i];
        s.hi[i] = lp.hi[i] * s.R[i];
    }
    if (bound_obj_rescale) {
        s.bscale = finite_bound_norm(s.lo, s.hi) + 1.0;
        s.cscale = norm2(s.c) + 1.0;
        for (auto& v : s.c) v /= s.cscale;
        for (auto& v : s.l) v /= s.bscale;
        for (auto& v : s.u) v /= s.bscale;
        for (auto& v : s.lo) v /= s.bscale;
        for (auto& v : s.hi) v /= s.bscale;
    }
    s.Rres.resize(lp.m);
    s.Cres.resize(lp.n);
    for (int i = 0; i < lp.m; ++i) s.Rres[i] = s.R[i] / s.bscale;
    for (int j = 0; j < lp.n; ++j) s.Cres[j] = s.C[j] / s.cscale;
    s.obj_const = sign * lp.obj_const;
    s.Q = empty_csr(lp.n, lp.n);
    if (lp.has_q()) {
        s.Q = lp.Q;
        for (int i = 0; i < lp.n; ++i)
            for (int p = s.Q.ptr[i]; p < s.Q.ptr[i + 1]; ++p)
                s.Q.val[p] *= sign * s.C[i] * s.C[s.Q.idx[p]] * s.bscale / s.cscale;
        std::vector<double> v(lp.n, 1.0), w(lp.n);
        const double nv = norm2(v);
        for (auto& t : v) t /= nv;
        double lam = 0.0;
        for (int it = 0; it < 60; ++it) {
            spmv_seq_rows(s.Q, v, w);
            const double nw = norm2(w);
            if (nw == 0.0) break;
            lam = nw;
            for (int i = 0; i < lp.n; ++i) v[i] = w[i] / nw;
        }
        s.normQ = lam * 1.01;
    }
    s.bnorm = finite_bound_norm(lp.lo, lp.hi);
    s.cnorm = norm2(lp.c);
}

namespace {
void hip_chk(int rc) {
    if (rc != GANIT_OK) throw std::runtime_error(std::string("HIP set-up: ") + ganit_hip_last_error());
}
}  // namespace

ScaledLP scale_lp_device(const LP& lp, int ruiz_iters, bool pock_chambolle, ganit_hip_ctx* ctx, double* norm_a,
                         int power_iters, bool bound_obj_rescale) {
    ScaledLP s;
    s.m = lp.m;
    s.n = lp.n;
    s.A = lp.A;
    s.maximize = lp.maximize;
    s.R.resize(lp.m);
    s.C.resize(lp.n);
    s.AT = empty_csr(lp.n, lp.m);
    s.AT.idx.resize(s.A.val.size());
    s.AT.val.resize(s.A.val.size());
    ganit_hip_setup* h = nullptr;
    hip_chk(ganit_hip_setup_create(ctx, lp.m, lp.n, lp.A.ptr.data(), lp.A.idx.data(), lp.A.val.data(), ruiz_iters,
                                   pock_chambolle ? 1 : 0, &h));
    try {
        hip_chk(ganit_hip_setup_result(h, s.R.data(), s.C.data(), s.A.val.data(), s.AT.ptr.data(), s.AT.idx.data(),
                                       s.AT.val.data()));
        if (norm_a) hip_chk(ganit_hip_setup_norm(h, power_iters, norm_a));
    } catch (...) {
        ganit_hip_setup_destroy(h);
        throw;
    }
    ganit_hip_setup_destroy(h);
    scale_vectors(lp, bound_obj_rescale, s);
    return s;
}

double estimate_norm(const ScaledLP& s, int iters) {
    if (s.A.nnz() == 0) return 1.0;
    std::vector<double> v(s.n, 1.0), w(s.m), z(s.n);
    const double nv = norm2(v);
    for (auto& x : v) x /= nv;
    double lambda = 0.0;
    for (int it = 0; it < iters; ++it) {
        spmv_seq_rows(s.A, v, w);
        spmv_seq_rows(s.AT, w, z);
        const double nz = norm2(z);
        if (nz == 0.0) break;
        lambda = nz;  // ||v|| = 1, so ||A'^T A' v|| -> sigma_max^2
#pragma omp parallel for schedule(static)
        for (int j = 0; j < s.n; ++j) v[j] = z[j] / nz;
    }
    return std::sqrt(lambda) * 1.01;
}

}  // namespace ganit
