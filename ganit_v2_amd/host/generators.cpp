// Synthetic workload generators (see generators.hpp for the models).
#include "generators.hpp"

#include <algorithm>
#include <cmath>
#include <string>

namespace ganit {

uint64_t Rng::next() {  // SplitMix64
    s += 0x9e3779b97f4a7c15ULL;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

double Rng::uniform() { return (double)(next() >> 11) * 0x1.0p-53; }

double Rng::normal() {  // Box-Muller, both outputs used
    if (have_spare) {
        have_spare = false;
        return spare;
    }
    const double u1 = 1.0 - uniform();  // (0, 1]
    const double u2 = uniform();
    const double r = std::sqrt(-2.0 * std::log(u1));
    const double th = 2.0 * M_PI * u2;
    spare = r * std::sin(th);
    have_spare = true;
    return r * std::cos(th);
}

uint64_t Rng::below(uint64_t n) {  // rejection sampling, unbiased
    if (n <= 1) return 0;
    const uint64_t lim = UINT64_MAX - (UINT64_MAX % n);
    uint64_t v;
    do v = next(); while (v >= lim);
    return v % n;
}

namespace {
void names(LP& lp, const char* rp, const char* cp) {
    lp.row_names.resize(lp.m);
    lp.col_names.resize(lp.n);
    for (int i = 0; i < lp.m; ++i) lp.row_names[i] = rp + std::to_string(i);
    for (int j = 0; j < lp.n; ++j) lp.col_names[j] = cp + std::to_string(j);
    lp.is_int.assign(lp.n, 0);
}
}  // namespace

LP gen_random_lp(const RandomLpSpec& sp, uint64_t seed, std::vector<double>* x0_out, std::vector<double>* y0_out) {
    Rng rng(seed);
    const int m = sp.m, n = sp.n;
    LP lp;
    lp.name = "RANDLP_M" + std::to_string(m) + "_N" + std::to_string(n) + "_S" + std::to_string(seed);
    lp.m = m;
    lp.n = n;
    std::vector<int> tr, tc;
    std::vector<double> tv;
    std::vector<int> rows;
    std::vector<char> used(m, 0);
    for (int j = 0; j < n; ++j) {
        const int span = std::max(0, sp.nnz_max - sp.nnz_min);
        int k = sp.nnz_min + (int)rng.below((uint64_t)span + 1);
        k = std::min(k, m);
        rows.clear();
        while ((int)rows.size() < k) {
            const int i = (int)rng.below((uint64_t)m);
            if (!used[i]) { used[i] = 1; rows.push_back(i); }
        }
        for (int i : rows) {
            used[i] = 0;
            tr.push_back(i);
            tc.push_back(j);
            tv.push_back(rng.normal());
        }
    }
    lp.A = csr_from_triplets(m, n, tr, tc, tv);
    std::vector<double> x0(n), y0(m), s0(n, 0.0);
    for (int j = 0; j < n; ++j) x0[j] = rng.uniform() < 0.7 ? rng.uniform(0.0, 10.0) : 0.0;
    lp.lo.assign(m, 0.0);
    lp.hi.assign(m, 0.0);
    for (int i = 0; i < m; ++i) {
        double ax = 0.0;
        for (int p = lp.A.ptr[i]; p < lp.A.ptr[i + 1]; ++p) ax += lp.A.val[p] * x0[lp.A.idx[p]];
        if (i < sp.m_eq) { lp.lo[i] = ax; lp.hi[i] = ax; }
        else { lp.lo[i] = -kInf; lp.hi[i] = ax + rng.uniform(); }
    }
    for (int i = 0; i < m; ++i) y0[i] = i < sp.m_eq ? rng.normal() : -rng.uniform();
    for (int j = 0; j < n; ++j) if (x0[j] == 0.0) s0[j] = rng.uniform();
    const Csr at = transpose(lp.A);
    lp.c.assign(n, 0.0);
    for (int j = 0; j < n; ++j) {
        double t = 0.0;
        for (int p = at.ptr[j]; p < at.ptr[j + 1]; ++p) t += at.val[p] * y0[at.idx[p]];
        lp.c[j] = t + s0[j];
    }
    lp.l.assign(n, 0.0);
    lp.u.assign(n, 10.0);
    if (sp.tighten_frac > 0.0) {
        for (int j = 0; j < n; ++j) {
            if (rng.uniform() >= sp.tighten_frac) continue;
            if (rng.uniform() < 0.5) lp.u[j] = std::floor(x0[j]);
            else lp.l[j] = std::ceil(x0[j]);
        }
    }
    lp.Q = empty_csr(n, n);
    names(lp, "R", "C");
    if (x0_out) *x0_out = x0;
    if (y0_out) *y0_out = y0;
    return lp;
}

LP gen_batch_instance(const BatchLpSpec& b, uint64_t seed, std::vector<double>* x0, std::vector<double>* y0) {
    // sizes from a stream separate from the instance's own (seed mixed with a fixed odd constant)
    Rng r(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull);
    auto log_uniform = [&](int lo, int hi) {
        const double v = std::exp(r.uniform(std::log((double)lo), std::log((double)hi)));
        return std::min(hi, std::max(lo, (int)std::lround(v)));
    };
    RandomLpSpec sp;
    sp.m = log_uniform(b.m_lo, b.m_hi);
    sp.n = log_uniform(b.n_lo, b.n_hi);
    sp.m_eq = sp.m / 2;
    sp.nnz_min = b.nnz_min;
    sp.nnz_max = b.nnz_max;
    sp.tighten_frac = b.tighten_frac;
    LP lp = gen_random_lp(sp, seed, x0, y0);
    lp.name = "BATCHLP_S" + std::to_string(seed);
    return lp;
}

LP gen_blending(int C, int P, int K, int T, uint64_t seed) {
    Rng rng(seed);
    LP lp;
    lp.name = "BLEND_C" + std::to_string(C) + "_P" + std::to_string(P) + "_K" + std::to_string(K) + "_T" +
              std::to_string(T);
    // quality data and spec limits
    std::vector<double> q((size_t)C * K);
    for (auto& v : q) v = rng.uniform(0.5, 1.5);
    const int kle = (K + 1) / 2;  // specs [0, kle) are maxima (<=), the rest minima (>=)
    std::vector<double> lim(K);
    for (int k = 0; k < K; ++k) {
        std::vector<double> col(C);
        for (int c = 0; c < C; ++c) col[c] = q[(size_t)c * K + k];
        std::sort(col.begin(), col.end());
        const double f = k < kle ? 0.6 : 0.4;
        lim[k] = col[(size_t)std::floor(f * (C - 1))];
    }
    // per-period layout
    const int nv = C * P + 2 * C + P;           // variables per period
    const int nr = P + 2 * C + P * K + 1;       // rows per period
    lp.n = nv * T;
    lp.m = nr * T;
    auto X = [&](int c, int p, int t) { return t * nv + c * P + p; };
    auto Rv = [&](int c, int t) { return t * nv + C * P + c; };
    auto CI = [&](int c, int t) { return t * nv + C * P + C + c; };
    auto PI = [&](int p, int t) { return t * nv + C * P + 2 * C + p; };
    auto DEM = [&](int p, int t) { return t * nr + p; };
    auto AV = [&](int c, int t) { return t * nr + P + c; };
    auto CB = [&](int c, int t) { return t * nr + P + C + c; };
    auto SP = [&](int p, int k, int t) { return t * nr + P + 2 * C + p * K + k; };
    auto CAP = [&](int t) { return t * nr + P + 2 * C + P * K; };
    std::vector<int> tr, tc;
    std::vector<double> tv;
    auto add = [&](int r, int c, double v) { tr.push_back(r); tc.push_back(c); tv.push_back(v); };
    lp.lo.assign(lp.m, 0.0);
    lp.hi.assign(lp.m, 0.0);
    lp.c.assign(lp.n, 0.0);
    lp.l.assign(lp.n, 0.0);
    lp.u.assign(lp.n, kInf);
    double slack = 0.0;  // cumulative spare capacity carried as product inventory
    for (int t = 0; t < T; ++t) {
        double dem_t = 0.0;
        for (int p = 0; p < P; ++p) {
            const double d = rng.uniform(50.0, 150.0);
            lp.lo[DEM(p, t)] = d;
            lp.hi[DEM(p, t)] = d;
            dem_t += d;
        }
        for (int c = 0; c < C; ++c) {
            lp.lo[AV(c, t)] = -kInf;
            lp.hi[AV(c, t)] = rng.uniform(100.0, 300.0);
            // CB rows are equalities with zero right-hand side (already 0)
        }
        for (int p = 0; p < P; ++p)
            for (int k = 0; k < K; ++k) {
                if (k < kle) { lp.lo[SP(p, k, t)] = -kInf; lp.hi[SP(p, k, t)] = 0.0; }
                else { lp.lo[SP(p, k, t)] = 0.0; lp.hi[SP(p, k, t)] = kInf; }
            }
        const double cap = std::max(rng.uniform(0.95, 1.25) * dem_t, dem_t - slack + 1.0);
        slack += cap - dem_t;
        lp.lo[CAP(t)] = -kInf;
        lp.hi[CAP(t)] = cap;
        for (int c = 0; c < C; ++c)
            for (int p = 0; p < P; ++p) {
                const int j = X(c, p, t);
                add(DEM(p, t), j, 1.0);
                add(CB(c, t), j, -1.0);
                for (int k = 0; k < K; ++k) add(SP(p, k, t), j, q[(size_t)c * K + k] - lim[k]);
                add(CAP(t), j, 1.0);
            }
        for (int c = 0; c < C; ++c) {
            add(AV(c, t), Rv(c, t), 1.0);
            add(CB(c, t), Rv(c, t), 1.0);
            add(CB(c, t), CI(c, t), -1.0);
            if (t + 1 < T) add(CB(c, t + 1), CI(c, t), 1.0);
        }
        for (int p = 0; p < P; ++p) {
            add(DEM(p, t), PI(p, t), -1.0);
            if (t + 1 < T) add(DEM(p, t + 1), PI(p, t), 1.0);
        }
    }
    for (int j = 0; j < lp.n; ++j) lp.c[j] = rng.uniform(1.0, 10.0);
    lp.A = csr_from_triplets(lp.m, lp.n, tr, tc, tv);
    lp.Q = empty_csr(lp.n, lp.n);
    names(lp, "R", "C");
    return lp;
}

LP gen_portfolio(int N, int F, double gamma, uint64_t seed) {
    Rng rng(seed);
    LP lp;
    lp.name = "PORT_N" + std::to_string(N) + "_F" + std::to_string(F);
    lp.n = N + F;
    lp.m = F + 1;
    std::vector<int> tr, tc;
    std::vector<double> tv;
    tr.reserve((size_t)N * (F + 1) + F);
    tc.reserve(tr.capacity());
    tv.reserve(tr.capacity());
    lp.c.assign(lp.n, 0.0);
    lp.l.assign(lp.n, 0.0);
    lp.u.assign(lp.n, 0.01);
    std::vector<int> qi;
    std::vector<double> qv;
    for (int i = 0; i < N; ++i) {
        for (int k = 0; k < F; ++k) {
            tr.push_back(k);
            tc.push_back(i);
            tv.push_back(-0.1 * rng.normal());
        }
        tr.push_back(F);
        tc.push_back(i);
        tv.push_back(1.0);
        const double d = rng.uniform(0.01, 0.05);
        lp.c[i] = -(0.05 + 0.02 * rng.normal());
        qi.push_back(i);
        qv.push_back(2.0 * gamma * d);
    }
    for (int k = 0; k < F; ++k) {
        const int j = N + k;
        tr.push_back(k);
        tc.push_back(j);
        tv.push_back(1.0);
        lp.l[j] = -kInf;
        lp.u[j] = kInf;
        qi.push_back(j);
        qv.push_back(2.0 * gamma);
    }
    lp.lo.assign(lp.m, 0.0);
    lp.hi.assign(lp.m, 0.0);
    lp.lo[F] = 1.0;
    lp.hi[F] = 1.0;
    lp.A = csr_from_triplets(lp.m, lp.n, tr, tc, tv);
    lp.Q = csr_from_triplets(lp.n, lp.n, qi, qi, qv);
    names(lp, "R", "C");
    return lp;
}

LP gen_diff_qp(int n, int m_eq, int m_le, double density, double qdensity, uint64_t seed, std::vector<double>* x0_out) {
    Rng rng(seed);
    LP lp;
    lp.name = "DQP_N" + std::to_string(n) + "_S" + std::to_string(seed);
    lp.n = n;
    lp.m = m_eq + m_le;
    std::vector<int> tr, tc;
    std::vector<double> tv;
    for (int i = 0; i < lp.m; ++i)
        for (int j = 0; j < n; ++j)
            if (rng.uniform() < density) {
                tr.push_back(i);
                tc.push_back(j);
                tv.push_back(rng.normal());
            }
    lp.A = csr_from_triplets(lp.m, n, tr, tc, tv);
    std::vector<double> diag(n), off_abs(n, 0.0);
    for (int j = 0; j < n; ++j) diag[j] = rng.uniform(0.1, 1.0);
    std::vector<int> qr, qc;
    std::vector<double> qv;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (rng.uniform() < qdensity) {
                const double v = 0.01 * rng.normal();
                qr.push_back(i); qc.push_back(j); qv.push_back(v);
                qr.push_back(j); qc.push_back(i); qv.push_back(v);
                off_abs[i] += std::fabs(v);
                off_abs[j] += std::fabs(v);
            }
    for (int j = 0; j < n; ++j) {
        if (diag[j] <= off_abs[j]) diag[j] = off_abs[j] + 0.01;
        qr.push_back(j); qc.push_back(j); qv.push_back(diag[j]);
    }
    lp.Q = csr_from_triplets(n, n, qr, qc, qv);
    std::vector<double> x0(n);
    for (auto& v : x0) v = rng.uniform();
    lp.lo.assign(lp.m, 0.0);
    lp.hi.assign(lp.m, 0.0);
    for (int i = 0; i < lp.m; ++i) {
        double ax = 0.0;
        for (int p = lp.A.ptr[i]; p < lp.A.ptr[i + 1]; ++p) ax += lp.A.val[p] * x0[lp.A.idx[p]];
        if (i < m_eq) { lp.lo[i] = ax; lp.hi[i] = ax; }
        else { lp.lo[i] = -kInf; lp.hi[i] = ax + rng.uniform(); }
    }
    lp.c.resize(n);
    for (auto& v : lp.c) v = rng.normal();
    lp.l.assign(n, -5.0);
    lp.u.assign(n, 5.0);
    names(lp, "R", "C");
    if (x0_out) *x0_out = x0;
    return lp;
}

}  // namespace ganit
