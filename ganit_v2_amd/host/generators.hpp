// Seeded synthetic LP/QP generators for the benchmark and test workloads.
// All randomness comes from an in-repo SplitMix64 stream with 53-bit uniforms and Box-Muller normals, so an
// instance is a pure function of (sizes, seed) on every platform that shares the C math library.
#pragma once
#include <cstdint>
#include <vector>

#include "lp.hpp"

namespace ganit {

struct Rng {
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next();
    double uniform();                  // [0, 1), 53 bits
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal();                   // N(0, 1)
    uint64_t below(uint64_t n);        // uniform integer in [0, n)
    uint64_t s;
    bool have_spare = false;
    double spare = 0.0;
};

// Random sparse LP with a planted primal point x0 (the "seeded random LP" workload):
//   rows [0, m_eq) are equalities A x = A x0, rows [m_eq, m) are A x <= A x0 + U[0,1];
//   each column has k ~ U{nnz_min..nnz_max} distinct rows (uniform), values N(0,1);
//   x0_j ~ U[0,10] with probability 0.7, else 0; bounds 0 <= x <= 10;
//   c = A^T y0 + s0 with y0 ~ N(0,1) on equality rows, -U[0,1] on inequality rows, s0_j ~ U[0,1] where
//   x0_j = 0 and 0 elsewhere -- so the LP is feasible and bounded (dual_obj(y0) <= opt <= c'x0).
//   tighten_frac > 0 picks each variable with that probability and sets u = floor(x0) or l = ceil(x0)
//   (equal odds), like a branch-and-bound node relaxation.
struct RandomLpSpec {
    int m = 2000, n = 5000, m_eq = 1000;
    int nnz_min = 10, nnz_max = 10;
    double tighten_frac = 0.0;
};
LP gen_random_lp(const RandomLpSpec& spec, uint64_t seed, std::vector<double>* x0 = nullptr,
                 std::vector<double>* y0 = nullptr);

// One instance of the batched workload (uneven node relaxations): sizes drawn from the seed's own stream,
// m log-uniform in [m_lo, m_hi], n log-uniform in [n_lo, n_hi] (rounded), half the rows equalities, then
// gen_random_lp with column counts U{nnz_min..nnz_max}, the given tightening fraction and the same seed.
struct BatchLpSpec {
    int m_lo = 200, m_hi = 4000, n_lo = 500, n_hi = 10000;
    int nnz_min = 5, nnz_max = 15;
    double tighten_frac = 0.05;
};
LP gen_batch_instance(const BatchLpSpec& spec, uint64_t seed, std::vector<double>* x0 = nullptr,
                      std::vector<double>* y0 = nullptr);

// Multi-period crude-blending LP (C crudes, P products, K quality specs, T periods).  Per period t:
//   variables  x[c,p,t] >= 0 blend flow, r[c,t] >= 0 crude received, ci[c,t] >= 0 crude inventory,
//              pi[p,t] >= 0 product inventory;
//   rows       (a) demand balance   pi[p,t-1] + sum_c x[c,p,t] - pi[p,t] = d[p,t],  d ~ U[50,150]
//              (b) availability     r[c,t] <= a[c,t],                            a ~ U[100,300]
//              (c) crude carry-over ci[c,t-1] + r[c,t] - sum_p x[c,p,t] - ci[c,t] = 0
//              (d) quality specs    sum_c (q[c,k] - lim[k]) x[c,p,t] <= 0 (k < ceil(K/2)) or >= 0 (others),
//                                   q ~ U[0.5,1.5], lim = 60th / 40th percentile of q[.,k] over crudes
//              (e) unit capacity    sum_{c,p} x[c,p,t] <= cap[t]  (keeps cumulative capacity >= demand)
//   every variable costs U[1,10].  Percentiles are order statistics, so each spec row has one exact-zero
//   coefficient that the CSR build drops.
LP gen_blending(int crudes, int products, int specs, int periods, uint64_t seed);

// Factor-model portfolio QP: min gamma (sum d_i x_i^2 + sum_k f_k^2) - mu'x  (0.5 x'Qx form: Q = 2 gamma diag)
//   s.t. f_k - sum_i F_ik x_i = 0 (k < factors), sum_i x_i = 1, 0 <= x_i <= 0.01, f free;
//   F ~ N(0, 0.1^2) dense, d ~ U[0.01,0.05], mu ~ N(0.05, 0.02^2).
LP gen_portfolio(int assets, int factors, double gamma, uint64_t seed);

// Sparse convex QP for the differentiable layer: n variables, m_eq equality + m_le <= rows, A entries present
// with probability `density`, values N(0,1); Q = diag U[0.1,1] + symmetric off-diagonal part present with
// probability `qdensity` per pair, values N(0, 0.01^2), diagonal raised where needed to stay strictly
// diagonally dominant; x0 ~ U[0,1]; b = A x0 (+ U[0,1] slack on <= rows); -5 <= x <= 5; c ~ N(0,1).
LP gen_diff_qp(int n, int m_eq, int m_le, double density, double qdensity, uint64_t seed,
               std::vector<double>* x0 = nullptr);

}  // namespace ganit
