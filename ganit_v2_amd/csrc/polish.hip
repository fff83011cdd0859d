// libganit_hip.so, part 6: polishing a PDHG point of an LP onto the face its active set names -- the
// "crossover / polish" the reference lists as planned (README.md:195).
//
// From the FP64 iterate (x, y) of the scaled LP
//     min c'x  s.t.  lo <= A x <= hi,  l <= x <= u,
// the active set is read off with a tolerance: a variable within tol of a bound is fixed there (the rest are
// free, B), a row is active (T) when it is an equality, when its dual is not negligible, or when its activity
// lies within tol of one of its bounds (its target b_i is that bound).  Then two least-squares corrections, each an
// LSQR solve whose products are SpMVs with A and A^T and masks:
//   primal:  x_B <- x_B + d,  d = argmin ||A_TB d - (b_T - A_T x)|| of least norm   (the nearest point of the face)
//   dual:    y_T <- y_T + e,  e = argmin ||A_TB' e - (c_B - A_TB' y_T)|| of least norm, y = 0 off T
// so that x satisfies the active rows exactly and the free variables' reduced costs vanish.  When the active set
// is the optimal one, (x, y) is then optimal to rounding; whether it is, the caller decides from an FP64 KKT
// evaluation of the result.  All arithmetic is FP64; reductions are deterministic (block partials, fixed fold).
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.hpp"
#include "pdhg_state.hpp"

namespace ganit_hip {

namespace {

constexpr int kPolLanes = 8;  // lanes per row of the polish SpMVs

unsigned pgrid(int64_t n) { return std::min<unsigned>(grid_for(std::max<int64_t>(n, 1)), 2048u); }

// y = mask_out o (a x)   (mask_out may be null)
__global__ __launch_bounds__(kBlock) void k_pol_spmv(CsrView<double> a, const double* __restrict__ x,
                                                     const int* __restrict__ mask_out, double* __restrict__ y) {
    const int64_t gt = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int row = (int)(gt / kPolLanes), lane = (int)(gt % kPolLanes);
    if (row >= a.rows) return;
    const double t = row_dot<kPolLanes>(a, row, lane, x);
    if (lane == 0) y[row] = (mask_out && !mask_out[row]) ? 0.0 : t;
}

// Active set of the iterate.  Column j is fixed at a bound (xs_j = that bound) when x_j lies within
// tol (1 + |bound|) of it and its reduced cost rc_j = c_j - (A'y)_j does not contradict it (rc_j >= -tol at a lower
// bound, rc_j <= tol at an upper one); otherwise it is free (free[j] = 1, xs_j = x_j).  Row i is active (act[i] = 1,
// b[i] = its target) when it is an equality, when its dual names a bound (y_i > tol: lower, y_i < -tol: upper; the
// dual step keeps these signs), or when its activity lies within tol (1 + |bound|) of a finite bound whose side the
// dual does not contradict.
__global__ __launch_bounds__(kBlock) void k_pol_cols(int n, const double* __restrict__ x, const double* __restrict__ rc,
                                                     const double* __restrict__ l, const double* __restrict__ u,
                                                     double tol, double* __restrict__ xs, int* __restrict__ freev) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    const double xj = x[j], lj = l[j], uj = u[j], r = rc[j];
    const bool at_l = isfinite(lj) && xj <= lj + tol * (1.0 + fabs(lj)) && r >= -tol;
    const bool at_u = isfinite(uj) && xj >= uj - tol * (1.0 + fabs(uj)) && r <= tol;
    double v = xj;
    if (at_l && at_u) v = (xj - lj <= uj - xj) ? lj : uj;
    else if (at_l) v = lj;
    else if (at_u) v = uj;
    xs[j] = v;
    freev[j] = (at_l || at_u) ? 0 : 1;
}

__global__ __launch_bounds__(kBlock) void k_pol_rows(int m, const double* __restrict__ ax, const double* __restrict__ y,
                                                     const double* __restrict__ lo, const double* __restrict__ hi,
                                                     double tol, int* __restrict__ act, double* __restrict__ b) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= m) return;
    const double a = ax[i], l = lo[i], h = hi[i], yi = y[i];
    int on = 0;
    double t = 0.0;
    if (l == h) {
        on = 1;
        t = l;
    } else if (yi > tol && isfinite(l)) {
        on = 1;
        t = l;
    } else if (yi < -tol && isfinite(h)) {
        on = 1;
        t = h;
    } else {
        const bool near_l = isfinite(l) && fabs(a - l) <= tol * (1.0 + fabs(l)) && yi >= -tol;
        const bool near_h = isfinite(h) && fabs(a - h) <= tol * (1.0 + fabs(h)) && yi <= tol;
        if (near_l || near_h) {
            on = 1;
            t = (near_l && near_h) ? (fabs(a - l) <= fabs(a - h) ? l : h) : (near_l ? l : h);
        }
    }
    act[i] = on;
    b[i] = on ? t : 0.0;
}

// d = mask o (a - b)  (mask, b may be null)
__global__ __launch_bounds__(kBlock) void k_pol_masked_diff(int n, const int* __restrict__ mask,
                                                            const double* __restrict__ a, const double* __restrict__ b,
                                                            double* __restrict__ d) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    d[i] = (!mask || mask[i]) ? a[i] - (b ? b[i] : 0.0) : 0.0;
}

// y += alpha x  (mask: only where set; mask may be null)
__global__ __launch_bounds__(kBlock) void k_pol_axpy(int n, double alpha, const double* __restrict__ x,
                                                     const int* __restrict__ mask, double* __restrict__ y) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    if (!mask || mask[i]) y[i] += alpha * x[i];
}

// x = a x
__global__ __launch_bounds__(kBlock) void k_pol_scale(int n, double a, double* __restrict__ x) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) x[i] *= a;
}

// v = mask o v
__global__ __launch_bounds__(kBlock) void k_pol_mask(int n, const int* __restrict__ mask, double* __restrict__ v) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n && !mask[i]) v[i] = 0.0;
}

// partial[b] = sum over this block's share of a_i * b_i; then one block folds the partials in order
__global__ __launch_bounds__(kBlock) void k_pol_dot(int n, const double* __restrict__ a, const double* __restrict__ b,
                                                    double* __restrict__ part) {
    __shared__ double lds[kBlock / kWave];
    double s[1] = {0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) s[0] += a[i] * b[i];
    block_partial<1>(s, ReduceSlot{part, nullptr}, lds);
}
__global__ __launch_bounds__(kBlock) void k_pol_fold(const double* __restrict__ part, int nparts,
                                                     double* __restrict__ out) {
    __shared__ double lds[kBlock / kWave];
    sum_partials<1>(part, nparts, out, lds);
}

// Scratch of one polish call on one problem.
struct Work {
    Ctx* ctx;
    int m, n;
    std::vector<void*> bufs;
    double* part = nullptr;
    double* out = nullptr;
    double* host = nullptr;
    Work(Ctx* c, int m_, int n_) : ctx(c), m(m_), n(n_) {
        part = dbl(2048);
        out = dbl(1);
        GH_CHECK(hipHostMalloc((void**)&host, sizeof(double), hipHostMallocDefault));
    }
    ~Work() {
        for (void* p : bufs) (void)hipFree(p);
        if (host) (void)hipHostFree(host);
    }
    double* dbl(int64_t k) {
        void* p = nullptr;
        GH_CHECK(hipMalloc(&p, sizeof(double) * std::max<int64_t>(k, 2)));
        bufs.push_back(p);
        return static_cast<double*>(p);
    }
    int* ints(int64_t k) {
        void* p = nullptr;
        GH_CHECK(hipMalloc(&p, sizeof(int) * std::max<int64_t>(k, 2)));
        bufs.push_back(p);
        return static_cast<int*>(p);
    }
    double dot(int k, const double* a, const double* b) {
        hipStream_t st = ctx->stream;
        const unsigned g = pgrid(k);
        hipLaunchKernelGGL(k_pol_dot, dim3(g), dim3(kBlock), 0, st, k, a, b, part);
        hipLaunchKernelGGL(k_pol_fold, dim3(1), dim3(kBlock), 0, st, (const double*)part, (int)g, out);
        GH_CHECK(hipMemcpyAsync(host, out, sizeof(double), hipMemcpyDeviceToHost, st));
        GH_CHECK(hipStreamSynchronize(st));
        return *host;
    }
};

// LSQR (Paige and Saunders, ACM TOMS 8(1), 1982) for min ||K d - rhs|| from d = 0: the minimum-norm solution when
// the system is consistent, the least-squares one when it is not, without forming K'K (so a rank-deficient or
// ill-conditioned K does not blow up as conjugate gradients on the normal equations do).  K maps ncol entries to
// nrow; apply(v, out) = K v, apply_t(u, out) = K' u.  Stops when the residual or the normal-equations residual
// falls below tol relative to ||rhs|| (resp. ||K|| ||r||), or after max_it products.  Returns the products and the
// final ||r|| / ||rhs||.
template <class Apply, class ApplyT>
std::pair<int, double> lsqr(Work& w, int nrow, int ncol, Apply apply, ApplyT apply_t, const double* rhs, double* d,
                            double tol, int max_it) {
    hipStream_t st = w.ctx->stream;
    double* u = w.dbl(nrow);
    double* v = w.dbl(ncol);
    double* wv = w.dbl(ncol);
    double* tu = w.dbl(nrow);
    double* tv = w.dbl(ncol);
    auto scale = [&](int k, double* x, double a) {
        if (k > 0) hipLaunchKernelGGL(k_pol_scale, dim3(grid_for(k)), dim3(kBlock), 0, st, k, a, x);
    };
    auto axpy = [&](int k, double a, const double* x, double* y) {
        if (k > 0)
            hipLaunchKernelGGL(k_pol_axpy, dim3(grid_for(k)), dim3(kBlock), 0, st, k, a, x, (const int*)nullptr, y);
    };
    GH_CHECK(hipMemsetAsync(d, 0, sizeof(double) * std::max(ncol, 1), st));
    GH_CHECK(hipMemcpyAsync(u, rhs, sizeof(double) * std::max(nrow, 1), hipMemcpyDeviceToDevice, st));
    double beta = std::sqrt(w.dot(nrow, u, u));
    const double bnorm = beta;
    if (beta == 0.0) return {0, 0.0};
    scale(nrow, u, 1.0 / beta);
    apply_t(u, v);
    double alpha = std::sqrt(w.dot(ncol, v, v));
    if (alpha == 0.0) return {1, 1.0};
    scale(ncol, v, 1.0 / alpha);
    GH_CHECK(hipMemcpyAsync(wv, v, sizeof(double) * ncol, hipMemcpyDeviceToDevice, st));
    double phibar = beta, rhobar = alpha, anorm2 = alpha * alpha;
    int it = 1;
    for (; it < max_it; ++it) {
        // bidiagonalisation: beta u = K v - alpha u, alpha v = K' u - beta v
        apply(v, tu);
        scale(nrow, u, -alpha);
        axpy(nrow, 1.0, tu, u);
        beta = std::sqrt(w.dot(nrow, u, u));
        if (beta > 0) scale(nrow, u, 1.0 / beta);
        apply_t(u, tv);
        scale(ncol, v, -beta);
        axpy(ncol, 1.0, tv, v);
        alpha = std::sqrt(w.dot(ncol, v, v));
        if (alpha > 0) scale(ncol, v, 1.0 / alpha);
        anorm2 += alpha * alpha + beta * beta;
        // plane rotation
        const double rho = std::sqrt(rhobar * rhobar + beta * beta);
        const double c = rhobar / rho, sn = beta / rho;
        const double theta = sn * alpha;
        rhobar = -c * alpha;
        const double phi = c * phibar;
        phibar = sn * phibar;
        axpy(ncol, phi / rho, wv, d);                                    // d += (phi / rho) w
        scale(ncol, wv, -theta / rho);                                   // w = v - (theta / rho) w
        axpy(ncol, 1.0, v, wv);
        const double rnorm = phibar, arnorm = phibar * alpha * std::fabs(c);
        if (rnorm <= tol * bnorm || arnorm <= tol * std::sqrt(anorm2) * rnorm || alpha == 0.0 || beta == 0.0) break;
    }
    GH_CHECK(hipGetLastError());
    return {it, phibar / bnorm};
}

}  // namespace

}  // namespace ganit_hip

using namespace ganit_hip;

extern "C" {

int ganit_hip_pdhg_polish(ganit_hip_pdhg* h, const double* params, double* info) {
    return guarded([&] {
        auto* pd = reinterpret_cast<Pdhg*>(h);
        GH_REQUIRE(pd && params && info, "null argument");
        GH_REQUIRE(!pd->hasq, "polish is for LPs (the problem has a quadratic term)");
        const double tol = params[0], cg_tol = params[1];
        const int max_it = (int)params[2];
        GH_REQUIRE(tol > 0 && cg_tol > 0 && max_it > 0, "polish needs positive tolerances and iteration limit");
        VecSet& s = pd->set(GANIT_F64);
        const int m = pd->m, n = pd->n;
        hipStream_t st = pd->ctx->stream;
        const CsrView<double> A = pd->A->view<double>(), AT = pd->AT->view<double>();
        double* x = (double*)s.X[s.cur];
        double* y = (double*)s.Y[s.cur];
        Work w(pd->ctx, m, n);
        int* freev = w.ints(n);
        int* act = w.ints(m);
        double* xs = w.dbl(n);
        double* b = w.dbl(m);
        double* ax = w.dbl(m);
        double* tn = w.dbl(n);
        double* tm = w.dbl(m);
        double* rhs_m = w.dbl(m);
        double* rhs_n = w.dbl(n);
        double* lam = w.dbl(m);
        double* mu = w.dbl(n);
        const unsigned gm = grid_for((int64_t)m * kPolLanes), gn = grid_for((int64_t)n * kPolLanes);
        auto spmv_A = [&](const double* v, const int* mask, double* out) {  // out = mask o (A v)
            if (m > 0) hipLaunchKernelGGL(k_pol_spmv, dim3(gm), dim3(kBlock), 0, st, A, v, mask, out);
        };
        auto spmv_AT = [&](const double* v, const int* mask, double* out) {  // out = mask o (A' v)
            if (n > 0) hipLaunchKernelGGL(k_pol_spmv, dim3(gn), dim3(kBlock), 0, st, AT, v, mask, out);
        };
        // ---- active set (the reduced costs c - A'y first)
        spmv_AT(y, nullptr, tn);
        if (n > 0) {
            hipLaunchKernelGGL(k_pol_masked_diff, dim3(grid_for(n)), dim3(kBlock), 0, st, n, (const int*)nullptr,
                               (const double*)s.c, (const double*)tn, rhs_n);
            hipLaunchKernelGGL(k_pol_cols, dim3(grid_for(n)), dim3(kBlock), 0, st, n, (const double*)x,
                               (const double*)rhs_n, (const double*)s.l, (const double*)s.u, tol, xs, freev);
        }
        spmv_A(xs, nullptr, ax);
        if (m > 0)
            hipLaunchKernelGGL(k_pol_rows, dim3(grid_for(m)), dim3(kBlock), 0, st, m, (const double*)ax,
                               (const double*)y, (const double*)s.lo, (const double*)s.hi, tol, act, b);
        GH_CHECK(hipGetLastError());
        // ---- primal: d_B = argmin ||A_TB d - (b_T - A_T xs)|| (minimum norm), x_B += d
        if (m > 0)
            hipLaunchKernelGGL(k_pol_masked_diff, dim3(grid_for(m)), dim3(kBlock), 0, st, m, (const int*)act,
                               (const double*)b, (const double*)ax, rhs_m);
        auto k_tb = [&](const double* v, double* out) {  // out = A_TB v  (v on B)
            spmv_A(v, act, out);
        };
        auto k_tb_t = [&](const double* u, double* out) {  // out = A_TB' u  (u on T)
            spmv_AT(u, freev, out);
        };
        // the operands are masked as they are produced: A_TB v reads v only where B is (v is kept zero off B
        // because every vector LSQR combines into it comes out of A_TB' masked to B)
        const auto pr = lsqr(w, m, n, k_tb, k_tb_t, rhs_m, mu, cg_tol, max_it);
        if (n > 0)
            hipLaunchKernelGGL(k_pol_axpy, dim3(grid_for(n)), dim3(kBlock), 0, st, n, 1.0, (const double*)mu,
                               (const int*)freev, xs);
        // ---- dual: y = 0 off T, then e_T = argmin ||A_TB' e - (c_B - A_TB' y_T)|| (minimum norm), y_T += e
        if (m > 0) hipLaunchKernelGGL(k_pol_mask, dim3(grid_for(m)), dim3(kBlock), 0, st, m, (const int*)act, y);
        spmv_AT(y, freev, tn);
        if (n > 0)
            hipLaunchKernelGGL(k_pol_masked_diff, dim3(grid_for(n)), dim3(kBlock), 0, st, n, (const int*)freev,
                               (const double*)s.c, (const double*)tn, rhs_n);
        const auto du = lsqr(w, n, m, k_tb_t, k_tb, rhs_n, lam, cg_tol, max_it);
        if (m > 0)
            hipLaunchKernelGGL(k_pol_axpy, dim3(grid_for(m)), dim3(kBlock), 0, st, m, 1.0, (const double*)lam,
                               (const int*)act, y);
        // ---- the polished point becomes the iterate (and the restart point, so a later check's ray is zero)
        if (n > 0) {
            GH_CHECK(hipMemcpyAsync(x, xs, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
            GH_CHECK(hipMemcpyAsync(s.xr, xs, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        }
        if (m > 0) GH_CHECK(hipMemcpyAsync(s.yr, y, sizeof(double) * m, hipMemcpyDeviceToDevice, st));
        GH_CHECK(hipGetLastError());
        // counts of free variables and active rows
        std::vector<int> fh(n), ah(m);
        if (n > 0) GH_CHECK(hipMemcpyAsync(fh.data(), freev, sizeof(int) * n, hipMemcpyDeviceToHost, st));
        if (m > 0) GH_CHECK(hipMemcpyAsync(ah.data(), act, sizeof(int) * m, hipMemcpyDeviceToHost, st));
        GH_CHECK(hipStreamSynchronize(st));
        info[0] = (double)std::count(fh.begin(), fh.end(), 1);
        info[1] = (double)std::count(ah.begin(), ah.end(), 1);
        info[2] = pr.first;
        info[3] = pr.second;
        info[4] = du.first;
        info[5] = du.second;
    });
}

}  // extern "C"
