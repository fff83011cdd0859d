// libganit_hip.so, part 4: the adjoint of the QP solution map, for differentiating through a solve.
//
// At a PDHG fixed point z* = (x*, y*) of T(z; theta), theta = (c, row bounds, diag Q), the implicit-function
// theorem gives  dL/dtheta = (d_theta T)^T lambda  with  lambda = J^T lambda + (dL/dx*, 0),  J = d_z T.
// With the projection derivatives frozen at z* -- D_j = 1 when l_j < x*_j < u_j (primal clip inactive),
// F_i = 1 when row i's bound is active (equality rows, or y*_i != 0) -- one application of J^T is a PDHG
// iteration run backwards, with A and A^T swapped:
//     mu = F o ly,   r = A^T mu,   p = D o (lx - 2 sigma r),
//     lx' = p - tau Q p + sigma r + g,   ly' = mu + tau A p.
// and the gradient is read off the fixed point:  dL/dc = -tau p,  dL/dQ_jj = -tau p_j x*_j,
// dL/d(active bound of row i) = sigma mu_i.  The three kernels reuse the row-block SpMV machinery of the
// forward solve (k_adj_at over A^T, k_adj_a over A, k_adj_x over Q's rows); blocks of iterations are
// replayed from a captured HIP graph, the last iteration of a block also measuring ||lambda' - lambda||.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "internal.hpp"
#include "pdhg_state.hpp"

namespace ganit_hip {

// Masks from the FP64 iterate: D_j = 1 unless x_j lies within mask_tol (1 + |bound|) of l_j or u_j (an iterate
// converted from FP32, or one a slow direction leaves an ulp inside its bound, is at the bound);
// F_i = (lo_i == hi_i) || y_i != 0 (inactive rows have y exactly 0: the dual prox returns q + (-q)).
__global__ __launch_bounds__(kBlock) void k_adj_masks(int n, int m, const double* __restrict__ x,
                                                      const double* __restrict__ l, const double* __restrict__ u,
                                                      const double* __restrict__ y, const double* __restrict__ lo,
                                                      const double* __restrict__ hi, double mask_tol,
                                                      uint8_t* __restrict__ D, uint8_t* __restrict__ F) {
    const int64_t total = (int64_t)n + m;
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < total; t += (int64_t)gridDim.x * kBlock) {
        if (t < n) {
            const double v = x[t], lj = l[t], uj = u[t];
            const bool at_l = isfinite(lj) && v - lj <= mask_tol * (1.0 + fabs(lj));
            const bool at_u = isfinite(uj) && uj - v <= mask_tol * (1.0 + fabs(uj));
            D[t] = (at_l || at_u) ? 0 : 1;
        } else {
            const int64_t i = t - n;
            F[i] = (lo[i] == hi[i] || y[i] != 0.0) ? 1 : 0;
        }
    }
}

template <class V>
__global__ __launch_bounds__(kBlock) void k_adj_init(int n, int m, const double* __restrict__ g64,
                                                     V* __restrict__ g, V* __restrict__ lx, V* __restrict__ ly) {
    const int64_t total = (int64_t)n + m;
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < total; t += (int64_t)gridDim.x * kBlock) {
        if (t < n) {
            g[t] = V(g64[t]);
            lx[t] = V(g64[t]);
        } else {
            ly[t - n] = V(0);
        }
    }
}

// r = A^T (F o ly), p = D o (lx - 2 sigma r)
template <class M, class V>
__global__ __launch_bounds__(kBlock) void k_adj_at(CsrView<M> AT, const int* __restrict__ blocks, SplitPlan sp,
                                                   const V* __restrict__ ly, const uint8_t* __restrict__ F,
                                                   const V* __restrict__ lx, const uint8_t* __restrict__ D,
                                                   V two_sigma, V* __restrict__ r, V* __restrict__ p) {
    __shared__ __align__(16) V lds[kTile];
    stream_block<M, V, 1>(
        AT, blocks, xcd_block_id(), [&](int, int i) { return F[i] ? ly[i] : V(0); },
        [&](int j, V (&t)[1]) {
            r[j] = t[0];
            p[j] = D[j] ? V(lx[j] - two_sigma * t[0]) : V(0);
        },
        lds, sp);
}

// (J^T lambda)_y = F o ly + tau A p.  RES: out = ly - (J^T lambda)_y (the operator I - J^T), else
// out = (J^T lambda)_y.  out may alias ly (each row reads and writes only its own entry).  MEASURE: partial
// sums of |F o (out - ly)|^2 and |F o out|^2 per workgroup.
template <class M, class V, bool MEASURE, bool RES>
__global__ __launch_bounds__(kBlock) void k_adj_a(CsrView<M> A, const int* __restrict__ blocks, SplitPlan sp,
                                                  const V* __restrict__ p, const uint8_t* __restrict__ F, V tau,
                                                  const V* ly, V* out, double* __restrict__ partials) {
    __shared__ __align__(16) V lds[kTile];
    __shared__ double red[(kBlock / kWave) * 2];
    double acc[2] = {0.0, 0.0};
    stream_block<M, V, 1>(
        A, blocks, xcd_block_id(), [&](int, int j) { return p[j]; },
        [&](int i, V (&t)[1]) {
            const V in = ly[i];
            const V old = F[i] ? in : V(0);
            const V nv = old + tau * t[0];
            out[i] = RES ? V(in - nv) : nv;
            if (MEASURE && F[i]) {
                const double d = double(nv) - double(old);
                acc[0] += d * d;
                acc[1] += double(nv) * double(nv);
            }
        },
        lds, sp);
    if (MEASURE) {
        __syncthreads();
        block_partial<2>(acc, ReduceSlot{partials, nullptr}, red);
    }
}

// (J^T lambda)_x + g = p - tau Q p + sigma r + g (Q row summed by the owning thread, as in the forward
// primal kernel).  RES: out = lx - (p - tau Q p + sigma r), no g.  out may alias lx.
template <class M, class V, bool HASQ, bool MEASURE, bool RES>
__global__ __launch_bounds__(kBlock) void k_adj_x(int n, CsrView<M> Q, const V* __restrict__ p,
                                                  const V* __restrict__ r, const V* __restrict__ g, V tau, V sigma,
                                                  const V* lx, V* out, double* __restrict__ partials) {
    __shared__ double red[(kBlock / kWave) * 2];
    double acc[2] = {0.0, 0.0};
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
        V qp = V(0);
        if (HASQ)
            for (int k = Q.ptr[j]; k < Q.ptr[j + 1]; ++k) qp += widen<V>(Q.val[k]) * p[Q.idx[k]];
        const V in = lx[j];
        const V jt = p[j] - tau * qp + sigma * r[j];
        const V nv = RES ? V(in - jt) : V(jt + g[j]);
        if (MEASURE) {
            const double d = double(nv) - double(in);
            acc[0] += d * d;
            acc[1] += double(nv) * double(nv);
        }
        out[j] = nv;
    }
    if (MEASURE) block_partial<2>(acc, ReduceSlot{partials, nullptr}, red);
}

// out[0..1] = sums of the x partials, out[2..3] = sums of the y partials (one block per output).
__global__ __launch_bounds__(kBlock) void k_adj_finish(const double* __restrict__ px, int nx,
                                                       const double* __restrict__ py, int ny,
                                                       double* __restrict__ out) {
    __shared__ double lds[kBlock / kWave];
    const int k = blockIdx.x;
    if (k < 2) sum_partials_one(px, 2, k, nx, out + k, lds);
    else sum_partials_one(py, 2, k - 2, ny, out + k, lds);
}

template <class V>
__global__ __launch_bounds__(kBlock) void k_adj_out(int n, int m, const V* __restrict__ p,
                                                    const V* __restrict__ ly, const uint8_t* __restrict__ F,
                                                    double* __restrict__ p64, double* __restrict__ mu64) {
    const int64_t total = (int64_t)n + m;
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < total; t += (int64_t)gridDim.x * kBlock) {
        if (t < n) p64[t] = double(p[t]);
        else mu64[t - n] = F[t - n] ? double(ly[t - n]) : 0.0;
    }
}

// ---------------------------------------------------------------- GMRES on (I - J^T) lambda = (g, 0)
// Krylov basis: ncols vectors of length L = n + m stored back to back (x part, then y part).
// part[k][b] = partial dot of basis vector k with w over workgroup b's chunk (grid: nb x ncols).
template <class V>
__global__ __launch_bounds__(kBlock) void k_gm_dots(int64_t L, const V* __restrict__ basis, const V* __restrict__ w,
                                                    double* __restrict__ part) {
    __shared__ double red[kBlock / kWave];
    const int k = blockIdx.y;
    const V* v = basis + (size_t)k * L;
    double s[1] = {0.0};
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < L; t += (int64_t)gridDim.x * kBlock)
        s[0] += double(v[t]) * double(w[t]);
    block_sum<1>(s, red);
    if (threadIdx.x == 0) part[(size_t)k * gridDim.x + blockIdx.x] = s[0];
}

// h[k] (+)= sum_b part[k][b], one block per k; ACC adds to h instead of overwriting.
template <bool ACC>
__global__ __launch_bounds__(kBlock) void k_gm_finish(int nb, const double* __restrict__ part,
                                                      double* __restrict__ h) {
    __shared__ double lds[kBlock / kWave];
    const int k = blockIdx.x;
    double s[1] = {0.0};
    for (int b = threadIdx.x; b < nb; b += blockDim.x) s[0] += part[(size_t)k * nb + b];
    block_sum<1>(s, lds);
    if (threadIdx.x == 0) h[k] = ACC ? h[k] + s[0] : s[0];
}

// w += sgn * sum_{k < ncols} h[k] basis_k
template <class V>
__global__ __launch_bounds__(kBlock) void k_gm_update(int64_t L, const V* __restrict__ basis, int ncols,
                                                      const double* __restrict__ h, double sgn, V* __restrict__ w) {
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < L; t += (int64_t)gridDim.x * kBlock) {
        double a = double(w[t]);
        for (int k = 0; k < ncols; ++k) a += sgn * h[k] * double(basis[(size_t)k * L + t]);
        w[t] = V(a);
    }
}

// out = w / sqrt(*nrm2) (out may alias w)
template <class V>
__global__ __launch_bounds__(kBlock) void k_gm_scale(int64_t L, const V* w, const double* __restrict__ nrm2, V* out) {
    const double s = *nrm2 > 0.0 ? 1.0 / sqrt(*nrm2) : 0.0;
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < L; t += (int64_t)gridDim.x * kBlock)
        out[t] = V(double(w[t]) * s);
}

// out = b - a, where b = (g, 0)
template <class V>
__global__ __launch_bounds__(kBlock) void k_gm_rhs_minus(int n, int64_t L, const V* __restrict__ g,
                                                         const V* __restrict__ a, V* __restrict__ out) {
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < L; t += (int64_t)gridDim.x * kBlock)
        out[t] = V((t < n ? double(g[t]) : 0.0) - double(a[t]));
}

namespace {

struct DevBuf {
    void* p = nullptr;
    DevBuf() = default;
    explicit DevBuf(size_t bytes) { GH_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 16))); }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

unsigned grid_n(int64_t n) { return std::min<unsigned>(grid_for(std::max<int64_t>(n, 1)), 2048u); }

}  // namespace

}  // namespace ganit_hip

using namespace ganit_hip;

extern "C" {

// params = {method (0 fixed-point iteration, 1 GMRES), max_iter, tol, check_every (fixed point) or restart
// length (GMRES), mask_tol}
int ganit_hip_pdhg_adjoint(ganit_hip_pdhg* h, int mdt, int vdt, const double* gx, double tau, double sigma,
                           const double* params, double* p_out, double* mu_out, double* info) {
    return guarded([&] {
        auto* pd = reinterpret_cast<Pdhg*>(h);
        GH_REQUIRE(pd && gx && params && p_out && mu_out, "null argument");
        const int method = (int)params[0];
        const int max_iter = (int)std::min(params[1], 2.0e9);
        const double tol = params[2];
        const int every = (int)params[3];
        const double mask_tol = params[4];
        GH_REQUIRE(method == 0 || method == 1, "adjoint method must be 0 (fixed point) or 1 (GMRES)");
        GH_REQUIRE(tau > 0 && sigma > 0 && max_iter > 0 && every > 0 && mask_tol >= 0, "bad adjoint parameters");
        GH_REQUIRE(method == 0 || every <= 1024, "GMRES restart length is limited to 1024");
        VecSet& s64 = pd->set(GANIT_F64);
        pd->set(vdt);
        Ctx* ctx = pd->ctx;
        hipStream_t st = ctx->stream;
        const int n = pd->n, m = pd->m;
        const int64_t L = (int64_t)n + m;
        const size_t e = dtype_size(vdt);
        const unsigned ga = pd->A->rb_grid(), gat = pd->AT->rb_grid();
        GH_REQUIRE((n == 0 || gat > 0) && (m == 0 || ga > 0), "row-block partition missing");
        DevBuf D(n), F(m), g(e * n), r(e * n), p(e * n), g64(8 * (size_t)n), lam(e * L);
        const unsigned gx_grid = grid_n(n), gl = grid_n(L);
        const unsigned nsa = (unsigned)pd->A->nsplit, nsat = (unsigned)pd->AT->nsplit;
        DevBuf part_x(sizeof(double) * 2 * gx_grid), part_y(sizeof(double) * 2 * std::max(ga + nsa, 1u)),
            out(sizeof(double) * 4), p64(8 * (size_t)n), mu64(8 * (size_t)std::max(m, 1));
        GH_CHECK(hipMemcpyAsync(g64.p, gx, 8 * (size_t)n, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_adj_masks, dim3(grid_n(L)), dim3(kBlock), 0, st, n, m, (const double*)s64.X[s64.cur],
                           (const double*)s64.l, (const double*)s64.u, (const double*)s64.Y[s64.cur],
                           (const double*)s64.lo, (const double*)s64.hi, mask_tol, D.as<uint8_t>(), F.as<uint8_t>());
        int iters = 0;
        double rel = INFINITY;
        with_mdt(mdt, [&](auto MT) {
            with_vdt(vdt, [&](auto VT) {
                using M = typename decltype(MT)::type;
                using V = typename decltype(VT)::type;
                V* lx = lam.as<V>();
                V* ly = lx + n;
                hipLaunchKernelGGL(k_adj_init<V>, dim3(grid_n(L)), dim3(kBlock), 0, st, n, m, g64.as<double>(),
                                   g.as<V>(), lx, ly);
                const CsrView<M> qv = pd->hasq ? pd->Q->view<M>() : CsrView<M>{0, nullptr, nullptr, nullptr};
                using T = std::true_type;
                using Fa = std::false_type;
                // out = J^T in + g (RES false) or out = in - J^T in (RES true); in / out are L-vectors
                auto apply = [&](const V* in, V* o, auto MS, auto RS) {
                    constexpr bool ms = decltype(MS)::value, rs = decltype(RS)::value;
                    // each row-block kernel: main pass, then the split-row finish pass (partials after the main's)
                    if (n > 0)
                        for (int pass = 0; pass < (nsat ? 2 : 1); ++pass)
                            hipLaunchKernelGGL((k_adj_at<M, V>), dim3(pass ? nsat : gat), dim3(kBlock), 0, st,
                                               pd->AT->view<M>(), pd->AT->blocks,
                                               pass ? pd->AT->split_finish() : pd->AT->split(), in + n,
                                               F.as<uint8_t>(), in, D.as<uint8_t>(), V(2.0 * sigma), r.as<V>(),
                                               p.as<V>());
                    if (m > 0)
                        for (int pass = 0; pass < (nsa ? 2 : 1); ++pass)
                            hipLaunchKernelGGL((k_adj_a<M, V, ms, rs>), dim3(pass ? nsa : ga), dim3(kBlock), 0, st,
                                               pd->A->view<M>(), pd->A->blocks,
                                               pass ? pd->A->split_finish() : pd->A->split(), p.as<V>(),
                                               F.as<uint8_t>(), V(tau), in + n, o + n,
                                               part_y.as<double>() + (pass ? 2 * (size_t)ga : 0));
                    if (n > 0) {
                        if (pd->hasq)
                            hipLaunchKernelGGL((k_adj_x<M, V, true, ms, rs>), dim3(gx_grid), dim3(kBlock), 0, st, n,
                                               qv, p.as<V>(), r.as<V>(), g.as<V>(), V(tau), V(sigma), in, o,
                                               part_x.as<double>());
                        else
                            hipLaunchKernelGGL((k_adj_x<M, V, false, ms, rs>), dim3(gx_grid), dim3(kBlock), 0, st, n,
                                               qv, p.as<V>(), r.as<V>(), g.as<V>(), V(tau), V(sigma), in, o,
                                               part_x.as<double>());
                    }
                };
                if (method == 0) {
                    // ---- plain fixed-point iteration lambda <- J^T lambda + (g, 0), in place
                    auto block = [&](int k) {
                        for (int i = 0; i + 1 < k; ++i) apply(lx, lx, Fa{}, Fa{});
                        apply(lx, lx, T{}, Fa{});
                        hipLaunchKernelGGL(k_adj_finish, dim3(4), dim3(kBlock), 0, st, part_x.as<double>(),
                                           n > 0 ? (int)gx_grid : 0, part_y.as<double>(), m > 0 ? (int)(ga + nsa) : 0,
                                           out.as<double>());
                    };
                    hipGraphExec_t ex = nullptr;
                    if (pd->use_graphs) {
                        hipGraph_t gr = nullptr;
                        GH_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
                        try {
                            block(every);
                        } catch (...) {
                            (void)hipStreamEndCapture(st, &gr);
                            if (gr) (void)hipGraphDestroy(gr);
                            throw;
                        }
                        GH_CHECK(hipStreamEndCapture(st, &gr));
                        const hipError_t ie = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
                        (void)hipGraphDestroy(gr);
                        GH_CHECK(ie);
                    }
                    try {
                        while (iters < max_iter) {
                            const int k = std::min(every, max_iter - iters);
                            if (ex && k == every) GH_CHECK(hipGraphLaunch(ex, st));
                            else block(k);
                            iters += k;
                            GH_CHECK(hipMemcpyAsync(ctx->host_out, out.p, sizeof(double) * 4, hipMemcpyDeviceToHost,
                                                    st));
                            GH_CHECK(hipStreamSynchronize(st));
                            const double* res = ctx->host_out;
                            const double num = res[0] + res[2], den = res[1] + res[3];
                            rel = den > 0 ? std::sqrt(num / den) : std::sqrt(num);
                            if (!std::isfinite(rel) || rel <= tol) break;
                        }
                    } catch (...) {
                        if (ex) (void)hipGraphExecDestroy(ex);
                        throw;
                    }
                    if (ex) (void)hipGraphExecDestroy(ex);
                } else {
                    // ---- restarted GMRES(k) on (I - J^T) lambda = (g, 0), lambda starting at (g, 0):
                    // Krylov acceleration of the same transposed PDHG operator.  Classical Gram-Schmidt with
                    // one re-orthogonalisation pass; dot products in FP64; the small Hessenberg least-squares
                    // problem on the host (Givens rotations).
                    const int k = every;
                    DevBuf basis(e * (size_t)L * (k + 1)), w(e * (size_t)L), hbuf(sizeof(double) * (2 * k + 4));
                    const unsigned nb = (unsigned)std::max<int64_t>(1, std::min<int64_t>(256, (L + 4095) / 4096));
                    DevBuf part(sizeof(double) * (size_t)nb * (k + 1));
                    V* Vb = basis.as<V>();
                    V* W = w.as<V>();
                    double* h1 = hbuf.as<double>();      // [k + 1] first Gram-Schmidt pass
                    double* h2 = h1 + (k + 1);           // [k + 1] second pass
                    double* nr2 = h1 + 2 * (k + 1);      // squared norm
                    auto dots = [&](const V* vecs, int ncols, const V* x, double* dst) {
                        hipLaunchKernelGGL(k_gm_dots<V>, dim3(nb, ncols), dim3(kBlock), 0, st, L, vecs, x,
                                           part.as<double>());
                        hipLaunchKernelGGL(k_gm_finish<false>, dim3(ncols), dim3(kBlock), 0, st, (int)nb,
                                           part.as<double>(), dst);
                    };
                    std::vector<double> hh(2 * k + 4);
                    auto fetch = [&]() {
                        GH_CHECK(hipMemcpyAsync(hh.data(), h1, sizeof(double) * (2 * k + 3), hipMemcpyDeviceToHost,
                                                st));
                        GH_CHECK(hipStreamSynchronize(st));
                    };
                    // ||(g, 0)||
                    hipLaunchKernelGGL(k_gm_dots<V>, dim3(nb, 1), dim3(kBlock), 0, st, (int64_t)n, g.as<V>(),
                                       g.as<V>(), part.as<double>());
                    hipLaunchKernelGGL(k_gm_finish<false>, dim3(1), dim3(kBlock), 0, st, (int)nb, part.as<double>(),
                                       nr2);
                    fetch();
                    const double bnorm = std::sqrt(hh[2 * k + 2]);
                    std::vector<double> H((size_t)k * k), cs(k), sn(k), sv(k + 1), y(k);
                    rel = bnorm > 0 ? INFINITY : 0.0;
                    double prev_beta = INFINITY;  // true residual at the previous restart
                    while (bnorm > 0 && iters < max_iter) {
                        // r0 = b - (I - J^T) lambda, into basis column 0
                        apply(lx, W, Fa{}, T{});
                        hipLaunchKernelGGL(k_gm_rhs_minus<V>, dim3(gl), dim3(kBlock), 0, st, n, L, g.as<V>(), W, Vb);
                        dots(Vb, 1, Vb, nr2);
                        fetch();
                        const double beta = std::sqrt(hh[2 * k + 2]);
                        rel = beta / bnorm;
                        if (!std::isfinite(rel) || rel <= tol) break;
                        // a whole restart cycle that gained less than 1% has met the working precision's floor
                        if (beta > 0.99 * prev_beta) break;
                        prev_beta = beta;
                        hipLaunchKernelGGL(k_gm_scale<V>, dim3(gl), dim3(kBlock), 0, st, L, Vb, nr2, Vb);
                        std::fill(sv.begin(), sv.end(), 0.0);
                        sv[0] = beta;
                        int cols = 0;
                        bool stop = false;
                        for (int j = 0; j < k && iters < max_iter && !stop; ++j) {
                            apply(Vb + (size_t)j * L, W, Fa{}, T{});  // W = (I - J^T) v_j
                            dots(Vb, j + 1, W, h1);
                            hipLaunchKernelGGL(k_gm_update<V>, dim3(gl), dim3(kBlock), 0, st, L, Vb, j + 1, h1, -1.0,
                                               W);
                            dots(Vb, j + 1, W, h2);
                            hipLaunchKernelGGL(k_gm_update<V>, dim3(gl), dim3(kBlock), 0, st, L, Vb, j + 1, h2, -1.0,
                                               W);
                            dots(W, 1, W, nr2);
                            hipLaunchKernelGGL(k_gm_scale<V>, dim3(gl), dim3(kBlock), 0, st, L, W, nr2,
                                               Vb + (size_t)(j + 1) * L);
                            fetch();
                            ++iters;
                            // Hessenberg column j, previous rotations, then a new one (Givens)
                            double* col = &H[(size_t)j * k];
                            for (int i = 0; i <= j; ++i) col[i] = hh[i] + hh[k + 1 + i];
                            double sub = std::sqrt(std::max(hh[2 * k + 2], 0.0));
                            for (int i = 0; i < j; ++i) {
                                const double t = cs[i] * col[i] + sn[i] * col[i + 1];
                                col[i + 1] = -sn[i] * col[i] + cs[i] * col[i + 1];
                                col[i] = t;
                            }
                            const double den = std::hypot(col[j], sub);
                            cs[j] = den > 0 ? col[j] / den : 1.0;
                            sn[j] = den > 0 ? sub / den : 0.0;
                            col[j] = den;
                            sv[j + 1] = -sn[j] * sv[j];
                            sv[j] = cs[j] * sv[j];
                            cols = j + 1;
                            rel = std::fabs(sv[j + 1]) / bnorm;
                            if (!std::isfinite(rel) || rel <= tol || sub <= 1e-300 * bnorm) stop = true;
                        }
                        // y = R^{-1} s, lambda += V y
                        for (int i = cols - 1; i >= 0; --i) {
                            double t = sv[i];
                            for (int l = i + 1; l < cols; ++l) t -= H[(size_t)l * k + i] * y[l];
                            y[i] = H[(size_t)i * k + i] != 0.0 ? t / H[(size_t)i * k + i] : 0.0;
                        }
                        GH_CHECK(hipMemcpyAsync(h1, y.data(), sizeof(double) * cols, hipMemcpyHostToDevice, st));
                        hipLaunchKernelGGL(k_gm_update<V>, dim3(gl), dim3(kBlock), 0, st, L, Vb, cols, h1, 1.0, lx);
                        GH_CHECK(hipStreamSynchronize(st));
                        if (!std::isfinite(rel)) break;
                    }
                }
                // p and mu of the final iterate
                if (n > 0)
                    for (int pass = 0; pass < (nsat ? 2 : 1); ++pass)
                        hipLaunchKernelGGL((k_adj_at<M, V>), dim3(pass ? nsat : gat), dim3(kBlock), 0, st,
                                           pd->AT->view<M>(), pd->AT->blocks,
                                           pass ? pd->AT->split_finish() : pd->AT->split(), ly, F.as<uint8_t>(), lx,
                                           D.as<uint8_t>(), V(2.0 * sigma), r.as<V>(), p.as<V>());
                hipLaunchKernelGGL(k_adj_out<V>, dim3(grid_n(L)), dim3(kBlock), 0, st, n, m, p.as<V>(), ly,
                                   F.as<uint8_t>(), p64.as<double>(), mu64.as<double>());
            });
        });
        GH_CHECK(hipGetLastError());
        GH_CHECK(hipMemcpyAsync(p_out, p64.p, 8 * (size_t)n, hipMemcpyDeviceToHost, st));
        if (m > 0) GH_CHECK(hipMemcpyAsync(mu_out, mu64.p, 8 * (size_t)m, hipMemcpyDeviceToHost, st));
        std::vector<uint8_t> dh(n), fh(m);
        if (n > 0) GH_CHECK(hipMemcpyAsync(dh.data(), D.p, n, hipMemcpyDeviceToHost, st));
        if (m > 0) GH_CHECK(hipMemcpyAsync(fh.data(), F.p, m, hipMemcpyDeviceToHost, st));
        GH_CHECK(hipStreamSynchronize(st));
        if (info) {
            info[0] = iters;
            info[1] = rel;
            info[2] = (double)std::count(dh.begin(), dh.end(), (uint8_t)1);
            info[3] = (double)std::count(fh.begin(), fh.end(), (uint8_t)1);
        }
    });
}

}  // extern "C"
