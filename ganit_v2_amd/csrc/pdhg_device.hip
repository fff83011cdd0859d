// libganit_hip.so, part 2: the device-resident PDHG problem and its fused kernels.
//
// Per iteration (src/pdhg_impl.hpp:133-139) two kernels run instead of the reference's four or five:
//   k_iter_primal : for each column j -- (A^T y)_j [+ (Q x)_j] by one row group of A^T (and Q), then the
//                   primal prox, extrapolation and running sum (src/backend_cpu.cpp:72-79) in the same
//                   thread: A^T y and Q x never touch HBM.
//   k_iter_dual   : for each row i -- (A xbar)_i by one row group of A, then the dual prox and running sum
//                   (src/backend_cpu.cpp:85-91).
// A check (src/pdhg_impl.hpp:144-147 + :110-118 + :179) is one averaging kernel, one pass over A^T/Q and one
// pass over A that evaluate the current point, the average and the infeasibility ray together, and reduce
// all 22 sums deterministically; the host then reads them with a single copy.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <tuple>
#include <type_traits>

#include "internal.hpp"
#include "pdhg_math.hpp"
#include "pdhg_state.hpp"

namespace ganit_hip {

constexpr int kCheckGridMax = 2048;
constexpr int64_t kGraphMaxNnz = 1000000;  // replay iteration blocks from graphs below this many nonzeros
// per-block sums of the two check passes: A^T pass outputs 0..11 and, with GANIT_CHECK_DRAY, c'd and the
// column cone excess (12, 13); A pass outputs 12..18 of the check and, with GANIT_CHECK_DRAY, the row cone
// excess (its slot 7)
constexpr int kNOutN = 14, kNOutM = 8;

// Step sizes on the device: prm[0] = tau, prm[1] = sigma (FP64), and the same two rounded to FP32 in the next
// slot, so an FP32 kernel reads its step size without a conversion (same rounding as converting in place).
constexpr int kPrmDoubles = 3;
template <class V>
__device__ __forceinline__ V step_as(const double* __restrict__ prm, int k);
template <>
__device__ __forceinline__ double step_as<double>(const double* __restrict__ prm, int k) { return prm[k]; }
template <>
__device__ __forceinline__ float step_as<float>(const double* __restrict__ prm, int k) {
    return reinterpret_cast<const float*>(prm + 2)[k];
}

// ================================================================ iteration kernels
template <int LPR, class M, class V, bool HASQ>
__global__ __launch_bounds__(kBlock) void k_iter_primal(CsrView<M> AT, CsrView<M> Q, const V* __restrict__ y,
                                                        const V* __restrict__ x, const V* __restrict__ c,
                                                        const V* __restrict__ l, const V* __restrict__ u,
                                                        const double* __restrict__ prm, V* __restrict__ xn,
                                                        V* __restrict__ xbar, V* __restrict__ sumx) {
    const int64_t gt = (int64_t)xcd_block_id() * kBlock + threadIdx.x;
    const int j = (int)(gt / LPR), lane = (int)(gt % LPR);
    if (j >= AT.rows) return;
    const V aty = row_dot<LPR>(AT, j, lane, y);
    const V qx = HASQ ? row_dot<LPR>(Q, j, lane, x) : V(0);
    if (lane == 0) {
        V s = sumx[j];
        primal_update<V>(x[j], aty, qx, c[j], l[j], u[j], V(prm[0]), xn[j], xbar[j], s);
        sumx[j] = s;
    }
}

template <int LPR, class M, class V>
__global__ __launch_bounds__(kBlock) void k_iter_dual(CsrView<M> A, const V* __restrict__ xbar,
                                                      const V* __restrict__ y, const V* __restrict__ lo,
                                                      const V* __restrict__ hi, const double* __restrict__ prm,
                                                      V* __restrict__ yn, V* __restrict__ sumy) {
    const int64_t gt = (int64_t)xcd_block_id() * kBlock + threadIdx.x;
    const int i = (int)(gt / LPR), lane = (int)(gt % LPR);
    if (i >= A.rows) return;
    const V axb = row_dot<LPR>(A, i, lane, xbar);
    if (lane == 0) {
        V s = sumy[i];
        dual_update<V>(y[i], axb, lo[i], hi[i], V(prm[1]), yn[i], s);
        sumy[i] = s;
    }
}

// The same two kernels on LDS-staged row blocks (one block per row block; see stream_block).
template <class M, class V, bool HASQ>
__global__ __launch_bounds__(kBlock) void k_iter_primal_rb(CsrView<M> AT, const int* __restrict__ blocks,
                                                           SplitPlan sp, CsrView<M> Q, const V* __restrict__ y,
                                                           const V* __restrict__ x, const V* __restrict__ c,
                                                           const V* __restrict__ l, const V* __restrict__ u,
                                                           V l0, V u0, const double* __restrict__ prm,
                                                           V* __restrict__ xn, V* __restrict__ xbar,
                                                           V* __restrict__ sumx) {
    __shared__ __align__(16) V lds[kTile + kWin];
    // the row's operands, and the step size with them: prm is written by the previous launch, so its load is
    // issued once the block is known, in flight together with the block's index / value loads (and it is
    // read in V already, so nothing waits for it before the row update).  l / u null: every column has the
    // bounds l0 / u0 (x >= 0 with no upper bounds, say), and the kernel streams neither vector.
    struct Ops {
        V x, c, l, u, s, tau;
    };
    stream_block_pf<M, V, 1, true>(
        AT, blocks, xcd_block_id(), [&](int, int i) { return y[i]; },
        [&](int j) { return Ops{x[j], c[j], l ? l[j] : l0, u ? u[j] : u0, sumx[j], step_as<V>(prm, 0)}; },
        [&](int j, V (&aty)[1], Ops& o) {
#pragma clang fp contract(off)
            V qx = V(0);
            if (HASQ) {  // Q row summed sequentially by the owning thread
                for (int p = Q.ptr[j]; p < Q.ptr[j + 1]; ++p) {
                    const V prod = widen<V>(Q.val[p]) * x[Q.idx[p]];
                    qx = qx + prod;
                }
            }
            V sm = o.s;
            primal_update<V>(o.x, aty[0], qx, o.c, o.l, o.u, o.tau, xn[j], xbar[j], sm);
            sumx[j] = sm;
        },
        lds, sp);
}

template <class M, class V>
__global__ __launch_bounds__(kBlock) void k_iter_dual_rb(CsrView<M> A, const int* __restrict__ blocks,
                                                         SplitPlan sp, const V* __restrict__ xbar, const V* __restrict__ y,
                                                         const V* __restrict__ lo, const V* __restrict__ hi,
                                                         const double* __restrict__ prm, V* __restrict__ yn,
                                                         V* __restrict__ sumy) {
    __shared__ __align__(16) V lds[kTile + kWin];
    struct Ops {
        V y, lo, hi, s, sigma;
    };
    stream_block_pf<M, V, 1, true>(
        A, blocks, xcd_block_id(), [&](int, int j) { return xbar[j]; },
        [&](int i) { return Ops{y[i], lo[i], hi[i], sumy[i], step_as<V>(prm, 1)}; },
        [&](int i, V (&axb)[1], Ops& o) {
            V sm = o.s;
            dual_update<V>(o.y, axb[0], o.lo, o.hi, o.sigma, yn[i], sm);
            sumy[i] = sm;
        },
        lds, sp);
}

// Streams a buffer once (grid-stride float4 loads); the store is never taken but keeps the loads.
__global__ __launch_bounds__(kBlock) void k_stream_read(const float4* __restrict__ a, int64_t n4,
                                                        float* __restrict__ sink) {
    float acc = 0.f;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n4; i += (int64_t)gridDim.x * kBlock) {
        const float4 v = a[i];
        acc += v.x + v.y + v.z + v.w;
    }
    if (acc == 1234.5f) sink[blockIdx.x] = acc;
}

__global__ void k_set_steps(double* prm, double tau, double sigma) {
    if (threadIdx.x == 0) {
        prm[0] = tau;
        prm[1] = sigma;
        float* f = reinterpret_cast<float*>(prm + 2);
        f[0] = (float)tau;
        f[1] = (float)sigma;
    }
}

// ================================================================ check kernels
// (xa, ya) = V(inv_k * double(sum)) over [0, n) then [0, m)  (src/backend_cuda.cu:135-139)
template <class V>
__global__ __launch_bounds__(kBlock) void k_average(int n, int m, double inv_k, const V* __restrict__ sumx,
                                                    const V* __restrict__ sumy, V* __restrict__ xa,
                                                    V* __restrict__ ya) {
    const int64_t total = (int64_t)n + m;
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < total; t += (int64_t)gridDim.x * kBlock) {
        if (t < n) xa[t] = V(inv_k * double(sumx[t]));
        else ya[t - n] = V(inv_k * double(sumy[t - n]));
    }
}

// Reduced-cost terms of one column (src/backend_cpu.cpp:120-127): z = cw c - aty + qx; a finite bound on
// z's side contributes z * bound to the dual objective, an infinite one contributes z to the residual.
__device__ __forceinline__ void dual_terms(double z, double lj, double uj, double cres, double& r2, double& ob) {
    double r = 0.0, o = 0.0;
    if (z > 0) { if (isfinite(lj)) o = z * lj; else r = z; }
    else if (z < 0) { if (isfinite(uj)) o = z * uj; else r = z; }
    r /= cres;
    r2 = r * r;
    ob = o;
}

// The distance of a ray component d to the recession cone of the interval [lo, hi]: {0} when both ends are
// finite, [0, inf) when only lo is, (-inf, 0] when only hi is, everything when neither (the dual-infeasibility
// test of a primal ray, for rows with d = (A d)_i and for variables with d = d_j).
__device__ __forceinline__ double cone_excess(double d, double lo, double hi) {
    const bool fl = isfinite(lo), fh = isfinite(hi);
    if (fl && fh) return d;
    if (fl) return d < 0 ? d : 0.0;
    if (fh) return d > 0 ? d : 0.0;
    return 0.0;
}

// Row-bound dual objective term (src/backend_cpu.cpp:137-139).
__device__ __forceinline__ double row_obj_term(double yi, double lw, double hg) {
    if (yi > 0 && isfinite(lw)) return yi * lw;
    if (yi < 0 && isfinite(hg)) return yi * hg;
    return 0.0;
}

// One pass over A^T (and Q): out[0..11] as documented for ganit_hip_pdhg_check.  Vectors are stored in V;
// the row sums are formed in C (C = V reproduces the level's own evaluation, C = double with FP64 matrix
// values is the FP64 evaluation of a reduced-precision iterate).
// Per-thread accumulator of the check sums: FP64, as the reference's reductions; or, with GANIT_CHECK_KAHAN,
// every term rounded to FP32 and added with Kahan compensation in FP32.  Block partials and the fold are
// FP64 either way: compensated FP32 within a thread, FP64 across threads (the oracle's Kahan reductions).
template <int N>
struct CheckAcc {
    double d[N];
    float s[N], c[N];
    bool kahan;
    __device__ explicit CheckAcc(bool k) : kahan(k) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            d[i] = 0.0;
            s[i] = 0.0f;
            c[i] = 0.0f;
        }
    }
    __device__ __forceinline__ void add(int i, double v) {
#pragma clang fp contract(off)
        if (kahan) {
            const float yv = float(v) - c[i];
            const float t = s[i] + yv;
            c[i] = (t - s[i]) - yv;
            s[i] = t;
        } else {
            d[i] += v;
        }
    }
    __device__ __forceinline__ void totals(double (&o)[N]) const {
#pragma unroll
        for (int i = 0; i < N; ++i) o[i] = kahan ? double(s[i]) : d[i];
    }
};

template <int LPR, class M, class V, class C, bool HASQ, class VI = V>
__global__ __launch_bounds__(kBlock) void k_check_n(CsrView<M> AT, CsrView<M> Q, const VI* __restrict__ x,
                                                    const VI* __restrict__ y, const VI* __restrict__ xa,
                                                    const VI* __restrict__ ya, const VI* __restrict__ xr,
                                                    const VI* __restrict__ yr, const V* __restrict__ c,
                                                    const V* __restrict__ l, const V* __restrict__ u,
                                                    const V* __restrict__ cres, int flags, ReduceSlot slot) {
    __shared__ double lds[(kBlock / kWave) * kNOutN];
    CheckAcc<kNOutN> acc((flags & GANIT_CHECK_KAHAN) != 0);
    const bool avg = flags & GANIT_CHECK_AVG, ray = flags & GANIT_CHECK_RAY, dray = flags & GANIT_CHECK_DRAY;
    constexpr int kGroups = kBlock / LPR;
    const int lane = threadIdx.x % LPR;
    for (int64_t g = (int64_t)xcd_block_id() * kGroups + threadIdx.x / LPR; g < AT.rows;
         g += (int64_t)gridDim.x * kGroups) {
        const int j = (int)g;
        C sc = C(0), sa = C(0), sr = C(0), qc = C(0), qa = C(0);
        for (int p = AT.ptr[j] + lane; p < AT.ptr[j + 1]; p += LPR) {
            const C a = widen<C>(AT.val[p]);
            const int i = AT.idx[p];
            sc += a * C(y[i]);
            if (avg) sa += a * C(ya[i]);
            if (ray) sr += a * C(V(V(y[i]) - V(yr[i])));
        }
        if (HASQ) {
            for (int p = Q.ptr[j] + lane; p < Q.ptr[j + 1]; p += LPR) {
                const C a = widen<C>(Q.val[p]);
                const int i = Q.idx[p];
                qc += a * C(x[i]);
                if (avg) qa += a * C(xa[i]);
            }
        }
        sc = group_sum<LPR>(sc);
        if (avg) sa = group_sum<LPR>(sa);
        if (ray) sr = group_sum<LPR>(sr);
        if (HASQ) {
            qc = group_sum<LPR>(qc);
            if (avg) qa = group_sum<LPR>(qa);
        }
        if (lane == 0) {
            const double cj = double(c[j]), lj = double(l[j]), uj = double(u[j]), cr = double(cres[j]);
            const double xj = double(x[j]);
            double r2, ob;
            dual_terms(cj - double(sc) + (HASQ ? double(qc) : 0.0), lj, uj, cr, r2, ob);
            acc.add(0, r2);
            acc.add(1, ob);
            acc.add(2, cj * xj);
            if (HASQ) acc.add(3, xj * double(qc));
            if (avg) {
                const double xaj = double(xa[j]), xrj = double(xr[j]);
                dual_terms(cj - double(sa) + (HASQ ? double(qa) : 0.0), lj, uj, cr, r2, ob);
                acc.add(4, r2);
                acc.add(5, ob);
                acc.add(6, cj * xaj);
                if (HASQ) acc.add(7, xaj * double(qa));
                const double d0 = xj - xrj, d1 = xaj - xrj;
                acc.add(10, d0 * d0);
                acc.add(11, d1 * d1);
            }
            if (ray) {  // homogeneous dual: cw = 0, no Qx (src/pdhg_impl.hpp:114)
                dual_terms(0.0 * cj - double(sr) + 0.0, lj, uj, cr, r2, ob);
                acc.add(8, r2);
                acc.add(9, ob);
            }
            if (dray) {  // the primal ray d = x - xr: c'd and d's excess over the bounds' recession cone
                const double d = xj - double(xr[j]);
                acc.add(12, cj * d);
                const double e = cone_excess(d, lj, uj) * cr;
                acc.add(13, e * e);
            }
        }
    }
    {
        double tot[kNOutN];
        acc.totals(tot);
        block_partial<kNOutN>(tot, slot, lds);
    }
}

// One pass over A: out[12..18] as documented for ganit_hip_pdhg_check (sums in C as for k_check_n).
template <int LPR, class M, class V, class C, class VI = V>
__global__ __launch_bounds__(kBlock) void k_check_m(CsrView<M> A, const VI* __restrict__ x,
                                                    const VI* __restrict__ xa, const VI* __restrict__ y,
                                                    const VI* __restrict__ ya, const VI* __restrict__ yr,
                                                    const VI* __restrict__ xr, const V* __restrict__ lo,
                                                    const V* __restrict__ hi, const V* __restrict__ rres, int flags,
                                                    ReduceSlot slot) {
    __shared__ double lds[(kBlock / kWave) * kNOutM];
    CheckAcc<kNOutM> acc((flags & GANIT_CHECK_KAHAN) != 0);
    const bool avg = flags & GANIT_CHECK_AVG, ray = flags & GANIT_CHECK_RAY, dray = flags & GANIT_CHECK_DRAY;
    constexpr int kGroups = kBlock / LPR;
    const int lane = threadIdx.x % LPR;
    for (int64_t g = (int64_t)xcd_block_id() * kGroups + threadIdx.x / LPR; g < A.rows;
         g += (int64_t)gridDim.x * kGroups) {
        const int i = (int)g;
        C sc = C(0), sa = C(0), sd = C(0);
        for (int p = A.ptr[i] + lane; p < A.ptr[i + 1]; p += LPR) {
            const C a = widen<C>(A.val[p]);
            const int j = A.idx[p];
            sc += a * C(x[j]);
            if (avg) sa += a * C(xa[j]);
            if (dray) sd += a * C(V(V(x[j]) - V(xr[j])));
        }
        sc = group_sum<LPR>(sc);
        if (avg) sa = group_sum<LPR>(sa);
        if (dray) sd = group_sum<LPR>(sd);
        if (lane == 0) {
            const double lw = double(lo[i]), hg = double(hi[i]), rr = double(rres[i]), yi = double(y[i]);
            auto pres = [&](double a) {  // src/backend_cpu.cpp:106-109
                const double p = fmin(fmax(a, lw), hg);
                const double r = (a - p) / rr;
                return r * r;
            };
            acc.add(0, pres(double(sc)));
            acc.add(1, row_obj_term(yi, lw, hg));
            if (avg) {
                const double yai = double(ya[i]), yri = double(yr[i]);
                acc.add(2, pres(double(sa)));
                acc.add(3, row_obj_term(yai, lw, hg));
                const double d0 = yi - yri, d1 = yai - yri;
                acc.add(5, d0 * d0);
                acc.add(6, d1 * d1);
            }
            if (ray) acc.add(4, row_obj_term(double(V(V(y[i]) - V(yr[i]))), lw, hg));
            if (dray) {  // (A d)_i's excess over the row bounds' recession cone, in the residual's units
                const double e = cone_excess(double(sd), lw, hg) / rr;
                acc.add(7, e * e);
            }
        }
    }
    {
        double tot[kNOutM];
        acc.totals(tot);
        block_partial<kNOutM>(tot, slot, lds);
    }
}

// The two check passes on LDS-staged row blocks, both in one launch: sums in C as above.  The averages
// (xa, ya) = (sumx, sumy) * inv_k are formed where they are read -- each gathered entry, and each owning row's
// own entry, which that row's thread also stores for a later adopt/restart -- with k_average's arithmetic, so
// the values are the same bits and no averaging launch precedes the passes.
template <class M, class V, class VI>
struct CheckArgs {
    CsrView<M> AT, A, Q;
    const int *bn, *bm;  // row-block descriptors of A^T and A
    const VI *x, *y, *sumx, *sumy, *xr, *yr;
    VI *xa, *ya;
    const V *c, *l, *u, *cres, *lo, *hi, *rres;
    double inv_k;
    int flags;
    double *part_n, *part_m;  // [blocks][kNOutN], [blocks][kNOutM]
};

// One block of the pass over A^T (and Q): per column, the reduced-cost terms of the current point, the
// average and the ray y - yr, the objective terms and the distance terms; the block's sums go to partial `at`.
template <class M, class V, class C, bool HASQ, class VI>
__device__ __forceinline__ void check_n_block(const CheckArgs<M, V, VI>& g, int b, const SplitPlan& sp, int at,
                                              C* lds, int cap) {
    CheckAcc<kNOutN> acc((g.flags & GANIT_CHECK_KAHAN) != 0);
    const int flags = g.flags;
    const bool avg = flags & GANIT_CHECK_AVG, ray = flags & GANIT_CHECK_RAY, dray = flags & GANIT_CHECK_DRAY;
    const double ik = g.inv_k;
    const VI *__restrict__ x = g.x, *__restrict__ y = g.y, *__restrict__ sumx = g.sumx, *__restrict__ sumy = g.sumy,
             *__restrict__ yr = g.yr;
    // the column's operands, loaded before the sweep; an unrequested iterate reads x in its place (a valid
    // address, so the loads need no branch)
    const VI* __restrict__ sxp = avg ? sumx : x;
    const VI* __restrict__ xrp = (avg || dray) ? g.xr : x;
    struct Ops {
        V c, l, u, cres;
        VI x, sx, xr;
    };
    stream_block_seq_pf<M, C, 3, true>(
        g.AT, g.bn, b,
        [&](int k, int i) {
            return k == 0   ? C(y[i])
                   : k == 1 ? (avg ? C(VI(ik * double(sumy[i]))) : C(0))
                            : (ray ? C(V(V(y[i]) - V(yr[i]))) : C(0));
        },
        [&](int j) { return Ops{g.c[j], g.l[j], g.u[j], g.cres[j], x[j], sxp[j], xrp[j]}; },
        [&](int j, C (&t)[3], Ops& o) {
#pragma clang fp contract(off)
            C qc = C(0), qa = C(0);
            if (HASQ) {
                for (int p = g.Q.ptr[j]; p < g.Q.ptr[j + 1]; ++p) {
                    const C a = widen<C>(g.Q.val[p]);
                    const int q = g.Q.idx[p];
                    const C pc = a * C(x[q]);
                    qc = qc + pc;
                    if (avg) {
                        const C pa = a * C(VI(ik * double(sumx[q])));
                        qa = qa + pa;
                    }
                }
            }
            const double cj = double(o.c), lj = double(o.l), uj = double(o.u), cr = double(o.cres);
            const double xj = double(o.x);
            double r2, ob;
            dual_terms(cj - double(t[0]) + (HASQ ? double(qc) : 0.0), lj, uj, cr, r2, ob);
            acc.add(0, r2);
            acc.add(1, ob);
            acc.add(2, cj * xj);
            if (HASQ) acc.add(3, xj * double(qc));
            if (avg) {
                const VI xav = VI(ik * double(o.sx));
                g.xa[j] = xav;
                const double xaj = double(xav), xrj = double(o.xr);
                dual_terms(cj - double(t[1]) + (HASQ ? double(qa) : 0.0), lj, uj, cr, r2, ob);
                acc.add(4, r2);
                acc.add(5, ob);
                acc.add(6, cj * xaj);
                if (HASQ) acc.add(7, xaj * double(qa));
                const double d0 = xj - xrj, d1 = xaj - xrj;
                acc.add(10, d0 * d0);
                acc.add(11, d1 * d1);
            }
            if (ray) {
                dual_terms(0.0 * cj - double(t[2]) + 0.0, lj, uj, cr, r2, ob);
                acc.add(8, r2);
                acc.add(9, ob);
            }
            if (dray) {  // the primal ray d = x - xr: c'd and d's excess over the bounds' recession cone
                const double d = xj - double(o.xr);
                acc.add(12, cj * d);
                const double e = cone_excess(d, lj, uj) * cr;
                acc.add(13, e * e);
            }
        },
        lds, sp, 1u | (avg ? 2u : 0u) | (ray ? 4u : 0u));
    __syncthreads();
    double tot[kNOutN];
    acc.totals(tot);
    block_partial_wide<kNOutN>(tot, ReduceSlot{g.part_n, nullptr}, reinterpret_cast<double*>(lds), cap, at);
}

// One block of the pass over A: per row, the primal residual of the current point and the average, the row
// objective terms of y, ya and the ray, the distance terms and the primal-ray cone excess.
template <class M, class V, class C, class VI>
__device__ __forceinline__ void check_m_block(const CheckArgs<M, V, VI>& g, int b, const SplitPlan& sp, int at,
                                              C* lds, int cap) {
    CheckAcc<kNOutM> acc((g.flags & GANIT_CHECK_KAHAN) != 0);
    const int flags = g.flags;
    const bool avg = flags & GANIT_CHECK_AVG, ray = flags & GANIT_CHECK_RAY, dray = flags & GANIT_CHECK_DRAY;
    const double ik = g.inv_k;
    const VI *__restrict__ x = g.x, *__restrict__ y = g.y, *__restrict__ sumx = g.sumx, *__restrict__ xr = g.xr;
    // the row's operands, loaded before the sweep (an unrequested iterate reads y in its place)
    const VI* __restrict__ syp = avg ? g.sumy : y;
    const VI* __restrict__ yrp = (avg || ray) ? g.yr : y;
    struct Ops {
        V lo, hi, rres;
        VI y, sy, yr;
    };
    stream_block_seq_pf<M, C, 3, true>(
        g.A, g.bm, b,
        [&](int k, int j) {
            return k == 0   ? C(x[j])
                   : k == 1 ? (avg ? C(VI(ik * double(sumx[j]))) : C(0))
                            : (dray ? C(V(V(x[j]) - V(xr[j]))) : C(0));
        },
        [&](int i) { return Ops{g.lo[i], g.hi[i], g.rres[i], y[i], syp[i], yrp[i]}; },
        [&](int i, C (&t)[3], Ops& o) {
            const double lw = double(o.lo), hg = double(o.hi), rr = double(o.rres), yi = double(o.y);
            auto pres = [&](double a) {
                const double p = fmin(fmax(a, lw), hg);
                const double r = (a - p) / rr;
                return r * r;
            };
            acc.add(0, pres(double(t[0])));
            acc.add(1, row_obj_term(yi, lw, hg));
            if (avg) {
                const VI yav = VI(ik * double(o.sy));
                g.ya[i] = yav;
                const double yai = double(yav), yri = double(o.yr);
                acc.add(2, pres(double(t[1])));
                acc.add(3, row_obj_term(yai, lw, hg));
                const double d0 = yi - yri, d1 = yai - yri;
                acc.add(5, d0 * d0);
                acc.add(6, d1 * d1);
            }
            if (ray) acc.add(4, row_obj_term(double(V(V(o.y) - V(o.yr))), lw, hg));
            if (dray) {  // (A d)_i's excess over the row bounds' recession cone, in the residual's units
                const double e = cone_excess(double(t[2]), lw, hg) / rr;
                acc.add(7, e * e);
            }
        },
        lds, sp, 1u | (avg ? 2u : 0u) | (dray ? 4u : 0u));
    __syncthreads();
    double tot[kNOutM];
    acc.totals(tot);
    block_partial_wide<kNOutM>(tot, ReduceSlot{g.part_m, nullptr}, reinterpret_cast<double*>(lds), cap, at);
}

// Both row-block passes in one grid: logical blocks [0, gn) are blocks of A^T (plan spn, partials from
// at_n), the rest blocks of A (plan spm, partials from at_m).  With the main plans this is the whole pass;
// with the split_finish plans, the per-split-row folds that follow it.
template <class M, class V, class C, bool HASQ, class VI>
__global__ __launch_bounds__(kBlock) void k_check_rb(CheckArgs<M, V, VI> g, SplitPlan spn, SplitPlan spm, int gn,
                                                     int at_n, int at_m) {
    // the staging tile and window; after the sweep, the scratch of the block reduction (hence 16-byte aligned)
    __shared__ __align__(16) C lds[kTile + kWin];
    constexpr int cap = (int)(sizeof(lds) / sizeof(double));
    const int id = xcd_block_id();
    if (id < gn) check_n_block<M, V, C, HASQ, VI>(g, id, spn, at_n + id, lds, cap);
    else check_m_block<M, V, C, VI>(g, id - gn, spm, at_m + id - gn, lds, cap);
}

// Folds the per-block partials of both check passes, one block per output, each in block order.
__global__ __launch_bounds__(kBlock) void k_check_finish(const double* __restrict__ part_n, int nb_n,
                                                         const double* __restrict__ part_m, int nb_m,
                                                         double* __restrict__ out) {
    __shared__ double lds[kBlock / kWave];
    // outputs 0..11: A^T pass 0..11; 12..18: A pass 0..6; 19: A pass 7; 20, 21: A^T pass 12, 13
    const int k = blockIdx.x;  // one block per output
    if (k < 12) sum_partials_one(part_n, kNOutN, k, nb_n, out + k, lds);
    else if (k < 20) sum_partials_one(part_m, kNOutM, k - 12, nb_m, out + k, lds);
    else sum_partials_one(part_n, kNOutN, k - 8, nb_n, out + k, lds);
}

// ================================================================ restart / hand-over kernels
template <class V>
__global__ __launch_bounds__(kBlock) void k_restart(int n, int m, int adopt, V* __restrict__ x,
                                                    const V* __restrict__ xa, V* __restrict__ xr,
                                                    V* __restrict__ sumx, V* __restrict__ y, const V* __restrict__ ya,
                                                    V* __restrict__ yr, V* __restrict__ sumy, int mode) {
    // mode 0: restart (adopt?, then restart point = iterate, clear sums); mode 1: adopt only
    const int64_t total = (int64_t)n + m;
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < total; t += (int64_t)gridDim.x * kBlock) {
        if (t < n) {
            const V v = adopt ? xa[t] : x[t];
            x[t] = v;
            if (mode == 0) {
                xr[t] = v;
                sumx[t] = V(0);
            }
        } else {
            const int64_t i = t - n;
            const V v = adopt ? ya[i] : y[i];
            y[i] = v;
            if (mode == 0) {
                yr[i] = v;
                sumy[i] = V(0);
            }
        }
    }
}

template <class S, class D>
__global__ __launch_bounds__(kBlock) void k_handover(int64_t n, const S* __restrict__ s, D* __restrict__ d) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        d[i] = D(s[i]);
}

// ================================================================ host object (pdhg_state.hpp)
void upload_as(Ctx* ctx, int vdt, const double* h, int64_t n, void* d) {
    if (n == 0) return;
    if (vdt == GANIT_F64) {
        GH_CHECK(hipMemcpyAsync(d, h, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
        GH_CHECK(hipStreamSynchronize(ctx->stream));
    } else {
        std::vector<float> t(h, h + n);
        GH_CHECK(hipMemcpyAsync(d, t.data(), sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
        GH_CHECK(hipStreamSynchronize(ctx->stream));
    }
}

void download_as(Ctx* ctx, int vdt, const void* d, int64_t n, double* h) {
    if (n == 0 || !h) return;
    if (vdt == GANIT_F64) {
        GH_CHECK(hipMemcpyAsync(h, d, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
        GH_CHECK(hipStreamSynchronize(ctx->stream));
    } else {
        std::vector<float> t(n);
        GH_CHECK(hipMemcpyAsync(t.data(), d, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
        GH_CHECK(hipStreamSynchronize(ctx->stream));
        for (int64_t i = 0; i < n; ++i) h[i] = t[i];
    }
}

template <class V>
__global__ __launch_bounds__(kBlock) void k_zero2(int n, int m, V* __restrict__ a, V* __restrict__ b) {
    const int64_t total = (int64_t)n + m;
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < total; t += (int64_t)gridDim.x * kBlock) {
        if (t < n) a[t] = V(0);
        else b[t - n] = V(0);
    }
}

unsigned ew_grid2(int64_t n) { return std::min<unsigned>(grid_for(n), 8192u); }

template <class V>
void clear_sums(Pdhg* pd, VecSet& s) {
    hipLaunchKernelGGL(k_zero2<V>, dim3(ew_grid2((int64_t)pd->n + pd->m)), dim3(kBlock), 0, pd->ctx->stream, pd->n,
                       pd->m, (V*)s.sumx, (V*)s.sumy);
    GH_CHECK(hipGetLastError());
}

// Launch one of the two fused kernels of an iteration (0: primal over the columns, 1: dual over the rows)
// on the current iterate, without advancing the iterate parity.
template <class M, class V>
void launch_one(Pdhg* pd, VecSet& s, int which) {
    hipStream_t st = pd->ctx->stream;
    const V* x = (const V*)s.X[s.cur];
    V* xn = (V*)s.X[1 - s.cur];
    const V* y = (const V*)s.Y[s.cur];
    V* yn = (V*)s.Y[1 - s.cur];
    if (which == 0 && pd->n > 0 && pd->AT->lanes == 0) {
        const CsrView<M> at = pd->AT->view<M>();
        // main pass, then (split rows only) the finish pass that folds the chunk partials
        for (int pass = 0; pass < (pd->AT->nsplit > 0 ? 2 : 1); ++pass) {
            const dim3 g(pass ? (unsigned)pd->AT->nsplit : pd->AT->rb_grid());
            const SplitPlan sp = pass ? pd->AT->split_finish() : pd->AT->split();
            const V* l = pd->l_uniform ? nullptr : (const V*)s.l;
            const V* u = pd->u_uniform ? nullptr : (const V*)s.u;
            const V l0 = V(pd->l_value), u0 = V(pd->u_value);
            if (pd->hasq)
                hipLaunchKernelGGL((k_iter_primal_rb<M, V, true>), g, dim3(kBlock), 0, st, at, pd->AT->blocks, sp,
                                   pd->Q->view<M>(), y, x, (const V*)s.c, l, u, l0, u0, pd->prm, xn, (V*)s.xbar,
                                   (V*)s.sumx);
            else
                hipLaunchKernelGGL((k_iter_primal_rb<M, V, false>), g, dim3(kBlock), 0, st, at, pd->AT->blocks, sp, at,
                                   y, x, (const V*)s.c, l, u, l0, u0, pd->prm, xn, (V*)s.xbar, (V*)s.sumx);
        }
    } else if (which == 0 && pd->n > 0) {
        const CsrView<M> at = pd->AT->view<M>();
        with_lanes(pd->AT->lanes, [&](auto L) {
            constexpr int LPR = decltype(L)::value;
            const unsigned g = grid_for((int64_t)pd->n * LPR);
            if (pd->hasq)
                hipLaunchKernelGGL((k_iter_primal<LPR, M, V, true>), dim3(g), dim3(kBlock), 0, st, at,
                                   pd->Q->view<M>(), y, x, (const V*)s.c, (const V*)s.l, (const V*)s.u, pd->prm, xn,
                                   (V*)s.xbar, (V*)s.sumx);
            else
                hipLaunchKernelGGL((k_iter_primal<LPR, M, V, false>), dim3(g), dim3(kBlock), 0, st, at, at, y, x,
                                   (const V*)s.c, (const V*)s.l, (const V*)s.u, pd->prm, xn, (V*)s.xbar, (V*)s.sumx);
        });
    }
    if (which == 1 && pd->m > 0 && pd->A->lanes == 0) {
        for (int pass = 0; pass < (pd->A->nsplit > 0 ? 2 : 1); ++pass)
            hipLaunchKernelGGL((k_iter_dual_rb<M, V>), dim3(pass ? (unsigned)pd->A->nsplit : pd->A->rb_grid()),
                               dim3(kBlock), 0, st, pd->A->view<M>(), pd->A->blocks,
                               pass ? pd->A->split_finish() : pd->A->split(), (const V*)s.xbar, y, (const V*)s.lo,
                               (const V*)s.hi, pd->prm, yn, (V*)s.sumy);
    } else if (which == 1 && pd->m > 0) {
        const CsrView<M> a = pd->A->view<M>();
        with_lanes(pd->A->lanes, [&](auto L) {
            constexpr int LPR = decltype(L)::value;
            hipLaunchKernelGGL((k_iter_dual<LPR, M, V>), dim3(grid_for((int64_t)pd->m * LPR)), dim3(kBlock), 0, st,
                               a, (const V*)s.xbar, y, (const V*)s.lo, (const V*)s.hi, pd->prm, yn, (V*)s.sumy);
        });
    }
    GH_CHECK(hipGetLastError());
}

template <class M, class V>
void launch_iteration(Pdhg* pd, VecSet& s) {
    launch_one<M, V>(pd, s, 0);
    launch_one<M, V>(pd, s, 1);
    s.cur ^= 1;
}

// Iterates (x, y, averages, restart point) of type VI from working set `it`, problem data of type V from
// working set `s` (the same set unless an FP32 iterate is evaluated against the FP64 data: VI = float,
// V = double, each iterate entry widened exactly where it is read).
template <class M, class V, class C, class VI = V>
void run_check(Pdhg* pd, VecSet& it, const VecSet& s, int flags, double inv_k, double* host_out) {
    hipStream_t st = pd->ctx->stream;
    const VI* x = (const VI*)it.X[it.cur];
    const VI* y = (const VI*)it.Y[it.cur];
    const VI *xa = (const VI*)it.xa, *ya = (const VI*)it.ya, *xr = (const VI*)it.xr, *yr = (const VI*)it.yr;
    // matrices on row blocks go through the fused pass (which forms the averages itself); a lanes-per-row
    // matrix has its own pass, which reads the averages from memory
    const bool rb_n = pd->n > 0 && pd->AT->lanes == 0, rb_m = pd->m > 0 && pd->A->lanes == 0;
    const bool lpr_n = pd->n > 0 && !rb_n, lpr_m = pd->m > 0 && !rb_m;
    if ((flags & GANIT_CHECK_AVG) && (lpr_n || lpr_m)) {
        hipLaunchKernelGGL(k_average<VI>, dim3(ew_grid2((int64_t)pd->n + pd->m)), dim3(kBlock), 0, st, pd->n, pd->m,
                           inv_k, (const VI*)it.sumx, (const VI*)it.sumy, (VI*)it.xa, (VI*)it.ya);
    }
    unsigned nb_n = 0, nb_m = 0;  // partials written by the passes over A^T and A
    const CsrView<M> at = pd->AT->view<M>();
    const CsrView<M> a = pd->A->view<M>();
    // the row-block passes: one launch over both matrices' blocks (rb_grid partials each), then, if either
    // has split rows, one launch of their folds (one partial per split row)
    const unsigned gn = rb_n ? pd->AT->rb_grid() : 0u, gm = rb_m ? pd->A->rb_grid() : 0u;
    const unsigned fn = rb_n ? (unsigned)pd->AT->nsplit : 0u, fm = rb_m ? (unsigned)pd->A->nsplit : 0u;
    GH_REQUIRE(gn + fn <= pd->part_cap && gm + fm <= pd->part_cap, "check grid exceeds the partial buffers");
    if (gn + gm > 0) {
        CheckArgs<M, V, VI> g{at, a, pd->hasq ? pd->Q->view<M>() : at, pd->AT->blocks, pd->A->blocks, x, y,
                              (const VI*)it.sumx, (const VI*)it.sumy, xr, yr, (VI*)it.xa, (VI*)it.ya,
                              (const V*)s.c, (const V*)s.l, (const V*)s.u, (const V*)s.cres, (const V*)s.lo,
                              (const V*)s.hi, (const V*)s.rres, inv_k, flags, pd->part_n, pd->part_m};
        auto launch = [&](unsigned grid, SplitPlan spn, SplitPlan spm, unsigned g_n, unsigned at_n, unsigned at_m) {
            if (pd->hasq)
                hipLaunchKernelGGL((k_check_rb<M, V, C, true, VI>), dim3(grid), dim3(kBlock), 0, st, g, spn, spm,
                                   (int)g_n, (int)at_n, (int)at_m);
            else
                hipLaunchKernelGGL((k_check_rb<M, V, C, false, VI>), dim3(grid), dim3(kBlock), 0, st, g, spn, spm,
                                   (int)g_n, (int)at_n, (int)at_m);
        };
        launch(gn + gm, pd->AT->split(), pd->A->split(), gn, 0, 0);
        if (fn + fm > 0) launch(fn + fm, pd->AT->split_finish(), pd->A->split_finish(), fn, gn, gm);
        nb_n = gn + fn;
        nb_m = gm + fm;
    }
    if (lpr_n) with_lanes(pd->AT->lanes, [&](auto L) {
        constexpr int LPR = decltype(L)::value;
        const unsigned g = (unsigned)std::min<int64_t>(((int64_t)pd->n * LPR + kBlock - 1) / kBlock, kCheckGridMax);
        nb_n = g;
        const ReduceSlot sn{pd->part_n, nullptr};
        if (pd->hasq)
            hipLaunchKernelGGL((k_check_n<LPR, M, V, C, true, VI>), dim3(g), dim3(kBlock), 0, st, at,
                               pd->Q->view<M>(), x, y, xa, ya, xr, yr, (const V*)s.c, (const V*)s.l, (const V*)s.u,
                               (const V*)s.cres, flags, sn);
        else
            hipLaunchKernelGGL((k_check_n<LPR, M, V, C, false, VI>), dim3(g), dim3(kBlock), 0, st, at, at, x, y, xa,
                               ya, xr, yr, (const V*)s.c, (const V*)s.l, (const V*)s.u, (const V*)s.cres, flags, sn);
    });
    if (lpr_m) with_lanes(pd->A->lanes, [&](auto L) {
        constexpr int LPR = decltype(L)::value;
        const unsigned g = (unsigned)std::min<int64_t>(((int64_t)pd->m * LPR + kBlock - 1) / kBlock, kCheckGridMax);
        nb_m = g;
        const ReduceSlot sm{pd->part_m, nullptr};
        hipLaunchKernelGGL((k_check_m<LPR, M, V, C, VI>), dim3(g), dim3(kBlock), 0, st, a, x, xa, y, ya, yr, xr,
                           (const V*)s.lo, (const V*)s.hi, (const V*)s.rres, flags, sm);
    });
    GH_REQUIRE(nb_n <= pd->part_cap && nb_m <= pd->part_cap, "check grid exceeds the partial buffers");
    // the fold writes the sums straight into host-coherent pinned memory: no copy follows it.  The entries are
    // set to NaN first, so a value the GPU did not write can never pass for a result.
    double* res = pd->check_host;
    std::fill(res, res + GANIT_CHECK_NOUT, std::numeric_limits<double>::quiet_NaN());
    hipLaunchKernelGGL(k_check_finish, dim3(GANIT_CHECK_NOUT), dim3(kBlock), 0, st, pd->part_n, (int)nb_n,
                       pd->part_m, (int)nb_m, res);
    GH_CHECK(hipGetLastError());
    GH_CHECK(hipStreamSynchronize(st));
    std::copy(res, res + GANIT_CHECK_NOUT, host_out);
}

}  // namespace ganit_hip

using namespace ganit_hip;

extern "C" {

int ganit_hip_pdhg_create(ganit_hip_ctx* h, const ganit_hip_lp_desc* lp, unsigned mdt_mask, unsigned vdt_mask,
                          ganit_hip_pdhg** out) {
    return guarded([&] {
        auto* ctx = reinterpret_cast<Ctx*>(h);
        GH_REQUIRE(ctx && lp && out, "null argument");
        GH_REQUIRE(lp->m >= 0 && lp->n >= 0, "negative problem dimension");
        GH_REQUIRE((mdt_mask & 7u) && (vdt_mask & 3u), "empty dtype mask");
        GH_REQUIRE(lp->c && lp->l && lp->u && lp->cres && (lp->m == 0 || (lp->lo && lp->hi && lp->rres)),
                   "null problem vector");
        auto* pd = new Pdhg;
        try {
            pd->ctx = ctx;
            pd->m = lp->m;
            pd->n = lp->n;
            pd->A = csr_create(ctx, lp->m, lp->n, lp->a_ptr, lp->a_idx, lp->a_val, mdt_mask);
            pd->AT = csr_create(ctx, lp->n, lp->m, lp->at_ptr, lp->at_idx, lp->at_val, mdt_mask);
            GH_REQUIRE(pd->A->nnz == pd->AT->nnz, "A and its transpose differ in nnz");
            pd->hasq = lp->q_ptr && lp->q_ptr[lp->n] > 0;
            if (pd->hasq) pd->Q = csr_create(ctx, lp->n, lp->n, lp->q_ptr, lp->q_idx, lp->q_val, mdt_mask);
            if (pd->hasq) pd->Q->lanes = pd->AT->lanes;
            // column bounds that are the same bits in every column go to the primal kernel as two scalars
            auto uniform = [&](const double* v, bool& flag, double& value) {
                flag = lp->n > 0 && std::all_of(v, v + lp->n, [&](double t) {
                    return std::memcmp(&t, v, sizeof(double)) == 0;
                });
                value = lp->n > 0 ? v[0] : 0.0;
            };
            uniform(lp->l, pd->l_uniform, pd->l_value);
            uniform(lp->u, pd->u_uniform, pd->u_value);
            for (int vdt = 0; vdt < 2; ++vdt) {
                if (!(vdt_mask & (1u << vdt))) continue;
                VecSet& s = pd->vs[vdt];
                s.on = true;
                s.vdt = vdt;
                const size_t e = dtype_size(vdt);
                const int64_t n = lp->n, m = lp->m;
                s.c = pd->alloc(e * n); s.l = pd->alloc(e * n); s.u = pd->alloc(e * n); s.cres = pd->alloc(e * n);
                s.lo = pd->alloc(e * m); s.hi = pd->alloc(e * m); s.rres = pd->alloc(e * m);
                upload_as(ctx, vdt, lp->c, n, s.c); upload_as(ctx, vdt, lp->l, n, s.l);
                upload_as(ctx, vdt, lp->u, n, s.u); upload_as(ctx, vdt, lp->cres, n, s.cres);
                upload_as(ctx, vdt, lp->lo, m, s.lo); upload_as(ctx, vdt, lp->hi, m, s.hi);
                upload_as(ctx, vdt, lp->rres, m, s.rres);
                for (int k = 0; k < 2; ++k) { s.X[k] = pd->alloc(e * n); s.Y[k] = pd->alloc(e * m); }
                s.xbar = pd->alloc(e * n); s.sumx = pd->alloc(e * n); s.xa = pd->alloc(e * n); s.xr = pd->alloc(e * n);
                s.sumy = pd->alloc(e * m); s.ya = pd->alloc(e * m); s.yr = pd->alloc(e * m);
            }
            pd->prm = (double*)pd->alloc(sizeof(double) * kPrmDoubles);
            GH_CHECK(hipHostMalloc((void**)&pd->check_host, sizeof(double) * GANIT_CHECK_NOUT, hipHostMallocCoherent));
            GH_CHECK(hipStreamSynchronize(ctx->stream));
            // Graph replay pays for itself when an iteration is short (it saves the host launch cost, ~3 us per
            // iteration on C0-sized problems); on large matrices plain stream launches keep up with the GPU and
            // were measured ~2% faster than replay.  GANIT_HIP_NO_GRAPH=1 / GANIT_HIP_GRAPHS=1 force either.
            const char* ng = std::getenv("GANIT_HIP_NO_GRAPH");
            const char* fg = std::getenv("GANIT_HIP_GRAPHS");
            pd->use_graphs = (fg && fg[0] == '1') || (!(ng && ng[0] == '1') && pd->A->nnz < kGraphMaxNnz);
            const char* at = std::getenv("GANIT_HIP_AUTOTUNE");
            if (!(at && at[0] == '0')) {
                // Per-matrix variant, timed on the fused iteration kernels that will run, at the lowest resident
                // precision (the hot level) on a zeroed working set, in solve order: each launch of the candidate
                // is followed by the other fused kernel, which evicts the candidate's data from L2 as it does in a
                // solve.  (Timed back to back, a row-group variant re-reads what the previous launch left in L2
                // and can win the trial yet run at half the row blocks' speed in the solve.)
                const int mdt = (mdt_mask & 4u) ? GANIT_BF16 : (mdt_mask & 2u) ? GANIT_F32 : GANIT_F64;
                const int vdt = (vdt_mask & 2u) ? GANIT_F32 : GANIT_F64;
                VecSet& s = pd->set(vdt);
                with_vdt(vdt, [&](auto VT) {
                    using V = typename decltype(VT)::type;
                    hipLaunchKernelGGL(k_zero2<V>, dim3(ew_grid2((int64_t)pd->n + pd->m)), dim3(kBlock), 0, ctx->stream,
                                       pd->n, pd->m, (V*)s.X[0], (V*)s.Y[0]);
                    hipLaunchKernelGGL(k_zero2<V>, dim3(ew_grid2((int64_t)pd->n + pd->m)), dim3(kBlock), 0, ctx->stream,
                                       pd->n, pd->m, (V*)s.X[1], (V*)s.Y[1]);
                    hipLaunchKernelGGL(k_zero2<V>, dim3(ew_grid2((int64_t)pd->n + pd->m)), dim3(kBlock), 0, ctx->stream,
                                       pd->n, pd->m, (V*)s.xbar, (V*)s.sumy);
                    hipLaunchKernelGGL(k_zero2<V>, dim3(ew_grid2((int64_t)pd->n)), dim3(kBlock), 0, ctx->stream, pd->n,
                                       0, (V*)s.sumx, (V*)nullptr);
                });
                hipLaunchKernelGGL(k_set_steps, dim3(1), dim3(64), 0, ctx->stream, pd->prm, 0.0, 0.0);
                GH_CHECK(hipGetLastError());
                // while A^T is tuned, A's kernel (the other one of each pair) runs on its row blocks, the usual
                // winner, rather than on its lanes-per-row guess from the row lengths
                for (Csr* a : {pd->AT, pd->A})
                    if (a->nblocks > 0) a->lanes = 0;
                for (int which = 0; which < 2; ++which) {
                    Csr* a = which == 0 ? pd->AT : pd->A;
                    if ((which == 0 && pd->n == 0) || (which == 1 && pd->m == 0)) continue;
                    autotune_with(ctx, a, 12, [&](int k) {
                        if (pd->hasq) pd->Q->lanes = pd->AT->lanes;
                        with_mdt(mdt, [&](auto MT) {
                            with_vdt(vdt, [&](auto VT) {
                                using M = typename decltype(MT)::type;
                                using V = typename decltype(VT)::type;
                                for (int r = 0; r < k; ++r) {
                                    launch_one<M, V>(pd, s, which);
                                    launch_one<M, V>(pd, s, 1 - which);
                                }
                            });
                        });
                    });
                }
                if (pd->hasq) pd->Q->lanes = pd->AT->lanes;
            }
            // partials for the widest check grid of the chosen variants
            const size_t gmax = (size_t)std::max({(unsigned)kCheckGridMax, pd->A->rb_grid() + (unsigned)pd->A->nsplit,
                                                  pd->AT->rb_grid() + (unsigned)pd->AT->nsplit});
            pd->part_cap = gmax;
            pd->part_n = (double*)pd->alloc(sizeof(double) * gmax * kNOutN);
            pd->part_m = (double*)pd->alloc(sizeof(double) * gmax * kNOutM);
        } catch (...) {
            delete pd;
            throw;
        }
        *out = reinterpret_cast<ganit_hip_pdhg*>(pd);
    });
}

int ganit_hip_pdhg_destroy(ganit_hip_pdhg* h) {
    return guarded([&] { delete reinterpret_cast<Pdhg*>(h); });
}

int ganit_hip_pdhg_plan(const ganit_hip_pdhg* h, int which, int* lanes, int64_t* hist) {
    return guarded([&] {
        auto* pd = reinterpret_cast<const Pdhg*>(h);
        GH_REQUIRE(pd && which >= 0 && which <= 2, "bad plan query");
        const Csr* a = which == 0 ? pd->A : which == 1 ? pd->AT : pd->Q;
        GH_REQUIRE(a, "problem has no Q");
        if (lanes) *lanes = a->lanes;
        if (hist) std::copy(a->hist, a->hist + 9, hist);
    });
}

int ganit_hip_pdhg_index_bytes(const ganit_hip_pdhg* h, int which, int* bytes) {
    return guarded([&] {
        auto* pd = reinterpret_cast<const Pdhg*>(h);
        GH_REQUIRE(pd && bytes && which >= 0 && which <= 2, "bad index-width query");
        const Csr* a = which == 0 ? pd->A : which == 1 ? pd->AT : pd->Q;
        GH_REQUIRE(a, "problem has no Q");
        // the average width, rounded down (long rows or chunks too wide for 16-bit offsets read 4 bytes)
        *bytes = (a->lanes == 0 && a->idx16 && a->nnz > 0)
                     ? (int)((2 * (a->nnz - a->idx32_entries) + 4 * a->idx32_entries) / a->nnz)
                     : 4;
    });
}

int ganit_hip_pdhg_window_layout(const ganit_hip_pdhg* h, int which, int* swizzled) {
    return guarded([&] {
        auto* pd = reinterpret_cast<const Pdhg*>(h);
        GH_REQUIRE(pd && swizzled && which >= 0 && which <= 2, "bad window-layout query");
        const Csr* a = which == 0 ? pd->A : which == 1 ? pd->AT : pd->Q;
        GH_REQUIRE(a, "problem has no Q");
        *swizzled = (a->idx16 && a->bcol) ? a->wswz : 0;
    });
}

int ganit_hip_pdhg_set_lanes(ganit_hip_pdhg* h, int which, int lanes) {
    return guarded([&] {
        auto* pd = reinterpret_cast<Pdhg*>(h);
        GH_REQUIRE(pd && which >= 0 && which <= 2, "bad plan override");
        GH_REQUIRE(valid_variant(lanes), "lanes per row must be 0 (row blocks) or a power of two in 1..64");
        Csr* a = which == 0 ? pd->A : which == 1 ? pd->AT : pd->Q;
        GH_REQUIRE(a, "problem has no Q");
        a->lanes = lanes;
        if (which == 1 && pd->Q) pd->Q->lanes = lanes;  // Q rows share the A^T row group
        GH_CHECK(hipStreamSynchronize(pd->ctx->stream));
        for (auto& kv : pd->graphs) (void)hipGraphExecDestroy(kv.second);
        pd->graphs.clear();
    });
}

int ganit_hip_pdhg_load(ganit_hip_pdhg* h, int vdt, const double* x, const double* y, const double* xr,
                        const double* yr) {
    return guarded([&] {
        auto* pd = reinterpret_cast<Pdhg*>(h);
        GH_REQUIRE(pd && (pd->n == 0 || x) && (pd->m == 0 || y), "null iterate");
        VecSet& s = pd->set(vdt);
        upload_as(pd->ctx, vdt, x, pd->n, s.X[s.cur]);
        upload_as(pd->ctx, vdt, y, pd->m, s.Y[s.cur]);
        upload_as(pd->ctx, vdt, xr ? xr : x, pd->n, s.xr);
        upload_as(pd->ctx, vdt, yr ? yr : y, pd->m, s.yr);
        with_vdt(vdt, [&](auto T) { clear_sums<typename decltype(T)::type>(pd, s); });
    });
}

int ganit_hip_pdhg_handover(ganit_hip_pdhg* h, int from, int to) {
    return guarded([&] {
        auto* pd = reinterpret_cast<Pdhg*>(h);
        GH_REQUIRE(pd, "null problem");
        VecSet& a = pd->set(from);
        VecSet& b = pd->set(to);
        with_vdt(from, [&](auto S) {
            with_vdt(to, [&](auto D) {
                using SV = typename decltype(S)::type;
                using DV = typename decltype(D)::type;
                auto conv = [&](const void* src, void* dst, int64_t k) {
                    if (k == 0 || src == dst) return;
                    hipLaunchKernelGGL((k_handover<SV, DV>), dim3(ew_grid2(k)), dim3(kBlock), 0, pd->ctx->stream, k,
                                       (const SV*)src, (DV*)dst);
                };
                conv(a.X[a.cur], b.X[b.cur], pd->n);
                conv(a.Y[a.cur], b.Y[b.cur], pd->m);
                conv(a.xr, b.xr, pd->n);
                conv(a.yr, b.yr, pd->m);
                clear_sums<DV>(pd, b);
            });
        });
        GH_CHECK(hipGetLastError());
    });
}

int ganit_hip_pdhg_iterate(ganit_hip_pdhg* h, int mdt, int vdt, int iters, double tau, double sigma) {
    return guarded([&] {
        auto* pd = reinterpret_cast<Pdhg*>(h);
        GH_REQUIRE(pd && iters >= 0, "bad iterate arguments");
        VecSet& s = pd->set(vdt);
        if (iters == 0) return;
        hipStream_t st = pd->ctx->stream;
        hipLaunchKernelGGL(k_set_steps, dim3(1), dim3(64), 0, st, pd->prm, tau, sigma);
        GH_CHECK(hipGetLastError());
        auto run = [&](int k) {
            with_mdt(mdt, [&](auto MT) {
                with_vdt(vdt, [&](auto VT) {
                    using M = typename decltype(MT)::type;
                    using V = typename decltype(VT)::type;
                    for (int t = 0; t < k; ++t) launch_iteration<M, V>(pd, s);
                });
            });
        };
        if (!pd->use_graphs || iters < 4) {
            run(iters);
            return;
        }
        const auto key = std::make_tuple(mdt, vdt, iters, s.cur);
        auto it = pd->graphs.find(key);
        if (it == pd->graphs.end()) {
            const int cur0 = s.cur;
            hipGraph_t g = nullptr;
            GH_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
            try {
                run(iters);
            } catch (...) {
                (void)hipStreamEndCapture(st, &g);
                if (g) (void)hipGraphDestroy(g);
                s.cur = cur0;
                throw;
            }
            GH_CHECK(hipStreamEndCapture(st, &g));
            hipGraphExec_t ex = nullptr;
            const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            GH_CHECK(e);
            s.cur = cur0;  // capture recorded the launches without running them
            it = pd->graphs.emplace(key, ex).first;
        }
        GH_CHECK(hipGraphLaunch(it->second, st));
        if (iters & 1) s.cur ^= 1;
    });
}

int ganit_hip_pdhg_check(ganit_hip_pdhg* h, int mdt, int vdt, int flags, double inv_k, double* out) {
    return guarded([&] {
        auto* pd = reinterpret_cast<Pdhg*>(h);
        GH_REQUIRE(pd && out, "null argument");
        GH_REQUIRE((flags & ~(GANIT_CHECK_AVG | GANIT_CHECK_RAY | GANIT_CHECK_FP64 | GANIT_CHECK_KAHAN |
                              GANIT_CHECK_DRAY)) == 0,
                   "unknown check flag");
        VecSet& s = pd->set(vdt);
        if ((flags & GANIT_CHECK_FP64) && vdt == GANIT_F32) {
            // FP64 evaluation of the FP32 iterate: the average is formed at FP32 (as the iteration would adopt
            // it); the check passes read the FP32 iterate, average and restart point, widen each entry as they
            // read it (exact), and evaluate against the FP64 problem data.
            const VecSet& d = pd->set(GANIT_F64);
            with_mdt(mdt, [&](auto MT) {
                run_check<typename decltype(MT)::type, double, double, float>(pd, s, d, flags, inv_k, out);
            });
        } else {
            with_mdt(mdt, [&](auto MT) {
                with_vdt(vdt, [&](auto VT) {
                    using M = typename decltype(MT)::type;
                    using V = typename decltype(VT)::type;
                    run_check<M, V, V>(pd, s, s, flags, inv_k, out);
                });
            });
        }
        if (!(flags & GANIT_CHECK_AVG)) for (int k : {4, 5, 6, 7, 10, 11, 14, 15, 17, 18}) out[k] = 0.0;
        if (!(flags & GANIT_CHECK_RAY)) for (int k : {8, 9, 16}) out[k] = 0.0;
        if (!(flags & GANIT_CHECK_DRAY)) for (int k : {19, 20, 21}) out[k] = 0.0;
    });
}

int ganit_hip_pdhg_restart(ganit_hip_pdhg* h, int vdt, int adopt) {
    return guarded([&] {
        auto* pd = reinterpret_cast<Pdhg*>(h);
        GH_REQUIRE(pd, "null problem");
        VecSet& s = pd->set(vdt);
        with_vdt(vdt, [&](auto T) {
            using V = typename decltype(T)::type;
            hipLaunchKernelGGL(k_restart<V>, dim3(ew_grid2((int64_t)pd->n + pd->m)), dim3(kBlock), 0, pd->ctx->stream,
                               pd->n, pd->m, adopt ? 1 : 0, (V*)s.X[s.cur], (const V*)s.xa, (V*)s.xr, (V*)s.sumx,
                               (V*)s.Y[s.cur], (const V*)s.ya, (V*)s.yr, (V*)s.sumy, 0);
        });
        GH_CHECK(hipGetLastError());
    });
}

int ganit_hip_pdhg_adopt_average(ganit_hip_pdhg* h, int vdt) {
    return guarded([&] {
        auto* pd = reinterpret_cast<Pdhg*>(h);
        GH_REQUIRE(pd, "null problem");
        VecSet& s = pd->set(vdt);
        with_vdt(vdt, [&](auto T) {
            using V = typename decltype(T)::type;
            hipLaunchKernelGGL(k_restart<V>, dim3(ew_grid2((int64_t)pd->n + pd->m)), dim3(kBlock), 0, pd->ctx->stream,
                               pd->n, pd->m, 1, (V*)s.X[s.cur], (const V*)s.xa, (V*)s.xr, (V*)s.sumx, (V*)s.Y[s.cur],
                               (const V*)s.ya, (V*)s.yr, (V*)s.sumy, 1);
        });
        GH_CHECK(hipGetLastError());
    });
}

int ganit_hip_pdhg_store(ganit_hip_pdhg* h, int vdt, double* x, double* y, double* xr, double* yr) {
    return guarded([&] {
        auto* pd = reinterpret_cast<Pdhg*>(h);
        GH_REQUIRE(pd, "null problem");
        VecSet& s = pd->set(vdt);
        download_as(pd->ctx, vdt, s.X[s.cur], pd->n, x);
        download_as(pd->ctx, vdt, s.Y[s.cur], pd->m, y);
        download_as(pd->ctx, vdt, s.xr, pd->n, xr);
        download_as(pd->ctx, vdt, s.yr, pd->m, yr);
    });
}

int ganit_hip_pdhg_time_kernels(ganit_hip_pdhg* h, int mdt, int vdt, int reps, double tau, double sigma,
                                double* ms) {
    return guarded([&] {
        auto* pd = reinterpret_cast<Pdhg*>(h);
        GH_REQUIRE(pd && ms && reps > 0, "bad timing arguments");
        VecSet& s = pd->set(vdt);
        hipStream_t st = pd->ctx->stream;
        constexpr int kEv = 7;
        hipEvent_t ev[kEv] = {};
        float4* sweep = nullptr;
        float* sink = nullptr;
        auto release = [&] {
            for (auto& e : ev)
                if (e) (void)hipEventDestroy(e);
            (void)hipFree(sweep);
            (void)hipFree(sink);
        };
        try {
            for (auto& e : ev) GH_CHECK(hipEventCreate(&e));
            // a read of this many bytes passes every XCD's L2 several times over: after it, a kernel finds
            // none of its own data there, as in a solve, where the two kernels alternate
            const int64_t sweep_bytes = int64_t(64) << 20;
            const unsigned sweep_grid = 2048;
            GH_CHECK(hipMalloc(&sweep, sweep_bytes));
            GH_CHECK(hipMalloc(&sink, sizeof(float) * sweep_grid));
            GH_CHECK(hipMemsetAsync(sweep, 0, sweep_bytes, st));
            auto sweep_l2 = [&] {
                hipLaunchKernelGGL(k_stream_read, dim3(sweep_grid), dim3(kBlock), 0, st, (const float4*)sweep,
                                   sweep_bytes / 16, sink);
            };
            hipLaunchKernelGGL(k_set_steps, dim3(1), dim3(64), 0, st, pd->prm, tau, sigma);
            with_mdt(mdt, [&](auto MT) {
                with_vdt(vdt, [&](auto VT) {
                    using M = typename decltype(MT)::type;
                    using V = typename decltype(VT)::type;
                    launch_iteration<M, V>(pd, s);  // warm-up, and leaves xbar defined
                    sweep_l2();
                    GH_CHECK(hipEventRecord(ev[0], st));
                    for (int r = 0; r < reps; ++r) sweep_l2();  // the sweep alone
                    GH_CHECK(hipEventRecord(ev[1], st));
                    for (int r = 0; r < reps; ++r) {  // each kernel after a sweep: L2 holds none of its data
                        launch_one<M, V>(pd, s, 0);
                        sweep_l2();
                    }
                    GH_CHECK(hipEventRecord(ev[2], st));
                    for (int r = 0; r < reps; ++r) {
                        launch_one<M, V>(pd, s, 1);
                        sweep_l2();
                    }
                    GH_CHECK(hipEventRecord(ev[3], st));
                    // whole iterations in the order a solve launches them, events only at the ends
                    for (int r = 0; r < reps; ++r) launch_iteration<M, V>(pd, s);
                    GH_CHECK(hipEventRecord(ev[4], st));
                    // each kernel back to back: its own data partly stays in L2 from one launch to the next
                    for (int r = 0; r < reps; ++r) launch_one<M, V>(pd, s, 0);
                    GH_CHECK(hipEventRecord(ev[5], st));
                    for (int r = 0; r < reps; ++r) launch_one<M, V>(pd, s, 1);
                    GH_CHECK(hipEventRecord(ev[6], st));
                });
            });
            GH_CHECK(hipEventSynchronize(ev[6]));
            float t[kEv - 1];
            for (int k = 0; k + 1 < kEv; ++k) {
                GH_CHECK(hipEventElapsedTime(&t[k], ev[k], ev[k + 1]));
                t[k] /= reps;
            }
            ms[5] = t[0];
            ms[0] = t[1] - t[0];
            ms[1] = t[2] - t[0];
            ms[2] = t[3];
            ms[3] = t[4];
            ms[4] = t[5];
        } catch (...) {
            release();
            throw;
        }
        release();
    });
}

}  // extern "C"

