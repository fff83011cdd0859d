// libganit_hip.so, part 3: many uneven LPs and convex QPs solved together (branch-and-bound node relaxations,
// strong-branching probes, a batch of differentiable QP layers).
//
// Layout: the instances are packed into one block-diagonal problem -- A_all and its transpose are the
// instances' matrices placed on the diagonal (global column indices), and so is Q when an instance has one
// (an LP's rows of it are empty); every vector is the concatenation of
// the instances' vectors -- and the LDS-staged row-block partition is built per instance, so a workgroup
// never mixes instances and carries its instance id.  Every kernel therefore runs over the whole batch
// with full-GPU parallelism, whatever the instance sizes, instead of one instance per workgroup where the
// largest instance would set the pace.
//
// State: the restart / step-size / convergence state of every instance lives on the device
// (BatchState); after each block of check_every iterations a check pass produces 19 sums per instance
// (per-block partials folded per instance in block order: deterministic) and k_bt_decide applies the
// reference's check logic (src/pdhg_impl.hpp:144-189) to each instance in one thread.  Workgroups of
// instances that have stopped return at once.  A block of iterations plus its check is one captured HIP
// graph, replayed until no instance is running; the host reads a single counter per replay.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <limits>
#include <map>
#include <tuple>
#include <type_traits>
#include <vector>

#include "internal.hpp"
#include "pdhg_decide.hpp"
#include "pdhg_math.hpp"

namespace ganit_hip {

constexpr int kBN = 12, kBM = 7;

// ------------------------------------------------------------------ iteration kernels
// A listed workgroup {row block, instance}: its block's head, and whether it works.  The head's loads are issued
// with the instance state's and the test reads both (a listed block always has a row; `&`, not `&&`), so the two
// round trips overlap instead of one following the other.
template <class M>
__device__ __forceinline__ bool bt_head(const CsrView<M>& a, const int* __restrict__ blocks, int2 e,
                                        const BatchState* __restrict__ st, BlockHead& h) {
    h = block_head(a, blocks, e.x);
    return (st[e.y].active != 0) & (h.d.y > h.d.x);
}

template <class M, class V, bool HASQ>
__global__ __launch_bounds__(kBlock) void k_bt_primal(CsrView<M> AT, const int* __restrict__ blocks,
                                                      const int2* __restrict__ list, CsrView<M> Q,
                                                      const BatchState* __restrict__ st, const V* __restrict__ y,
                                                      const V* __restrict__ x, const V* __restrict__ c,
                                                      const V* __restrict__ l, const V* __restrict__ u,
                                                      V* __restrict__ xn, V* __restrict__ xbar,
                                                      V* __restrict__ sumx) {
    __shared__ __align__(16) V lds[kTile + kWin];
    const int2 e = list[xcd_block_id()];  // a workgroup of a running instance: {row block, instance}
    BlockHead h;
    if (!bt_head(AT, blocks, e, st, h)) return;
    const V tau = V(st[e.y].tau);
    struct Ops {
        V x, c, l, u, s;
    };
    stream_head_pf<M, V, 1, true>(
        AT, h, [&](int, int i) { return y[i]; },
        [&](int j) { return Ops{x[j], c[j], l[j], u[j], sumx[j]}; },
        [&](int j, V (&aty)[1], Ops& o) {
#pragma clang fp contract(off)
            V qx = V(0);
            if (HASQ) {  // Q row summed sequentially by the owning thread, as in the single-instance kernel
                for (int p = Q.ptr[j]; p < Q.ptr[j + 1]; ++p) {
                    const V prod = widen<V>(Q.val[p]) * x[Q.idx[p]];
                    qx = qx + prod;
                }
            }
            V sm = o.s;
            primal_update<V>(o.x, aty[0], qx, o.c, o.l, o.u, tau, xn[j], xbar[j], sm);
            sumx[j] = sm;
        },
        lds);
}

template <class M, class V>
__global__ __launch_bounds__(kBlock) void k_bt_dual(CsrView<M> A, const int* __restrict__ blocks,
                                                    const int2* __restrict__ list, const BatchState* __restrict__ st,
                                                    const V* __restrict__ xbar, const V* __restrict__ y,
                                                    const V* __restrict__ lo, const V* __restrict__ hi,
                                                    V* __restrict__ yn, V* __restrict__ sumy) {
    __shared__ __align__(16) V lds[kTile + kWin];
    const int2 e = list[xcd_block_id()];  // a workgroup of a running instance: {row block, instance}
    BlockHead h;
    if (!bt_head(A, blocks, e, st, h)) return;
    const V sigma = V(st[e.y].sigma);
    struct Ops {
        V y, lo, hi, s;
    };
    stream_head_pf<M, V, 1, true>(
        A, h, [&](int, int j) { return xbar[j]; },
        [&](int i) { return Ops{y[i], lo[i], hi[i], sumy[i]}; },
        [&](int i, V (&axb)[1], Ops& o) {
            V sm = o.s;
            dual_update<V>(o.y, axb[0], o.lo, o.hi, sigma, yn[i], sm);
            sumy[i] = sm;
        },
        lds);
}

// A launch-bound tail: at most kTailInstances instances listed, with at most kTailNnz nonzeros between them.  (Sixteen
// small LPs with 130K nonzeros ran 7% slower lanes-per-row; a few slow instances left of 512 run 25% faster.)
constexpr int kTailInstances = 8;
constexpr int64_t kTailNnz = int64_t(1) << 18;

// The rows of a tail, passed by value: instance q's rows r0[q] .. r1[q]-1 go to workgroups wg0[q] .. wg0[q+1]-1,
// kBlock / LPR consecutive rows each.  A workgroup finds its rows from its index and these arguments, so no list
// entry is loaded ahead of the rows' own loads.
struct TailArgs {
    int n;                        // instances (each with at least one row)
    int wg0[kTailInstances + 1];  // wg0[n] = the grid
    int r0[kTailInstances], r1[kTailInstances], inst[kTailInstances];
};

struct TailRows {
    int row, end, inst;
};

// This workgroup's instance and row group (uniform: a select chain over constant indices, kept in scalar registers).
template <int LPR>
__device__ __forceinline__ TailRows tail_rows(const TailArgs& t) {
    const int w = blockIdx.x;
    int r0 = t.r0[0], r1 = t.r1[0], w0 = t.wg0[0], in = t.inst[0];
#pragma unroll
    for (int q = 1; q < kTailInstances; ++q)
        if (q < t.n && w >= t.wg0[q]) {
            r0 = t.r0[q];
            r1 = t.r1[q];
            w0 = t.wg0[q];
            in = t.inst[q];
        }
    return TailRows{r0 + (w - w0) * (kBlock / LPR) + (int)threadIdx.x / LPR, r1, in};
}

// The same two kernels for a launch-bound tail: LPR lanes per row (row_dot), a workgroup over kBlock / LPR
// consecutive rows of one instance (TailArgs).  Rows are summed by a fixed lane tree, not in row order (kernels
// over a few small instances are bound by their launches and dependent loads, which the row-block staging
// lengthens).  The instance state and the row's own operands are loaded first, beside the row's pointer loads,
// so the dependent chain is pointers, entries, gathered values.  A stopped instance's workgroups load but store
// nothing.
template <int LPR, class M, class V, bool HASQ>
__global__ __launch_bounds__(kBlock) void k_bt_primal_lpr(CsrView<M> AT, TailArgs t, CsrView<M> Q,
                                                          const BatchState* __restrict__ st,
                                                          const V* __restrict__ y, const V* __restrict__ x,
                                                          const V* __restrict__ c, const V* __restrict__ l,
                                                          const V* __restrict__ u, V* __restrict__ xn,
                                                          V* __restrict__ xbar, V* __restrict__ sumx) {
    const TailRows r = tail_rows<LPR>(t);
    const int j = r.row, lane = (int)threadIdx.x % LPR;
    if (j >= r.end) return;  // whole row groups leave together
    const int active = st[r.inst].active;
    const V tau = V(st[r.inst].tau);
    const V xj = x[j], cj = c[j], lj = l[j], uj = u[j], sj = sumx[j];
    const V aty = row_dot<LPR>(AT, j, lane, y);
    const V qx = HASQ ? row_dot<LPR>(Q, j, lane, x) : V(0);
    if (lane == 0 && active) {
        V sm = sj;
        primal_update<V>(xj, aty, qx, cj, lj, uj, tau, xn[j], xbar[j], sm);
        sumx[j] = sm;
    }
}

template <int LPR, class M, class V>
__global__ __launch_bounds__(kBlock) void k_bt_dual_lpr(CsrView<M> A, TailArgs t, const BatchState* __restrict__ st,
                                                        const V* __restrict__ xbar, const V* __restrict__ y,
                                                        const V* __restrict__ lo, const V* __restrict__ hi,
                                                        V* __restrict__ yn, V* __restrict__ sumy) {
    const TailRows r = tail_rows<LPR>(t);
    const int i = r.row, lane = (int)threadIdx.x % LPR;
    if (i >= r.end) return;
    const int active = st[r.inst].active;
    const V sigma = V(st[r.inst].sigma);
    const V yi = y[i], li = lo[i], hi_ = hi[i], si = sumy[i];
    const V axb = row_dot<LPR>(A, i, lane, xbar);
    if (lane == 0 && active) {
        V sm = si;
        dual_update<V>(yi, axb, li, hi_, sigma, yn[i], sm);
        sumy[i] = sm;
    }
}

// ------------------------------------------------------------------ check kernels
template <class V>
__global__ __launch_bounds__(kBlock) void k_bt_average(int N, int M_, const int* __restrict__ col_inst,
                                                       const int* __restrict__ row_inst,
                                                       const BatchState* __restrict__ st, const V* __restrict__ sumx,
                                                       const V* __restrict__ sumy, V* __restrict__ xa,
                                                       V* __restrict__ ya) {
    const int64_t total = (int64_t)N + M_;
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < total; t += (int64_t)gridDim.x * kBlock) {
        if (t < N) {
            const BatchState& s = st[col_inst[t]];
            if (s.active) xa[t] = V(s.inv_k * double(sumx[t]));
        } else {
            const int64_t i = t - N;
            const BatchState& s = st[row_inst[i]];
            if (s.active) ya[i] = V(s.inv_k * double(sumy[i]));
        }
    }
}

__device__ __forceinline__ void bt_dual_terms(double z, double lj, double uj, double cr, double& r2, double& ob) {
    double r = 0.0, o = 0.0;
    if (z > 0) { if (isfinite(lj)) o = z * lj; else r = z; }
    else if (z < 0) { if (isfinite(uj)) o = z * uj; else r = z; }
    r /= cr;
    r2 = r * r;
    ob = o;
}
__device__ __forceinline__ double bt_row_obj(double yi, double lw, double hg) {
    if (yi > 0 && isfinite(lw)) return yi * lw;
    if (yi < 0 && isfinite(hg)) return yi * hg;
    return 0.0;
}

// Same sums as check_n_block of the single-instance path, one partial row per workgroup.
// V: the problem vectors and the arithmetic of the row sums; VI: the iterate's working set (VI = float with
// V = double is the FP64 evaluation of an FP32 iterate: each entry widened as it is read, exactly).
template <class M, class V, bool HASQ, class VI = V>
__global__ __launch_bounds__(kBlock) void k_bt_check_n(CsrView<M> AT, const int* __restrict__ blocks,
                                                       const int2* __restrict__ list, CsrView<M> Q,
                                                       const BatchState* __restrict__ st, const VI* __restrict__ x,
                                                       const VI* __restrict__ y, const VI* __restrict__ xa,
                                                       const VI* __restrict__ ya, const VI* __restrict__ xr,
                                                       const VI* __restrict__ yr, const V* __restrict__ c,
                                                       const V* __restrict__ l, const V* __restrict__ u,
                                                       const V* __restrict__ cres, int avg,
                                                       double* __restrict__ partials) {
    __shared__ __align__(16) V lds[kTile];
    __shared__ double red[(kBlock / kWave) * kBN];
    const int2 e = list[xcd_block_id()];  // a workgroup of a running instance: {row block, instance}
    const int b = e.x;
    BlockHead h;
    if (!bt_head(AT, blocks, e, st, h)) return;
    double acc[kBN];
#pragma unroll
    for (int k = 0; k < kBN; ++k) acc[k] = 0.0;
    stream_head_seq_pf<M, V, 3, false>(
        AT, h,
        [&](int k, int i) { return k == 0 ? V(y[i]) : k == 1 ? (avg ? V(ya[i]) : V(0)) : V(V(y[i]) - V(yr[i])); },
        [](int) { return NoRowOperands{}; },
        [&](int j, V (&t)[3], NoRowOperands&) {
#pragma clang fp contract(off)
            V qc = V(0), qa = V(0);
            if (HASQ) {
                for (int p = Q.ptr[j]; p < Q.ptr[j + 1]; ++p) {
                    const V a = widen<V>(Q.val[p]);
                    const V pc = a * V(x[Q.idx[p]]);
                    qc = qc + pc;
                    if (avg) {
                        const V pa = a * V(xa[Q.idx[p]]);
                        qa = qa + pa;
                    }
                }
            }
            const double cj = double(c[j]), lj = double(l[j]), uj = double(u[j]), cr = double(cres[j]);
            const double xj = double(x[j]);
            double r2, ob;
            bt_dual_terms(cj - double(t[0]) + (HASQ ? double(qc) : 0.0), lj, uj, cr, r2, ob);
            acc[0] += r2;
            acc[1] += ob;
            acc[2] += cj * xj;
            if (HASQ) acc[3] += xj * double(qc);
            if (avg) {
                const double xaj = double(xa[j]), xrj = double(xr[j]);
                bt_dual_terms(cj - double(t[1]) + (HASQ ? double(qa) : 0.0), lj, uj, cr, r2, ob);
                acc[4] += r2;
                acc[5] += ob;
                acc[6] += cj * xaj;
                if (HASQ) acc[7] += xaj * double(qa);
                const double d0 = xj - xrj, d1 = xaj - xrj;
                acc[10] += d0 * d0;
                acc[11] += d1 * d1;
            }
            bt_dual_terms(0.0 * cj - double(t[2]) + 0.0, lj, uj, cr, r2, ob);
            acc[8] += r2;
            acc[9] += ob;
        },
        lds, ~0u);
    __syncthreads();
    block_sum<kBN>(acc, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < kBN; ++k) partials[(size_t)b * kBN + k] = acc[k];
    }
}

template <class M, class V, class VI = V>
__global__ __launch_bounds__(kBlock) void k_bt_check_m(CsrView<M> A, const int* __restrict__ blocks,
                                                       const int2* __restrict__ list,
                                                       const BatchState* __restrict__ st, const VI* __restrict__ x,
                                                       const VI* __restrict__ xa, const VI* __restrict__ y,
                                                       const VI* __restrict__ ya, const VI* __restrict__ yr,
                                                       const V* __restrict__ lo, const V* __restrict__ hi,
                                                       const V* __restrict__ rres, int avg,
                                                       double* __restrict__ partials) {
    __shared__ __align__(16) V lds[kTile];
    __shared__ double red[(kBlock / kWave) * kBM];
    const int2 e = list[xcd_block_id()];  // a workgroup of a running instance: {row block, instance}
    const int b = e.x;
    BlockHead h;
    if (!bt_head(A, blocks, e, st, h)) return;
    double acc[kBM];
#pragma unroll
    for (int k = 0; k < kBM; ++k) acc[k] = 0.0;
    stream_head_seq_pf<M, V, 2, false>(
        A, h, [&](int k, int j) { return k == 0 ? V(x[j]) : (avg ? V(xa[j]) : V(0)); },
        [](int) { return NoRowOperands{}; },
        [&](int i, V (&t)[2], NoRowOperands&) {
            const double lw = double(lo[i]), hg = double(hi[i]), rr = double(rres[i]), yi = double(y[i]);
            auto pres = [&](double a) {
                const double p = fmin(fmax(a, lw), hg);
                const double r = (a - p) / rr;
                return r * r;
            };
            acc[0] += pres(double(t[0]));
            acc[1] += bt_row_obj(yi, lw, hg);
            if (avg) {
                const double yai = double(ya[i]), yri = double(yr[i]);
                acc[2] += pres(double(t[1]));
                acc[3] += bt_row_obj(yai, lw, hg);
                const double d0 = yi - yri, d1 = yai - yri;
                acc[5] += d0 * d0;
                acc[6] += d1 * d1;
            }
            acc[4] += bt_row_obj(double(V(V(y[i]) - V(yr[i]))), lw, hg);
        },
        lds, ~0u);
    __syncthreads();
    block_sum<kBM>(acc, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < kBM; ++k) partials[(size_t)b * kBM + k] = acc[k];
    }
}

// One wave per instance: fold its workgroups' partials in block order into sums[inst][19].
__global__ __launch_bounds__(kWave) void k_bt_reduce(int ninst, const int* __restrict__ first_n,
                                                     const int* __restrict__ first_m,
                                                     const BatchState* __restrict__ st,
                                                     const double* __restrict__ part_n,
                                                     const double* __restrict__ part_m, double* __restrict__ sums) {
    const int i = blockIdx.x;
    if (i >= ninst || !st[i].active) return;
    double a[kBN], bm[kBM];
#pragma unroll
    for (int k = 0; k < kBN; ++k) a[k] = 0.0;
#pragma unroll
    for (int k = 0; k < kBM; ++k) bm[k] = 0.0;
    for (int q = first_n[i] + (int)threadIdx.x; q < first_n[i + 1]; q += kWave) {
#pragma unroll
        for (int k = 0; k < kBN; ++k) a[k] += part_n[(size_t)q * kBN + k];
    }
    for (int q = first_m[i] + (int)threadIdx.x; q < first_m[i + 1]; q += kWave) {
#pragma unroll
        for (int k = 0; k < kBM; ++k) bm[k] += part_m[(size_t)q * kBM + k];
    }
#pragma unroll
    for (int k = 0; k < kBN; ++k) {
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) a[k] += __shfl_down(a[k], off, kWave);
    }
#pragma unroll
    for (int k = 0; k < kBM; ++k) {
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) bm[k] += __shfl_down(bm[k], off, kWave);
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < kBN; ++k) sums[(size_t)i * kBOut + k] = a[k];
#pragma unroll
        for (int k = 0; k < kBM; ++k) sums[(size_t)i * kBOut + kBN + k] = bm[k];
    }
}

// One thread per instance; block 0 also counts the instances still running into *running.
__global__ __launch_bounds__(1024) void k_bt_decide(int ninst, const BatchInst* __restrict__ inst,
                                                    BatchState* __restrict__ st, const double* __restrict__ sums,
                                                    BatchParams P, int mode, int* __restrict__ running) {
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    int mine = 0;
    for (int i = threadIdx.x; i < ninst; i += blockDim.x) {
        BatchState s = st[i];
        const BatchInst in = inst[i];
        const double* o = sums + (size_t)i * kBOut;
        if (decide_one(in, s, o, P, mode)) ++mine;
        st[i] = s;
    }
    atomicAdd(&cnt, mine);
    __syncthreads();
    if (threadIdx.x == 0) *running = cnt;
}

// Apply each instance's decision: adopt 0/1 = restart (1: from the average), 2 = adopt the average as the
// phase result; then instances whose phase ended stop taking part.
template <class V>
__global__ __launch_bounds__(kBlock) void k_bt_restart(int N, int M_, const int* __restrict__ col_inst,
                                                       const int* __restrict__ row_inst,
                                                       const BatchState* __restrict__ st, V* __restrict__ x,
                                                       const V* __restrict__ xa, V* __restrict__ xr,
                                                       V* __restrict__ sumx, V* __restrict__ y,
                                                       const V* __restrict__ ya, V* __restrict__ yr,
                                                       V* __restrict__ sumy) {
    const int64_t total = (int64_t)N + M_;
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < total; t += (int64_t)gridDim.x * kBlock) {
        const bool col = t < N;
        const int64_t e = col ? t : t - N;
        const int a = st[col ? col_inst[e] : row_inst[e]].adopt;
        if (a < 0) continue;
        V* v = col ? x : y;
        const V* va = col ? xa : ya;
        const V nv = (a >= 1) ? va[e] : v[e];
        v[e] = nv;
        if (a <= 1) {
            (col ? xr : yr)[e] = nv;
            (col ? sumx : sumy)[e] = V(0);
        }
    }
}

// The clock ran out: every running instance ends its phase with the candidate of its latest check
// (src/pdhg_impl.hpp:162).
__global__ void k_bt_stop(int ninst, BatchState* __restrict__ st) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ninst; i += gridDim.x * blockDim.x) {
        BatchState& s = st[i];
        if (!s.active || s.status != kRunning) continue;
        s.status = kTimeLimit;
        s.best_avg = s.cand_avg;
        for (int k = 0; k < 6; ++k) s.best[k] = s.cand[k];
        s.adopt = s.cand_avg ? 2 : -1;
    }
}

// Next phase: instances flagged in `sel` run again (status running), the others stay stopped.
__global__ void k_bt_select(int ninst, BatchState* __restrict__ st, const int* __restrict__ sel) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ninst; i += gridDim.x * blockDim.x) {
        BatchState& s = st[i];
        s.active = sel[i] ? 1 : 0;
        if (sel[i]) s.status = kRunning;
        s.adopt = -1;
    }
}

template <class S, class D>
__global__ __launch_bounds__(kBlock) void k_bt_convert(int64_t n, const S* __restrict__ s, D* __restrict__ d) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        d[i] = D(s[i]);
}

__global__ void k_bt_deactivate(int ninst, BatchState* __restrict__ st) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ninst; i += gridDim.x * blockDim.x) {
        if (st[i].status != kRunning) st[i].active = 0;
        st[i].adopt = -1;
    }
}

// ------------------------------------------------------------------ host object
struct BatchVecs {
    void *c = nullptr, *l = nullptr, *u = nullptr, *lo = nullptr, *hi = nullptr, *rres = nullptr, *cres = nullptr;
    void *X[2] = {nullptr, nullptr}, *Y[2] = {nullptr, nullptr};
    void *xbar = nullptr, *sumx = nullptr, *sumy = nullptr, *xa = nullptr, *ya = nullptr, *xr = nullptr, *yr = nullptr;
};

struct Batch {
    Ctx* ctx = nullptr;
    int ninst = 0, N = 0, Mr = 0;
    Csr *A = nullptr, *AT = nullptr;
    Csr* Q = nullptr;  // block-diagonal Q when an instance is a QP
    int *first_a = nullptr, *first_at = nullptr;  // per instance, its first row block (A, A^T)
    int *col_inst = nullptr, *row_inst = nullptr;
    BatchInst* inst = nullptr;
    BatchState* st = nullptr;
    double *part_n = nullptr, *part_m = nullptr, *sums = nullptr;
    int* running = nullptr;
    int* host_running = nullptr;  // pinned
    // host copies for re-partitioning (autotuner)
    std::vector<int> aptr_h, atptr_h, aidx_h, atidx_h;
    std::vector<BatchInst> inst_h;
    int part_rows = kBlock, part_nnz = kTile;
    size_t part_n_cap = 0, part_m_cap = 0;
    // launch lists: the row blocks of the instances taking part (kernels map blockIdx through them)
    std::vector<int> first_a_h, first_at_h;
    int2 *list_a = nullptr, *list_at = nullptr;  // {row block, instance} of each launched workgroup
    int nact_a = 0, nact_at = 0;
    // launch-bound tail: when the listed instances are few and small (bt_tail_nnz), the iteration kernels run
    // lanes-per-row over the tail's rows (TailArgs; the checks keep the row-block lists)
    // bit 0: lanes-per-row tail; bit 1: a reduced-precision phase ends once its running set is a tail (its instances
    // leave as stalled, so the ladder carries them on in FP64); 0: row blocks to the end (row-order sums)
    int tail_mode = 1;
    int tail_lpr_a = 0, tail_lpr_at = 0;  // lanes per row of the current tail lists; 0: no tail lists
    TailArgs targ_a{}, targ_at{};  // the tail's rows of A and of A^T
    int tact_a = 0, tact_at = 0;    // their grids
    int64_t tail_iters = 0;  // iterations run on tail lists (all phases)
    std::vector<char> listed;  // instances in the current lists
    BatchVecs vs[2];
    bool on[2] = {false, false};
    int cur = 0;
    BatchParams P{};
    std::vector<void*> owned;
    // A captured block bakes in its kernels' arguments: the dtypes, the parameters passed by value to
    // k_bt_decide and the iterate buffers X[cur] / Y[cur].  All of them are in the key, so a phase with other
    // tolerances, limits or stall rules, or an iterate reloaded into the other buffer, captures its own block.
    using GraphKey = std::tuple<int, int, int, int, double, double, int, int, int, int, int>;
    std::map<GraphKey, hipGraphExec_t> graphs;
    void* alloc(size_t bytes) {
        void* p = nullptr;
        GH_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 16)));
        owned.push_back(p);
        return p;
    }
    void drop_graphs() {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }
    ~Batch() {
        if (ctx) (void)hipStreamSynchronize(ctx->stream);
        drop_graphs();
        for (void* p : owned) (void)hipFree(p);
        for (void* p : {(void*)first_a, (void*)first_at, (void*)part_n, (void*)part_m, (void*)list_a, (void*)list_at})
            if (p) (void)hipFree(p);
        if (host_running) (void)hipHostFree(host_running);
        csr_destroy(A);
        csr_destroy(AT);
        csr_destroy(Q);
    }
};

unsigned bt_grid(int64_t n) { return std::min<unsigned>(grid_for(n), 8192u); }

// Nonzeros of A held by the instances flagged in `on`, when they are a tail (at most kTailInstances instances
// with at most kTailNnz nonzeros, at least one); 0 otherwise.
int64_t bt_tail_nnz(const Batch* bt, const std::vector<char>& on) {
    int64_t nz = 0;
    int count = 0;
    for (int k = 0; k < bt->ninst; ++k)
        if (on[k]) {
            const BatchInst& in = bt->inst_h[k];
            nz += bt->aptr_h[in.row0 + in.m] - bt->aptr_h[in.row0];
            ++count;
        }
    return (nz == 0 || nz > kTailNnz || count > kTailInstances) ? 0 : nz;
}

// Tail rows for the instances flagged in `on` when they are few and small enough (tail_mode, bt_tail_nnz): per
// matrix, lanes per row from the listed rows' nnz-weighted median length (8, 16 or 32), and each instance's rows
// in groups of kBlock / lanes, one workgroup each (TailArgs).  Otherwise none (tail_lpr_* = 0).
void bt_tail_lists(Batch* bt, const std::vector<char>& on) {
    bt->tail_lpr_a = bt->tail_lpr_at = 0;
    if (!(bt->tail_mode & 1)) return;
    const int64_t nz = bt_tail_nnz(bt, on);
    if (nz == 0) return;
    auto build = [&](const std::vector<int>& ptr, bool rows_of_a, TailArgs& t, int& grid) {
        std::vector<int> len;
        for (int k = 0; k < bt->ninst; ++k)
            if (on[k]) {
                const BatchInst& in = bt->inst_h[k];
                const int r0 = rows_of_a ? in.row0 : in.col0, nr = rows_of_a ? in.m : in.n;
                for (int r = r0; r < r0 + nr; ++r) len.push_back(ptr[r + 1] - ptr[r]);
            }
        const int lpr = std::min(32, std::max(8, choose_lanes_from_lengths(len, nz)));
        const int per = kBlock / lpr;
        t = TailArgs{};
        grid = 0;
        for (int k = 0; k < bt->ninst; ++k)
            if (on[k]) {
                const BatchInst& in = bt->inst_h[k];
                const int r0 = rows_of_a ? in.row0 : in.col0, nr = rows_of_a ? in.m : in.n;
                if (nr == 0) continue;
                GH_REQUIRE(t.n < kTailInstances, "tail holds too many instances");
                t.r0[t.n] = r0;
                t.r1[t.n] = r0 + nr;
                t.inst[t.n] = k;
                t.wg0[t.n] = grid;
                grid += (nr + per - 1) / per;
                ++t.n;
            }
        t.wg0[t.n] = grid;
        return lpr;
    };
    bt->tail_lpr_a = build(bt->aptr_h, true, bt->targ_a, bt->tact_a);
    bt->tail_lpr_at = build(bt->atptr_h, false, bt->targ_at, bt->tact_at);
}

// Launch lists for the instances flagged in `on`: the workgroups of instances that have stopped are left out
// of the grid instead of being launched to return at once, which matters when a few slow instances remain.
// Changing the lists changes grid sizes, so captured graphs are dropped.
void bt_list(Batch* bt, const std::vector<char>& on) {
    if (on == bt->listed) return;
    auto fill = [&](const std::vector<int>& first, int2* dev, int& count) {
        std::vector<int2> l;
        for (int k = 0; k < bt->ninst; ++k)
            if (on[k])
                for (int b = first[k]; b < first[k + 1]; ++b) l.push_back(make_int2(b, k));
        count = (int)l.size();
        if (count > 0)
            GH_CHECK(hipMemcpyAsync(dev, l.data(), sizeof(int2) * l.size(), hipMemcpyHostToDevice, bt->ctx->stream));
        GH_CHECK(hipStreamSynchronize(bt->ctx->stream));
    };
    fill(bt->first_a_h, bt->list_a, bt->nact_a);
    fill(bt->first_at_h, bt->list_at, bt->nact_at);
    bt_tail_lists(bt, on);
    bt->listed = on;
    bt->drop_graphs();
}

// The instances whose state says active.
std::vector<char> bt_active(Batch* bt) {
    std::vector<BatchState> s(bt->ninst);
    GH_CHECK(hipMemcpyAsync(s.data(), bt->st, sizeof(BatchState) * bt->ninst, hipMemcpyDeviceToHost, bt->ctx->stream));
    GH_CHECK(hipStreamSynchronize(bt->ctx->stream));
    std::vector<char> on(bt->ninst);
    for (int k = 0; k < bt->ninst; ++k) on[k] = s[k].active ? 1 : 0;
    return on;
}

// Row-block partition of A and A^T, built per instance (at most `rows` rows and `nnz` nonzeros per block;
// rows longer than kLong get a block of their own), so a workgroup never mixes instances.  Replaces the
// previous partition and drops captured graphs, whose kernel arguments name the old arrays.
void bt_partition(Batch* bt, int rows, int nnz) {
    Ctx* ctx = bt->ctx;
    const int ninst = bt->ninst;
    auto build = [&](Csr* a, const std::vector<int>& ptr, const std::vector<int>& idx, bool rows_of_a,
                     int** first, std::vector<int>& first_h, int2** list) {
        std::vector<int> blocks{0}, bi, fi{0};
        std::vector<int> lp;
        for (int k = 0; k < ninst; ++k) {
            const BatchInst& in = bt->inst_h[k];
            const int r0 = rows_of_a ? in.row0 : in.col0;
            const int nr = rows_of_a ? in.m : in.n;
            lp.resize(nr + 1);
            for (int r = 0; r <= nr; ++r) lp[r] = ptr[r0 + r] - ptr[r0];
            const std::vector<int> rb = row_blocks(lp.data(), nr, rows, nnz);
            for (size_t q = 1; q < rb.size(); ++q) {
                blocks.push_back(r0 + rb[q]);
                bi.push_back(k);
            }
            fi.push_back((int)bi.size());
        }
        if (a->blocks) GH_CHECK(hipFree(a->blocks));
        a->blocks = nullptr;
        a->blocks = upload_row_blocks(ctx, ptr.data(), blocks);
        build_compact_indices(ctx, a, ptr.data(), idx.data(), blocks);
        a->nblocks = (int)blocks.size() - 1;
        a->lanes = 0;
        a->rb_rows = rows;
        a->rb_nnz = nnz;
        if (*first) GH_CHECK(hipFree(*first));
        *first = nullptr;
        GH_CHECK(hipMalloc(first, sizeof(int) * fi.size()));
        GH_CHECK(hipMemcpyAsync(*first, fi.data(), sizeof(int) * fi.size(), hipMemcpyHostToDevice, ctx->stream));
        GH_CHECK(hipStreamSynchronize(ctx->stream));
        first_h = fi;
        if (*list) GH_CHECK(hipFree(*list));
        *list = nullptr;
        GH_CHECK(hipMalloc(list, sizeof(int2) * std::max<size_t>(bi.size(), 1)));
    };
    GH_CHECK(hipStreamSynchronize(ctx->stream));
    bt->drop_graphs();
    build(bt->A, bt->aptr_h, bt->aidx_h, true, &bt->first_a, bt->first_a_h, &bt->list_a);
    build(bt->AT, bt->atptr_h, bt->atidx_h, false, &bt->first_at, bt->first_at_h, &bt->list_at);
    bt->listed.assign(ninst, 0);
    bt_list(bt, std::vector<char>(ninst, 1));
    const size_t need_n = (size_t)kBN * std::max(bt->AT->nblocks, 1), need_m = (size_t)kBM * std::max(bt->A->nblocks, 1);
    if (need_n > bt->part_n_cap) {
        if (bt->part_n) GH_CHECK(hipFree(bt->part_n));
        bt->part_n = nullptr;
        GH_CHECK(hipMalloc(&bt->part_n, sizeof(double) * need_n));
        bt->part_n_cap = need_n;
    }
    // the partial buffer of A also stages the instance selection (one int per instance)
    const size_t need_m2 = std::max(need_m, (size_t)(ninst + 1) / 2 + 1);
    if (need_m2 > bt->part_m_cap) {
        if (bt->part_m) GH_CHECK(hipFree(bt->part_m));
        bt->part_m = nullptr;
        GH_CHECK(hipMalloc(&bt->part_m, sizeof(double) * need_m2));
        bt->part_m_cap = need_m2;
    }
    bt->part_rows = rows;
    bt->part_nnz = nnz;
}

// Lanes-per-row iteration over the tail lists (bt_tail_lists).
template <int LPR, class M, class V>
void bt_primal_lpr(Batch* bt, BatchVecs& v, int c) {
    hipStream_t s = bt->ctx->stream;
    if (bt->tact_at == 0) return;
    const CsrView<M> q = bt->Q ? bt->Q->view<M>() : bt->AT->view<M>();
    if (bt->Q)
        hipLaunchKernelGGL((k_bt_primal_lpr<LPR, M, V, true>), dim3(bt->tact_at), dim3(kBlock), 0, s,
                           bt->AT->view<M>(), bt->targ_at, q, bt->st, (const V*)v.Y[c], (const V*)v.X[c],
                           (const V*)v.c, (const V*)v.l, (const V*)v.u, (V*)v.X[1 - c], (V*)v.xbar, (V*)v.sumx);
    else
        hipLaunchKernelGGL((k_bt_primal_lpr<LPR, M, V, false>), dim3(bt->tact_at), dim3(kBlock), 0, s,
                           bt->AT->view<M>(), bt->targ_at, q, bt->st, (const V*)v.Y[c], (const V*)v.X[c],
                           (const V*)v.c, (const V*)v.l, (const V*)v.u, (V*)v.X[1 - c], (V*)v.xbar, (V*)v.sumx);
}
template <int LPR, class M, class V>
void bt_dual_lpr(Batch* bt, BatchVecs& v, int c) {
    if (bt->tact_a == 0) return;
    hipLaunchKernelGGL((k_bt_dual_lpr<LPR, M, V>), dim3(bt->tact_a), dim3(kBlock), 0, bt->ctx->stream,
                       bt->A->view<M>(), bt->targ_a, bt->st, (const V*)v.xbar, (const V*)v.Y[c], (const V*)v.lo,
                       (const V*)v.hi, (V*)v.Y[1 - c], (V*)v.sumy);
}

template <class M, class V>
void bt_iteration(Batch* bt, int mdt_unused) {
    (void)mdt_unused;
    hipStream_t s = bt->ctx->stream;
    BatchVecs& v = bt->vs[std::is_same<V, double>::value ? 0 : 1];
    const int c = bt->cur;
    if (bt->tail_lpr_a > 0) {
        switch (bt->tail_lpr_at) {
            case 8: bt_primal_lpr<8, M, V>(bt, v, c); break;
            case 16: bt_primal_lpr<16, M, V>(bt, v, c); break;
            default: bt_primal_lpr<32, M, V>(bt, v, c); break;
        }
        switch (bt->tail_lpr_a) {
            case 8: bt_dual_lpr<8, M, V>(bt, v, c); break;
            case 16: bt_dual_lpr<16, M, V>(bt, v, c); break;
            default: bt_dual_lpr<32, M, V>(bt, v, c); break;
        }
        bt->cur ^= 1;
        return;
    }
    if (bt->nact_at > 0) {
        if (bt->Q)
            hipLaunchKernelGGL((k_bt_primal<M, V, true>), dim3(bt->nact_at), dim3(kBlock), 0, s, bt->AT->view<M>(),
                               bt->AT->blocks, bt->list_at, bt->Q->view<M>(), bt->st,
                               (const V*)v.Y[c], (const V*)v.X[c], (const V*)v.c, (const V*)v.l, (const V*)v.u,
                               (V*)v.X[1 - c], (V*)v.xbar, (V*)v.sumx);
        else
            hipLaunchKernelGGL((k_bt_primal<M, V, false>), dim3(bt->nact_at), dim3(kBlock), 0, s, bt->AT->view<M>(),
                               bt->AT->blocks, bt->list_at, bt->AT->view<M>(), bt->st,
                               (const V*)v.Y[c], (const V*)v.X[c], (const V*)v.c, (const V*)v.l, (const V*)v.u,
                               (V*)v.X[1 - c], (V*)v.xbar, (V*)v.sumx);
    }
    if (bt->nact_a > 0)
        hipLaunchKernelGGL((k_bt_dual<M, V>), dim3(bt->nact_a), dim3(kBlock), 0, s, bt->A->view<M>(), bt->A->blocks,
                           bt->list_a, bt->st, (const V*)v.xbar, (const V*)v.Y[c], (const V*)v.lo,
                           (const V*)v.hi, (V*)v.Y[1 - c], (V*)v.sumy);
    bt->cur ^= 1;
}

// One check of the running instances: the A^T and A passes (current point, and the average with avg), the
// per-instance fold, the decisions and (mode kCheck) the restarts.  The iterate lives in working set V.  E64: the
// passes evaluate it in FP64 -- FP64 matrix values and problem vectors, each iterate entry widened as it is read
// (exact) -- the batched form of the single solve's FP64 candidate evaluation (GANIT_CHECK_FP64); the average is
// still formed in V, as the iteration would adopt it.
template <class M, class V, bool E64 = false>
void bt_check(Batch* bt, int avg, int mode) {
    using PM = std::conditional_t<E64, double, M>;  // matrix values of the passes
    using PV = std::conditional_t<E64, double, V>;  // problem vectors and the passes' arithmetic
    hipStream_t s = bt->ctx->stream;
    BatchVecs& v = bt->vs[std::is_same<V, double>::value ? 0 : 1];
    BatchVecs& pv = bt->vs[std::is_same<PV, double>::value ? 0 : 1];
    const int c = bt->cur;
    if (avg)
        hipLaunchKernelGGL(k_bt_average<V>, dim3(bt_grid((int64_t)bt->N + bt->Mr)), dim3(kBlock), 0, s, bt->N, bt->Mr,
                           bt->col_inst, bt->row_inst, bt->st, (const V*)v.sumx, (const V*)v.sumy, (V*)v.xa,
                           (V*)v.ya);
    if (bt->nact_at > 0) {
        const CsrView<PM> q = bt->Q ? bt->Q->view<PM>() : bt->AT->view<PM>();
        if (bt->Q)
            hipLaunchKernelGGL((k_bt_check_n<PM, PV, true, V>), dim3(bt->nact_at), dim3(kBlock), 0, s,
                               bt->AT->view<PM>(), bt->AT->blocks, bt->list_at, q, bt->st, (const V*)v.X[c],
                               (const V*)v.Y[c], (const V*)v.xa, (const V*)v.ya, (const V*)v.xr, (const V*)v.yr,
                               (const PV*)pv.c, (const PV*)pv.l, (const PV*)pv.u, (const PV*)pv.cres, avg, bt->part_n);
        else
            hipLaunchKernelGGL((k_bt_check_n<PM, PV, false, V>), dim3(bt->nact_at), dim3(kBlock), 0, s,
                               bt->AT->view<PM>(), bt->AT->blocks, bt->list_at, q, bt->st, (const V*)v.X[c],
                               (const V*)v.Y[c], (const V*)v.xa, (const V*)v.ya, (const V*)v.xr, (const V*)v.yr,
                               (const PV*)pv.c, (const PV*)pv.l, (const PV*)pv.u, (const PV*)pv.cres, avg, bt->part_n);
    }
    if (bt->nact_a > 0)
        hipLaunchKernelGGL((k_bt_check_m<PM, PV, V>), dim3(bt->nact_a), dim3(kBlock), 0, s, bt->A->view<PM>(),
                           bt->A->blocks, bt->list_a, bt->st, (const V*)v.X[c], (const V*)v.xa, (const V*)v.Y[c],
                           (const V*)v.ya, (const V*)v.yr, (const PV*)pv.lo, (const PV*)pv.hi, (const PV*)pv.rres, avg,
                           bt->part_m);
    hipLaunchKernelGGL(k_bt_reduce, dim3(bt->ninst), dim3(kWave), 0, s, bt->ninst, bt->first_at, bt->first_a, bt->st,
                       bt->part_n, bt->part_m, bt->sums);
    hipLaunchKernelGGL(k_bt_decide, dim3(1), dim3(1024), 0, s, bt->ninst, bt->inst, bt->st, bt->sums, bt->P, mode,
                       bt->running);
    if (mode == kCheck)
        hipLaunchKernelGGL(k_bt_restart<V>, dim3(bt_grid((int64_t)bt->N + bt->Mr)), dim3(kBlock), 0, s, bt->N, bt->Mr,
                           bt->col_inst, bt->row_inst, bt->st, (V*)v.X[c], (const V*)v.xa, (V*)v.xr, (V*)v.sumx,
                           (V*)v.Y[c], (const V*)v.ya, (V*)v.yr, (V*)v.sumy);
    GH_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------ adjoint of every instance at once
// The adjoint of the single-instance solve (adjoint.hip: restarted GMRES on (I - J^T) lambda = (g, 0), J^T one
// transposed PDHG iteration with the projections frozen at the solution) for every instance of the batch in
// lockstep.  The block-diagonal A, A^T and Q apply all instances' J^T in one set of launches, each row and
// column with its own instance's tau and sigma; the Krylov basis holds concatenated vectors; dot products are
// per instance (one workgroup per instance and basis vector, fixed order); each instance's Hessenberg system is
// solved on the host.  An instance that meets its tolerance stops extending its system, and its coefficients
// stay zero while the others go on.

// D_j = 1 unless x_j is within mtol (1 + |bound|) of a finite l_j or u_j (mtol of column j's instance);
// F_i = (lo_i == hi_i) || y_i != 0.
__global__ __launch_bounds__(kBlock) void k_ba_masks(int N, int M_, const int* __restrict__ col_inst,
                                                     const double* __restrict__ mtol, const double* __restrict__ x,
                                                     const double* __restrict__ l, const double* __restrict__ u,
                                                     const double* __restrict__ y, const double* __restrict__ lo,
                                                     const double* __restrict__ hi, uint8_t* __restrict__ D,
                                                     uint8_t* __restrict__ F) {
    const int64_t total = (int64_t)N + M_;
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < total; t += (int64_t)gridDim.x * kBlock) {
        if (t < N) {
            const double v = x[t], lj = l[t], uj = u[t], tl = mtol[col_inst[t]];
            const bool at_l = isfinite(lj) && v - lj <= tl * (1.0 + fabs(lj));
            const bool at_u = isfinite(uj) && uj - v <= tl * (1.0 + fabs(uj));
            D[t] = (at_l || at_u) ? 0 : 1;
        } else {
            const int64_t i = t - N;
            F[i] = (lo[i] == hi[i] || y[i] != 0.0) ? 1 : 0;
        }
    }
}

// g = V(g64), lambda = (g, 0)
template <class V>
__global__ __launch_bounds__(kBlock) void k_ba_init(int N, int M_, const double* __restrict__ g64,
                                                    V* __restrict__ g, V* __restrict__ lam) {
    const int64_t total = (int64_t)N + M_;
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < total; t += (int64_t)gridDim.x * kBlock) {
        if (t < N) {
            g[t] = V(g64[t]);
            lam[t] = V(g64[t]);
        } else {
            lam[t] = V(0);
        }
    }
}

// r = A^T (F o ly), p = D o (lx - 2 sigma_k r) over every row block of A^T (one instance per block)
template <class M, class V>
__global__ __launch_bounds__(kBlock) void k_ba_at(CsrView<M> AT, const int* __restrict__ blocks,
                                                  const int* __restrict__ col_inst, const V* __restrict__ two_sigma,
                                                  const V* __restrict__ ly, const uint8_t* __restrict__ F,
                                                  const V* __restrict__ lx, const uint8_t* __restrict__ D,
                                                  V* __restrict__ r, V* __restrict__ p) {
    __shared__ __align__(16) V lds[kTile];
    const SplitPlan none{0x7fffffff, nullptr, nullptr, nullptr, 0};
    stream_block<M, V, 1>(
        AT, blocks, xcd_block_id(), [&](int, int i) { return F[i] ? ly[i] : V(0); },
        [&](int j, V (&t)[1]) {
            r[j] = t[0];
            p[j] = D[j] ? V(lx[j] - two_sigma[col_inst[j]] * t[0]) : V(0);
        },
        lds, none);
}

// y part of (I - J^T) lambda: out = ly - (F o ly + tau_k A p)
template <class M, class V>
__global__ __launch_bounds__(kBlock) void k_ba_a(CsrView<M> A, const int* __restrict__ blocks,
                                                 const int* __restrict__ row_inst, const V* __restrict__ tau,
                                                 const V* __restrict__ p, const uint8_t* __restrict__ F,
                                                 const V* __restrict__ ly, V* __restrict__ out) {
    __shared__ __align__(16) V lds[kTile];
    const SplitPlan none{0x7fffffff, nullptr, nullptr, nullptr, 0};
    stream_block<M, V, 1>(
        A, blocks, xcd_block_id(), [&](int, int j) { return p[j]; },
        [&](int i, V (&t)[1]) {
            const V in = ly[i];
            const V old = F[i] ? in : V(0);
            const V nv = old + tau[row_inst[i]] * t[0];
            out[i] = V(in - nv);
        },
        lds, none);
}

// x part of (I - J^T) lambda: out = lx - (p - tau_k Q p + sigma_k r), the Q row summed by the owning thread
template <class M, class V, bool HASQ>
__global__ __launch_bounds__(kBlock) void k_ba_x(int N, CsrView<M> Q, const int* __restrict__ col_inst,
                                                 const V* __restrict__ tau, const V* __restrict__ sigma,
                                                 const V* __restrict__ p, const V* __restrict__ r,
                                                 const V* __restrict__ lx, V* __restrict__ out) {
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < N; j += gridDim.x * kBlock) {
        const int k = col_inst[j];
        V qp = V(0);
        if (HASQ)
            for (int q = Q.ptr[j]; q < Q.ptr[j + 1]; ++q) qp += widen<V>(Q.val[q]) * p[Q.idx[q]];
        const V jt = p[j] - tau[k] * qp + sigma[k] * r[j];
        out[j] = V(lx[j] - jt);
    }
}

// h[c * K + k] = <basis_c, w> over instance k's entries (its columns, then its rows unless xonly), FP64, one
// workgroup per (instance, basis vector): grid (K, ncols)
template <class V>
__global__ __launch_bounds__(kBlock) void k_bgm_dots(int N, int64_t L, const BatchInst* __restrict__ inst, int K,
                                                     const V* __restrict__ basis, const V* __restrict__ w, int xonly,
                                                     double* __restrict__ h) {
    __shared__ double red[kBlock / kWave];
    const int k = blockIdx.x, c = blockIdx.y;
    const V* __restrict__ v = basis + (size_t)c * L;
    const int c0 = inst[k].col0, n = inst[k].n;
    const int64_t r0 = (int64_t)N + inst[k].row0;
    const int m = xonly ? 0 : inst[k].m;
    double s[1] = {0.0};
    for (int q = threadIdx.x; q < n; q += kBlock) s[0] += double(v[c0 + q]) * double(w[c0 + q]);
    for (int q = threadIdx.x; q < m; q += kBlock) s[0] += double(v[r0 + q]) * double(w[r0 + q]);
    block_sum<1>(s, red);
    if (threadIdx.x == 0) h[(size_t)c * K + k] = s[0];
}

// w += sgn * sum_{c < ncols} h[c * K + k(t)] basis_c, k(t) the instance of entry t
template <class V>
__global__ __launch_bounds__(kBlock) void k_bgm_update(int N, int64_t L, const int* __restrict__ col_inst,
                                                       const int* __restrict__ row_inst, int K,
                                                       const V* __restrict__ basis, int ncols,
                                                       const double* __restrict__ h, double sgn, V* __restrict__ w) {
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < L; t += (int64_t)gridDim.x * kBlock) {
        const int k = t < N ? col_inst[t] : row_inst[t - N];
        double a = double(w[t]);
        for (int c = 0; c < ncols; ++c) a += sgn * h[(size_t)c * K + k] * double(basis[(size_t)c * L + t]);
        w[t] = V(a);
    }
}

// out = w / sqrt(nrm2[k(t)]), 0 where that norm is 0 (out may alias w)
template <class V>
__global__ __launch_bounds__(kBlock) void k_bgm_scale(int N, int64_t L, const int* __restrict__ col_inst,
                                                      const int* __restrict__ row_inst, const V* w,
                                                      const double* __restrict__ nrm2, V* out) {
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < L; t += (int64_t)gridDim.x * kBlock) {
        const double q = nrm2[t < N ? col_inst[t] : row_inst[t - N]];
        const double s = q > 0.0 ? 1.0 / sqrt(q) : 0.0;
        out[t] = V(double(w[t]) * s);
    }
}

// out = (g, 0) - a
template <class V>
__global__ __launch_bounds__(kBlock) void k_bgm_rhs_minus(int N, int64_t L, const V* __restrict__ g,
                                                          const V* __restrict__ a, V* __restrict__ out) {
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < L; t += (int64_t)gridDim.x * kBlock)
        out[t] = V((t < N ? double(g[t]) : 0.0) - double(a[t]));
}

// p64 = p, mu64 = F o ly
template <class V>
__global__ __launch_bounds__(kBlock) void k_ba_out(int N, int M_, const V* __restrict__ p, const V* __restrict__ ly,
                                                   const uint8_t* __restrict__ F, double* __restrict__ p64,
                                                   double* __restrict__ mu64) {
    const int64_t total = (int64_t)N + M_;
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < total; t += (int64_t)gridDim.x * kBlock) {
        if (t < N) p64[t] = double(p[t]);
        else mu64[t - N] = F[t - N] ? double(ly[t - N]) : 0.0;
    }
}

// ---- the GMRES bookkeeping of every instance on the device, one thread per instance (so the host does not wait
// for each Arnoldi step): Hessenberg columns (kk x kk per instance, column-major), Givens rotations, the rotated
// right-hand side, and the activity flags the host polls now and then.
struct GmState {
    double *H, *cs, *sn, *s;          // K kk kk, K kk, K kk, K (kk + 1)
    double *bnorm, *prev_beta, *rel;  // K each
    int *done, *act, *cols, *prods;   // K each
};

__global__ void k_gm_init(int K, const double* __restrict__ nr2, GmState g) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const double b = sqrt(nr2[k]);
    g.bnorm[k] = b;
    g.rel[k] = b > 0 ? INFINITY : 0.0;
    g.done[k] = !(b > 0);
    g.prev_beta[k] = INFINITY;
    g.prods[k] = 0;
    g.act[k] = 0;
    g.cols[k] = 0;
}

// A restart cycle begins: beta_k = ||r0_k||; an instance is done when converged or when a whole cycle gained less
// than 1% (the working precision's floor); the others start a new system with s = (beta, 0, ...).
__global__ void k_gm_cycle(int K, int kk, double tol, const double* __restrict__ nr2, GmState g) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    g.act[k] = 0;
    g.cols[k] = 0;
    if (g.done[k]) return;
    const double beta = sqrt(nr2[k]);
    const double rel = beta / g.bnorm[k];
    g.rel[k] = rel;
    if (!isfinite(rel) || rel <= tol || beta > 0.99 * g.prev_beta[k]) {
        g.done[k] = 1;
        return;
    }
    g.prev_beta[k] = beta;
    g.act[k] = 1;
    double* sk = g.s + (size_t)k * (kk + 1);
    for (int i = 0; i <= kk; ++i) sk[i] = 0.0;
    sk[0] = beta;
}

constexpr int kGmMax = 1024;  // longest restart cycle (GMRES columns per instance)

// Arnoldi step j of every active instance (one wave per instance): column j of H from the two Gram-Schmidt passes,
// the earlier rotations applied to it, a new rotation that zeroes the subdiagonal sqrt(nr2), the residual estimate
// |s[j+1]| / ||b||.  The column and the rotations are staged in LDS by the wave; one lane then runs the (inherently
// sequential) rotations with the running entry in a register, its loads independent of its arithmetic.
__global__ __launch_bounds__(kWave) void k_gm_step(int K, int kk, int j, double tol, const double* __restrict__ h1,
                                                   const double* __restrict__ h2, const double* __restrict__ nr2,
                                                   GmState g) {
    __shared__ double col[kGmMax + 1], cl[kGmMax], sl[kGmMax];
    const int k = blockIdx.x;
    if (!g.act[k]) return;  // the whole wave
    double* __restrict__ colg = g.H + ((size_t)k * kk + j) * kk;
    double* __restrict__ c = g.cs + (size_t)k * kk;
    double* __restrict__ sn = g.sn + (size_t)k * kk;
    for (int i = threadIdx.x; i <= j; i += kWave) col[i] = h1[(size_t)i * K + k] + h2[(size_t)i * K + k];
    for (int i = threadIdx.x; i < j; i += kWave) {
        cl[i] = c[i];
        sl[i] = sn[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ++g.prods[k];
        double* sk = g.s + (size_t)k * (kk + 1);
        const double sub = sqrt(fmax(nr2[k], 0.0));
        double cur = col[0];
        for (int i = 0; i < j; ++i) {
            const double nxt = col[i + 1], ci = cl[i], si = sl[i];
            col[i] = ci * cur + si * nxt;
            cur = -si * cur + ci * nxt;
        }
        const double den = hypot(cur, sub);
        const double cj = den > 0 ? cur / den : 1.0, sj = den > 0 ? sub / den : 0.0;
        c[j] = cj;
        sn[j] = sj;
        col[j] = den;
        sk[j + 1] = -sj * sk[j];
        sk[j] = cj * sk[j];
        g.cols[k] = j + 1;
        const double rel = fabs(sk[j + 1]) / g.bnorm[k];
        g.rel[k] = rel;
        if (!isfinite(rel) || rel <= tol || sub <= 1e-300 * g.bnorm[k]) g.act[k] = 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i <= j; i += kWave) colg[i] = col[i];
}

// End of a cycle (one workgroup per instance): y_k = R_k^{-1} s_k by columns -- y_l = s_l / R_ll, then s_i -= R_il y_l
// for every i < l in parallel -- into yc[i * K + k] for i < ncols (0 past the instance's own columns), for the
// update lambda_k += V y_k; an instance whose estimate is not finite is done.
__global__ __launch_bounds__(kBlock) void k_gm_solve(int K, int kk, int ncols, GmState g, double* __restrict__ yc) {
    __shared__ double sh[kGmMax];
    const int k = blockIdx.x;
    const int nc = g.cols[k];
    const double* __restrict__ Hk = g.H + (size_t)k * kk * kk;
    const double* __restrict__ sk = g.s + (size_t)k * (kk + 1);
    for (int i = threadIdx.x; i < ncols; i += kBlock) {
        if (i < nc) sh[i] = sk[i];
        else yc[(size_t)i * K + k] = 0.0;
    }
    for (int l = nc - 1; l >= 0; --l) {
        __syncthreads();
        const double d = Hk[(size_t)l * kk + l];
        const double yl = d != 0.0 ? sh[l] / d : 0.0;
        for (int i = threadIdx.x; i < l; i += kBlock) sh[i] -= Hk[(size_t)l * kk + i] * yl;
        if (threadIdx.x == 0) yc[(size_t)l * K + k] = yl;
    }
    if (threadIdx.x == 0 && !isfinite(g.rel[k])) g.done[k] = 1;
}

// A device buffer owned by one call.
struct BtBuf {
    void* p = nullptr;
    explicit BtBuf(size_t bytes) { GH_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 16))); }
    ~BtBuf() {
        if (p) (void)hipFree(p);
    }
    BtBuf(const BtBuf&) = delete;
    BtBuf& operator=(const BtBuf&) = delete;
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

template <class F>
void bt_types(int mdt, int vdt, F&& f) {
    with_mdt(mdt, [&](auto MT) {
        with_vdt(vdt, [&](auto VT) { f(MT, VT); });
    });
}

}  // namespace ganit_hip

using namespace ganit_hip;

extern "C" {

int ganit_hip_batch_create(ganit_hip_ctx* h, int ninst, const ganit_hip_lp_desc* const* lps,
                           const double* const* inst_scalars, unsigned mdt_mask, unsigned vdt_mask,
                           ganit_hip_batch** out) {
    return guarded([&] {
        auto* ctx = reinterpret_cast<Ctx*>(h);
        GH_REQUIRE(ctx && out && ninst > 0 && lps && inst_scalars, "bad batch arguments");
        GH_REQUIRE((vdt_mask & 3u) && (mdt_mask & 7u), "empty dtype mask");
        auto* bt = new Batch;
        try {
            bt->ctx = ctx;
            bt->ninst = ninst;
            // ---- pack the block-diagonal A and A^T with global indices
            std::vector<int> aptr{0}, aidx, atptr{0}, atidx, qptr{0}, qidx;
            std::vector<double> aval, atval, qval;
            bool anyq = false;
            std::vector<BatchInst> inst(ninst);
            std::vector<int> col_inst, row_inst;
            for (int k = 0; k < ninst; ++k) {
                const ganit_hip_lp_desc& d = *lps[k];
                GH_REQUIRE(d.m >= 0 && d.n >= 0, "negative problem dimension");
                BatchInst& in = inst[k];
                in.m = d.m; in.n = d.n; in.row0 = bt->Mr; in.col0 = bt->N;
                // eta, norm_a, obj_scale, obj_const, bnorm, cnorm, w0 (and ||Q'|| for a QP)
                const double* sc = inst_scalars[k];
                in.eta = sc[0]; in.norm_a = sc[1]; in.obj_scale = sc[2]; in.obj_const = sc[3];
                in.bnorm = sc[4]; in.cnorm = sc[5]; in.w0 = sc[6];
                in.hasq = d.q_ptr != nullptr;
                in.norm_q = in.hasq ? sc[7] : 0.0;
                anyq = anyq || in.hasq;
                for (int j = 0; j < d.n; ++j) {  // this instance's rows of the block-diagonal Q
                    if (in.hasq)
                        for (int p = d.q_ptr[j]; p < d.q_ptr[j + 1]; ++p) {
                            GH_REQUIRE(d.q_idx[p] >= 0 && d.q_idx[p] < d.n, "Q column index out of range");
                            qidx.push_back(in.col0 + d.q_idx[p]);
                            qval.push_back(d.q_val[p]);
                        }
                    qptr.push_back((int)qidx.size());
                }
                for (int i = 0; i < d.m; ++i) {
                    GH_REQUIRE(d.a_ptr[i + 1] - d.a_ptr[i] <= kSplit, "batched rows are limited to 8192 entries");
                    for (int p = d.a_ptr[i]; p < d.a_ptr[i + 1]; ++p) {
                        aidx.push_back(in.col0 + d.a_idx[p]);
                        aval.push_back(d.a_val[p]);
                    }
                    aptr.push_back((int)aidx.size());
                    row_inst.push_back(k);
                }
                for (int j = 0; j < d.n; ++j) {
                    GH_REQUIRE(d.at_ptr[j + 1] - d.at_ptr[j] <= kSplit, "batched columns are limited to 8192 entries");
                    for (int p = d.at_ptr[j]; p < d.at_ptr[j + 1]; ++p) {
                        atidx.push_back(in.row0 + d.at_idx[p]);
                        atval.push_back(d.at_val[p]);
                    }
                    atptr.push_back((int)atidx.size());
                    col_inst.push_back(k);
                }
                bt->Mr += d.m;
                bt->N += d.n;
                GH_REQUIRE(aidx.size() < (size_t)std::numeric_limits<int>::max(),
                           "batched nonzeros exceed 32-bit indexing");
            }
            bt->A = csr_create(ctx, bt->Mr, bt->N, aptr.data(), aidx.data(), aval.data(), mdt_mask);
            bt->AT = csr_create(ctx, bt->N, bt->Mr, atptr.data(), atidx.data(), atval.data(), mdt_mask);
            if (anyq) bt->Q = csr_create(ctx, bt->N, bt->N, qptr.data(), qidx.data(), qval.data(), mdt_mask);
            bt->aptr_h = std::move(aptr);
            bt->atptr_h = std::move(atptr);
            bt->aidx_h = std::move(aidx);
            bt->atidx_h = std::move(atidx);
            bt->inst_h = inst;
            bt_partition(bt, kBlock, kTile);
            bt->col_inst = (int*)bt->alloc(sizeof(int) * std::max<size_t>(col_inst.size(), 1));
            bt->row_inst = (int*)bt->alloc(sizeof(int) * std::max<size_t>(row_inst.size(), 1));
            if (!col_inst.empty())
                GH_CHECK(hipMemcpyAsync(bt->col_inst, col_inst.data(), sizeof(int) * col_inst.size(), hipMemcpyHostToDevice, ctx->stream));
            if (!row_inst.empty())
                GH_CHECK(hipMemcpyAsync(bt->row_inst, row_inst.data(), sizeof(int) * row_inst.size(), hipMemcpyHostToDevice, ctx->stream));
            bt->inst = (BatchInst*)bt->alloc(sizeof(BatchInst) * ninst);
            GH_CHECK(hipMemcpyAsync(bt->inst, inst.data(), sizeof(BatchInst) * ninst, hipMemcpyHostToDevice, ctx->stream));
            bt->st = (BatchState*)bt->alloc(sizeof(BatchState) * ninst);
            bt->sums = (double*)bt->alloc(sizeof(double) * kBOut * ninst);
            bt->running = (int*)bt->alloc(sizeof(int));
            GH_CHECK(hipHostMalloc((void**)&bt->host_running, sizeof(int), hipHostMallocDefault));
            // ---- vectors (concatenated)
            std::vector<double> cat_n(bt->N), cat_m(bt->Mr);
            for (int vdt = 0; vdt < 2; ++vdt) {
                if (!(vdt_mask & (1u << vdt))) continue;
                const int slot = vdt == GANIT_F64 ? 0 : 1;
                bt->on[slot] = true;
                BatchVecs& v = bt->vs[slot];
                const size_t e = dtype_size(vdt);
                auto upload_cat = [&](void** dst, bool is_col, const double* ganit_hip_lp_desc::*field) {
                    std::vector<double>& cat = is_col ? cat_n : cat_m;
                    for (int k = 0; k < ninst; ++k) {
                        const int len = is_col ? inst[k].n : inst[k].m, off = is_col ? inst[k].col0 : inst[k].row0;
                        const double* src = lps[k]->*field;
                        for (int q = 0; q < len; ++q) cat[off + q] = src[q];
                    }
                    *dst = bt->alloc(e * cat.size());
                    if (vdt == GANIT_F64) {
                        GH_CHECK(hipMemcpyAsync(*dst, cat.data(), 8 * cat.size(), hipMemcpyHostToDevice, ctx->stream));
                        GH_CHECK(hipStreamSynchronize(ctx->stream));
                    } else {
                        std::vector<float> f(cat.begin(), cat.end());
                        GH_CHECK(hipMemcpyAsync(*dst, f.data(), 4 * f.size(), hipMemcpyHostToDevice, ctx->stream));
                        GH_CHECK(hipStreamSynchronize(ctx->stream));
                    }
                };
                upload_cat(&v.c, true, &ganit_hip_lp_desc::c);
                upload_cat(&v.l, true, &ganit_hip_lp_desc::l);
                upload_cat(&v.u, true, &ganit_hip_lp_desc::u);
                upload_cat(&v.cres, true, &ganit_hip_lp_desc::cres);
                upload_cat(&v.lo, false, &ganit_hip_lp_desc::lo);
                upload_cat(&v.hi, false, &ganit_hip_lp_desc::hi);
                upload_cat(&v.rres, false, &ganit_hip_lp_desc::rres);
                for (int q = 0; q < 2; ++q) { v.X[q] = bt->alloc(e * bt->N); v.Y[q] = bt->alloc(e * bt->Mr); }
                v.xbar = bt->alloc(e * bt->N); v.sumx = bt->alloc(e * bt->N); v.xa = bt->alloc(e * bt->N);
                v.xr = bt->alloc(e * bt->N);
                v.sumy = bt->alloc(e * bt->Mr); v.ya = bt->alloc(e * bt->Mr); v.yr = bt->alloc(e * bt->Mr);
            }
            GH_CHECK(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            delete bt;
            throw;
        }
        *out = reinterpret_cast<ganit_hip_batch*>(bt);
    });
}

int ganit_hip_batch_destroy(ganit_hip_batch* h) {
    return guarded([&] { delete reinterpret_cast<Batch*>(h); });
}

int ganit_hip_batch_dims(const ganit_hip_batch* h, int64_t* dims) {
    return guarded([&] {
        auto* bt = reinterpret_cast<const Batch*>(h);
        GH_REQUIRE(bt && dims, "null argument");
        dims[0] = bt->ninst; dims[1] = bt->Mr; dims[2] = bt->N; dims[3] = bt->A->nnz;
        dims[4] = bt->A->nblocks; dims[5] = bt->AT->nblocks;
    });
}

int ganit_hip_batch_load(ganit_hip_batch* h, int vdt, const double* x, const double* y, const double* w) {
    return guarded([&] {
        auto* bt = reinterpret_cast<Batch*>(h);
        GH_REQUIRE(bt && x && y && w, "null argument");
        const int slot = vdt == GANIT_F64 ? 0 : 1;
        GH_REQUIRE(bt->on[slot], "vector dtype not resident");
        BatchVecs& v = bt->vs[slot];
        bt->cur = 0;
        hipStream_t st = bt->ctx->stream;
        auto up = [&](const double* src, void* dst, int64_t k) {
            if (k == 0) return;
            if (vdt == GANIT_F64) {
                GH_CHECK(hipMemcpyAsync(dst, src, 8 * k, hipMemcpyHostToDevice, st));
            } else {
                std::vector<float> f(src, src + k);
                GH_CHECK(hipMemcpyAsync(dst, f.data(), 4 * k, hipMemcpyHostToDevice, st));
            }
            GH_CHECK(hipStreamSynchronize(st));
        };
        std::vector<double> zn(bt->N, 0.0), zm(bt->Mr, 0.0);
        up(x, v.X[0], bt->N); up(y, v.Y[0], bt->Mr); up(x, v.xr, bt->N); up(y, v.yr, bt->Mr);
        up(zn.data(), v.sumx, bt->N); up(zm.data(), v.sumy, bt->Mr);
        std::vector<BatchState> s(bt->ninst);
        for (int k = 0; k < bt->ninst; ++k) {
            std::memset(&s[k], 0, sizeof(BatchState));
            s[k].w = w[k];
            s[k].restart_err = -1.0;
            s[k].status = kRunning;
            s[k].active = 1;
            s[k].adopt = -1;
        }
        GH_CHECK(hipMemcpyAsync(bt->st, s.data(), sizeof(BatchState) * bt->ninst, hipMemcpyHostToDevice, st));
        GH_CHECK(hipStreamSynchronize(st));
        bt_list(bt, std::vector<char>(bt->ninst, 1));
    });
}

int ganit_hip_batch_set_bounds(ganit_hip_batch* h, int k, const double* l, const double* u) {
    return guarded([&] {
        auto* bt = reinterpret_cast<Batch*>(h);
        GH_REQUIRE(bt && l && u, "null argument");
        GH_REQUIRE(k >= 0 && k < bt->ninst, "instance index out of range");
        const BatchInst& in = bt->inst_h[k];
        for (int j = 0; j < in.n; ++j)
            GH_REQUIRE(!(l[j] > u[j]) && !std::isnan(l[j]) && !std::isnan(u[j]),
                       "variable bounds must satisfy l <= u and not be NaN");
        hipStream_t st = bt->ctx->stream;
        for (int slot = 0; slot < 2; ++slot) {
            if (!bt->on[slot] || in.n == 0) continue;
            BatchVecs& v = bt->vs[slot];
            if (slot == 0) {
                GH_CHECK(hipMemcpyAsync((double*)v.l + in.col0, l, 8 * (size_t)in.n, hipMemcpyHostToDevice, st));
                GH_CHECK(hipMemcpyAsync((double*)v.u + in.col0, u, 8 * (size_t)in.n, hipMemcpyHostToDevice, st));
                GH_CHECK(hipStreamSynchronize(st));
            } else {
                std::vector<float> fl(l, l + in.n), fu(u, u + in.n);
                GH_CHECK(hipMemcpyAsync((float*)v.l + in.col0, fl.data(), 4 * (size_t)in.n, hipMemcpyHostToDevice, st));
                GH_CHECK(hipMemcpyAsync((float*)v.u + in.col0, fu.data(), 4 * (size_t)in.n, hipMemcpyHostToDevice, st));
                GH_CHECK(hipStreamSynchronize(st));
            }
        }
    });
}

int ganit_hip_batch_handover(ganit_hip_batch* h, int from, int to) {
    return guarded([&] {
        auto* bt = reinterpret_cast<Batch*>(h);
        GH_REQUIRE(bt, "null batch");
        const int sa = from == GANIT_F64 ? 0 : 1, sb = to == GANIT_F64 ? 0 : 1;
        GH_REQUIRE(bt->on[sa] && bt->on[sb], "vector dtype not resident");
        BatchVecs &a = bt->vs[sa], &b = bt->vs[sb];
        hipStream_t st = bt->ctx->stream;
        with_vdt(from, [&](auto S) {
            with_vdt(to, [&](auto D) {
                using SV = typename decltype(S)::type;
                using DV = typename decltype(D)::type;
                auto conv = [&](void* src, void* dst, int64_t k) {
                    if (k > 0 && src != dst)
                        hipLaunchKernelGGL((k_bt_convert<SV, DV>), dim3(bt_grid(k)), dim3(kBlock), 0, st, k,
                                           (const SV*)src, (DV*)dst);
                };
                conv(a.X[bt->cur], b.X[bt->cur], bt->N);
                conv(a.Y[bt->cur], b.Y[bt->cur], bt->Mr);
                conv(a.xr, b.xr, bt->N);
                conv(a.yr, b.yr, bt->Mr);
            });
        });
        // fresh running sums for the next phase
        with_vdt(to, [&](auto D) {
            using DV = typename decltype(D)::type;
            std::vector<DV> zz(std::max(bt->N, bt->Mr), DV(0));
            GH_CHECK(hipMemcpyAsync(b.sumx, zz.data(), sizeof(DV) * bt->N, hipMemcpyHostToDevice, st));
            GH_CHECK(hipMemcpyAsync(b.sumy, zz.data(), sizeof(DV) * bt->Mr, hipMemcpyHostToDevice, st));
            GH_CHECK(hipStreamSynchronize(st));
        });
        GH_CHECK(hipGetLastError());
    });
}

int ganit_hip_batch_select(ganit_hip_batch* h, const int* sel) {
    return guarded([&] {
        auto* bt = reinterpret_cast<Batch*>(h);
        GH_REQUIRE(bt && sel, "null argument");
        int* d = (int*)bt->part_m;  // the partial buffer is free between phases
        GH_REQUIRE((size_t)bt->ninst * sizeof(int) <= sizeof(double) * bt->part_m_cap,
                   "selection does not fit the staging buffer");
        GH_CHECK(hipMemcpyAsync(d, sel, sizeof(int) * bt->ninst, hipMemcpyHostToDevice, bt->ctx->stream));
        hipLaunchKernelGGL(k_bt_select, dim3(std::min(1024, (bt->ninst + 255) / 256)), dim3(256), 0, bt->ctx->stream,
                           bt->ninst, bt->st, d);
        GH_CHECK(hipGetLastError());
        GH_CHECK(hipStreamSynchronize(bt->ctx->stream));
    });
}

// params: tol, infeas_tol, max_iter, check_every (even), stall_checks, stall_exit, time_limit_s
int ganit_hip_batch_phase(ganit_hip_batch* h, int mdt, int vdt, const double* params, int* running_left) {
    return guarded([&] {
        auto* bt = reinterpret_cast<Batch*>(h);
        GH_REQUIRE(bt && params, "null argument");
        const int slot = vdt == GANIT_F64 ? 0 : 1;
        GH_REQUIRE(bt->on[slot], "vector dtype not resident");
        bt->P.tol = params[0];
        bt->P.infeas_tol = params[1];
        bt->P.max_iter = (int)std::min(params[2], 2.0e9);
        bt->P.check_every = (int)params[3];
        bt->P.stall_checks = (int)params[4];
        bt->P.stall_exit = params[5] != 0 ? 1 : 0;
        bt->P.end_phase = 0;
        const double time_limit = params[6];
        // FP64 evaluation of a reduced-precision level's checks (the single solve's GANIT_CHECK_FP64)
        const bool e64 = params[7] != 0 && (mdt != GANIT_F64 || vdt != GANIT_F64);
        GH_REQUIRE(!e64 || (bt->on[0] && bt->A->v64), "FP64 evaluation needs the FP64 problem resident");
        GH_REQUIRE(bt->P.check_every >= 2 && bt->P.check_every % 2 == 0, "batched check interval must be even");
        GH_REQUIRE(std::isfinite(bt->P.tol) && std::isfinite(bt->P.infeas_tol), "batched tolerances must be finite");
        hipStream_t st = bt->ctx->stream;
        const auto t0 = std::chrono::steady_clock::now();
        bt_types(mdt, vdt, [&](auto MT, auto VT) {
            using M = typename decltype(MT)::type;
            using V = typename decltype(VT)::type;
            bt_list(bt, bt_active(bt));  // launch only the selected instances' workgroups
            // tail_mode bit 1: a reduced-precision phase ends once its listed instances are a tail (launch-bound:
            // FP32 saves nothing there), its running instances leaving as stalled at the next check
            const bool reduced = mdt != GANIT_F64 || vdt != GANIT_F64;
            auto at_tail = [&]() {
                return (bt->tail_mode & 2) && reduced && bt->P.stall_exit && bt_tail_nnz(bt, bt->listed) > 0;
            };
            bt->P.end_phase = at_tail() ? 1 : 0;
            auto check = [&](int avg, int mode) {
                if (e64) bt_check<M, V, true>(bt, avg, mode);
                else bt_check<M, V>(bt, avg, mode);
            };
            check(0, kInit);  // k_restart of every running instance
            auto make_key = [&]() {
                return Batch::GraphKey{mdt, vdt, bt->cur, bt->P.check_every, bt->P.tol, bt->P.infeas_tol,
                                       bt->P.max_iter, bt->P.stall_checks, bt->P.stall_exit, e64 ? 1 : 0,
                                       bt->P.end_phase};
            };
            Batch::GraphKey key = make_key();
            auto block = [&]() {
                for (int k = 0; k < bt->P.check_every; ++k) bt_iteration<M, V>(bt, mdt);
                check(1, kCheck);
                hipLaunchKernelGGL(k_bt_deactivate, dim3(std::min(1024, (bt->ninst + 255) / 256)), dim3(256), 0, st,
                                   bt->ninst, bt->st);
            };
            const char* ng = std::getenv("GANIT_HIP_NO_GRAPH");
            const bool graphs = !(ng && ng[0] == '1');
            // the graph of one block (check_every iterations + check) for the current launch lists
            auto graph = [&]() -> hipGraphExec_t {
                auto it = bt->graphs.find(key);
                if (it != bt->graphs.end()) return it->second;
                hipGraph_t g = nullptr;
                const int cur0 = bt->cur;
                GH_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
                try {
                    block();
                } catch (...) {
                    (void)hipStreamEndCapture(st, &g);
                    if (g) (void)hipGraphDestroy(g);
                    bt->cur = cur0;
                    throw;
                }
                GH_CHECK(hipStreamEndCapture(st, &g));
                hipGraphExec_t ex = nullptr;
                const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                GH_CHECK(e);
                bt->cur = cur0;
                bt->graphs.emplace(key, ex);
                return ex;
            };
            GH_CHECK(hipMemcpyAsync(bt->host_running, bt->running, sizeof(int), hipMemcpyDeviceToHost, st));
            GH_CHECK(hipStreamSynchronize(st));
            int running = *bt->host_running;
            int listed = running;  // instances in the launch lists
            while (running > 0) {
                if (graphs) GH_CHECK(hipGraphLaunch(graph(), st));
                else block();
                if (bt->tail_lpr_a > 0) bt->tail_iters += bt->P.check_every;
                GH_CHECK(hipMemcpyAsync(bt->host_running, bt->running, sizeof(int), hipMemcpyDeviceToHost, st));
                GH_CHECK(hipStreamSynchronize(st));
                running = *bt->host_running;
                const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                if (running > 0 && el >= time_limit) {
                    hipLaunchKernelGGL(k_bt_stop, dim3(std::min(1024, (bt->ninst + 255) / 256)), dim3(256), 0, st,
                                       bt->ninst, bt->st);
                    BatchVecs& v = bt->vs[slot];
                    hipLaunchKernelGGL(k_bt_restart<V>, dim3(bt_grid((int64_t)bt->N + bt->Mr)), dim3(kBlock), 0, st,
                                       bt->N, bt->Mr, bt->col_inst, bt->row_inst, bt->st, (V*)v.X[bt->cur],
                                       (const V*)v.xa, (V*)v.xr, (V*)v.sumx, (V*)v.Y[bt->cur], (const V*)v.ya,
                                       (V*)v.yr, (V*)v.sumy);
                    hipLaunchKernelGGL(k_bt_deactivate, dim3(std::min(1024, (bt->ninst + 255) / 256)), dim3(256), 0,
                                       st, bt->ninst, bt->st);
                    GH_CHECK(hipStreamSynchronize(st));
                    break;
                }
                // shrink the grids once half the listed instances have stopped (a few recaptures per phase)
                if (running > 0 && 2 * running <= listed) {
                    bt_list(bt, bt_active(bt));
                    listed = running;
                    if (!bt->P.end_phase && at_tail()) {
                        bt->P.end_phase = 1;
                        key = make_key();
                    }
                }
            }
            if (running_left) *running_left = running;
        });
        GH_CHECK(hipGetLastError());
    });
}

// Autotuner: time `reps` batched iterations (primal + dual kernel) at (mdt, vdt) for every row-block
// partition candidate, twice, and keep the fastest.  Runs on whatever iterate is loaded and advances it.
int ganit_hip_batch_autotune(ganit_hip_batch* h, int mdt, int vdt, int reps, double* ms, int* chosen) {
    return guarded([&] {
        auto* bt = reinterpret_cast<Batch*>(h);
        GH_REQUIRE(bt && reps > 0, "bad autotune arguments");
        const int slot = vdt == GANIT_F64 ? 0 : 1;
        GH_REQUIRE(bt->on[slot], "vector dtype not resident");
        std::vector<std::pair<int, int>> cand;
        for (const Variant& v : variant_candidates())
            if (v.lanes == 0) cand.emplace_back(v.rows, v.nnz);
        GH_REQUIRE(!cand.empty(), "no row-block candidates");
        hipStream_t st = bt->ctx->stream;
        hipEvent_t e0, e1;
        GH_CHECK(hipEventCreate(&e0));
        GH_CHECK(hipEventCreate(&e1));
        std::vector<double> best(cand.size(), std::numeric_limits<double>::infinity());
        try {
            bt_types(mdt, vdt, [&](auto MT, auto VT) {
                using M = typename decltype(MT)::type;
                using V = typename decltype(VT)::type;
                // every instance takes part while timing
                std::vector<BatchState> keep(bt->ninst);
                GH_CHECK(hipMemcpyAsync(keep.data(), bt->st, sizeof(BatchState) * bt->ninst, hipMemcpyDeviceToHost, st));
                GH_CHECK(hipStreamSynchronize(st));
                std::vector<BatchState> tmp(keep);
                for (int k = 0; k < bt->ninst; ++k) {
                    tmp[k].active = 1;
                    tmp[k].sigma = bt->inst_h[k].eta * tmp[k].w;
                    tmp[k].tau = bt->inst_h[k].eta / tmp[k].w;
                }
                GH_CHECK(hipMemcpyAsync(bt->st, tmp.data(), sizeof(BatchState) * bt->ninst, hipMemcpyHostToDevice, st));
                for (int round = 0; round < 2; ++round) {
                    for (size_t c = 0; c < cand.size(); ++c) {
                        bt_partition(bt, cand[c].first, cand[c].second);
                        for (int k = 0; k < 4; ++k) bt_iteration<M, V>(bt, mdt);
                        GH_CHECK(hipEventRecord(e0, st));
                        for (int k = 0; k < reps; ++k) bt_iteration<M, V>(bt, mdt);
                        GH_CHECK(hipEventRecord(e1, st));
                        GH_CHECK(hipEventSynchronize(e1));
                        float t = 0.f;
                        GH_CHECK(hipEventElapsedTime(&t, e0, e1));
                        best[c] = std::min(best[c], (double)t / reps);
                    }
                }
                GH_CHECK(hipMemcpyAsync(bt->st, keep.data(), sizeof(BatchState) * bt->ninst, hipMemcpyHostToDevice, st));
                GH_CHECK(hipStreamSynchronize(st));
            });
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        const size_t pick = (size_t)(std::min_element(best.begin(), best.end()) - best.begin());
        bt_partition(bt, cand[pick].first, cand[pick].second);
        if (ms)
            for (size_t c = 0; c < cand.size(); ++c) ms[c] = best[c];
        if (chosen) *chosen = (int)pick;
    });
}

// Launch-bound tail switch: mode bit 0 (the default) lets the iteration kernels of a small running set run
// lanes-per-row (0: row blocks, row-order sums, to the end), bit 1 ends a reduced-precision phase once its running
// set is such a tail; -1 leaves the mode; tail_iters (may be null) = iterations run on tail lists so far.
int ganit_hip_batch_tail(ganit_hip_batch* h, int mode, int64_t* tail_iters) {
    return guarded([&] {
        auto* bt = reinterpret_cast<Batch*>(h);
        GH_REQUIRE(bt && mode >= -1 && mode <= 3, "bad tail arguments");
        if (mode >= 0 && mode != bt->tail_mode) {
            GH_CHECK(hipStreamSynchronize(bt->ctx->stream));
            bt->tail_mode = mode;
            const std::vector<char> on = bt->listed;
            bt->listed.assign(bt->ninst, 0);  // rebuild the lists (and drop the graphs) under the new mode
            if (std::count(on.begin(), on.end(), (char)1) > 0) {
                bt_list(bt, on);
            } else {
                bt_tail_lists(bt, on);
                bt->drop_graphs();
            }
        }
        if (tail_iters) *tail_iters = bt->tail_iters;
    });
}

// FP64 KKT of the FP64 working set's iterate of every instance (the certification, src/pdhg_impl.hpp:301).
int ganit_hip_batch_certify(ganit_hip_batch* h, int mdt) {
    return guarded([&] {
        auto* bt = reinterpret_cast<Batch*>(h);
        GH_REQUIRE(bt && bt->on[0], "FP64 working set not resident");
        hipStream_t st = bt->ctx->stream;
        // evaluate every instance: mark all active, run the check in FINAL mode, restore the flags
        std::vector<BatchState> keep(bt->ninst);
        GH_CHECK(hipMemcpyAsync(keep.data(), bt->st, sizeof(BatchState) * bt->ninst, hipMemcpyDeviceToHost, st));
        GH_CHECK(hipStreamSynchronize(st));
        std::vector<BatchState> tmp(keep);
        for (auto& s : tmp) s.active = 1;
        GH_CHECK(hipMemcpyAsync(bt->st, tmp.data(), sizeof(BatchState) * bt->ninst, hipMemcpyHostToDevice, st));
        bt_list(bt, std::vector<char>(bt->ninst, 1));
        with_mdt(mdt, [&](auto MT) { bt_check<typename decltype(MT)::type, double>(bt, 0, kFinal); });
        GH_CHECK(hipMemcpyAsync(tmp.data(), bt->st, sizeof(BatchState) * bt->ninst, hipMemcpyDeviceToHost, st));
        GH_CHECK(hipStreamSynchronize(st));
        for (int k = 0; k < bt->ninst; ++k) {
            std::memcpy(keep[k].kkt64, tmp[k].kkt64, sizeof(keep[k].kkt64));
        }
        GH_CHECK(hipMemcpyAsync(bt->st, keep.data(), sizeof(BatchState) * bt->ninst, hipMemcpyHostToDevice, st));
        GH_CHECK(hipStreamSynchronize(st));
    });
}

// ints[5k..5k+4] = status, iterations, restarts, best_is_avg, active; dbl[13k..13k+12] = phase result KKT
// (pobj, dobj, rel_p, rel_d, rel_gap, err), FP64 KKT of the same six, primal weight.
int ganit_hip_batch_results(ganit_hip_batch* h, int* ints, double* dbl) {
    return guarded([&] {
        auto* bt = reinterpret_cast<Batch*>(h);
        GH_REQUIRE(bt && ints && dbl, "null argument");
        std::vector<BatchState> s(bt->ninst);
        GH_CHECK(hipMemcpyAsync(s.data(), bt->st, sizeof(BatchState) * bt->ninst, hipMemcpyDeviceToHost, bt->ctx->stream));
        GH_CHECK(hipStreamSynchronize(bt->ctx->stream));
        for (int k = 0; k < bt->ninst; ++k) {
            ints[5 * k + 0] = s[k].status;
            ints[5 * k + 1] = s[k].it;
            ints[5 * k + 2] = s[k].restarts;
            ints[5 * k + 3] = s[k].best_avg;
            ints[5 * k + 4] = s[k].active;
            for (int q = 0; q < 6; ++q) dbl[13 * k + q] = s[k].best[q];
            for (int q = 0; q < 6; ++q) dbl[13 * k + 6 + q] = s[k].kkt64[q];
            dbl[13 * k + 12] = s[k].w;
        }
    });
}

int ganit_hip_batch_store(ganit_hip_batch* h, int vdt, double* x, double* y) {
    return guarded([&] {
        auto* bt = reinterpret_cast<Batch*>(h);
        GH_REQUIRE(bt, "null batch");
        const int slot = vdt == GANIT_F64 ? 0 : 1;
        GH_REQUIRE(bt->on[slot], "vector dtype not resident");
        BatchVecs& v = bt->vs[slot];
        hipStream_t st = bt->ctx->stream;
        auto down = [&](const void* src, double* dst, int64_t k) {
            if (!dst || k == 0) return;
            if (vdt == GANIT_F64) {
                GH_CHECK(hipMemcpyAsync(dst, src, 8 * k, hipMemcpyDeviceToHost, st));
                GH_CHECK(hipStreamSynchronize(st));
            } else {
                std::vector<float> f(k);
                GH_CHECK(hipMemcpyAsync(f.data(), src, 4 * k, hipMemcpyDeviceToHost, st));
                GH_CHECK(hipStreamSynchronize(st));
                for (int64_t i = 0; i < k; ++i) dst[i] = f[i];
            }
        };
        down(v.X[bt->cur], x, bt->N);
        down(v.Y[bt->cur], y, bt->Mr);
    });
}

// params = {max products, tol, restart length}; tau, sigma, mask_tol: one per instance; gx (N) in the scaled space.
int ganit_hip_batch_adjoint(ganit_hip_batch* h, int mdt, int vdt, const double* gx, const double* tau,
                            const double* sigma, const double* mask_tol, const double* params, double* p_out,
                            double* mu_out, double* info) {
    return guarded([&] {
        auto* bt = reinterpret_cast<Batch*>(h);
        GH_REQUIRE(bt && gx && tau && sigma && mask_tol && params && p_out && mu_out && info, "null argument");
        const int max_iter = (int)std::min(params[0], 2.0e9);
        const double tol = params[1];
        const int kk = (int)params[2];
        GH_REQUIRE(max_iter > 0 && kk > 0 && kk <= kGmMax, "bad adjoint parameters (restart length 1..1024)");
        const int slot = vdt == GANIT_F64 ? 0 : 1;
        GH_REQUIRE(bt->on[0] && bt->on[slot], "vector dtype not resident");
        const int K = bt->ninst, N = bt->N, Mr = bt->Mr;
        for (int k = 0; k < K; ++k)
            GH_REQUIRE(tau[k] > 0 && sigma[k] > 0 && mask_tol[k] >= 0, "bad step sizes or mask tolerance");
        const int64_t L = (int64_t)N + Mr;
        const size_t e = dtype_size(vdt);
        hipStream_t st = bt->ctx->stream;
        const BatchVecs& v64 = bt->vs[0];
        BtBuf D(N), F(Mr), g(e * N), r(e * N), p(e * N), g64(8 * (size_t)N), lam(e * L), w(e * L),
            basis(e * (size_t)L * (kk + 1)), steps(e * 3 * (size_t)K), mt(8 * (size_t)K),
            hbuf(sizeof(double) * ((size_t)(2 * (kk + 1) + 1) * K)), p64(8 * (size_t)N), mu64(8 * (size_t)Mr);
        GH_CHECK(hipMemcpyAsync(g64.p, gx, 8 * (size_t)N, hipMemcpyHostToDevice, st));
        GH_CHECK(hipMemcpyAsync(mt.p, mask_tol, 8 * (size_t)K, hipMemcpyHostToDevice, st));
        const unsigned gl = bt_grid(L), gn = bt_grid(std::max(N, 1));
        hipLaunchKernelGGL(k_ba_masks, dim3(gl), dim3(kBlock), 0, st, N, Mr, bt->col_inst, mt.as<double>(),
                           (const double*)v64.X[bt->cur], (const double*)v64.l, (const double*)v64.u,
                           (const double*)v64.Y[bt->cur], (const double*)v64.lo, (const double*)v64.hi,
                           D.as<uint8_t>(), F.as<uint8_t>());
        std::vector<int> prods(K, 0);
        std::vector<double> rel(K, 0.0);
        bt_types(mdt, vdt, [&](auto MT, auto VT) {
            using M = typename decltype(MT)::type;
            using V = typename decltype(VT)::type;
            // per-instance step sizes in V: tau, sigma, 2 sigma
            std::vector<V> sv(3 * (size_t)K);
            for (int k = 0; k < K; ++k) {
                sv[k] = V(tau[k]);
                sv[K + k] = V(sigma[k]);
                sv[2 * K + k] = V(2.0 * sigma[k]);
            }
            GH_CHECK(hipMemcpyAsync(steps.p, sv.data(), e * sv.size(), hipMemcpyHostToDevice, st));
            GH_CHECK(hipStreamSynchronize(st));
            const V* tv = steps.as<V>();
            const V* sg = tv + K;
            const V* s2 = tv + 2 * K;
            V* lx = lam.as<V>();
            V* W = w.as<V>();
            V* Vb = basis.as<V>();
            hipLaunchKernelGGL(k_ba_init<V>, dim3(gl), dim3(kBlock), 0, st, N, Mr, g64.as<double>(), g.as<V>(), lx);
            const bool hasq = bt->Q != nullptr;
            const CsrView<M> qv = hasq ? bt->Q->view<M>() : CsrView<M>{0, nullptr, nullptr, nullptr};
            // out = (I - J^T) in, for every instance
            auto apply = [&](const V* in, V* out) {
                if (N > 0 && bt->AT->nblocks > 0)
                    hipLaunchKernelGGL((k_ba_at<M, V>), dim3(bt->AT->nblocks), dim3(kBlock), 0, st,
                                       bt->AT->view<M>(), bt->AT->blocks, bt->col_inst, s2, in + N, F.as<uint8_t>(),
                                       in, D.as<uint8_t>(), r.as<V>(), p.as<V>());
                if (Mr > 0 && bt->A->nblocks > 0)
                    hipLaunchKernelGGL((k_ba_a<M, V>), dim3(bt->A->nblocks), dim3(kBlock), 0, st, bt->A->view<M>(),
                                       bt->A->blocks, bt->row_inst, tv, p.as<V>(), F.as<uint8_t>(), in + N, out + N);
                if (N > 0) {
                    if (hasq)
                        hipLaunchKernelGGL((k_ba_x<M, V, true>), dim3(gn), dim3(kBlock), 0, st, N, qv, bt->col_inst,
                                           tv, sg, p.as<V>(), r.as<V>(), in, out);
                    else
                        hipLaunchKernelGGL((k_ba_x<M, V, false>), dim3(gn), dim3(kBlock), 0, st, N, qv, bt->col_inst,
                                           tv, sg, p.as<V>(), r.as<V>(), in, out);
                }
            };
            double* h1 = hbuf.as<double>();            // [(kk + 1) K] first Gram-Schmidt pass
            double* h2 = h1 + (size_t)(kk + 1) * K;    // [(kk + 1) K] second pass
            double* nr2 = h2 + (size_t)(kk + 1) * K;   // [K] squared norms
            auto dots = [&](const V* vecs, int ncols, const V* x, int xonly, double* dst) {
                hipLaunchKernelGGL(k_bgm_dots<V>, dim3(K, ncols), dim3(kBlock), 0, st, N, L, bt->inst, K, vecs, x,
                                   xonly, dst);
            };
            auto update = [&](const V* vecs, int ncols, const double* hc, double sgn, V* dst) {
                hipLaunchKernelGGL(k_bgm_update<V>, dim3(gl), dim3(kBlock), 0, st, N, L, bt->col_inst, bt->row_inst, K,
                                   vecs, ncols, hc, sgn, dst);
            };
            auto scale = [&](const V* src, V* dst) {
                hipLaunchKernelGGL(k_bgm_scale<V>, dim3(gl), dim3(kBlock), 0, st, N, L, bt->col_inst, bt->row_inst,
                                   src, nr2, dst);
            };
            // per-instance GMRES bookkeeping on the device (k_gm_*); the host polls the activity flags once per
            // restart cycle and every kPoll Arnoldi steps (steps past the last instance's convergence change nothing:
            // an inactive instance's column count, rotations and estimate are frozen)
            constexpr int kPoll = 8;
            const size_t kk1 = (size_t)kk + 1;
            BtBuf gH(8 * (size_t)K * kk * kk), gcs(8 * (size_t)K * kk), gsn(8 * (size_t)K * kk), gs(8 * (size_t)K * kk1),
                gd(8 * 3 * (size_t)K), gi(4 * 4 * (size_t)K);
            GmState gm{gH.as<double>(), gcs.as<double>(), gsn.as<double>(), gs.as<double>(), gd.as<double>(),
                       gd.as<double>() + K, gd.as<double>() + 2 * (size_t)K, gi.as<int>(), gi.as<int>() + K,
                       gi.as<int>() + 2 * (size_t)K, gi.as<int>() + 3 * (size_t)K};
            const unsigned gk = (unsigned)((K + 255) / 256);
            std::vector<int> acth(K);
            auto running_now = [&]() {
                GH_CHECK(hipMemcpyAsync(acth.data(), gm.act, sizeof(int) * K, hipMemcpyDeviceToHost, st));
                GH_CHECK(hipStreamSynchronize(st));
                return (int)std::count_if(acth.begin(), acth.end(), [](int a) { return a != 0; });
            };
            // ||(g_k, 0)|| per instance
            dots(g.as<V>(), 1, g.as<V>(), 1, nr2);
            hipLaunchKernelGGL(k_gm_init, dim3(gk), dim3(256), 0, st, K, (const double*)nr2, gm);
            int iters = 0;
            while (iters < max_iter) {
                // r0 = b - (I - J^T) lambda into basis column 0; beta_k = ||r0_k||
                apply(lx, W);
                hipLaunchKernelGGL(k_bgm_rhs_minus<V>, dim3(gl), dim3(kBlock), 0, st, N, L, g.as<V>(), W, Vb);
                dots(Vb, 1, Vb, 0, nr2);
                hipLaunchKernelGGL(k_gm_cycle, dim3(gk), dim3(256), 0, st, K, kk, tol, (const double*)nr2, gm);
                if (running_now() == 0) break;
                scale(Vb, Vb);
                int steps = 0;
                for (int j = 0; j < kk && iters < max_iter; ++j) {
                    V* vj1 = Vb + (size_t)(j + 1) * L;
                    apply(Vb + (size_t)j * L, W);
                    dots(Vb, j + 1, W, 0, h1);
                    update(Vb, j + 1, h1, -1.0, W);
                    dots(Vb, j + 1, W, 0, h2);
                    update(Vb, j + 1, h2, -1.0, W);
                    dots(W, 1, W, 0, nr2);
                    scale(W, vj1);
                    hipLaunchKernelGGL(k_gm_step, dim3(K), dim3(kWave), 0, st, K, kk, j, tol, (const double*)h1,
                                       (const double*)h2, (const double*)nr2, gm);
                    ++iters;
                    ++steps;
                    if ((j + 1) % kPoll == 0 && running_now() == 0) break;
                }
                // y_k = R_k^{-1} s_k, lambda_k += V y_k (coefficient c of instance k at h1[c * K + k])
                hipLaunchKernelGGL(k_gm_solve, dim3(K), dim3(kBlock), 0, st, K, kk, steps, gm, h1);
                update(Vb, steps, h1, 1.0, lx);
            }
            GH_CHECK(hipMemcpyAsync(prods.data(), gm.prods, sizeof(int) * K, hipMemcpyDeviceToHost, st));
            GH_CHECK(hipMemcpyAsync(rel.data(), gm.rel, sizeof(double) * K, hipMemcpyDeviceToHost, st));
            GH_CHECK(hipStreamSynchronize(st));
            // p and mu at the final lambda
            if (N > 0 && bt->AT->nblocks > 0)
                hipLaunchKernelGGL((k_ba_at<M, V>), dim3(bt->AT->nblocks), dim3(kBlock), 0, st, bt->AT->view<M>(),
                                   bt->AT->blocks, bt->col_inst, s2, lx + N, F.as<uint8_t>(), lx, D.as<uint8_t>(),
                                   r.as<V>(), p.as<V>());
            hipLaunchKernelGGL(k_ba_out<V>, dim3(gl), dim3(kBlock), 0, st, N, Mr, p.as<V>(), lx + N, F.as<uint8_t>(),
                               p64.as<double>(), mu64.as<double>());
        });
        GH_CHECK(hipGetLastError());
        if (N > 0) GH_CHECK(hipMemcpyAsync(p_out, p64.p, 8 * (size_t)N, hipMemcpyDeviceToHost, st));
        if (Mr > 0) GH_CHECK(hipMemcpyAsync(mu_out, mu64.p, 8 * (size_t)Mr, hipMemcpyDeviceToHost, st));
        std::vector<uint8_t> dh(N), fh(Mr);
        if (N > 0) GH_CHECK(hipMemcpyAsync(dh.data(), D.p, N, hipMemcpyDeviceToHost, st));
        if (Mr > 0) GH_CHECK(hipMemcpyAsync(fh.data(), F.p, Mr, hipMemcpyDeviceToHost, st));
        GH_CHECK(hipStreamSynchronize(st));
        for (int k = 0; k < K; ++k) {
            const BatchInst& in = bt->inst_h[k];
            info[4 * k + 0] = prods[k];
            info[4 * k + 1] = rel[k];
            info[4 * k + 2] = (double)std::count(dh.begin() + in.col0, dh.begin() + in.col0 + in.n, (uint8_t)1);
            info[4 * k + 3] = (double)std::count(fh.begin() + in.row0, fh.begin() + in.row0 + in.m, (uint8_t)1);
        }
    });
}

}  // extern "C"
