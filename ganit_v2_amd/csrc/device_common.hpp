// Device-side building blocks shared by the PDHG kernels (gfx950, wave64).
//
//   * value types: FP64 / FP32 / BF16 matrix values, FP64 / FP32 vectors;
//   * row groups: LPR lanes of one wavefront cooperate on one CSR row (LPR in {1..64});
//   * deterministic reductions: per-block tree sums in a fixed order, partials in HBM folded in block
//     order by a one-block follow-up launch -- no FP atomics, so a reduction returns the same bits on
//     every run;
//   * XCD-aware block numbering: the 8 XCDs of an MI355X receive workgroups round-robin, so consecutive
//     hardware block ids are remapped to give each XCD one contiguous slab of rows, keeping the slab's
//     gathers of x inside that XCD's own L2.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace ganit_hip {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kNumXcd = 8;
// LDS-staged row blocks ("stream" variant): a block covers at most kBlock rows and kTile nonzeros; a row
// longer than kLong is a block of its own and is reduced by the whole block instead.
constexpr int kTile = 2048;
constexpr int kLong = 256;
// Rows longer than kSplit are cut into chunks of kChunk nonzeros, one workgroup each; a second launch (one
// workgroup per split row) combines the chunk partials in chunk order (so very long rows spread over many CUs).
constexpr int kSplit = 8192;
constexpr int kChunk = 4096;
// A staged block whose columns lie within kWin consecutive entries of the gathered vector may stage that window
// of the vector in LDS with coalesced loads and gather from LDS instead of global memory (windowed blocks).
constexpr int kWin = 2048;
static_assert(kTile % kBlock == 0 && kWin % kBlock == 0 && kLong <= kTile, "staging is whole slots per thread");
static_assert((kWin & (kWin - 1)) == 0 && kWin >= 16, "the window swizzle permutes [0, kWin)");

// Window layout.  Gathers from a window hit LDS banks at the stride between a row's columns, and rows whose
// columns lie a fixed stride apart (a blending LP's quality rows: one column per crude, 16 apart) make 10-20
// lanes of a wave wait on one bank.  A swizzled window keeps entry w at slot w ^ (w >> 4) -- a permutation of
// [0, kWin) that moves stride-2, -8 or -16 runs onto distinct banks and maps each aligned run of 16 entries
// onto an aligned run of 16 slots (the four runs a wave fills landing on distinct banks), so the coalesced
// fill stays conflict-free.  Each matrix says which layout
// its windows use (CsrView::wswz, chosen at upload from the gathers' bank pattern); the values read are the
// same either way.
template <bool SW>
__device__ __forceinline__ int win_slot(int w) {
    if constexpr (SW) return w ^ (w >> 4);
    else return w;
}

// ---------------------------------------------------------------- BF16 storage
struct bf16_t {
    uint16_t bits;
};

__device__ __forceinline__ float bf16_to_float(bf16_t v) { return __uint_as_float(((uint32_t)v.bits) << 16); }

// Widen a stored matrix value to the vector type V (BF16 goes through float, exactly as on the host).
template <class V>
__device__ __forceinline__ V widen(double v) { return (V)v; }
template <class V>
__device__ __forceinline__ V widen(float v) { return (V)v; }
template <class V>
__device__ __forceinline__ V widen(bf16_t v) { return (V)bf16_to_float(v); }

// ---------------------------------------------------------------- CSR view
template <class M>
struct CsrView {
    int rows;
    const int* __restrict__ ptr;
    const int* __restrict__ idx;
    const M* __restrict__ val;
    // optional compact indices: idx = bcol[2 * block] + idx16 (null: use idx); bcol[2 * block + 1] = the
    // block's window length (columns bcol[2b] .. + length - 1), 0 when not windowed, -1 for a long row whose
    // span needs 32-bit indices (split-row chunks carry their own bases, SplitPlan::cbase)
    const uint16_t* __restrict__ idx16 = nullptr;
    const int* __restrict__ bcol = nullptr;
    int wswz = 0;  // 1: windows are staged swizzled (win_slot<true>)
};

// ---------------------------------------------------------------- block numbering
// Hardware dispatches block b to XCD b % 8.  Return a logical block id such that the blocks living on one
// XCD cover a contiguous range of logical ids (a bijection on [0, gridDim.x)).
__device__ __forceinline__ int xcd_block_id() {
    const int G = (int)gridDim.x, b = (int)blockIdx.x;
    const int q = G / kNumXcd, r = G % kNumXcd;
    const int xcd = b % kNumXcd, k = b / kNumXcd;
    return xcd * q + (xcd < r ? xcd : r) + k;
}

// ---------------------------------------------------------------- wave / group sums
// Sum over the LPR lanes of a row group; the group's lane 0 holds the total.  Fixed tree order.
template <int LPR, class T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) v += __shfl_down(v, off, LPR);
    return v;
}

// Dot product of CSR row `row` with x, computed by the LPR lanes of a group, accumulated in V.
template <int LPR, class M, class V>
__device__ __forceinline__ V row_dot(const CsrView<M>& a, int row, int lane, const V* __restrict__ x) {
    const int beg = a.ptr[row], end = a.ptr[row + 1];
    V t = V(0);
    for (int p = beg + lane; p < end; p += LPR) t += widen<V>(a.val[p]) * x[a.idx[p]];
    return group_sum<LPR>(t);
}

// ---------------------------------------------------------------- row sums in row order
// s[k] += lds[k * kTile + q] for q = beg .. end-1, in that order (the scalar CPU row loop's additions), with
// the aligned middle read 16 bytes at a time: a quarter of the LDS reads, and rows that start a fixed stride
// apart (8 entries for most columns of A^T) no longer meet on one bank.  lds must be 16-byte aligned.
template <class V, int NX>
__device__ __forceinline__ void ordered_row_sums(const V* lds, int beg, int end, V (&s)[NX]) {
    constexpr int W = 16 / (int)sizeof(V);
    typedef V Wide __attribute__((ext_vector_type(W)));
    static_assert(kTile % W == 0, "tiles start 16-byte aligned");
    int q = beg;
    const int head = min(end, (beg + W - 1) & ~(W - 1));
    for (; q < head; ++q) {
#pragma unroll
        for (int k = 0; k < NX; ++k) s[k] += lds[k * kTile + q];
    }
#pragma unroll 8
    for (; q + W <= end; q += W) {
#pragma unroll
        for (int k = 0; k < NX; ++k) {
            const Wide w = *reinterpret_cast<const Wide*>(lds + k * kTile + q);
#pragma unroll
            for (int e = 0; e < W; ++e) s[k] += w[e];
        }
    }
    for (; q < end; ++q) {
#pragma unroll
        for (int k = 0; k < NX; ++k) s[k] += lds[k * kTile + q];
    }
}

// ---------------------------------------------------------------- block sums in the vector type
// Sum NX values of type V over the block; valid in thread 0.  Fixed order (wave tree, then waves in order).
template <class V, int NX>
__device__ __forceinline__ void block_sum_v(V (&v)[NX], V* lds /* (kBlock / kWave) * NX */) {
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < NX; ++k) {
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, kWave);
    }
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NX; ++k) lds[wave * NX + k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NX; ++k) {
            V s = lds[k];
            for (int w = 1; w < kBlock / kWave; ++w) s += lds[w * NX + k];
            v[k] = s;
        }
    }
}

// ---------------------------------------------------------------- split rows
// chunk[4c .. 4c+3] = (row, first nonzero, end nonzero, split-row index s); first[s] .. first[s+1] are the
// chunks of split row s; part holds NX partial sums per chunk (stored as double, formed in the vector type).
// A kernel over a matrix with split rows runs twice: the main pass (finish = 0) over the row blocks and the
// chunks, which only write their partials, then the finish pass (finish = 1, one workgroup per split row)
// that folds each row's partials in a fixed order and applies the row function.  The kernel boundary
// orders the two, so no workgroup waits for or fences against another.
struct SplitPlan {
    int nblocks;  // main pass: workgroups [0, nblocks) are row blocks, [nblocks, nblocks + nchunks) chunks
    const int* __restrict__ chunk;
    const int* __restrict__ first;
    double* part;
    int finish;   // 1: workgroup b folds split row b
    // compact indices of the chunks: chunk c's columns = cbase[c] + idx16; cbase[c] <= -2: the chunk's columns are
    // consecutive from -2 - cbase[c] (no index is read); -1 or null: 32-bit
    const int* __restrict__ cbase = nullptr;
};

// ---------------------------------------------------------------- LDS-staged row blocks
// A row block's head, the first thing a kernel reads of it: the descriptor (first row, end row, first nonzero,
// end nonzero; one 16-byte load) and, with compact indices, the block's base column and window length (an
// independent load, in flight at the same time).  wraw: window length, 0 none, -1 a long row with 32-bit indices.
// A batched kernel loads it beside its instance's state, so the two round trips overlap.
struct BlockHead {
    int4 d;
    int cb, wraw;
};
template <class M>
__device__ __forceinline__ BlockHead block_head(const CsrView<M>& a, const int* __restrict__ blocks, int b) {
    BlockHead h;
    h.d = reinterpret_cast<const int4*>(blocks)[b];
    h.cb = 0;
    h.wraw = -1;
    if (a.idx16) {
        const int2 bw = reinterpret_cast<const int2*>(a.bcol)[b];
        h.cb = bw.x;
        h.wraw = bw.y;
    }
    return h;
}

template <class M, class V, int NX, bool WIN, class Gather, class Pref, class RowFn>
__device__ __forceinline__ void stream_head_pf(const CsrView<M>& a, const BlockHead& h, Gather gather, Pref pref,
                                               RowFn row_fn, V* lds);

// One block of the stream variant: rows [r0, r1) of `a` whose nonzeros fit one LDS tile.  Every product
// val * x_k[idx] (k < NX gathered vectors, `gather(k, i)`) is formed by a coalesced sweep over the block's
// nonzeros and stored in LDS (rounded to V there); then thread t adds up row r0 + t sequentially in row
// order -- the same operation sequence as a scalar CPU row loop, so these row sums are bit-identical to it.
// A single row longer than kLong is instead swept by all threads and reduced with a fixed tree.
// `row_fn(r, sums, ops)` runs once per row, in the thread that owns it (thread 0 for a long row), with
// ops = pref(r): the row's own operands, loaded early (stream_block_pf); stream_block has none.
// WIN (NX = 1 only): windowed blocks gather from a copy of their window of the vector in LDS; lds then holds
// kTile + kWin elements.
template <class M, class V, int NX, bool WIN = false, class Gather, class Pref, class RowFn>
__device__ __forceinline__ void stream_block_pf(const CsrView<M>& a, const int* __restrict__ blocks, int b,
                                                Gather gather, Pref pref, RowFn row_fn,
                                                V* lds /* NX * kTile (+ kWin with WIN) */, const SplitPlan& sp) {
    static_assert(!WIN || NX == 1, "windowed gathers stage one vector");
    if (sp.finish) {  // finish pass: fold split row b's chunk partials, fixed-order tree
        const int q0 = sp.first[b], nch = sp.first[b + 1] - q0;
        // the row and its operands are loaded before the fold (by every thread, unconditionally: one cache
        // line, and no branch that would force a wait), so their latency overlaps it
        const int row = sp.chunk[4 * q0];
        auto pv = pref(row);
        V tot[NX];
#pragma unroll
        for (int k = 0; k < NX; ++k) tot[k] = V(0);
        for (int q = threadIdx.x; q < nch; q += kBlock) {
#pragma unroll
            for (int k = 0; k < NX; ++k) tot[k] += (V)sp.part[(size_t)(q0 + q) * NX + k];
        }
        block_sum_v<V, NX>(tot, lds);
        if (threadIdx.x == 0) row_fn(row, tot, pv);
        return;
    }
    if (b >= sp.nblocks) {  // one chunk of a split row: its partial only
        const int c = b - sp.nblocks;
        const int p0 = sp.chunk[4 * c + 1], p1 = sp.chunk[4 * c + 2];
        const int ccb = (a.idx16 && sp.cbase) ? sp.cbase[c] : -1;
        V s[NX];
#pragma unroll
        for (int k = 0; k < NX; ++k) s[k] = V(0);
        // unrolled so that several index / value / gather loads are in flight; each thread still adds its
        // products in increasing position order (the consecutive / compact / plain index choice is outside the
        // loops)
        if (ccb <= -2) {
            const int c0 = -2 - ccb - p0;  // column of entry p: c0 + p
#pragma unroll 8
            for (int p = p0 + threadIdx.x; p < p1; p += kBlock) {
                const V av = widen<V>(a.val[p]);
#pragma unroll
                for (int k = 0; k < NX; ++k) s[k] += av * gather(k, c0 + p);
            }
        } else if (ccb >= 0) {
#pragma unroll 8
            for (int p = p0 + threadIdx.x; p < p1; p += kBlock) {
                const V av = widen<V>(a.val[p]);
                const int i = ccb + (int)a.idx16[p];
#pragma unroll
                for (int k = 0; k < NX; ++k) s[k] += av * gather(k, i);
            }
        } else {
#pragma unroll 8
            for (int p = p0 + threadIdx.x; p < p1; p += kBlock) {
                const V av = widen<V>(a.val[p]);
                const int i = a.idx[p];
#pragma unroll
                for (int k = 0; k < NX; ++k) s[k] += av * gather(k, i);
            }
        }
        block_sum_v<V, NX>(s, lds);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < NX; ++k) sp.part[(size_t)c * NX + k] = (double)s[k];
        }
        return;
    }
    stream_head_pf<M, V, NX, WIN>(a, block_head(a, blocks, b), gather, pref, row_fn, lds);
}

// The row-block part of stream_block_pf, on a head already loaded (block_head).
template <class M, class V, int NX, bool WIN, class Gather, class Pref, class RowFn>
__device__ __forceinline__ void stream_head_pf(const CsrView<M>& a, const BlockHead& h, Gather gather, Pref pref,
                                               RowFn row_fn, V* lds /* NX * kTile (+ kWin with WIN) */) {
    static_assert(!WIN || NX == 1, "windowed gathers stage one vector");
    const int4 d = h.d;
    const uint16_t* __restrict__ i16 = a.idx16;
    const int cb = h.cb;
    const int wl = WIN ? max(h.wraw, 0) : 0;
    const int r0 = d.x, r1 = d.y;
    const int base = d.z;
    const int cnt = d.w - d.z;
    if (r1 - r0 == 1 && cnt > kSplit) return;  // a split row: its chunks do the work
    if (r1 - r0 == 1 && cnt > kLong) {
        // a long row: the whole block reduces it; the row's operands are loaded before the sweep (by every
        // thread, unconditionally) so that their latency overlaps it instead of following the reduction
        auto pv = pref(r0);
        V s[NX];
#pragma unroll
        for (int k = 0; k < NX; ++k) s[k] = V(0);
        if constexpr (WIN && NX == 1) {
            if (wl > 0 && cnt <= kTile) {
                // a windowed long row of at most kTile entries: its window (coalesced loads), offsets and
                // values are all loaded in one phase, the window goes to LDS (past the reduction slots) and the
                // products read it there -- the same products, added in the same per-thread order, as the
                // gather loop below
                constexpr int kPer = kTile / kBlock, kWPer = kWin / kBlock;
                V* win = lds + kTile;
                const int wlast = wl - 1, last = cnt - 1;
                V wv[kWPer];
                M mv[kPer];
                int off[kPer];
#pragma unroll
                for (int k = 0; k < kWPer; ++k) wv[k] = gather(0, cb + min((int)threadIdx.x + k * kBlock, wlast));
#pragma unroll
                for (int k = 0; k < kPer; ++k) {
                    const int q = min((int)threadIdx.x + k * kBlock, last);
                    mv[k] = a.val[base + q];
                    off[k] = (int)i16[base + q];
                }
                auto products = [&](auto swz) {
                    constexpr bool SW = decltype(swz)::value;
#pragma unroll
                    for (int k = 0; k < kWPer; ++k) win[win_slot<SW>((int)threadIdx.x + k * kBlock)] = wv[k];
                    __syncthreads();
#pragma unroll
                    for (int k = 0; k < kPer; ++k)
                        if ((int)threadIdx.x + k * kBlock < cnt) s[0] += widen<V>(mv[k]) * win[win_slot<SW>(off[k])];
                };
                if (a.wswz) products(std::true_type{});
                else products(std::false_type{});
                block_sum_v<V, NX>(s, lds);
                if (threadIdx.x == 0) row_fn(r0, s, pv);
                return;
            }
        }
        if (i16 && h.wraw >= 0) {
#pragma unroll 8
            for (int p = base + threadIdx.x; p < base + cnt; p += kBlock) {
                const V av = widen<V>(a.val[p]);
                const int i = cb + (int)i16[p];
#pragma unroll
                for (int k = 0; k < NX; ++k) s[k] += av * gather(k, i);
            }
        } else {
#pragma unroll 8
            for (int p = base + threadIdx.x; p < base + cnt; p += kBlock) {
                const V av = widen<V>(a.val[p]);
                const int i = a.idx[p];
#pragma unroll
                for (int k = 0; k < NX; ++k) s[k] += av * gather(k, i);
            }
        }
        block_sum_v<V, NX>(s, lds);
        if (threadIdx.x == 0) row_fn(r0, s, pv);
        return;
    }
    // the owning thread's row extent and row operands are loaded before the staging sweep, so their
    // latency overlaps it instead of following the barrier (the extents are only used after it).  Threads
    // without a row load row r0's (always a row of the block): unconditional loads leave no branch whose
    // merge would make the sweep wait for them.
    const int r = r0 + (int)threadIdx.x;
    const bool own = r < r1;
    const int rl = own ? r : r0;
    const int pbeg = a.ptr[rl], pend = a.ptr[rl + 1];
    auto pv = pref(rl);
    // Staging in two straight-line phases: every index/value load of the thread's (at most kPer) slots is
    // issued first, then every gather, so a block costs one index round trip and one gather round trip
    // instead of one of each per slot.  Slots past the block's end load its last entry (always in bounds)
    // and store into the tile's unused tail [cnt, kTile), which no row reads; the store is unconditional so
    // that the gathers are not sunk into a branch one slot at a time.  The products are the same as one slot
    // at a time, so the row sums are too.
    constexpr int kPer = kTile / kBlock;
    if (WIN && wl > 0) {
        // windowed block: the window's values, the offsets and the matrix values are all loaded in one
        // phase (the window loads are coalesced), the window goes to LDS, and the products read it there.
        // Offsets index the window directly; window slots past wl (<= kWin) are filled but never read.
        constexpr int kWPer = kWin / kBlock;
        V* win = lds + kTile;
        const int wlast = wl - 1, last = cnt - 1;
        V wv[kWPer];
        M mv[kPer];
        int off[kPer];
#pragma unroll
        for (int k = 0; k < kWPer; ++k) wv[k] = gather(0, cb + min((int)threadIdx.x + k * kBlock, wlast));
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int q = min((int)threadIdx.x + k * kBlock, last);
            mv[k] = a.val[base + q];
            off[k] = (int)i16[base + q];
        }
        auto products = [&](auto swz) {
            constexpr bool SW = decltype(swz)::value;
#pragma unroll
            for (int k = 0; k < kWPer; ++k) win[win_slot<SW>((int)threadIdx.x + k * kBlock)] = wv[k];
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kPer; ++k)
                lds[(int)threadIdx.x + k * kBlock] = widen<V>(mv[k]) * win[win_slot<SW>(off[k])];
        };
        if (a.wswz) products(std::true_type{});
        else products(std::false_type{});
    } else if (cnt > 0) {
        const int last = cnt - 1;
        M mv[kPer];
        int col[kPer];
        if (i16) {
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const int q = min((int)threadIdx.x + k * kBlock, last);
                mv[k] = a.val[base + q];
                col[k] = cb + (int)i16[base + q];
            }
        } else {
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const int q = min((int)threadIdx.x + k * kBlock, last);
                mv[k] = a.val[base + q];
                col[k] = a.idx[base + q];
            }
        }
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int q = (int)threadIdx.x + k * kBlock;  // < kTile
            const V av = widen<V>(mv[k]);
#pragma unroll
            for (int j = 0; j < NX; ++j) lds[j * kTile + q] = av * gather(j, col[k]);
        }
    }
    __syncthreads();
    if (own) {
        const int beg = pbeg - base, end = pend - base;
        V s[NX];
#pragma unroll
        for (int k = 0; k < NX; ++k) s[k] = V(0);
        ordered_row_sums<V, NX>(lds, beg, end, s);
        row_fn(r, s, pv);
    }
}

// stream_block for NX gathered vectors with ONE tile of LDS: on a staged block the indices and values are
// loaded once (into registers) and the NX products are staged and summed one vector after another, so a check
// pass needs kTile * sizeof(V) bytes of LDS (+ one window) instead of NX times that.  Same row sums, same bits.
// Long rows, chunks of split rows and the finish pass need only (kBlock / kWave) * NX elements of LDS and take
// the NX-wide path.

struct NoRowOperands {};

// stream_block without row operands: row_fn(r, sums).
template <class M, class V, int NX, bool WIN = false, class Gather, class RowFn>
__device__ __forceinline__ void stream_block(const CsrView<M>& a, const int* __restrict__ blocks, int b, Gather gather,
                                             RowFn row_fn, V* lds /* NX * kTile (+ kWin with WIN) */,
                                             const SplitPlan& sp) {
    stream_block_pf<M, V, NX, WIN>(
        a, blocks, b, gather, [](int) { return NoRowOperands{}; },
        [&](int r, V (&t)[NX], NoRowOperands&) { row_fn(r, t); }, lds, sp);
}

template <class M, class V, int NX, bool WIN, class Gather, class Pref, class RowFn>
__device__ __forceinline__ void stream_head_seq_pf(const CsrView<M>& a, const BlockHead& h, Gather gather, Pref pref,
                                                   RowFn row_fn, V* lds, unsigned act);

// act: bit v set = vector v is wanted; an unwanted vector's sums are 0 and, on a staged block, it costs no
// pass (the long-row and chunk paths still gather it, and the gather returns 0 for it).  pref(r) loads row r's
// operands before the sweep (as in stream_block_pf), and row_fn(r, sums, operands) runs after it.
template <class M, class V, int NX, bool WIN = false, class Gather, class Pref, class RowFn>
__device__ __forceinline__ void stream_block_seq_pf(const CsrView<M>& a, const int* __restrict__ blocks, int b,
                                                    Gather gather, Pref pref, RowFn row_fn,
                                                    V* lds /* kTile (+ kWin with WIN) */, const SplitPlan& sp,
                                                    unsigned act = ~0u) {
    if (sp.finish || b >= sp.nblocks) {
        stream_block_pf<M, V, NX>(a, blocks, b, gather, pref, row_fn, lds, sp);
        return;
    }
    stream_head_seq_pf<M, V, NX, WIN>(a, block_head(a, blocks, b), gather, pref, row_fn, lds, act);
}

// The row-block part of stream_block_seq_pf, on a head already loaded (block_head).
template <class M, class V, int NX, bool WIN, class Gather, class Pref, class RowFn>
__device__ __forceinline__ void stream_head_seq_pf(const CsrView<M>& a, const BlockHead& h, Gather gather, Pref pref,
                                                   RowFn row_fn, V* lds /* kTile (+ kWin with WIN) */,
                                                   unsigned act) {
    static_assert((kBlock / kWave) * NX <= kTile, "reduction slots exceed the tile");
    const int4 d = h.d;
    const int r0 = d.x, r1 = d.y, base = d.z, cnt = d.w - d.z;
    if (r1 - r0 == 1 && cnt > kLong) {  // long (or split) row: the NX-wide block-reduction path
        stream_head_pf<M, V, NX, false>(a, h, gather, pref, row_fn, lds);
        return;
    }
    // A staged block: its indices and values are loaded once, every gather of every wanted vector is issued
    // in one phase (the windows of a windowed block; the products of a plain one), then for each vector the
    // products go through the one tile and each row is summed in order -- the same row sums as NX separate
    // passes, with one gather round trip instead of NX.  The owning thread's row extent and row operands are
    // loaded first (unconditionally: threads without a row load row r0's) so their latency overlaps the sweep.
    const uint16_t* __restrict__ i16 = a.idx16;
    int cb = 0, wl = 0;
    if (i16) {
        cb = h.cb;
        wl = WIN ? h.wraw : 0;
    }
    const int r = r0 + (int)threadIdx.x;
    const bool own = r < r1;
    const int rl = own ? r : r0;
    const int pbeg = a.ptr[rl], pend = a.ptr[rl + 1];
    auto pv = pref(rl);
    constexpr int kPer = kTile / kBlock, kWPer = kWin / kBlock;
    const int last = cnt > 0 ? cnt - 1 : 0;
    M mv[kPer];
    int col[kPer];  // the window offset (windowed block) or the column
    if (cnt > 0 && i16) {
        const int add = (WIN && wl > 0) ? 0 : cb;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int q = min((int)threadIdx.x + k * kBlock, last);
            mv[k] = a.val[base + q];
            col[k] = (int)i16[base + q] + add;
        }
    } else if (cnt > 0) {
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int q = min((int)threadIdx.x + k * kBlock, last);
            mv[k] = a.val[base + q];
            col[k] = a.idx[base + q];
        }
    }
    V acc[NX];
    if (WIN && wl > 0) {
        V* win = lds + kTile;
        const int wlast = wl - 1;
        V wv[NX][kWPer];
#pragma unroll
        for (int v = 0; v < NX; ++v) {
            if (!((act >> v) & 1u)) continue;
#pragma unroll
            for (int k = 0; k < kWPer; ++k) wv[v][k] = gather(v, cb + min((int)threadIdx.x + k * kBlock, wlast));
        }
#pragma unroll
        for (int v = 0; v < NX; ++v) {
            if (!((act >> v) & 1u)) {
                acc[v] = V(0);
                continue;
            }
            auto products = [&](auto swz) {
                constexpr bool SW = decltype(swz)::value;
#pragma unroll
                for (int k = 0; k < kWPer; ++k) win[win_slot<SW>((int)threadIdx.x + k * kBlock)] = wv[v][k];
                __syncthreads();
#pragma unroll
                for (int k = 0; k < kPer; ++k)
                    lds[(int)threadIdx.x + k * kBlock] = widen<V>(mv[k]) * win[win_slot<SW>(col[k])];
            };
            if (a.wswz) products(std::true_type{});
            else products(std::false_type{});
            __syncthreads();
            V t[1] = {V(0)};
            if (own) ordered_row_sums<V, 1>(lds, pbeg - base, pend - base, t);
            acc[v] = t[0];
            __syncthreads();  // the next vector reuses the window and the tile
        }
    } else {
        V pr[NX][kPer];  // the products of every wanted vector, gathered in one phase
        if (cnt > 0) {
#pragma unroll
            for (int v = 0; v < NX; ++v) {
                if (!((act >> v) & 1u)) continue;
#pragma unroll
                for (int k = 0; k < kPer; ++k) pr[v][k] = widen<V>(mv[k]) * gather(v, col[k]);
            }
        }
#pragma unroll
        for (int v = 0; v < NX; ++v) {
            if (!((act >> v) & 1u)) {
                acc[v] = V(0);
                continue;
            }
            if (cnt > 0) {
#pragma unroll
                for (int k = 0; k < kPer; ++k) lds[(int)threadIdx.x + k * kBlock] = pr[v][k];
            }
            __syncthreads();
            V t[1] = {V(0)};
            if (own) ordered_row_sums<V, 1>(lds, pbeg - base, pend - base, t);
            acc[v] = t[0];
            __syncthreads();  // the next vector reuses the tile
        }
    }
    if (own) row_fn(r, acc, pv);
}

// stream_block_seq_pf without row operands: row_fn(r, sums).
template <class M, class V, int NX, bool WIN = false, class Gather, class RowFn>
__device__ __forceinline__ void stream_block_seq(const CsrView<M>& a, const int* __restrict__ blocks, int b,
                                                 Gather gather, RowFn row_fn, V* lds /* kTile (+ kWin with WIN) */,
                                                 const SplitPlan& sp, unsigned act = ~0u) {
    stream_block_seq_pf<M, V, NX, WIN>(
        a, blocks, b, gather, [](int) { return NoRowOperands{}; },
        [&](int r, V (&t)[NX], NoRowOperands&) { row_fn(r, t); }, lds, sp, act);
}

// ---------------------------------------------------------------- deterministic grid reduction
// Two launches, no atomics: every block of the producing kernel writes its N block sums to partials[b][.]
// (block_partial); one block of a follow-up kernel then adds them up in block order (sum_partials).  A
// single shared ticket counter would serialise ~2000 blocks on one address; this never does.
struct ReduceSlot {
    double* partials;  // gridDim.x * N doubles
    double* out;       // N doubles
};

// Block-wide sum of N values per thread; the result is valid in thread 0.  `lds` must hold
// (blockDim.x / 64) * N doubles.
template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N], double* lds) {
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave, nw = blockDim.x / kWave;
#pragma unroll
    for (int k = 0; k < N; ++k) {
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, kWave);
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) lds[wave * N + k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += lds[w * N + k];
            v[k] = s;
        }
    }
    __syncthreads();
}

template <int N>
__device__ __forceinline__ void block_partial(double (&v)[N], const ReduceSlot& slot, double* lds) {
    block_sum<N>(v, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) slot.partials[(size_t)blockIdx.x * N + k] = v[k];
    }
}

// block_partial for many sums through a large LDS scratch (`cap` doubles, free for the duration): every thread
// stores its N values, 16-entry slices are added in parallel, and thread k adds value k's 16 slice sums and
// writes partial k.  Fixed order, no shuffles: for N ~ 10 it replaces N six-step shuffle trees (two LDS
// permutes and one dependent add per step and value) with one store, 16 + 16 adds and three barriers per
// group of values.
template <int N>
__device__ __forceinline__ void block_partial_wide(const double (&v)[N], const ReduceSlot& slot, double* lds,
                                                   int cap, int at = -1 /* partial index; -1: blockIdx.x */) {
    const size_t row = at < 0 ? (size_t)blockIdx.x : (size_t)at;
    constexpr int kSlices = 16, kPerSlice = kBlock / kSlices;
    const int per = kBlock + kSlices;  // doubles per value: the block's values, then the slice sums
    const int group = min(N, cap / per);
    for (int k0 = 0; k0 < N; k0 += group) {
        const int g = min(group, N - k0);
#pragma unroll
        for (int k = 0; k < N; ++k)
            if (k >= k0 && k < k0 + g) lds[(k - k0) * per + threadIdx.x] = v[k];
        __syncthreads();
        if ((int)threadIdx.x < g * kSlices) {
            const int k = threadIdx.x / kSlices, j = threadIdx.x % kSlices;
            const double* src = lds + k * per + j * kPerSlice;
            double s = 0.0;
#pragma unroll
            for (int i = 0; i < kPerSlice; ++i) s += src[i];
            lds[k * per + kBlock + j] = s;
        }
        __syncthreads();
        if ((int)threadIdx.x < g) {
            const double* src = lds + threadIdx.x * per + kBlock;
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < kSlices; ++j) s += src[j];
            slot.partials[row * N + k0 + threadIdx.x] = s;
        }
        __syncthreads();  // the next group (or the caller) reuses the scratch
    }
}

// Run by every thread of one block: out[k] = sum over b < nparts of partials[b][k] for the single output k,
// fixed order (one block per output keeps the fold parallel).
__device__ __forceinline__ void sum_partials_one(const double* __restrict__ partials, int N, int k, int nparts,
                                                 double* out, double* lds) {
    // each thread adds its partials b, b + blockDim.x, ... in that order; eight loads are issued before
    // their adds so that a large grid's fold costs few memory round trips
    double s[1] = {0.0};
    const int step = blockDim.x;
    int b = threadIdx.x;
    for (; b + 7 * step < nparts; b += 8 * step) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = partials[(size_t)(b + u * step) * N + k];
#pragma unroll
        for (int u = 0; u < 8; ++u) s[0] += v[u];
    }
    for (; b < nparts; b += step) s[0] += partials[(size_t)b * N + k];
    block_sum<1>(s, lds);
    if (threadIdx.x == 0) *out = s[0];
}

// Run by every thread of one block: out[k] = sum over b < nparts of partials[b][k], fixed order.
template <int N>
__device__ __forceinline__ void sum_partials(const double* __restrict__ partials, int nparts, double* out,
                                             double* lds) {
    double s[N];
#pragma unroll
    for (int k = 0; k < N; ++k) s[k] = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
        const double* p = partials + (size_t)b * N;
#pragma unroll
        for (int k = 0; k < N; ++k) s[k] += p[k];
    }
    block_sum<N>(s, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) out[k] = s[k];
    }
}

}  // namespace ganit_hip
