// libganit_hip.so, part 1: context, device vectors, CSR upload with the row-length plan, and the
// backend-concept kernels (SpMV, primal / dual step, the FP64 reductions) behind the C ABI.
#include <algorithm>
#include <exception>
#include <mutex>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "internal.hpp"
#include "pdhg_math.hpp"

namespace ganit_hip {

namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }

// ================================================================ elementwise kernels
template <class V>
__global__ __launch_bounds__(kBlock) void k_fill(int64_t n, double a, V* __restrict__ d) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) d[i] = V(a);
}
template <class V>
__global__ __launch_bounds__(kBlock) void k_copy(int64_t n, const V* __restrict__ s, V* __restrict__ d) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) d[i] = s[i];
}
template <class V>
__global__ __launch_bounds__(kBlock) void k_sub(int64_t n, const V* __restrict__ a, const V* __restrict__ b,
                                                V* __restrict__ d) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        d[i] = a[i] - b[i];
}
// d = V(a * double(s)) -- the running-average formation (src/backend_cuda.cu:135-139)
template <class V>
__global__ __launch_bounds__(kBlock) void k_scale(int64_t n, double a, const V* __restrict__ s, V* __restrict__ d) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        d[i] = V(a * double(s[i]));
}
template <class S, class D>
__global__ __launch_bounds__(kBlock) void k_convert(int64_t n, const S* __restrict__ s, D* __restrict__ d) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        d[i] = D(s[i]);
}

unsigned ew_grid(int64_t n) { return std::min<unsigned>(grid_for(n), 8192u); }

// ================================================================ SpMV and the two steps
// FP32 and / or BF16 copies of FP64 matrix values, each rounded as bf16_bits_from_double / a C cast does on the
// host: to nearest, ties to even (BF16 through FP32; NaN kept quiet).
__global__ __launch_bounds__(kBlock) void k_round_values(int64_t n, const double* __restrict__ v64,
                                                         float* __restrict__ v32, uint16_t* __restrict__ v16) {
    for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < n; p += (int64_t)gridDim.x * kBlock) {
        const float f = (float)v64[p];
        if (v32) v32[p] = f;
        if (v16) {
            const uint32_t u = __float_as_uint(f);
            v16[p] = (u & 0x7fffffffu) > 0x7f800000u ? (uint16_t)((u >> 16) | 0x40)
                                                     : (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
        }
    }
}

template <int LPR, class M, class V>
__global__ __launch_bounds__(kBlock) void k_spmv(CsrView<M> a, const V* __restrict__ x, V* __restrict__ y) {
    const int64_t gt = (int64_t)xcd_block_id() * kBlock + threadIdx.x;
    const int row = (int)(gt / LPR), lane = (int)(gt % LPR);
    if (row >= a.rows) return;  // a whole row group leaves together (LPR divides the block)
    const V t = row_dot<LPR>(a, row, lane, x);
    if (lane == 0) y[row] = t;
}

// Stand-alone SpMV on LDS-staged row blocks (one block per row block).
template <class M, class V>
__global__ __launch_bounds__(kBlock) void k_spmv_rb(CsrView<M> a, const int* __restrict__ blocks, SplitPlan sp,
                                                    const V* __restrict__ x, V* __restrict__ y) {
    __shared__ __align__(16) V lds[kTile + kWin];
    stream_block<M, V, 1, true>(
        a, blocks, xcd_block_id(), [&](int, int j) { return x[j]; }, [&](int r, V (&t)[1]) { y[r] = t[0]; }, lds,
        sp);
}

template <class V, bool HASQ>
__global__ __launch_bounds__(kBlock) void k_primal_step(int n, const V* __restrict__ x, const V* __restrict__ aty,
                                                        const V* __restrict__ qx, const V* __restrict__ c,
                                                        const V* __restrict__ l, const V* __restrict__ u, V tau,
                                                        V* __restrict__ xn, V* __restrict__ xbar,
                                                        V* __restrict__ sumx) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    V s = sumx[j];
    primal_update<V>(x[j], aty[j], HASQ ? qx[j] : V(0), c[j], l[j], u[j], tau, xn[j], xbar[j], s);
    sumx[j] = s;
}

template <class V>
__global__ __launch_bounds__(kBlock) void k_dual_step(int m, const V* __restrict__ y, const V* __restrict__ axbar,
                                                      const V* __restrict__ lo, const V* __restrict__ hi, V sigma,
                                                      V* __restrict__ yn, V* __restrict__ sumy) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= m) return;
    V s = sumy[i];
    dual_update<V>(y[i], axbar[i], lo[i], hi[i], sigma, yn[i], s);
    sumy[i] = s;
}

// ================================================================ FP64 reductions (two outputs)
// Per-index contributions follow the reference's functors (src/backend_cuda.cu:199-249).
template <class F>
__global__ __launch_bounds__(kBlock) void k_reduce2(int64_t n, F f, ReduceSlot slot) {
    __shared__ double lds[(kBlock / kWave) * 2];
    double acc[2] = {0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        double a0, a1;
        f(i, a0, a1);
        acc[0] += a0;
        acc[1] += a1;
    }
    block_partial<2>(acc, slot, lds);
}

__global__ __launch_bounds__(kBlock) void k_finish2(const double* __restrict__ partials, int nparts,
                                                    double* __restrict__ out) {
    __shared__ double lds[(kBlock / kWave) * 2];
    sum_partials<2>(partials, nparts, out, lds);
}

template <class V>
struct PrimalResTerm {
    const V *ax, *lo, *hi, *rr;
    __device__ void operator()(int64_t i, double& a0, double& a1) const {
        const double a = double(ax[i]);
        const double p = fmin(fmax(a, double(lo[i])), double(hi[i]));
        const double r = (a - p) / double(rr[i]);
        a0 = r * r;
        a1 = 0.0;
    }
};
template <class V>
struct DualResTerm {
    const V *aty, *qx, *c, *l, *u, *cr;
    double cw;
    __device__ void operator()(int64_t j, double& a0, double& a1) const {
        const double z = cw * double(c[j]) - double(aty[j]) + (qx ? double(qx[j]) : 0.0);
        const double lj = double(l[j]), uj = double(u[j]);
        double r = 0.0, ob = 0.0;
        if (z > 0) { if (isfinite(lj)) ob = z * lj; else r = z; }
        else if (z < 0) { if (isfinite(uj)) ob = z * uj; else r = z; }
        r /= double(cr[j]);
        a0 = r * r;
        a1 = ob;
    }
};
template <class V>
struct RowDualObjTerm {
    const V *y, *lo, *hi;
    __device__ void operator()(int64_t i, double& a0, double& a1) const {
        const double yi = double(y[i]), lw = double(lo[i]), hg = double(hi[i]);
        a0 = 0.0;
        if (yi > 0 && isfinite(lw)) a0 = yi * lw;
        else if (yi < 0 && isfinite(hg)) a0 = yi * hg;
        a1 = 0.0;
    }
};
template <class V>
struct DotTerm {
    const V *a, *b;
    __device__ void operator()(int64_t i, double& a0, double& a1) const {
        a0 = double(a[i]) * double(b[i]);
        a1 = 0.0;
    }
};
template <class V>
struct Diff2Term {
    const V *a, *b;
    __device__ void operator()(int64_t i, double& a0, double& a1) const {
        const double d = double(a[i]) - double(b[i]);
        a0 = d * d;
        a1 = 0.0;
    }
};

// Grid of the fine-grained reductions: one workgroup per 256 entries, at most kMaxReduceGrid.
unsigned reduce_grid(int64_t n) {
    return std::min<unsigned>(grid_for(std::max<int64_t>(n, 1)), (unsigned)kMaxReduceGrid);
}

// ================================================================ allocation-free dot product
// The scratch of ganit_hip_dot_noalloc: the reduction grid's block partials (two doubles per workgroup, as
// k_reduce2 writes them), 16 KiB.  Every GPU the library runs on holds its own copy at its own address until
// the process ends; host code reaches it only through the address hipGetSymbolAddress returns for the
// current device (noalloc_scratch), never through this name.
constexpr int kNoAllocDoubles = kMaxReduceGrid * 2;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

// One output of k_reduce2's partials (the first of each pair), folded in the order k_finish2 folds it, so
// the value is bitwise the one ganit_hip_dot returns.
__global__ __launch_bounds__(kBlock) void k_finish_first(const double* __restrict__ partials, int nparts,
                                                         double* __restrict__ out) {
    __shared__ double lds[kBlock / kWave];
    sum_partials_one(partials, 2, 0, nparts, out, lds);
}

constexpr int kMaxNoAllocDevices = 64;

// Device address of `scratch` on `device` (which must be current), looked up once per device.
double* noalloc_scratch(int device) {
    static std::mutex mu;
    static double* addr[kMaxNoAllocDevices] = {};
    GH_REQUIRE(device >= 0 && device < kMaxNoAllocDevices, "device ordinal beyond the allocation-free entry's table");
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        void* p = nullptr;
        GH_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        GH_REQUIRE(p, "hipGetSymbolAddress returned no address for the reduction scratch");
        addr[device] = static_cast<double*>(p);
    }
    return addr[device];
}

template <class F>
void reduce2(Ctx* ctx, int64_t n, F f, double* o0, double* o1) {
    const unsigned g = reduce_grid(n);
    ReduceSlot slot{ctx->red_partials, ctx->red_out};
    hipLaunchKernelGGL(k_reduce2<F>, dim3(g), dim3(kBlock), 0, ctx->stream, n, f, slot);
    hipLaunchKernelGGL(k_finish2, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->red_partials, (int)g, ctx->red_out);
    GH_CHECK(hipGetLastError());
    GH_CHECK(hipMemcpyAsync(ctx->host_out, ctx->red_out, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    GH_CHECK(hipStreamSynchronize(ctx->stream));
    if (o0) *o0 = ctx->host_out[0];
    if (o1) *o1 = ctx->host_out[1];
}

// ================================================================ CSR upload + plan
uint16_t bf16_bits_from_double(double d) {
    const float f = (float)d;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // quiet NaN
    const uint32_t rounding = 0x7fffu + ((u >> 16) & 1u);                      // nearest, ties to even
    return (uint16_t)((u + rounding) >> 16);
}

// Lanes per row from the row-length distribution: the nnz-weighted median row length (half of all
// nonzeros sit in rows no longer than it), rounded up to a power of two and clamped to one wavefront.
int choose_lanes_from_lengths(const std::vector<int>& len, int64_t nnz) {
    if (nnz == 0) return 1;
    // nonzeros per row length, lengths above one wavefront in one bin (they all give kWave lanes): the
    // weighted median of the sorted lengths without sorting them
    std::vector<int64_t> w(kWave + 2, 0);
    for (int L : len) w[std::min(L, kWave + 1)] += L;
    int64_t acc = 0;
    int med = 1;
    for (int L = 0; L <= kWave + 1; ++L) {
        acc += w[L];
        if (w[L] > 0 && 2 * acc >= nnz) { med = L; break; }
    }
    int lpr = 1;
    while (lpr < med && lpr < kWave) lpr <<= 1;
    return lpr;
}

bool valid_variant(int lanes) { return lanes == 0 || (lanes >= 1 && lanes <= kWave && (lanes & (lanes - 1)) == 0); }

// Row blocks of the LDS-staged variant: consecutive rows while the block holds at most max_rows (<= kBlock)
// rows and max_nnz (<= kTile) nonzeros and every row has at most kLong entries; a longer row forms a block
// of its own.
std::vector<int> row_blocks(const int* ptr, int rows, int max_rows, int max_nnz) {
    std::vector<int> b{0};
    int r = 0;
    while (r < rows) {
        const int len = ptr[r + 1] - ptr[r];
        if (len > kLong) {
            ++r;
        } else {
            int cnt = 0, nz = 0;
            while (r < rows && cnt < max_rows) {
                const int l = ptr[r + 1] - ptr[r];
                if (l > kLong || nz + l > max_nnz) break;
                nz += l;
                ++cnt;
                ++r;
            }
        }
        b.push_back(r);
    }
    return b;
}

int* upload_row_blocks(Ctx* ctx, const int* ptr, const std::vector<int>& rb) {
    const size_t nb = rb.empty() ? 0 : rb.size() - 1;
    std::vector<int> d(4 * std::max<size_t>(nb, 1), 0);
    for (size_t q = 0; q < nb; ++q) {
        d[4 * q + 0] = rb[q];
        d[4 * q + 1] = rb[q + 1];
        d[4 * q + 2] = ptr[rb[q]];
        d[4 * q + 3] = ptr[rb[q + 1]];
    }
    int* dev = nullptr;
    GH_CHECK(hipMalloc(&dev, sizeof(int) * d.size()));
    GH_CHECK(hipMemcpyAsync(dev, d.data(), sizeof(int) * d.size(), hipMemcpyHostToDevice, ctx->stream));
    GH_CHECK(hipStreamSynchronize(ctx->stream));
    return dev;
}

// Window layout of a matrix (win_slot in device_common.hpp): 1 = swizzled.  A wave's gather from an LDS window
// takes as many cycles as the most distinct addresses that meet on one of the LDS's 64 four-byte banks; this
// sums that count over the gathers of up to 256 windowed blocks spread over the matrix, for entry w at slot w
// and at slot w ^ (w >> 4), and keeps the swizzled layout when it costs at most 3/4 of the plain one (the
// swizzle's index arithmetic is not free, so a matrix whose gathers rarely conflict stays plain).
int window_layout(const std::vector<int>& rb, const int* ptr, const std::vector<int>& bc,
                  const std::vector<uint16_t>& c16) {
    constexpr int kBanks = 64;
    const size_t nb = rb.size() - 1;
    const size_t step = std::max<size_t>(1, nb / 256);
    std::vector<int> stamp(kWin, -1);
    int64_t cost[2] = {0, 0};
    int gid = 0;
    for (size_t q = 0; q < nb; q += step) {
        if (bc[2 * q + 1] <= 0) continue;
        const int p0 = ptr[rb[q]], p1 = ptr[rb[q + 1]];
        for (int g = p0; g < p1; g += kWave) {  // one wave's lanes read 64 consecutive entries of the block
            const int ge = std::min(g + kWave, p1);
            for (int sw = 0; sw < 2; ++sw, ++gid) {
                int per_bank[kBanks] = {0};
                int worst = 0;
                for (int p = g; p < ge; ++p) {
                    const int w = c16[p];
                    const int slot = sw ? (w ^ (w >> 4)) : w;
                    if (stamp[slot] == gid) continue;  // lanes reading one address share the read
                    stamp[slot] = gid;
                    worst = std::max(worst, ++per_bank[slot % kBanks]);
                }
                cost[sw] += worst;
            }
        }
    }
    return (cost[0] > 0 && 4 * cost[1] <= 3 * cost[0]) ? 1 : 0;
}

void build_compact_indices(Ctx* ctx, Csr* a, const int* ptr, const int* idx, const std::vector<int>& rb) {
    if (a->idx16) GH_CHECK(hipFree(a->idx16));
    if (a->bcol) GH_CHECK(hipFree(a->bcol));
    if (a->chunk_base) GH_CHECK(hipFree(a->chunk_base));
    a->idx16 = nullptr;
    a->bcol = nullptr;
    a->chunk_base = nullptr;
    a->idx32_entries = 0;
    a->wswz = 0;
    const char* env = std::getenv("GANIT_HIP_COMPACT_IDX");  // 0: keep the 32-bit indices only
    if ((env && env[0] == '0') || a->nnz == 0 || rb.size() < 2) return;
    const char* wenv = std::getenv("GANIT_HIP_WINDOW");  // 0: no windowed blocks (every gather from HBM / L2)
    const bool windows = !(wenv && wenv[0] == '0');
    const size_t nb = rb.size() - 1;
    std::vector<int> bc(2 * nb, 0);  // per block: base column, window length (0: not windowed)
    std::vector<uint16_t> c16((size_t)a->nnz, 0);
    auto span_of = [&](int p0, int p1, int& mn) {
        int lo = idx[p0], hi = idx[p0];
        for (int p = p0 + 1; p < p1; ++p) {
            lo = std::min(lo, idx[p]);
            hi = std::max(hi, idx[p]);
        }
        mn = lo;
        return hi - lo;
    };
    auto fill = [&](int p0, int p1, int mn) {
        for (int p = p0; p < p1; ++p) c16[p] = (uint16_t)(idx[p] - mn);
    };
    // blocks are independent (each writes its own descriptor pair and offset range): a few host threads share them
    const int nt = (int)std::max<int64_t>(1, std::min<int64_t>({8, (int64_t)std::max(1u, std::thread::hardware_concurrency()),
                                                              a->nnz / 262144 + 1}));
    std::vector<int64_t> n32t(nt, 0);
    std::vector<char> too_wide(nt, 0);
    auto blocks_of = [&](int t) {
        for (size_t q = (size_t)t * nb / nt, qe = (size_t)(t + 1) * nb / nt; q < qe; ++q) {
            const int r0 = rb[q], r1 = rb[q + 1], p0 = ptr[r0], p1 = ptr[r1];
            if (p1 == p0) continue;
            int mn = 0;
            const int span = span_of(p0, p1, mn);
            if (r1 - r0 == 1 && p1 - p0 > kLong) {
                if (p1 - p0 > kSplit) continue;  // a split row: its chunks carry their own bases (below)
                // a long row: compact when its span fits 16 bits, else 32-bit indices for this row only; windowed
                // on the same terms as a staged block when it has at most kTile entries
                if (span > 65535) {
                    bc[2 * q + 1] = -1;
                    n32t[t] += p1 - p0;
                } else {
                    bc[2 * q] = mn;
                    bc[2 * q + 1] = (windows && p1 - p0 <= kTile && span + 1 <= kWin && span + 1 <= 2 * (p1 - p0))
                                        ? span + 1
                                        : 0;
                    fill(p0, p1, mn);
                }
                continue;
            }
            if (span > 65535) {  // a staged block too wide: keep the 32-bit path for this matrix
                too_wide[t] = 1;
                return;
            }
            bc[2 * q] = mn;
            // windowed when the span fits the LDS window and loading it costs at most two entries per nonzero
            bc[2 * q + 1] = (windows && span + 1 <= kWin && span + 1 <= 2 * (p1 - p0)) ? span + 1 : 0;
            fill(p0, p1, mn);
        }
    };
    {
        std::vector<std::thread> th;
        int started = 1;  // ranges 1 .. started-1 have a thread; a thread that cannot be created leaves the rest here
        try {
            for (; started < nt; ++started) th.emplace_back(blocks_of, started);
        } catch (...) {
        }
        for (int t = started; t < nt; ++t) blocks_of(t);
        blocks_of(0);
        for (auto& x : th) x.join();
    }
    if (std::count(too_wide.begin(), too_wide.end(), (char)1) > 0) return;
    int64_t n32 = 0;
    for (int64_t v : n32t) n32 += v;
    // chunks of the split rows, enumerated as the chunk plan enumerates them (rows ascending, kChunk apart)
    std::vector<int> cbase;
    for (int i = 0; i < a->rows; ++i) {
        if (ptr[i + 1] - ptr[i] <= kSplit) continue;
        for (int p0 = ptr[i]; p0 < ptr[i + 1]; p0 += kChunk) {
            const int p1 = std::min(p0 + kChunk, ptr[i + 1]);
            int mn = 0;
            bool dense = true;  // columns idx[p0], idx[p0] + 1, ...: no index needs reading
            for (int p = p0 + 1; p < p1 && dense; ++p) dense = idx[p] == idx[p0] + (p - p0);
            if (dense) {
                cbase.push_back(-2 - idx[p0]);
                fill(p0, p1, idx[p0]);  // (the compact offsets stay valid; the kernels do not read them)
            } else if (span_of(p0, p1, mn) > 65535) {
                cbase.push_back(-1);
                n32 += p1 - p0;
            } else {
                cbase.push_back(mn);
                fill(p0, p1, mn);
            }
        }
    }
    GH_CHECK(hipMalloc(&a->idx16, sizeof(uint16_t) * c16.size()));
    GH_CHECK(hipMalloc(&a->bcol, sizeof(int) * bc.size()));
    GH_CHECK(hipMemcpyAsync(a->idx16, c16.data(), sizeof(uint16_t) * c16.size(), hipMemcpyHostToDevice, ctx->stream));
    GH_CHECK(hipMemcpyAsync(a->bcol, bc.data(), sizeof(int) * bc.size(), hipMemcpyHostToDevice, ctx->stream));
    if (!cbase.empty()) {
        GH_CHECK(hipMalloc(&a->chunk_base, sizeof(int) * cbase.size()));
        GH_CHECK(hipMemcpyAsync(a->chunk_base, cbase.data(), sizeof(int) * cbase.size(), hipMemcpyHostToDevice,
                                ctx->stream));
    }
    a->idx32_entries = n32;
    const char* senv = std::getenv("GANIT_HIP_WINDOW_SWIZZLE");  // 0 / 1: force the plain / swizzled layout
    a->wswz = (senv && (senv[0] == '0' || senv[0] == '1')) ? senv[0] - '0' : window_layout(rb, ptr, bc, c16);
    GH_CHECK(hipStreamSynchronize(ctx->stream));
}

Csr* csr_create(Ctx* ctx, int rows, int cols, const int* ptr, const int* idx, const double* val, unsigned mask) {
    GH_REQUIRE(rows >= 0 && cols >= 0, "negative matrix dimension");
    GH_REQUIRE(rows == 0 || ptr, "null row pointer");
    GH_REQUIRE(mask & 7u, "dtype mask selects no value type");
    auto* a = new Csr;
    a->ctx = ctx;
    a->rows = rows;
    a->cols = cols;
    a->nnz = rows ? ptr[rows] : 0;
    try {
        GH_REQUIRE(a->nnz >= 0 && (rows == 0 || ptr[0] == 0), "malformed CSR row pointer");
        std::vector<int> len(rows);
        for (int i = 0; i < rows; ++i) {
            len[i] = ptr[i + 1] - ptr[i];
            GH_REQUIRE(len[i] >= 0, "CSR row pointer not monotone");
            int b = 0;
            while (b < 8 && len[i] > (1 << b)) ++b;
            a->hist[b]++;
        }
        for (int64_t p = 0; p < a->nnz; ++p) GH_REQUIRE(idx[p] >= 0 && idx[p] < cols, "CSR column index out of range");
        a->lanes = choose_lanes_from_lengths(len, a->nnz);
        const std::vector<int> rb = row_blocks(ptr, rows);
        a->nblocks = (int)rb.size() - 1;
        a->blocks = upload_row_blocks(ctx, ptr, rb);
        if (a->nnz) build_compact_indices(ctx, a, ptr, idx, rb);
        {  // chunks of the rows longer than kSplit
            std::vector<int> ch, first{0};
            for (int i = 0; i < rows; ++i) {
                if (len[i] <= kSplit) continue;
                const int s_idx = (int)first.size() - 1;
                for (int p0 = ptr[i]; p0 < ptr[i + 1]; p0 += kChunk) {
                    ch.insert(ch.end(), {i, p0, std::min(p0 + kChunk, ptr[i + 1]), s_idx});
                }
                first.push_back((int)ch.size() / 4);
            }
            a->nsplit = (int)first.size() - 1;
            a->nchunks = (int)ch.size() / 4;
            if (a->nsplit > 0) {
                GH_CHECK(hipMalloc(&a->chunk, sizeof(int) * ch.size()));
                GH_CHECK(hipMalloc(&a->split_first, sizeof(int) * first.size()));
                GH_CHECK(hipMalloc(&a->split_part, sizeof(double) * 3 * a->nchunks));
                GH_CHECK(hipMemcpyAsync(a->chunk, ch.data(), sizeof(int) * ch.size(), hipMemcpyHostToDevice, ctx->stream));
                GH_CHECK(hipMemcpyAsync(a->split_first, first.data(), sizeof(int) * first.size(), hipMemcpyHostToDevice,
                                        ctx->stream));
            }
        }
        // short rows (nnz-weighted median under 16 entries) start on the row-block variant
        if (a->lanes < 16 && a->nblocks > 0) a->lanes = 0;
        const size_t nz = (size_t)std::max<int64_t>(a->nnz, 1);
        GH_CHECK(hipMalloc(&a->ptr, sizeof(int) * (rows + 1)));
        GH_CHECK(hipMalloc(&a->idx, sizeof(int) * nz));
        if (rows) GH_CHECK(hipMemcpyAsync(a->ptr, ptr, sizeof(int) * (rows + 1), hipMemcpyHostToDevice, ctx->stream));
        else GH_CHECK(hipMemsetAsync(a->ptr, 0, sizeof(int), ctx->stream));
        if (a->nnz) GH_CHECK(hipMemcpyAsync(a->idx, idx, sizeof(int) * a->nnz, hipMemcpyHostToDevice, ctx->stream));
        if (mask & (1u << GANIT_F64)) {
            GH_CHECK(hipMalloc(&a->v64, sizeof(double) * nz));
            if (a->nnz) GH_CHECK(hipMemcpyAsync(a->v64, val, sizeof(double) * a->nnz, hipMemcpyHostToDevice, ctx->stream));
        }
        // the reduced-precision copies are rounded on the device from the FP64 values when those are resident
        // (the same rounding as on the host: nearest, ties to even), else on the host
        const bool on_device = a->v64 && a->nnz > 0;
        if (mask & (1u << GANIT_F32)) {
            GH_CHECK(hipMalloc(&a->v32, sizeof(float) * nz));
            if (on_device) {
                hipLaunchKernelGGL(k_round_values, dim3(grid_for(a->nnz)), dim3(kBlock), 0, ctx->stream, a->nnz,
                                   (const double*)a->v64, a->v32, (uint16_t*)nullptr);
            } else if (a->nnz) {
                std::vector<float> h(val, val + a->nnz);
                GH_CHECK(hipMemcpyAsync(a->v32, h.data(), sizeof(float) * a->nnz, hipMemcpyHostToDevice, ctx->stream));
                GH_CHECK(hipStreamSynchronize(ctx->stream));
            }
        }
        if (mask & (1u << GANIT_BF16)) {
            GH_CHECK(hipMalloc(&a->v16, sizeof(uint16_t) * nz));
            if (on_device) {
                hipLaunchKernelGGL(k_round_values, dim3(grid_for(a->nnz)), dim3(kBlock), 0, ctx->stream, a->nnz,
                                   (const double*)a->v64, (float*)nullptr, reinterpret_cast<uint16_t*>(a->v16));
            } else if (a->nnz) {
                std::vector<uint16_t> h(a->nnz);
                for (int64_t p = 0; p < a->nnz; ++p) h[p] = bf16_bits_from_double(val[p]);
                GH_CHECK(hipMemcpyAsync(a->v16, h.data(), sizeof(uint16_t) * a->nnz, hipMemcpyHostToDevice, ctx->stream));
                GH_CHECK(hipStreamSynchronize(ctx->stream));
            }
        }
        GH_CHECK(hipGetLastError());
        GH_CHECK(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        csr_destroy(a);
        throw;
    }
    return a;
}

void csr_destroy(Csr* a) {
    if (!a) return;
    (void)hipFree(a->ptr);
    (void)hipFree(a->idx);
    (void)hipFree(a->v64);
    (void)hipFree(a->v32);
    (void)hipFree(a->v16);
    (void)hipFree(a->blocks);
    (void)hipFree(a->idx16);
    (void)hipFree(a->bcol);
    (void)hipFree(a->chunk_base);
    (void)hipFree(a->chunk);
    (void)hipFree(a->split_first);
    (void)hipFree(a->split_part);
    delete a;
}

template <class M, class V>
void launch_spmv(Ctx* ctx, const Csr* a, const V* x, V* y) {
    if (a->rows == 0) return;
    const CsrView<M> v = a->view<M>();
    if (a->lanes == 0) {
        hipLaunchKernelGGL((k_spmv_rb<M, V>), dim3(a->rb_grid()), dim3(kBlock), 0, ctx->stream, v, a->blocks,
                           a->split(), x, y);
        if (a->nsplit > 0)
            hipLaunchKernelGGL((k_spmv_rb<M, V>), dim3(a->nsplit), dim3(kBlock), 0, ctx->stream, v, a->blocks,
                               a->split_finish(), x, y);
        GH_CHECK(hipGetLastError());
        return;
    }
    with_lanes(a->lanes, [&](auto L) {
        constexpr int LPR = decltype(L)::value;
        hipLaunchKernelGGL((k_spmv<LPR, M, V>), dim3(grid_for((int64_t)a->rows * LPR)), dim3(kBlock), 0,
                           ctx->stream, v, x, y);
    });
    GH_CHECK(hipGetLastError());
}

// Time the stand-alone SpMV of each candidate variant (the row-block variant and the row-group widths
// around the histogram choice) with HIP events and keep the fastest.  The fused kernels share the row
// traversal of the stand-alone one, so its ranking carries over.
void csr_set_partition(Ctx* ctx, Csr* a, int rows, int nnz) {
    GH_REQUIRE(rows >= 1 && rows <= kBlock && nnz >= kLong && nnz <= kTile, "bad row-block partition caps");
    if (a->rb_rows == rows && a->rb_nnz == nnz) return;
    std::vector<int> ptr(a->rows + 1);
    GH_CHECK(hipMemcpyAsync(ptr.data(), a->ptr, sizeof(int) * ptr.size(), hipMemcpyDeviceToHost, ctx->stream));
    GH_CHECK(hipStreamSynchronize(ctx->stream));
    const std::vector<int> rb = row_blocks(ptr.data(), a->rows, rows, nnz);
    int* blocks = upload_row_blocks(ctx, ptr.data(), rb);
    GH_CHECK(hipFree(a->blocks));
    a->blocks = blocks;
    std::vector<int> idx((size_t)a->nnz);
    if (a->nnz) {
        GH_CHECK(hipMemcpyAsync(idx.data(), a->idx, sizeof(int) * idx.size(), hipMemcpyDeviceToHost, ctx->stream));
        GH_CHECK(hipStreamSynchronize(ctx->stream));
    }
    build_compact_indices(ctx, a, ptr.data(), idx.data(), rb);
    a->nblocks = (int)rb.size() - 1;
    a->rb_rows = rows;
    a->rb_nnz = nnz;
}

const std::vector<Variant>& variant_candidates() {
    static const std::vector<Variant> v{{0, 256, 2048}, {0, 256, 1024}, {0, 128, 1024}, {0, 64, 512},
                                        {1, 0, 0},      {2, 0, 0},      {4, 0, 0},      {8, 0, 0},
                                        {16, 0, 0},     {32, 0, 0},     {64, 0, 0}};
    return v;
}

void apply_variant(Ctx* ctx, Csr* a, const Variant& v) {
    a->lanes = v.lanes;
    if (v.lanes == 0) csr_set_partition(ctx, a, v.rows, v.nnz);
}

constexpr float kRowGroupMargin = 0.9f;  // a row-group variant must be >= 10% faster than the row blocks

namespace {

// The device state of one row-block partition (what csr_set_partition builds).  While the autotuner compares
// partitions it keeps every plan it has built, so each partition is built once instead of once per round.
struct PartitionPlan {
    bool built = false;
    int rows = 0, nnz = 0, nblocks = 0;
    int* blocks = nullptr;
    uint16_t* idx16 = nullptr;
    int* bcol = nullptr;
    int* chunk_base = nullptr;
    int64_t idx32_entries = 0;
    int wswz = 0;
};

PartitionPlan plan_of(const Csr* a) {
    return {true,   a->rb_rows,   a->rb_nnz,     a->nblocks,         a->blocks,
            a->idx16, a->bcol, a->chunk_base, a->idx32_entries, a->wswz};
}

void attach_plan(Csr* a, const PartitionPlan& p) {
    a->rb_rows = p.rows;
    a->rb_nnz = p.nnz;
    a->nblocks = p.nblocks;
    a->blocks = p.blocks;
    a->idx16 = p.idx16;
    a->bcol = p.bcol;
    a->chunk_base = p.chunk_base;
    a->idx32_entries = p.idx32_entries;
    a->wswz = p.wswz;
}

// `a` lets go of its plan (its owner keeps the buffers); the next csr_set_partition builds a fresh one
void detach_plan(Csr* a) {
    a->blocks = nullptr;
    a->idx16 = nullptr;
    a->bcol = nullptr;
    a->chunk_base = nullptr;
    a->rb_rows = a->rb_nnz = 0;
}

void free_plan(PartitionPlan& p) {
    (void)hipFree(p.blocks);
    (void)hipFree(p.idx16);
    (void)hipFree(p.bcol);
    (void)hipFree(p.chunk_base);
    p = PartitionPlan{};
}

}  // namespace

void autotune_with(Ctx* ctx, Csr* a, int reps, const std::function<void(int)>& launch) {
    if (a->rows == 0 || a->nnz == 0) return;
    const auto& cand = variant_candidates();
    const size_t nc = cand.size();
    // plans[c]: row-block candidate c's partition once built; `entry` holds the plan `a` came with when it is
    // none of the candidates'.  Every plan belongs to this table until the end, when `a` keeps the chosen one.
    std::vector<PartitionPlan> plans(nc);
    PartitionPlan entry = plan_of(a);
    for (size_t c = 0; c < nc; ++c) {
        if (cand[c].lanes == 0 && cand[c].rows == a->rb_rows && cand[c].nnz == a->rb_nnz) {
            plans[c] = entry;
            entry = PartitionPlan{};
            break;
        }
    }
    auto use = [&](size_t c) {
        a->lanes = cand[c].lanes;
        if (cand[c].lanes != 0) return;
        if (!plans[c].built) {
            detach_plan(a);
            csr_set_partition(ctx, a, cand[c].rows, cand[c].nnz);
            plans[c] = plan_of(a);
        } else {
            attach_plan(a, plans[c]);
        }
    };
    // Finer row-block partitions only add parallelism.  When the matrix's own partition (the coarsest) already
    // gives every CU several blocks, a finer one moves the same bytes in more, smaller blocks and is slower
    // (on the 2.9M-nonzero blending LP: 7.8 and 14 us against 6.5 us per SpMV), so it is neither built nor
    // timed -- building a partition is host work over every nonzero.
    constexpr int kManyBlocks = 1024;
    const bool coarse_enough = a->rb_rows == kBlock && a->rb_nnz == kTile && a->nblocks >= kManyBlocks;
    // When most nonzeros sit in rows longer than kSplit (cut into chunks over many workgroups on the row blocks), a
    // row group of at most 64 lanes cannot compete (one such row took 36 ms at 1 lane and 3 ms at 16 lanes, against
    // 7 us split): those variants are not timed either.
    const bool split_bound = a->nchunks > 0 && (int64_t)a->nchunks * kChunk >= a->nnz / 2;
    std::vector<char> skip(nc, 0);
    for (size_t c = 0; c < nc; ++c)
        skip[c] = (coarse_enough && cand[c].lanes == 0 && !(cand[c].rows == a->rb_rows && cand[c].nnz == a->rb_nnz)) ||
                  (split_bound && cand[c].lanes != 0);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<float> best(nc, 1e30f);
    size_t bi = nc, bl = nc;
    std::exception_ptr err;
    try {
        GH_CHECK(hipEventCreate(&e0));
        GH_CHECK(hipEventCreate(&e1));
        // Warm the clocks and caches first.  A short screening round (a third of the launches) drops every
        // candidate more than twice as slow as the fastest (a lanes-per-row width far from the rows' lengths can
        // be 20x slower than the row blocks); the others are timed in two interleaved full rounds and keep their
        // best, so a candidate timed early is not penalised by a cold GPU.  The choice rests on the full rounds
        // only: the screening's short windows are too noisy to rank near-equal variants.
        launch(20);
        auto time_one = [&](size_t c, int k) {
            use(c);
            launch(2);
            GH_CHECK(hipEventRecord(e0, ctx->stream));
            launch(k);
            GH_CHECK(hipEventRecord(e1, ctx->stream));
            GH_CHECK(hipEventSynchronize(e1));
            float ms = 0;
            GH_CHECK(hipEventElapsedTime(&ms, e0, e1));
            return ms / (float)k;
        };
        // a single launch first: a candidate already 1.3x slower than the lead's average is not screened further
        // (to be kept, a row-group width has to beat the row blocks by kRowGroupMargin anyway)
        auto probe_one = [&](size_t c) {
            use(c);
            GH_CHECK(hipEventRecord(e0, ctx->stream));
            launch(1);
            GH_CHECK(hipEventRecord(e1, ctx->stream));
            GH_CHECK(hipEventSynchronize(e1));
            float ms = 0;
            GH_CHECK(hipEventElapsedTime(&ms, e0, e1));
            return ms;
        };
        const int screen = std::max(1, reps / 3);
        std::vector<float> first(nc, 1e30f);
        float lead0 = 1e30f;
        for (size_t c = 0; c < nc; ++c) {
            if (skip[c]) continue;
            const float one = lead0 < 1e30f ? probe_one(c) : 0.f;
            if (one > 1.3f * lead0) {
                skip[c] = 1;
                continue;
            }
            first[c] = time_one(c, screen);
            lead0 = std::min(lead0, first[c]);
        }
        const float lead = *std::min_element(first.begin(), first.end());
        for (size_t c = 0; c < nc; ++c) skip[c] = skip[c] || first[c] > 2.0f * lead;
        for (int round = 0; round < 2; ++round)
            for (size_t c = 0; c < nc; ++c)
                if (!skip[c]) best[c] = std::min(best[c], time_one(c, reps));
        // the fastest row-block partition, unless a row-group width beats it by a clear margin: the row
        // blocks' row sums are the CPU row loop's bits, and a short timing window can rank near-equal variants
        // either way
        for (size_t c = 0; c < nc; ++c) {
            size_t& slot = cand[c].lanes == 0 ? bi : bl;
            if (slot == nc || best[c] < best[slot]) slot = c;
        }
        if (bi == nc || (bl != nc && best[bl] < kRowGroupMargin * best[bi])) bi = bl;
        use(bi);
    } catch (...) {
        err = std::current_exception();
        bi = nc;  // keep whatever plan `a` holds now
    }
    // `a` keeps the chosen partition (a row-group winner keeps the first row-block plan built, for the
    // passes that still use row blocks); every other plan is freed
    if (bi < nc && cand[bi].lanes != 0) {
        const int lanes = a->lanes;
        for (size_t c = 0; c < nc; ++c) {
            if (plans[c].built) {
                attach_plan(a, plans[c]);
                break;
            }
        }
        a->lanes = lanes;
    }
    for (size_t c = 0; c < nc; ++c)
        if (plans[c].built && plans[c].blocks != a->blocks) free_plan(plans[c]);
    if (entry.built && entry.blocks != a->blocks) free_plan(entry);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (err) std::rethrow_exception(err);
}

void autotune_csr(Ctx* ctx, Csr* a, int mdt, int vdt, int reps) {
    if (a->rows == 0 || a->nnz == 0) return;
    const size_t vs = dtype_size(vdt);
    void *x = nullptr, *y = nullptr;
    GH_CHECK(hipMalloc(&x, vs * std::max(a->cols, 1)));
    GH_CHECK(hipMalloc(&y, vs * std::max(a->rows, 1)));
    try {
        GH_CHECK(hipMemsetAsync(x, 0, vs * std::max(a->cols, 1), ctx->stream));
        autotune_with(ctx, a, reps, [&](int k) {
            with_mdt(mdt, [&](auto MT) {
                with_vdt(vdt, [&](auto VT) {
                    using M = typename decltype(MT)::type;
                    using V = typename decltype(VT)::type;
                    for (int r = 0; r < k; ++r) launch_spmv<M, V>(ctx, a, (const V*)x, (V*)y);
                });
            });
        });
    } catch (...) {
        (void)hipFree(x);
        (void)hipFree(y);
        throw;
    }
    (void)hipFree(x);
    (void)hipFree(y);
}

}  // namespace ganit_hip

// ================================================================ C ABI
using namespace ganit_hip;

extern "C" {

const char* ganit_hip_last_error(void) { return g_last_error.c_str(); }
int ganit_hip_version(void) { return 100; }

int ganit_hip_ctx_create(int device, ganit_hip_ctx** out) {
    return guarded([&] {
        GH_REQUIRE(out, "null output pointer");
        int count = 0;
        GH_CHECK(hipGetDeviceCount(&count));
        GH_REQUIRE(device >= 0 && device < count, "no such HIP device");
        GH_CHECK(hipSetDevice(device));
        hipDeviceProp_t prop;
        GH_CHECK(hipGetDeviceProperties(&prop, device));
        auto* c = new Ctx;
        c->device = device;
        c->cus = prop.multiProcessorCount;
        c->name = prop.name;
        if (std::strlen(prop.gcnArchName)) c->name += std::string(" (") + prop.gcnArchName + ")";
        GH_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        GH_CHECK(hipMalloc(&c->red_partials, sizeof(double) * kMaxReduceGrid * 2));
        GH_CHECK(hipMalloc(&c->red_out, sizeof(double) * 2));
        GH_CHECK(hipHostMalloc((void**)&c->host_out, sizeof(double) * 64, hipHostMallocDefault));
        GH_CHECK(hipStreamSynchronize(c->stream));
        *out = reinterpret_cast<ganit_hip_ctx*>(c);
    });
}

int ganit_hip_ctx_destroy(ganit_hip_ctx* h) {
    return guarded([&] {
        auto* c = reinterpret_cast<Ctx*>(h);
        if (!c) return;
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(c->red_partials);
        (void)hipFree(c->red_out);
        (void)hipHostFree(c->host_out);
        (void)hipStreamDestroy(c->stream);
        delete c;
    });
}

int ganit_hip_ctx_sync(ganit_hip_ctx* h) {
    return guarded([&] {
        GH_REQUIRE(h, "null context");
        GH_CHECK(hipStreamSynchronize(reinterpret_cast<Ctx*>(h)->stream));
    });
}

int ganit_hip_ctx_stream(ganit_hip_ctx* h, void** s) {
    return guarded([&] {
        GH_REQUIRE(h && s, "null argument");
        *s = (void*)reinterpret_cast<Ctx*>(h)->stream;
    });
}

int ganit_hip_ctx_device_info(ganit_hip_ctx* h, char* name, int name_len, int* cus) {
    return guarded([&] {
        GH_REQUIRE(h, "null context");
        auto* c = reinterpret_cast<Ctx*>(h);
        if (name && name_len > 0) {
            std::strncpy(name, c->name.c_str(), (size_t)name_len - 1);
            name[name_len - 1] = 0;
        }
        if (cus) *cus = c->cus;
    });
}

int ganit_hip_alloc(ganit_hip_ctx* h, size_t bytes, void** d) {
    return guarded([&] {
        GH_REQUIRE(h && d, "null argument");
        GH_CHECK(hipMalloc(d, std::max<size_t>(bytes, 1)));
    });
}

int ganit_hip_free(ganit_hip_ctx* h, void* d) {
    return guarded([&] {
        GH_REQUIRE(h, "null context");
        if (d) {
            GH_CHECK(hipStreamSynchronize(reinterpret_cast<Ctx*>(h)->stream));
            GH_CHECK(hipFree(d));
        }
    });
}

int ganit_hip_upload(ganit_hip_ctx* h, int vdt, const double* host, int64_t n, void* dst) {
    return guarded([&] {
        auto* c = reinterpret_cast<Ctx*>(h);
        GH_REQUIRE(c && (n == 0 || (host && dst)) && n >= 0, "bad upload arguments");
        if (n == 0) return;
        if (vdt == GANIT_F64) {
            GH_CHECK(hipMemcpyAsync(dst, host, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
            GH_CHECK(hipStreamSynchronize(c->stream));
        } else if (vdt == GANIT_F32) {
            std::vector<float> t(host, host + n);
            GH_CHECK(hipMemcpyAsync(dst, t.data(), sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
            GH_CHECK(hipStreamSynchronize(c->stream));
        } else {
            throw Error(GANIT_EINVAL, "vector dtype must be GANIT_F64 or GANIT_F32");
        }
    });
}

int ganit_hip_download(ganit_hip_ctx* h, int vdt, const void* src, int64_t n, double* host) {
    return guarded([&] {
        auto* c = reinterpret_cast<Ctx*>(h);
        GH_REQUIRE(c && (n == 0 || (host && src)) && n >= 0, "bad download arguments");
        if (n == 0) return;
        if (vdt == GANIT_F64) {
            GH_CHECK(hipMemcpyAsync(host, src, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
            GH_CHECK(hipStreamSynchronize(c->stream));
        } else if (vdt == GANIT_F32) {
            std::vector<float> t(n);
            GH_CHECK(hipMemcpyAsync(t.data(), src, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
            GH_CHECK(hipStreamSynchronize(c->stream));
            for (int64_t i = 0; i < n; ++i) host[i] = t[i];
        } else {
            throw Error(GANIT_EINVAL, "vector dtype must be GANIT_F64 or GANIT_F32");
        }
    });
}

int ganit_hip_vec_fill(ganit_hip_ctx* h, int vdt, void* dst, int64_t n, double a) {
    return guarded([&] {
        auto* c = reinterpret_cast<Ctx*>(h);
        GH_REQUIRE(c && n >= 0 && (n == 0 || dst), "bad fill arguments");
        if (n == 0) return;
        with_vdt(vdt, [&](auto T) {
            using V = typename decltype(T)::type;
            hipLaunchKernelGGL(k_fill<V>, dim3(ew_grid(n)), dim3(kBlock), 0, c->stream, n, a, (V*)dst);
        });
        GH_CHECK(hipGetLastError());
    });
}

int ganit_hip_vec_copy(ganit_hip_ctx* h, int vdt, const void* src, void* dst, int64_t n) {
    return guarded([&] {
        auto* c = reinterpret_cast<Ctx*>(h);
        GH_REQUIRE(c && n >= 0 && (n == 0 || (src && dst)), "bad copy arguments");
        if (n == 0) return;
        with_vdt(vdt, [&](auto T) {
            using V = typename decltype(T)::type;
            hipLaunchKernelGGL(k_copy<V>, dim3(ew_grid(n)), dim3(kBlock), 0, c->stream, n, (const V*)src, (V*)dst);
        });
        GH_CHECK(hipGetLastError());
    });
}

int ganit_hip_vec_sub(ganit_hip_ctx* h, int vdt, const void* a, const void* b, void* d, int64_t n) {
    return guarded([&] {
        auto* c = reinterpret_cast<Ctx*>(h);
        GH_REQUIRE(c && n >= 0 && (n == 0 || (a && b && d)), "bad sub arguments");
        if (n == 0) return;
        with_vdt(vdt, [&](auto T) {
            using V = typename decltype(T)::type;
            hipLaunchKernelGGL(k_sub<V>, dim3(ew_grid(n)), dim3(kBlock), 0, c->stream, n, (const V*)a, (const V*)b,
                               (V*)d);
        });
        GH_CHECK(hipGetLastError());
    });
}

int ganit_hip_vec_scale_into(ganit_hip_ctx* h, int vdt, const void* src, double a, void* dst, int64_t n) {
    return guarded([&] {
        auto* c = reinterpret_cast<Ctx*>(h);
        GH_REQUIRE(c && n >= 0 && (n == 0 || (src && dst)), "bad scale arguments");
        if (n == 0) return;
        with_vdt(vdt, [&](auto T) {
            using V = typename decltype(T)::type;
            hipLaunchKernelGGL(k_scale<V>, dim3(ew_grid(n)), dim3(kBlock), 0, c->stream, n, a, (const V*)src, (V*)dst);
        });
        GH_CHECK(hipGetLastError());
    });
}

int ganit_hip_vec_convert(ganit_hip_ctx* h, int sdt, const void* src, int ddt, void* dst, int64_t n) {
    return guarded([&] {
        auto* c = reinterpret_cast<Ctx*>(h);
        GH_REQUIRE(c && n >= 0 && (n == 0 || (src && dst)), "bad convert arguments");
        if (n == 0) return;
        with_vdt(sdt, [&](auto S) {
            with_vdt(ddt, [&](auto D) {
                using SV = typename decltype(S)::type;
                using DV = typename decltype(D)::type;
                hipLaunchKernelGGL((k_convert<SV, DV>), dim3(ew_grid(n)), dim3(kBlock), 0, c->stream, n,
                                   (const SV*)src, (DV*)dst);
            });
        });
        GH_CHECK(hipGetLastError());
    });
}

int ganit_hip_csr_upload(ganit_hip_ctx* h, int rows, int cols, const int* ptr, const int* idx, const double* val,
                         unsigned mask, ganit_hip_csr** out) {
    return guarded([&] {
        GH_REQUIRE(h && out, "null argument");
        *out = reinterpret_cast<ganit_hip_csr*>(csr_create(reinterpret_cast<Ctx*>(h), rows, cols, ptr, idx, val, mask));
    });
}

int ganit_hip_csr_free(ganit_hip_csr* a) {
    return guarded([&] { csr_destroy(reinterpret_cast<Csr*>(a)); });
}

int ganit_hip_csr_plan(const ganit_hip_csr* h, int* lanes, int64_t* hist) {
    return guarded([&] {
        GH_REQUIRE(h, "null matrix");
        auto* a = reinterpret_cast<const Csr*>(h);
        if (lanes) *lanes = a->lanes;
        if (hist) std::copy(a->hist, a->hist + 9, hist);
    });
}

int ganit_hip_csr_set_lanes(ganit_hip_csr* h, int lanes) {
    return guarded([&] {
        GH_REQUIRE(h, "null matrix");
        GH_REQUIRE(valid_variant(lanes), "lanes per row must be 0 (row blocks) or a power of two in 1..64");
        reinterpret_cast<Csr*>(h)->lanes = lanes;
    });
}

int ganit_hip_csr_autotune(ganit_hip_ctx* h, ganit_hip_csr* ah, int mdt, int vdt, int reps) {
    return guarded([&] {
        GH_REQUIRE(h && ah && reps > 0, "bad autotune arguments");
        autotune_csr(reinterpret_cast<Ctx*>(h), reinterpret_cast<Csr*>(ah), mdt, vdt, reps);
    });
}

int ganit_hip_spmv(ganit_hip_ctx* h, const ganit_hip_csr* ah, int mdt, int vdt, const void* x, void* y) {
    return guarded([&] {
        auto* c = reinterpret_cast<Ctx*>(h);
        auto* a = reinterpret_cast<const Csr*>(ah);
        GH_REQUIRE(c && a && (a->rows == 0 || (x && y)), "bad spmv arguments");
        with_mdt(mdt, [&](auto MT) {
            with_vdt(vdt, [&](auto VT) {
                using M = typename decltype(MT)::type;
                using V = typename decltype(VT)::type;
                launch_spmv<M, V>(c, a, (const V*)x, (V*)y);
            });
        });
    });
}

int ganit_hip_primal_step(ganit_hip_ctx* h, int vdt, int n, const void* x, const void* aty, const void* qx,
                          const void* cc, const void* l, const void* u, double tau, void* xn, void* xbar,
                          void* sumx) {
    return guarded([&] {
        auto* c = reinterpret_cast<Ctx*>(h);
        GH_REQUIRE(c && n >= 0, "bad primal_step arguments");
        if (n == 0) return;
        GH_REQUIRE(x && aty && cc && l && u && xn && xbar && sumx, "null vector");
        with_vdt(vdt, [&](auto T) {
            using V = typename decltype(T)::type;
            if (qx)
                hipLaunchKernelGGL((k_primal_step<V, true>), dim3(grid_for(n)), dim3(kBlock), 0, c->stream, n,
                                   (const V*)x, (const V*)aty, (const V*)qx, (const V*)cc, (const V*)l, (const V*)u,
                                   V(tau), (V*)xn, (V*)xbar, (V*)sumx);
            else
                hipLaunchKernelGGL((k_primal_step<V, false>), dim3(grid_for(n)), dim3(kBlock), 0, c->stream, n,
                                   (const V*)x, (const V*)aty, (const V*)nullptr, (const V*)cc, (const V*)l,
                                   (const V*)u, V(tau), (V*)xn, (V*)xbar, (V*)sumx);
        });
        GH_CHECK(hipGetLastError());
    });
}

int ganit_hip_dual_step(ganit_hip_ctx* h, int vdt, int m, const void* y, const void* axbar, const void* lo,
                        const void* hi, double sigma, void* yn, void* sumy) {
    return guarded([&] {
        auto* c = reinterpret_cast<Ctx*>(h);
        GH_REQUIRE(c && m >= 0, "bad dual_step arguments");
        if (m == 0) return;
        GH_REQUIRE(y && axbar && lo && hi && yn && sumy, "null vector");
        with_vdt(vdt, [&](auto T) {
            using V = typename decltype(T)::type;
            hipLaunchKernelGGL(k_dual_step<V>, dim3(grid_for(m)), dim3(kBlock), 0, c->stream, m, (const V*)y,
                               (const V*)axbar, (const V*)lo, (const V*)hi, V(sigma), (V*)yn, (V*)sumy);
        });
        GH_CHECK(hipGetLastError());
    });
}

int ganit_hip_primal_res2(ganit_hip_ctx* h, int vdt, int m, const void* ax, const void* lo, const void* hi,
                          const void* rres, double* out) {
    return guarded([&] {
        auto* c = reinterpret_cast<Ctx*>(h);
        GH_REQUIRE(c && out && m >= 0 && (m == 0 || (ax && lo && hi && rres)), "bad primal_res2 arguments");
        with_vdt(vdt, [&](auto T) {
            using V = typename decltype(T)::type;
            reduce2(c, m, PrimalResTerm<V>{(const V*)ax, (const V*)lo, (const V*)hi, (const V*)rres}, out, nullptr);
        });
    });
}

int ganit_hip_dual_res(ganit_hip_ctx* h, int vdt, int n, const void* aty, const void* qx, const void* cc,
                       const void* l, const void* u, const void* cres, int with_c, double* rd2, double* vobj) {
    return guarded([&] {
        auto* c = reinterpret_cast<Ctx*>(h);
        GH_REQUIRE(c && rd2 && vobj && n >= 0 && (n == 0 || (aty && cc && l && u && cres)), "bad dual_res arguments");
        with_vdt(vdt, [&](auto T) {
            using V = typename decltype(T)::type;
            reduce2(c, n,
                    DualResTerm<V>{(const V*)aty, (const V*)qx, (const V*)cc, (const V*)l, (const V*)u, (const V*)cres,
                                   with_c ? 1.0 : 0.0},
                    rd2, vobj);
        });
    });
}

int ganit_hip_row_dual_obj(ganit_hip_ctx* h, int vdt, int m, const void* y, const void* lo, const void* hi,
                           double* out) {
    return guarded([&] {
        auto* c = reinterpret_cast<Ctx*>(h);
        GH_REQUIRE(c && out && m >= 0 && (m == 0 || (y && lo && hi)), "bad row_dual_obj arguments");
        with_vdt(vdt, [&](auto T) {
            using V = typename decltype(T)::type;
            reduce2(c, m, RowDualObjTerm<V>{(const V*)y, (const V*)lo, (const V*)hi}, out, nullptr);
        });
    });
}

int ganit_hip_dot(ganit_hip_ctx* h, int vdt, int64_t n, const void* a, const void* b, double* out) {
    return guarded([&] {
        auto* c = reinterpret_cast<Ctx*>(h);
        GH_REQUIRE(c && out && n >= 0 && (n == 0 || (a && b)), "bad dot arguments");
        with_vdt(vdt, [&](auto T) {
            using V = typename decltype(T)::type;
            reduce2(c, n, DotTerm<V>{(const V*)a, (const V*)b}, out, nullptr);
        });
    });
}

int ganit_hip_dot_noalloc(void* hip_stream, int vdt, int64_t n, const void* a, const void* b, double* out) {
    return guarded([&] {
        GH_REQUIRE(out && n >= 0 && (n == 0 || (a && b)), "bad dot_noalloc arguments");
        const unsigned g = reduce_grid(n);
        GH_REQUIRE(2 * (int64_t)g <= kNoAllocDoubles,
                   "dot_noalloc: the reduction grid needs more partials than its 16 KiB scratch holds");
        int device = -1;
        GH_CHECK(hipGetDevice(&device));
        auto stream = static_cast<hipStream_t>(hip_stream);
        if (stream) {
            hipDevice_t sdev = -1;
            GH_CHECK(hipStreamGetDevice(stream, &sdev));
            GH_REQUIRE(sdev == device, "dot_noalloc: the stream belongs to another device than the current one");
        }
        double* partials = noalloc_scratch(device);
        with_vdt(vdt, [&](auto T) {
            using V = typename decltype(T)::type;
            using F = DotTerm<V>;
            ReduceSlot slot{partials, nullptr};
            hipLaunchKernelGGL(k_reduce2<F>, dim3(g), dim3(kBlock), 0, stream, n, F{(const V*)a, (const V*)b}, slot);
        });
        hipLaunchKernelGGL(k_finish_first, dim3(1), dim3(kBlock), 0, stream, (const double*)partials, (int)g, out);
        GH_CHECK(hipGetLastError());
    });
}

int ganit_hip_diff2(ganit_hip_ctx* h, int vdt, int64_t n, const void* a, const void* b, double* out) {
    return guarded([&] {
        auto* c = reinterpret_cast<Ctx*>(h);
        GH_REQUIRE(c && out && n >= 0 && (n == 0 || (a && b)), "bad diff2 arguments");
        with_vdt(vdt, [&](auto T) {
            using V = typename decltype(T)::type;
            reduce2(c, n, Diff2Term<V>{(const V*)a, (const V*)b}, out, nullptr);
        });
    });
}

}  // extern "C"
