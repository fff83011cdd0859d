// libganit_hip.so, part 5: the problem set-up on the device -- the matrix half of scale_lp (10 Ruiz passes and
// the Pock-Chambolle pass, reference src/scaling.cpp:36-70) and the SpMVs of the norm estimate's power
// iteration (reference src/scaling.cpp:131-155), on a matrix uploaded once.
//
// Every value is the host's to the bit: maxima are exact in any order; a row's (or a column's) 1-norm is summed
// by one thread in ascending position order, which for a column is ascending row order -- the reference's
// row-major accumulation; each scaling factor is the same 1 / sqrt of the same value, and each matrix entry is
// multiplied by the same products in the same order (IEEE multiplication commutes, so A^T's copy of an entry is
// scaled alike); the power iteration's SpMVs form each product, then sum each row sequentially, with no
// contraction into FMA.  The one sequential reduction over a whole vector -- the norm of A'^T A' v -- is the
// reference's serial sum, taken on the host from a pinned copy of z each step.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"

namespace ganit_hip {

namespace {

struct DevMem {
    void* p = nullptr;
    DevMem() = default;
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    ~DevMem() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t bytes) { GH_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 16))); }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// row_of[p] = the row of entry p
__global__ __launch_bounds__(kBlock) void k_row_of(int rows, const int* __restrict__ ptr, int* __restrict__ row_of) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= rows) return;
    for (int p = ptr[i]; p < ptr[i + 1]; ++p) row_of[p] = i;
}

__global__ __launch_bounds__(kBlock) void k_iota(int64_t n, int* __restrict__ v) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) v[i] = (int)i;
}

// A^T's row pointers from the column-sorted keys: at_ptr[j] = first position whose column is >= j
__global__ __launch_bounds__(kBlock) void k_col_starts(int cols, int64_t nnz, const int* __restrict__ sorted_col,
                                                       int* __restrict__ at_ptr) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j > cols) return;
    int64_t lo = 0, hi = nnz;
    while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (sorted_col[mid] < j) lo = mid + 1;
        else hi = mid;
    }
    at_ptr[j] = (int)lo;
}

// A^T's entries: q-th entry of the column-sorted order is A entry perm[q]
__global__ __launch_bounds__(kBlock) void k_gather_transpose(int64_t nnz, const int* __restrict__ perm,
                                                             const int* __restrict__ row_of,
                                                             const double* __restrict__ a_val, int* __restrict__ at_idx,
                                                             double* __restrict__ at_val) {
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < nnz; q += (int64_t)gridDim.x * kBlock) {
        const int p = perm[q];
        at_idx[q] = row_of[p];
        at_val[q] = a_val[p];
    }
}

// f[i] = 1 / sqrt(max |a_ip|), 1 for an empty or all-zero row (reference src/scaling.cpp:48-55): LPR lanes per
// row, each taking every LPR-th entry, then a max over the group (maxima are exact in any order)
template <int LPR>
__global__ __launch_bounds__(kBlock) void k_row_absmax_factor(int rows, const int* __restrict__ ptr,
                                                              const double* __restrict__ val, double* __restrict__ f) {
    const int64_t gt = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int i = (int)(gt / LPR), lane = (int)(gt % LPR);
    double t = 0.0;
    if (i < rows)
        for (int p = ptr[i] + lane; p < ptr[i + 1]; p += LPR) t = fmax(t, fabs(val[p]));
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) t = fmax(t, __shfl_xor(t, off, LPR));
    if (i < rows && lane == 0) f[i] = t > 0 ? 1.0 / sqrt(t) : 1.0;
}

// f[i] = 1 / sqrt(sum |a_ip|) with the row summed by one thread in position order (reference
// src/scaling.cpp:61-68; for A^T's rows that is ascending row order, as the reference accumulates column sums)
__global__ __launch_bounds__(kBlock) void k_row_abssum_factor(int rows, const int* __restrict__ ptr,
                                                              const double* __restrict__ val, double* __restrict__ f) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= rows) return;
    double t = 0.0;
    for (int p = ptr[i]; p < ptr[i + 1]; ++p) t = t + fabs(val[p]);
    f[i] = t > 0 ? 1.0 / sqrt(t) : 1.0;
}

// a_p *= rf[row(p)] * cf[col(p)] (reference src/scaling.cpp:27-31; for A^T row and column swap roles, and the
// product of the two factors is the same either way round)
__global__ __launch_bounds__(kBlock) void k_scale_entries(int64_t nnz, const int* __restrict__ row_of,
                                                          const int* __restrict__ idx, double* __restrict__ val,
                                                          const double* __restrict__ rf,
                                                          const double* __restrict__ cf) {
#pragma clang fp contract(off)
    for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * kBlock)
        val[p] *= rf[row_of[p]] * cf[idx[p]];
}

// tot[i] *= f[i]
__global__ __launch_bounds__(kBlock) void k_scale_vec(int n, const double* __restrict__ f, double* __restrict__ tot) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) tot[i] *= f[i];
}

// The power iteration's SpMV in two steps with the host loop's roundings: every product val * x[idx] (rounded,
// no FMA), then each row's products added by one thread in position order
__global__ __launch_bounds__(kBlock) void k_products(int64_t nnz, const int* __restrict__ idx,
                                                     const double* __restrict__ val, const double* __restrict__ x,
                                                     double* __restrict__ prod) {
#pragma clang fp contract(off)
    for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * kBlock)
        prod[p] = val[p] * x[idx[p]];
}
__global__ __launch_bounds__(kBlock) void k_row_sums(int rows, const int* __restrict__ ptr,
                                                     const double* __restrict__ prod, double* __restrict__ y) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= rows) return;
    double t = 0.0;
    for (int p = ptr[i]; p < ptr[i + 1]; ++p) t = t + prod[p];
    y[i] = t;
}

// v = z / nz, or (first step) v = 1 / nz
__global__ __launch_bounds__(kBlock) void k_normalise(int n, const double* __restrict__ z, double nz, int first,
                                                      double* __restrict__ v) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    v[j] = (first ? 1.0 : z[j]) / nz;
}

// sqrt of the sum of squares in index order, no FMA: the reference's norm2 (src/scaling.cpp:8-12)
double serial_norm(const double* v, int n) {
#pragma clang fp contract(off)
    double s = 0.0;
    for (int j = 0; j < n; ++j) s += v[j] * v[j];
    return std::sqrt(s);
}

unsigned rows_grid(int64_t rows) { return grid_for(std::max<int64_t>(rows, 1)); }

}  // namespace

struct Setup {
    Ctx* ctx = nullptr;
    int m = 0, n = 0;
    int64_t nnz = 0;
    // A, its transpose, the row of each A entry and the row of each A^T entry (= its column in A)
    DevMem a_ptr, a_idx, a_val, at_ptr, at_idx, at_val, row_of, col_of, R, C, rf, cf;
    double* z_host = nullptr;  // pinned
    ~Setup() {
        if (z_host) (void)hipHostFree(z_host);
    }
};

}  // namespace ganit_hip

using namespace ganit_hip;

extern "C" {

int ganit_hip_setup_create(ganit_hip_ctx* h, int m, int n, const int* a_ptr, const int* a_idx, const double* a_val,
                           int ruiz_iters, int pock_chambolle, ganit_hip_setup** out) {
    return guarded([&] {
        auto* ctx = reinterpret_cast<Ctx*>(h);
        GH_REQUIRE(ctx && out && m >= 0 && n >= 0 && a_ptr && ruiz_iters >= 0, "bad set-up arguments");
        const int64_t nnz = a_ptr[m];
        GH_REQUIRE(nnz >= 0 && (nnz == 0 || (a_idx && a_val)), "bad set-up matrix");
        GH_REQUIRE(nnz < (int64_t)1 << 31, "set-up needs fewer than 2^31 nonzeros");
        for (int i = 0; i < m; ++i) GH_REQUIRE(a_ptr[i] <= a_ptr[i + 1], "row pointers must not decrease");
        for (int64_t p = 0; p < nnz; ++p) GH_REQUIRE(a_idx[p] >= 0 && a_idx[p] < n, "column index out of range");
        auto* s = new Setup;
        try {
            s->ctx = ctx;
            s->m = m;
            s->n = n;
            s->nnz = nnz;
            hipStream_t st = ctx->stream;
            s->a_ptr.alloc(sizeof(int) * ((size_t)m + 1));
            s->a_idx.alloc(sizeof(int) * nnz);
            s->a_val.alloc(sizeof(double) * nnz);
            s->at_ptr.alloc(sizeof(int) * ((size_t)n + 1));
            s->at_idx.alloc(sizeof(int) * nnz);
            s->at_val.alloc(sizeof(double) * nnz);
            s->row_of.alloc(sizeof(int) * nnz);
            s->col_of.alloc(sizeof(int) * nnz);
            s->R.alloc(sizeof(double) * m);
            s->C.alloc(sizeof(double) * n);
            s->rf.alloc(sizeof(double) * m);
            s->cf.alloc(sizeof(double) * n);
            GH_CHECK(hipMemcpyAsync(s->a_ptr.p, a_ptr, sizeof(int) * ((size_t)m + 1), hipMemcpyHostToDevice, st));
            if (nnz > 0) {
                GH_CHECK(hipMemcpyAsync(s->a_idx.p, a_idx, sizeof(int) * nnz, hipMemcpyHostToDevice, st));
                GH_CHECK(hipMemcpyAsync(s->a_val.p, a_val, sizeof(double) * nnz, hipMemcpyHostToDevice, st));
            }
            {  // R = C = 1
                std::vector<double> ones((size_t)std::max(m, n), 1.0);
                GH_CHECK(hipMemcpyAsync(s->R.p, ones.data(), sizeof(double) * m, hipMemcpyHostToDevice, st));
                GH_CHECK(hipMemcpyAsync(s->C.p, ones.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
                GH_CHECK(hipStreamSynchronize(st));
            }
            const unsigned eg = std::min<unsigned>(rows_grid(nnz), 8192u);
            // ---- the stable transpose of the pattern: entries sorted by column, ascending position (= row)
            // within a column (a stable radix sort), as the reference's transpose orders them (src/lp.cpp:55-72);
            // the sorted keys are the column of each A^T entry
            {
                DevMem perm_in, perm_out, temp;
                perm_in.alloc(sizeof(int) * nnz);
                perm_out.alloc(sizeof(int) * nnz);
                if (nnz > 0) {
                    hipLaunchKernelGGL(k_iota, dim3(eg), dim3(kBlock), 0, st, nnz, perm_in.as<int>());
                    hipLaunchKernelGGL(k_row_of, dim3(rows_grid(m)), dim3(kBlock), 0, st, m, s->a_ptr.as<const int>(),
                                       s->row_of.as<int>());
                    int end_bit = 1;
                    while (end_bit < 31 && ((int64_t)1 << end_bit) < (int64_t)n) ++end_bit;
                    size_t temp_bytes = 0;
                    GH_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, s->a_idx.as<const int>(),
                                                                s->col_of.as<int>(), perm_in.as<const int>(),
                                                                perm_out.as<int>(), (int)nnz, 0, end_bit, st));
                    temp.alloc(temp_bytes);
                    GH_CHECK(hipcub::DeviceRadixSort::SortPairs(temp.p, temp_bytes, s->a_idx.as<const int>(),
                                                                s->col_of.as<int>(), perm_in.as<const int>(),
                                                                perm_out.as<int>(), (int)nnz, 0, end_bit, st));
                    hipLaunchKernelGGL(k_gather_transpose, dim3(eg), dim3(kBlock), 0, st, nnz,
                                       perm_out.as<const int>(), s->row_of.as<const int>(), s->a_val.as<const double>(),
                                       s->at_idx.as<int>(), s->at_val.as<double>());
                }
                hipLaunchKernelGGL(k_col_starts, dim3(rows_grid((int64_t)n + 1)), dim3(kBlock), 0, st, n, nnz,
                                   s->col_of.as<const int>(), s->at_ptr.as<int>());
                GH_CHECK(hipGetLastError());
                GH_CHECK(hipStreamSynchronize(st));  // the temporaries are freed next
            }
            // ---- Ruiz passes, then Pock-Chambolle (reference src/scaling.cpp:45-70)
            auto pass = [&](bool sum) {
                if (sum) {
                    hipLaunchKernelGGL(k_row_abssum_factor, dim3(rows_grid(m)), dim3(kBlock), 0, st, m,
                                       s->a_ptr.as<const int>(), s->a_val.as<const double>(), s->rf.as<double>());
                    hipLaunchKernelGGL(k_row_abssum_factor, dim3(rows_grid(n)), dim3(kBlock), 0, st, n,
                                       s->at_ptr.as<const int>(), s->at_val.as<const double>(), s->cf.as<double>());
                } else {
                    hipLaunchKernelGGL(k_row_absmax_factor<kWave>, dim3(rows_grid((int64_t)m * kWave)), dim3(kBlock), 0,
                                       st, m, s->a_ptr.as<const int>(), s->a_val.as<const double>(), s->rf.as<double>());
                    hipLaunchKernelGGL(k_row_absmax_factor<8>, dim3(rows_grid((int64_t)n * 8)), dim3(kBlock), 0, st, n,
                                       s->at_ptr.as<const int>(), s->at_val.as<const double>(), s->cf.as<double>());
                }
                if (nnz > 0) {
                    hipLaunchKernelGGL(k_scale_entries, dim3(eg), dim3(kBlock), 0, st, nnz, s->row_of.as<const int>(),
                                       s->a_idx.as<const int>(), s->a_val.as<double>(), s->rf.as<const double>(),
                                       s->cf.as<const double>());
                    hipLaunchKernelGGL(k_scale_entries, dim3(eg), dim3(kBlock), 0, st, nnz, s->col_of.as<const int>(),
                                       s->at_idx.as<const int>(), s->at_val.as<double>(), s->cf.as<const double>(),
                                       s->rf.as<const double>());
                }
                hipLaunchKernelGGL(k_scale_vec, dim3(rows_grid(m)), dim3(kBlock), 0, st, m, s->rf.as<const double>(),
                                   s->R.as<double>());
                hipLaunchKernelGGL(k_scale_vec, dim3(rows_grid(n)), dim3(kBlock), 0, st, n, s->cf.as<const double>(),
                                   s->C.as<double>());
            };
            for (int it = 0; it < ruiz_iters; ++it) pass(false);
            if (pock_chambolle) pass(true);
            GH_CHECK(hipGetLastError());
            GH_CHECK(hipStreamSynchronize(st));
        } catch (...) {
            delete s;
            throw;
        }
        *out = reinterpret_cast<ganit_hip_setup*>(s);
    });
}

int ganit_hip_setup_result(const ganit_hip_setup* h, double* R, double* C, double* a_val, int* at_ptr, int* at_idx,
                           double* at_val) {
    return guarded([&] {
        auto* s = reinterpret_cast<const Setup*>(h);
        GH_REQUIRE(s, "null set-up");
        hipStream_t st = s->ctx->stream;
        auto down = [&](void* dst, const DevMem& src, size_t bytes) {
            if (dst && bytes) GH_CHECK(hipMemcpyAsync(dst, src.p, bytes, hipMemcpyDeviceToHost, st));
        };
        down(R, s->R, sizeof(double) * s->m);
        down(C, s->C, sizeof(double) * s->n);
        down(a_val, s->a_val, sizeof(double) * s->nnz);
        down(at_ptr, s->at_ptr, sizeof(int) * ((size_t)s->n + 1));
        down(at_idx, s->at_idx, sizeof(int) * s->nnz);
        down(at_val, s->at_val, sizeof(double) * s->nnz);
        GH_CHECK(hipStreamSynchronize(st));
    });
}

int ganit_hip_setup_norm(ganit_hip_setup* h, int iters, double* norm) {
    return guarded([&] {
        auto* s = reinterpret_cast<Setup*>(h);
        GH_REQUIRE(s && norm && iters >= 0, "bad norm arguments");
        if (s->nnz == 0) {  // reference src/scaling.cpp:132
            *norm = 1.0;
            return;
        }
        hipStream_t st = s->ctx->stream;
        const int m = s->m, n = s->n;
        DevMem v, w, z, prod;
        v.alloc(sizeof(double) * n);
        w.alloc(sizeof(double) * m);
        z.alloc(sizeof(double) * n);
        prod.alloc(sizeof(double) * s->nnz);
        if (!s->z_host) GH_CHECK(hipHostMalloc((void**)&s->z_host, sizeof(double) * std::max(n, 1), hipHostMallocDefault));
        const unsigned eg = std::min<unsigned>(rows_grid(s->nnz), 8192u);
        // v = ones / ||ones||, then w = A' v, z = A'^T w, lambda = ||z||, v = z / lambda (src/scaling.cpp:133-152)
        std::vector<double> ones((size_t)n, 1.0);
        double nz = serial_norm(ones.data(), n), lambda = 0.0;
        for (int it = 0; it < iters; ++it) {
            hipLaunchKernelGGL(k_normalise, dim3(rows_grid(n)), dim3(kBlock), 0, st, n, z.as<const double>(), nz,
                               it == 0 ? 1 : 0, v.as<double>());
            hipLaunchKernelGGL(k_products, dim3(eg), dim3(kBlock), 0, st, s->nnz, s->a_idx.as<const int>(),
                               s->a_val.as<const double>(), v.as<const double>(), prod.as<double>());
            hipLaunchKernelGGL(k_row_sums, dim3(rows_grid(m)), dim3(kBlock), 0, st, m, s->a_ptr.as<const int>(),
                               prod.as<const double>(), w.as<double>());
            hipLaunchKernelGGL(k_products, dim3(eg), dim3(kBlock), 0, st, s->nnz, s->at_idx.as<const int>(),
                               s->at_val.as<const double>(), w.as<const double>(), prod.as<double>());
            hipLaunchKernelGGL(k_row_sums, dim3(rows_grid(n)), dim3(kBlock), 0, st, n, s->at_ptr.as<const int>(),
                               prod.as<const double>(), z.as<double>());
            GH_CHECK(hipGetLastError());
            GH_CHECK(hipMemcpyAsync(s->z_host, z.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
            GH_CHECK(hipStreamSynchronize(st));
            nz = serial_norm(s->z_host, n);
            if (nz == 0.0) break;
            lambda = nz;  // ||v|| = 1, so ||A'^T A' v|| -> sigma_max^2
        }
        *norm = std::sqrt(lambda) * 1.01;
    });
}

int ganit_hip_setup_destroy(ganit_hip_setup* h) {
    return guarded([&] {
        auto* s = reinterpret_cast<Setup*>(h);
        if (!s) return;
        (void)hipStreamSynchronize(s->ctx->stream);
        delete s;
    });
}

}  // extern "C"
