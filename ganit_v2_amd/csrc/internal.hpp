// Host-side internals of libganit_hip.so shared by its translation units: error plumbing, the context and
// matrix handle layouts, and dtype / lanes-per-row dispatch.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/ganit_hip.h"
#include "device_common.hpp"

namespace ganit_hip {

// ---------------------------------------------------------------- errors
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define GH_CHECK(call)                                                                                   \
    do {                                                                                                 \
        hipError_t e__ = (call);                                                                         \
        if (e__ != hipSuccess)                                                                           \
            throw ::ganit_hip::Error(e__ == hipErrorOutOfMemory ? GANIT_ENOMEM : GANIT_EHIP,            \
                                     std::string("HIP error: ") + hipGetErrorString(e__) + " at " +     \
                                         __FILE__ + ":" + std::to_string(__LINE__));                     \
    } while (0)

#define GH_REQUIRE(cond, msg)                                                  \
    do {                                                                       \
        if (!(cond)) throw ::ganit_hip::Error(GANIT_EINVAL, std::string(msg)); \
    } while (0)

// Wrap a C-ABI body: exceptions become status codes plus a per-thread message.
template <class F>
int guarded(F&& f) {
    try {
        f();
        return GANIT_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return GANIT_ENOMEM;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return GANIT_EINTERNAL;
    }
}

// ---------------------------------------------------------------- handles
constexpr int kMaxReduceGrid = 1024;  // grid-stride reductions never use more blocks than this

struct Ctx {
    int device = 0;
    int cus = 0;
    std::string name;
    hipStream_t stream = nullptr;
    // scratch of the fine-grained reductions: partials [kMaxReduceGrid][2], out[2]
    double* red_partials = nullptr;
    double* red_out = nullptr;
    double* host_out = nullptr;  // pinned, 64 doubles
};

struct Csr {
    Ctx* ctx = nullptr;
    int rows = 0, cols = 0;
    int64_t nnz = 0;
    int* ptr = nullptr;
    int* idx = nullptr;
    double* v64 = nullptr;
    float* v32 = nullptr;
    bf16_t* v16 = nullptr;
    int lanes = 8;           // row-group variant: lanes per row (1..64); 0 selects the LDS-staged row blocks
    int64_t hist[9] = {0};
    uint16_t* idx16 = nullptr;  // compact indices (null when a staged block spans >= 2^16 columns); bcol[2b] =
    int* bcol = nullptr;        // block b's smallest column, bcol[2b + 1] = its window length (-1: long row, 32-bit)
    int* chunk_base = nullptr;  // per split-row chunk: its smallest column, or -1 (32-bit indices)
    int64_t idx32_entries = 0;  // entries still read through 32-bit indices when idx16 is present
    int wswz = 0;               // 1: windows staged swizzled (chosen from the gathers' LDS bank pattern)
    int* blocks = nullptr;   // row-block descriptors of the LDS-staged variant: 4 ints per block (first row,
                             // end row, first nonzero, end nonzero), device; see upload_row_blocks
    int nblocks = 0;
    int rb_rows = kBlock, rb_nnz = kTile;  // row-block caps the partition was built with
    // rows longer than kSplit nonzeros, cut into kChunk chunks (device arrays; see SplitPlan)
    int nsplit = 0, nchunks = 0;
    int* chunk = nullptr;
    int* split_first = nullptr;
    double* split_part = nullptr;  // 3 doubles per chunk
    SplitPlan split() const { return {nblocks, chunk, split_first, split_part, 0, chunk_base}; }
    // second launch of a row-block kernel when nsplit > 0: grid nsplit, folds the chunk partials
    SplitPlan split_finish() const { return {nblocks, chunk, split_first, split_part, 1, chunk_base}; }
    unsigned rb_grid() const { return (unsigned)(nblocks + nchunks); }
    template <class M>
    CsrView<M> view() const;
};

template <>
inline CsrView<double> Csr::view<double>() const {
    GH_REQUIRE(rows == 0 || v64, "matrix has no FP64 values resident");
    return {rows, ptr, idx, v64, idx16, bcol, wswz};
}
template <>
inline CsrView<float> Csr::view<float>() const {
    GH_REQUIRE(rows == 0 || v32, "matrix has no FP32 values resident");
    return {rows, ptr, idx, v32, idx16, bcol, wswz};
}
template <>
inline CsrView<bf16_t> Csr::view<bf16_t>() const {
    GH_REQUIRE(rows == 0 || v16, "matrix has no BF16 values resident");
    return {rows, ptr, idx, v16, idx16, bcol, wswz};
}

Csr* csr_create(Ctx* ctx, int rows, int cols, const int* ptr, const int* idx, const double* val, unsigned mask);
void csr_destroy(Csr* a);
int choose_lanes_from_lengths(const std::vector<int>& len, int64_t nnz);
bool valid_variant(int lanes);
std::vector<int> row_blocks(const int* ptr, int rows, int max_rows = kBlock, int max_nnz = kTile);
// Device array of block descriptors {r0, r1, ptr[r0], ptr[r1]} for the row-block boundaries rb (host ptr).
int* upload_row_blocks(Ctx* ctx, const int* ptr, const std::vector<int>& rb);
// (Re)build a->idx16 / a->bcol for the partition rb when every staged block spans fewer than 2^16 columns;
// otherwise leave them null (the kernels then read the 32-bit indices).
void build_compact_indices(Ctx* ctx, Csr* a, const int* ptr, const int* idx, const std::vector<int>& rb);
// Rebuild the row-block partition with at most `rows` rows and `nnz` nonzeros per block.
void csr_set_partition(Ctx* ctx, Csr* a, int rows, int nnz);
// Candidate variants: (lanes, rb_rows, rb_nnz); lanes 0 = row blocks with that partition.
struct Variant {
    int lanes, rows, nnz;
};
const std::vector<Variant>& variant_candidates();
void apply_variant(Ctx* ctx, Csr* a, const Variant& v);
// Time every candidate variant of `a` with `launch(k)` (k back-to-back launches of the kernel that will
// use it) and keep the fastest.
void autotune_with(Ctx* ctx, Csr* a, int reps, const std::function<void(int)>& launch);
// The same with the stand-alone SpMV in (mdt, vdt).
void autotune_csr(Ctx* ctx, Csr* a, int mdt, int vdt, int reps);
uint16_t bf16_bits_from_double(double d);

// ---------------------------------------------------------------- dispatch
template <class T>
struct Tag {
    using type = T;
};

inline size_t dtype_size(int dt) { return dt == GANIT_F64 ? 8 : dt == GANIT_F32 ? 4 : 2; }

template <class F>
void with_vdt(int vdt, F&& f) {
    if (vdt == GANIT_F64) f(Tag<double>{});
    else if (vdt == GANIT_F32) f(Tag<float>{});
    else throw Error(GANIT_EINVAL, "vector dtype must be GANIT_F64 or GANIT_F32");
}

template <class F>
void with_mdt(int mdt, F&& f) {
    if (mdt == GANIT_F64) f(Tag<double>{});
    else if (mdt == GANIT_F32) f(Tag<float>{});
    else if (mdt == GANIT_BF16) f(Tag<bf16_t>{});
    else throw Error(GANIT_EINVAL, "matrix dtype must be GANIT_F64, GANIT_F32 or GANIT_BF16");
}

// Dispatch on the row-group width; 0 (the LDS-staged variant) must be handled by the caller first.
template <class F>
void with_lanes(int lanes, F&& f) {
    switch (lanes) {
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 4: f(std::integral_constant<int, 4>{}); break;
        case 8: f(std::integral_constant<int, 8>{}); break;
        case 16: f(std::integral_constant<int, 16>{}); break;
        case 32: f(std::integral_constant<int, 32>{}); break;
        case 64: f(std::integral_constant<int, 64>{}); break;
        default: throw Error(GANIT_EINVAL, "lanes per row must be 0 (row blocks) or a power of two in 1..64");
    }
}

inline unsigned grid_for(int64_t threads) {
    int64_t g = (threads + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    GH_REQUIRE(g <= 0x7fffffffLL, "launch too large");
    return (unsigned)g;
}

}  // namespace ganit_hip
