// Device-resident state of one PDHG problem (the ganit_hip_pdhg handle), shared by the translation units
// that drive it: the solver kernels (pdhg_device.hip) and the adjoint (adjoint.hip).
#pragma once
#include <map>
#include <tuple>
#include <vector>

#include "internal.hpp"

namespace ganit_hip {

struct VecSet {
    bool on = false;
    int vdt = 0;
    void *c = nullptr, *l = nullptr, *u = nullptr, *lo = nullptr, *hi = nullptr, *rres = nullptr, *cres = nullptr;
    void *X[2] = {nullptr, nullptr}, *Y[2] = {nullptr, nullptr};
    void *xbar = nullptr, *sumx = nullptr, *sumy = nullptr, *xa = nullptr, *ya = nullptr, *xr = nullptr,
         *yr = nullptr;
    int cur = 0;
};

struct Pdhg {
    Ctx* ctx = nullptr;
    int m = 0, n = 0;
    bool hasq = false;
    // every column's lower (upper) bound is the same bits: the primal iteration kernel takes the value, not
    // the vector
    bool l_uniform = false, u_uniform = false;
    double l_value = 0.0, u_value = 0.0;
    Csr *A = nullptr, *AT = nullptr, *Q = nullptr;
    VecSet vs[2];
    double* prm = nullptr;
    double* part_n = nullptr;
    double* part_m = nullptr;
    size_t part_cap = 0;          // blocks the partial buffers hold
    double* check_host = nullptr;  // [GANIT_CHECK_NOUT] check sums, host-coherent pinned memory
    std::vector<void*> owned;
    std::map<std::tuple<int, int, int, int>, hipGraphExec_t> graphs;
    bool use_graphs = true;

    void* alloc(size_t bytes) {
        void* p = nullptr;
        GH_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 16)));
        owned.push_back(p);
        return p;
    }
    VecSet& set(int vdt) {
        GH_REQUIRE(vdt == GANIT_F64 || vdt == GANIT_F32, "vector dtype must be GANIT_F64 or GANIT_F32");
        GH_REQUIRE(vs[vdt].on, "vector dtype not resident in this problem (vdt_mask)");
        return vs[vdt];
    }
    ~Pdhg() {
        if (ctx) (void)hipStreamSynchronize(ctx->stream);
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        for (void* p : owned) (void)hipFree(p);
        if (check_host) (void)hipHostFree(check_host);
        csr_destroy(A);
        csr_destroy(AT);
        csr_destroy(Q);
    }
};


void upload_as(Ctx* ctx, int vdt, const double* h, int64_t n, void* d);
void download_as(Ctx* ctx, int vdt, const void* d, int64_t n, double* h);

}  // namespace ganit_hip
