// HipBackend<M, V>: the reference's compile-time backend concept implemented over the C ABI of
// libganit_hip.so, so the reference's own driver (src/pdhg_impl.hpp: pdhg_phase<B>, pdhg_evaluate<B>,
// run_pdhg<B1,B2,Bd>) can run on an MI355X without modification.
//
// Member for member it replaces CudaBackend<M, V> (src/backend_cuda.cu:286-378); the concept's members are
// listed in the order the reference's CPU backend declares them (src/backend_cpu.cpp:39-156).  Every method
// is a thin call into an extern "C" entry point of include/ganit_hip.h; errors become std::runtime_error,
// as GANIT_CUDA_CHECK does in the reference (src/backend_cuda.cu:22-29).
//
// Header-only and templated on the scaled-problem type, so it compiles against the reference's ScaledLP
// (src/scaling.hpp:8-23) as well as this library's mirror of it.
#pragma once
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ganit_hip.h"

namespace ganit_hipbe {

inline void check(int rc) {
    if (rc != GANIT_OK) throw std::runtime_error(std::string("HIP backend: ") + ganit_hip_last_error());
}

// One context (device + stream) per process, created on first use; GANIT_HIP_DEVICE picks the device.
inline ganit_hip_ctx* context() {
    static std::unique_ptr<ganit_hip_ctx, int (*)(ganit_hip_ctx*)> ctx = [] {
        const char* d = std::getenv("GANIT_HIP_DEVICE");
        ganit_hip_ctx* c = nullptr;
        check(ganit_hip_ctx_create(d ? std::atoi(d) : 0, &c));
        return std::unique_ptr<ganit_hip_ctx, int (*)(ganit_hip_ctx*)>(c, ganit_hip_ctx_destroy);
    }();
    return ctx.get();
}

template <class T>
constexpr int dtype_of() {
    return sizeof(T) == 8 ? GANIT_F64 : GANIT_F32;
}

// Move-only device vector (the reference's DevBuf<V>, src/backend_cuda.cu:44-74).
template <class T>
class DevVec {
   public:
    DevVec() = default;
    explicit DevVec(size_t n) : n_(n) {
        if (n_) check(ganit_hip_alloc(context(), n_ * sizeof(T), &p_));
    }
    ~DevVec() {
        if (p_) ganit_hip_free(context(), p_);
    }
    DevVec(const DevVec&) = delete;
    DevVec& operator=(const DevVec&) = delete;
    DevVec(DevVec&& o) noexcept : p_(o.p_), n_(o.n_) {
        o.p_ = nullptr;
        o.n_ = 0;
    }
    DevVec& operator=(DevVec&& o) noexcept {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
        return *this;
    }
    void* get() const { return p_; }
    size_t size() const { return n_; }

   private:
    void* p_ = nullptr;
    size_t n_ = 0;
};

// A device CSR matrix with values in M.
template <class M>
class DevMat {
   public:
    DevMat() = default;
    template <class Csr>
    explicit DevMat(const Csr& a) : rows_(a.rows) {
        if (a.rows == 0 && a.val.empty()) return;
        ganit_hip_csr* h = nullptr;
        std::vector<int> ptr(a.ptr);
        if (ptr.empty()) ptr.assign(a.rows + 1, 0);
        check(ganit_hip_csr_upload(context(), a.rows, a.cols, ptr.data(), a.idx.data(), a.val.data(),
                                   1u << dtype_of<M>(), &h));
        h_.reset(h);
    }
    ganit_hip_csr* get() const { return h_.get(); }
    int rows() const { return rows_; }

   private:
    struct Free {
        void operator()(ganit_hip_csr* h) const { ganit_hip_csr_free(h); }
    };
    std::unique_ptr<ganit_hip_csr, Free> h_;
    int rows_ = 0;
};

template <class M, class V>
struct HipBackend {
    using Vec = DevVec<V>;
    static constexpr int vdt = dtype_of<V>();
    static constexpr int mdt = dtype_of<M>();
    int m, n;
    DevMat<M> A, AT, Q;
    Vec c, l, u, lo, hi, R, C;

    template <class ScaledLP>
    explicit HipBackend(const ScaledLP& s) : m(s.m), n(s.n), A(s.A), AT(s.AT), Q(s.Q) {
        c = from_host(s.c); l = from_host(s.l); u = from_host(s.u);
        lo = from_host(s.lo); hi = from_host(s.hi);
        R = from_host(s.Rres); C = from_host(s.Cres);
    }

    Vec zeros(int k) {
        Vec v(k);
        fill(v, 0.0);
        return v;
    }
    Vec from_host(const std::vector<double>& h) {
        Vec v(h.size());
        if (!h.empty()) check(ganit_hip_upload(context(), vdt, h.data(), (int64_t)h.size(), v.get()));
        return v;
    }
    std::vector<double> to_host(const Vec& v) {
        std::vector<double> h(v.size());
        if (v.size()) check(ganit_hip_download(context(), vdt, v.get(), (int64_t)v.size(), h.data()));
        return h;
    }
    void copy(const Vec& src, Vec& dst) {
        if (src.size()) check(ganit_hip_vec_copy(context(), vdt, src.get(), dst.get(), (int64_t)src.size()));
    }
    void fill(Vec& v, double a) {
        if (v.size()) check(ganit_hip_vec_fill(context(), vdt, v.get(), (int64_t)v.size(), a));
    }
    void spmv(const DevMat<M>& a, const Vec& x, Vec& y) {
        if (a.get() && a.rows()) check(ganit_hip_spmv(context(), a.get(), mdt, vdt, x.get(), y.get()));
    }
    void spmv_A(const Vec& x, Vec& ax) { spmv(A, x, ax); }
    void spmv_AT(const Vec& y, Vec& aty) { spmv(AT, y, aty); }
    void spmv_Q(const Vec& x, Vec& qx) { spmv(Q, x, qx); }
    void primal_step(const Vec& x, const Vec& aty, const Vec* qx, double tau, Vec& xn, Vec& xbar, Vec& sumx) {
        check(ganit_hip_primal_step(context(), vdt, n, x.get(), aty.get(), qx ? qx->get() : nullptr, c.get(),
                                    l.get(), u.get(), tau, xn.get(), xbar.get(), sumx.get()));
    }
    void dual_step(const Vec& y, const Vec& axbar, double sigma, Vec& yn, Vec& sumy) {
        check(ganit_hip_dual_step(context(), vdt, m, y.get(), axbar.get(), lo.get(), hi.get(), sigma, yn.get(),
                                  sumy.get()));
    }
    void sub(const Vec& a, const Vec& b, Vec& d) {
        if (a.size()) check(ganit_hip_vec_sub(context(), vdt, a.get(), b.get(), d.get(), (int64_t)a.size()));
    }
    void scale_into(const Vec& src, double a, Vec& dst) {
        if (src.size())
            check(ganit_hip_vec_scale_into(context(), vdt, src.get(), a, dst.get(), (int64_t)src.size()));
    }
    double primal_res2(const Vec& ax) {
        double r = 0;
        check(ganit_hip_primal_res2(context(), vdt, m, ax.get(), lo.get(), hi.get(), R.get(), &r));
        return r;
    }
    void dual_res(const Vec& aty, const Vec* qx, double& rd2, double& vobj, bool with_c = true) {
        check(ganit_hip_dual_res(context(), vdt, n, aty.get(), qx ? qx->get() : nullptr, c.get(), l.get(), u.get(),
                                 C.get(), with_c ? 1 : 0, &rd2, &vobj));
    }
    double row_dual_obj(const Vec& y) {
        double r = 0;
        check(ganit_hip_row_dual_obj(context(), vdt, m, y.get(), lo.get(), hi.get(), &r));
        return r;
    }
    double dot(const Vec& a, const Vec& b) {
        double r = 0;
        check(ganit_hip_dot(context(), vdt, (int64_t)a.size(), a.get(), b.get(), &r));
        return r;
    }
    double dot_c(const Vec& x) { return dot(c, x); }
    double diff2(const Vec& a, const Vec& b) {
        double r = 0;
        check(ganit_hip_diff2(context(), vdt, (int64_t)a.size(), a.get(), b.get(), &r));
        return r;
    }
};

}  // namespace ganit_hipbe
