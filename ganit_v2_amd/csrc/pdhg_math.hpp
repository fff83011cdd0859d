// Element formulas of one PDHG iteration, shared by the stand-alone step kernels and the fused ones.
#pragma once
#include "device_common.hpp"

namespace ganit_hip {

// The reference's element formulas (src/backend_cpu.cpp:72-79 and :85-91), evaluated without FMA
// contraction so that, fed the same operands, they reproduce the CPU bits.
template <class V>
__device__ __forceinline__ void primal_update(V xj, V aty, V qx, V cj, V lj, V uj, V tau, V& xn, V& xb, V& sum) {
#pragma clang fp contract(off)
    const V g = cj - aty + qx;
    V v = xj - tau * g;
    v = (v < lj) ? lj : v;
    v = (uj < v) ? uj : v;
    xn = v;
    xb = V(2) * v - xj;
    sum = sum + v;
}

template <class V>
__device__ __forceinline__ void dual_update(V yi, V axb, V loi, V hii, V sigma, V& yn, V& sum) {
#pragma clang fp contract(off)
    const V q = yi - sigma * axb;
    const V a = sigma * loi, b = sigma * hii;
    V t = -q;
    t = (t < a) ? a : t;
    t = (b < t) ? b : t;
    const V v = q + t;
    yn = v;
    sum = sum + v;
}

}  // namespace ganit_hip
