// Drop-in for the reference's src/backend_cuda.cu: defines ganit::solve_pdhg_gpu (reference
// include/ganit/pdhg.hpp:45-47) by instantiating the reference's OWN driver, run_pdhg<B1, B2, Bd>
// (src/pdhg_impl.hpp:243-328), with HipBackend over libganit_hip.so -- exactly as backend_cuda.cu:382-393
// instantiates it with CudaBackend.  A maintainer adds this file (and backend_hip.hpp) to the reference's
// build in place of backend_cuda.cu and defines GANIT_HAVE_CUDA; `ganit model.mps --device gpu` then runs
// unchanged on an MI355X.  This file is compiled only against the reference's headers.
#include <string>

#include "backend_hip.hpp"
#include "pdhg_impl.hpp"

namespace ganit {

Result solve_pdhg_gpu(const LP& lp, const Options& opt) {
    char name[256] = {0};
    int cus = 0;
    ganit_hipbe::check(ganit_hip_ctx_device_info(ganit_hipbe::context(), name, sizeof(name), &cus));
    const std::string dev = std::string("gpu:") + name;
    return run_pdhg<ganit_hipbe::HipBackend<float, float>, ganit_hipbe::HipBackend<float, double>,
                    ganit_hipbe::HipBackend<double, double>>(lp, opt, dev.c_str());
}

}  // namespace ganit
