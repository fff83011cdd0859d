// ganit -- command-line driver for the MI355X PDHG engine.  Flags, progress output, --sol file, the --json
// summary line and the exit codes follow the reference's src/main.cpp:41-211 for the PDHG path; keys after
// "nnz" in the JSON line are additions.
//   ganit model.mps [--device gpu] [--tol 1e-4] [--precision double|mixed|dynamic] [--time S] [--iters N]
//                   [--values f32|bf16|bf16f32] [--eval level|fp64] [--kahan] [--detect-unbounded]
//                   [--setup auto|host|gpu] [--polish] [--scale-q] [--gpu K] [--sol FILE] [--json] [--quiet]
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>

#include "lp.hpp"
#include "pdhg.hpp"

using namespace ganit;

static void usage() {
    std::printf(
        "GANIT LP/QP solver (restarted PDHG on AMD Instinct MI355X)\n"
        "usage: ganit <model.mps[.gz]> [options]\n"
        "  --device gpu       compute device (gpu only in this build)\n"
        "  --tol X            relative KKT tolerance (default 1e-4)\n"
        "  --precision double|mixed|dynamic  PDHG arithmetic; mixed: FP32->FP64,\n"
        "                     dynamic: FP32 -> FP32 matrix/FP64 vectors -> FP64 (always FP64-certified)\n"
        "  --values f32|bf16|bf16f32  matrix values of the reduced-precision levels (default f32; bf16f32:\n"
        "                     BF16 until --stall-checks checks without progress, then FP32 until --stall-checks-l2)\n"
        "  --eval level|fp64  KKT checks of the reduced-precision levels at their own precision (default)\n"
        "                     or in FP64 with the FP64 matrix\n"
        "  --kahan            L1 checks: Kahan-compensated FP32 sums (instead of FP64 sums)\n"
        "  --detect-unbounded LPs: also test for a dual-infeasibility (unboundedness) ray\n"
        "  --polish           LPs: project an optimal PDHG point onto its active face and keep it when its\n"
        "                     FP64 KKT errors are all <= 1e-9 (a vertex-accurate objective)\n"
        "  --scale-q          QPs: equilibrate Q together with A in the column scaling\n"
        "  --setup auto|host|gpu  where scaling and the norm estimate run (same results; auto: the GPU\n"
        "                     from a million nonzeros)\n"
        "  --time S           time limit in seconds (default 3600)\n"
        "  --iters N          iteration limit\n"
        "  --stall-checks K   checks without progress before leaving L1 (default 12)\n"
        "  --stall-checks-l2 K  same for L2 (default 24)\n"
        "  --gpu K            HIP device ordinal (default 0)\n"
        "  --sol FILE         write primal solution\n"
        "  --relax            solve the LP relaxation of a model with integer markers\n"
        "  --json             print one JSON summary line at the end\n"
        "  --quiet            no progress log\n");
}

int main(int argc, char** argv) {
    if (argc < 2) { usage(); return 1; }
    std::string path, device = "gpu", solfile, method = "pdhg";
    Options opt;
    bool json = false, relax = false;
    for (int a = 1; a < argc; ++a) {
        const std::string s = argv[a];
        auto next = [&]() -> std::string {
            if (a + 1 >= argc) { std::fprintf(stderr, "missing value for %s\n", s.c_str()); std::exit(1); }
            return argv[++a];
        };
        if (s == "--device") device = next();
        else if (s == "--method") method = next();
        else if (s == "--relax") relax = true;
        else if (s == "--tol") opt.tol = std::stod(next());
        else if (s == "--stall-checks") opt.stall_checks = std::stoi(next());
        else if (s == "--stall-checks-l2") opt.stall_checks_l2 = std::stoi(next());
        else if (s == "--precision") {
            const std::string p = next();
            if (p == "mixed") opt.precision = Precision::Mixed;
            else if (p == "double") opt.precision = Precision::Double;
            else if (p == "dynamic") opt.precision = Precision::Dynamic;
            else { std::fprintf(stderr, "--precision must be double, mixed or dynamic\n"); return 1; }
        } else if (s == "--values") {
            const std::string v = next();
            if (v == "f32") opt.values = Values::F32;
            else if (v == "bf16") opt.values = Values::BF16;
            else if (v == "bf16f32") opt.values = Values::BF16_F32;
            else { std::fprintf(stderr, "--values must be f32, bf16 or bf16f32\n"); return 1; }
        }
        else if (s == "--eval") {
            const std::string v = next();
            if (v == "level") opt.eval_fp64 = false;
            else if (v == "fp64") opt.eval_fp64 = true;
            else { std::fprintf(stderr, "--eval must be level or fp64\n"); return 1; }
        }
        else if (s == "--kahan") opt.kahan = true;
        else if (s == "--detect-unbounded") opt.detect_unbounded = true;
        else if (s == "--polish") opt.polish = true;
        else if (s == "--scale-q") opt.scale_q = true;
        else if (s == "--setup") {
            const std::string v = next();
            if (v == "auto") opt.setup = Setup::Auto;
            else if (v == "host") opt.setup = Setup::Host;
            else if (v == "gpu") opt.setup = Setup::Device;
            else { std::fprintf(stderr, "--setup must be auto, host or gpu\n"); return 1; }
        }
        else if (s == "--time") opt.time_limit = std::stod(next());
        else if (s == "--iters") opt.max_iter = std::stol(next());
        else if (s == "--gpu") opt.device = std::stoi(next());
        else if (s == "--sol") solfile = next();
        else if (s == "--json") json = true;
        else if (s == "--quiet") opt.verbose = 0;
        else if (s == "-h" || s == "--help") { usage(); return 0; }
        else if (!s.empty() && s[0] == '-') { std::fprintf(stderr, "unknown option %s\n", s.c_str()); return 1; }
        else path = s;
    }
    if (path.empty()) { usage(); return 1; }
    if (method != "pdhg") {
        std::fprintf(stderr, "--method %s is not part of this build (restarted PDHG only)\n", method.c_str());
        return 1;
    }
    if (device != "gpu") {
        std::fprintf(stderr, "--device %s is not part of this build: it runs restarted PDHG on the GPU only\n",
                     device.c_str());
        return 1;
    }
    try {
        const auto t0 = std::chrono::steady_clock::now();
        LP lp = read_mps(path);
        const double read_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (opt.verbose) {
            std::printf("Read %s: %d rows, %d cols, %lld nnz in %.2fs\n", lp.name.c_str(), lp.m, lp.n,
                        (long long)lp.A.nnz(), read_s);
            if (lp.num_int() > 0 && relax)
                std::printf("Note: %d integer variables relaxed (LP relaxation solved)\n", lp.num_int());
        }
        if (lp.has_q() && opt.verbose)
            std::printf("Quadratic objective: %lld Hessian nonzeros (convex QP via PDHG)\n", (long long)lp.Q.nnz());
        if (lp.num_int() > 0 && !relax) {
            std::fprintf(stderr, "integer variables present: this build solves LP/QP relaxations only; use --relax\n");
            return 1;
        }
        const Result r = solve_pdhg_gpu(lp, opt);
        if (!solfile.empty()) {
            std::ofstream out(solfile);
            out.precision(17);
            out << "# status " << to_string(r.status) << "\n# objective " << r.pobj << "\n";
            for (int j = 0; j < lp.n; ++j) out << lp.col_names[j] << " " << r.x[j] << "\n";
        }
        if (json) {
            std::printf(
                "{\"model\":\"%s\",\"device\":\"%s\",\"problem\":\"%s\",\"status\":\"%s\",\"pobj\":%.12e,"
                "\"dobj\":%.12e,\"rel_p\":%.3e,\"rel_d\":%.3e,\"rel_gap\":%.3e,\"iters\":%ld,"
                "\"restarts\":%d,\"fp32_iters\":%ld,\"l2_iters\":%ld,\"read_s\":%.4f,\"setup_s\":%.4f,\"solve_s\":%.4f,"
                "\"rows\":%d,\"cols\":%d,\"nnz\":%lld,\"norm_a\":%.6e,\"upload_s\":%.4f,\"checks\":%ld}\n",
                lp.name.c_str(), r.device.c_str(), lp.has_q() ? "qp" : "lp", to_string(r.status), r.pobj, r.dobj,
                r.rel_primal, r.rel_dual, r.rel_gap, r.iterations, r.restarts, r.fp32_iterations, r.l2_iterations,
                read_s, r.setup_seconds, r.solve_seconds, lp.m, lp.n, (long long)lp.A.nnz(), r.norm_a,
                r.upload_seconds, r.checks);
        }
        return r.status == Status::Optimal ? 0 : 2;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
