// Problem scaling for PDHG (mirror of the reference's src/scaling.hpp:8-23 and scaling.cpp:8-155).
//
// Scaled problem: A' = diag(R) A diag(C), x = C x' * bscale, y = R y' * cscale; always a minimisation.
// The arithmetic is the reference's, operation for operation, so the scaled problem is bit-identical to
// the reference's; the loops that can run in parallel without changing any rounding (row / column
// maxima, per-row and per-column sums in their sequential order, row-wise SpMV) use OpenMP.
#pragma once
#include "lp.hpp"

struct ganit_hip_ctx;  // include/ganit_hip.h

namespace ganit {

// The problem PDHG iterates on.  Field names and layout are the reference's ScaledLP, which its driver reads.
struct ScaledLP {
    int m = 0, n = 0;
    Csr A, AT;                            // A' and its stored transpose
    std::vector<double> c, l, u, lo, hi;  // objective, variable bounds and row bounds after scaling
    std::vector<double> R, C;             // accumulated Ruiz (and Pock-Chambolle) factors of rows / columns
    double bscale = 1.0, cscale = 1.0;    // one factor for all bounds, one for the objective
    std::vector<double> Rres, Cres;       // R / bscale and C / cscale: residuals back in original units
    Csr Q;                                // sign * C Q C * bscale / cscale
    double normQ = 0.0;                   // ||Q'||_2 estimate (60 power steps, x1.01)
    double obj_const = 0.0;
    double bnorm = 0.0, cnorm = 0.0;      // ||row bounds||, ||c|| of the problem as given (before scaling)
    bool maximize = false;
    bool has_q() const { return Q.nnz() > 0; }
};

ScaledLP scale_lp(const LP& lp, int ruiz_iters, bool pock_chambolle, bool bound_obj_rescale = true,
                  bool q_columns = false);
// q_columns (extension, QPs; the reference scales columns by A alone): each Ruiz and Pock-Chambolle column factor
// also takes the column of the current C Q C into account (the max, resp. the 1-norm, over A's and Q's column), the
// equilibration of the block matrix [Q A'; A 0] -- it keeps C Q C's diagonal from spreading over orders of
// magnitude when A's column scales are uneven.
// scale_lp's second half (everything after the matrix and its factors R, C): c, l, u, lo, hi, the bound and
// objective rescaling, Rres, Cres, Q and ||Q'||, bnorm, cnorm.
void scale_vectors(const LP& lp, bool bound_obj_rescale, ScaledLP& s);
// scale_lp with the matrix half on a GPU (Ruiz / Pock-Chambolle passes and the transpose); with norm_a, also
// estimate_norm(result, power_iters) with the power iteration's SpMVs on the GPU.  Bitwise the same results.
ScaledLP scale_lp_device(const LP& lp, int ruiz_iters, bool pock_chambolle, ::ganit_hip_ctx* ctx,
                         double* norm_a, int power_iters = 50, bool bound_obj_rescale = true);
// ||A'||_2 by power iteration on A'^T A' (50 steps, x1.01).
double estimate_norm(const ScaledLP& s, int iters = 50);
double norm2(const std::vector<double>& v);
double finite_bound_norm(const std::vector<double>& lo, const std::vector<double>& hi);

}  // namespace ganit
