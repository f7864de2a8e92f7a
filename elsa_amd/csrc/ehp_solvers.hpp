// ehp_solvers.hpp -- the device-resident solvers over one operator or a batch of operators.
#pragma once

#include "ehp_kernels.hpp"

#include <functional>

namespace ehp
{
    // A block-diagonal operator for the solvers: nseg independent problems whose domain and
    // range vectors are packed back to back (problem j owns elements [doff[j], doff[j+1]) of
    // the domain and [roff[j], roff[j+1]) of the range). fwd / adj apply every problem at once
    // on stream s and return an EHP_* status.
    struct SolveOps {
        int nseg = 1;
        const int64_t* doff = nullptr;
        const int64_t* roff = nullptr;
        std::function<int(const void*, void*, hipStream_t)> fwd, adj;
    };

    int solve_cgls(const SolveOps& A, int dtype, const void* b, void* x, int use_x0, int iterations, double eps,
                   double tol, bool graph, double* residuals, int* iters_done, hipStream_t s);
    int solve_apgd(const SolveOps& A, int dtype, const void* b, void* x, int use_x0, int iterations, const double* mu,
                   double sigma, double epsilon, bool graph, double* grad_norms, int* iters_done, double* mu_used,
                   hipStream_t s);
    int solve_power_iterations(const SolveOps& A, int dtype, int niters, double* L, hipStream_t s);
    // ADMML2 (regularisation operator: the identity); tau > 0
    int solve_admml2(const SolveOps& A, int dtype, const void* b, void* x, int use_x0, int iterations, double tau,
                     double sigma, int ninner, bool graph, double* residuals, hipStream_t s);
    // CGLS with iterative refinement: FP32 operator Ahi, bfloat16 operator Alo, one problem
    int solve_cgls_refined(const SolveOps& Ahi, const SolveOps& Alo, const float* b, float* x, int use_x0, int outer,
                           int inner, bool graph, double* residuals, hipStream_t s);
} // namespace ehp
