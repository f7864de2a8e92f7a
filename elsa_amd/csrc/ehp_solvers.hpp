// ehp_solvers.hpp -- the device-resident solvers over one operator or a batch of operators.
#pragma once

#include "ehp_kernels.hpp"

#include <functional>
#include <vector>

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
        // per problem: number of volume axes and their sizes (dim, n0, n1, n2), for regularisers
        // defined on the volume grid
        std::vector<int> shapes;
    };

    // Forward finite differences over the packed volumes of nseg problems (ehp_fd.hip): segment
    // j maps its volume [vol_off, vol_off + nvox) to dim blocks of nvox at reg_off.
    struct FdSeg {
        int64_t vol_off, reg_off, nvox, blk0;
        int n[3];
        int dim;
    };
    struct FdPlan {
        FdSeg* segs = nullptr; // device
        int nseg = 0;
        int64_t blocks = 0, nreg = 0;
        std::vector<int64_t> reg_off; // host, nseg + 1
    };
    // shapes as in SolveOps; vol_off: nseg + 1 offsets. EHP_* status.
    int fd_plan_create(const int* shapes, const int64_t* vol_off, int nseg, FdPlan& plan);
    void fd_plan_destroy(FdPlan& plan);
    template <class T>
    hipError_t fd_apply(const FdPlan& plan, const T* in, T* out, bool adjoint, hipStream_t s);

    int solve_cgls(const SolveOps& A, int dtype, const void* b, void* x, int use_x0, int iterations, double eps,
                   double tol, bool graph, double* residuals, int* iters_done, hipStream_t s);
    int solve_apgd(const SolveOps& A, int dtype, const void* b, void* x, int use_x0, int iterations, const double* mu,
                   double sigma, double epsilon, bool graph, double* grad_norms, int* iters_done, double* mu_used,
                   hipStream_t s);
    int solve_power_iterations(const SolveOps& A, int dtype, int niters, double* L, hipStream_t s);
    // ADMML2 with the identity (fd == false) or forward finite differences (fd == true) as the
    // regularisation operator; tau <= 0: elsa's default 0.9 / powerIterations(A^T A) (per
    // problem, so a batch of several problems needs tau > 0 with finite differences); tau_used
    // receives the tau of problem 0
    int solve_admml2(const SolveOps& A, int dtype, const void* b, void* x, int use_x0, int iterations, double tau,
                     double sigma, int ninner, bool fd, bool graph, double* residuals, double* tau_used,
                     hipStream_t s);
    // CGLS with iterative refinement: FP32 operator Ahi, bfloat16 operator Alo, one problem
    int solve_cgls_refined(const SolveOps& Ahi, const SolveOps& Alo, const float* b, float* x, int use_x0, int outer,
                           int inner, bool graph, double* residuals, hipStream_t s);
} // namespace ehp
