// ehp_fd.hip -- forward finite differences on the device, the regularisation operator of the
// TV-regularised ADMML2.
//
// elsa's FiniteDifferences with DiffType::FORWARD and every dimension active
// (elsa/operators/FiniteDifferences.cpp, applyHelper / applyAdjointHelper): the range holds one
// block of nvox differences per axis, axis 0 first; block d at voxel i is -x[i] + x[i + e_d]
// (-x[i] alone on the last plane of axis d). The adjoint sums, axis by axis in that order,
// + y_d[i - e_d] (when i is not on the first plane) - y_d[i]. (elsa's adjoint reads the block of
// axis 1 of a 3D volume at offset n0 * n1 instead of nvox -- its `_coordDelta` -- which makes it
// no transpose in 3D; this is the transpose.) Many volumes of different shapes, packed back to
// back, go through one launch: one thread per voxel, each block within one volume.
#include "elsa_hip.h"
#include "ehp_solvers.hpp"

#include <vector>

namespace ehp
{
    namespace
    {
        constexpr int kThreads = 256;

        __device__ inline int fd_seg(const FdSeg* __restrict__ segs, int nseg)
        {
            int lo = 0, hi = nseg - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (segs[mid].blk0 <= (int64_t) blockIdx.x)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            return lo;
        }

        template <class T>
        __global__ void __launch_bounds__(kThreads)
            fd_forward_kernel(const FdSeg* __restrict__ segs, int nseg, const T* __restrict__ x, T* __restrict__ y)
        {
            const FdSeg S = segs[fd_seg(segs, nseg)];
            const int64_t id = ((int64_t) blockIdx.x - S.blk0) * kThreads + threadIdx.x;
            if (id >= S.nvox)
                return;
            const int64_t plane = (int64_t) S.n[0] * S.n[1];
            const int c[3] = {(int) (id % S.n[0]), (int) ((id / S.n[0]) % S.n[1]), (int) (id / plane)};
            const int64_t stride[3] = {1, S.n[0], plane};
            const T* xs = x + S.vol_off;
            for (int d = 0; d < S.dim; ++d) {
                T v = -xs[id];
                if (c[d] < S.n[d] - 1)
                    v += xs[id + stride[d]];
                y[S.reg_off + d * S.nvox + id] = v;
            }
        }

        template <class T>
        __global__ void __launch_bounds__(kThreads)
            fd_adjoint_kernel(const FdSeg* __restrict__ segs, int nseg, const T* __restrict__ y, T* __restrict__ x)
        {
            const FdSeg S = segs[fd_seg(segs, nseg)];
            const int64_t id = ((int64_t) blockIdx.x - S.blk0) * kThreads + threadIdx.x;
            if (id >= S.nvox)
                return;
            const int64_t plane = (int64_t) S.n[0] * S.n[1];
            const int c[3] = {(int) (id % S.n[0]), (int) ((id / S.n[0]) % S.n[1]), (int) (id / plane)};
            const int64_t stride[3] = {1, S.n[0], plane};
            T acc = T(0);
            for (int d = 0; d < S.dim; ++d) {
                const T* yd = y + S.reg_off + d * S.nvox;
                if (c[d] > 0)
                    acc += yd[id - stride[d]];
                acc -= yd[id];
            }
            x[S.vol_off + id] = acc;
        }
    } // namespace

    int fd_plan_create(const int* shapes, const int64_t* vol_off, int nseg, FdPlan& plan)
    {
        std::vector<FdSeg> h((size_t) nseg);
        int64_t reg = 0, blk = 0;
        for (int j = 0; j < nseg; ++j) {
            FdSeg& S = h[(size_t) j];
            S.dim = shapes[4 * j];
            if (S.dim < 1 || S.dim > 3)
                return EHP_ERR_ARGUMENT;
            S.nvox = 1;
            for (int i = 0; i < 3; ++i) {
                S.n[i] = i < S.dim ? shapes[4 * j + 1 + i] : 1;
                if (S.n[i] < 1)
                    return EHP_ERR_ARGUMENT;
                S.nvox *= S.n[i];
            }
            if (vol_off[j + 1] - vol_off[j] != S.nvox)
                return EHP_ERR_ARGUMENT;
            S.vol_off = vol_off[j];
            S.reg_off = reg;
            S.blk0 = blk;
            reg += (int64_t) S.dim * S.nvox;
            blk += (S.nvox + kThreads - 1) / kThreads;
        }
        if (blk > 0x7fffffff)
            return EHP_ERR_ARGUMENT;
        if (hipMalloc((void**) &plan.segs, (size_t) nseg * sizeof(FdSeg)) != hipSuccess)
            return EHP_ERR_ALLOC;
        if (hipMemcpy(plan.segs, h.data(), (size_t) nseg * sizeof(FdSeg), hipMemcpyHostToDevice) != hipSuccess) {
            (void) hipFree(plan.segs);
            plan.segs = nullptr;
            return EHP_ERR_HIP;
        }
        plan.nseg = nseg;
        plan.blocks = blk;
        plan.nreg = reg;
        plan.reg_off.resize((size_t) nseg + 1);
        for (int j = 0; j < nseg; ++j)
            plan.reg_off[(size_t) j] = h[(size_t) j].reg_off;
        plan.reg_off[(size_t) nseg] = reg;
        return EHP_OK;
    }

    void fd_plan_destroy(FdPlan& plan)
    {
        if (plan.segs)
            (void) hipFree(plan.segs);
        plan.segs = nullptr;
    }

    template <class T>
    hipError_t fd_apply(const FdPlan& plan, const T* in, T* out, bool adjoint, hipStream_t s)
    {
        if (plan.blocks == 0)
            return hipSuccess;
        if (adjoint)
            hipLaunchKernelGGL((fd_adjoint_kernel<T>), dim3((unsigned) plan.blocks), dim3(kThreads), 0, s, plan.segs,
                               plan.nseg, in, out);
        else
            hipLaunchKernelGGL((fd_forward_kernel<T>), dim3((unsigned) plan.blocks), dim3(kThreads), 0, s, plan.segs,
                               plan.nseg, in, out);
        return hipGetLastError();
    }

    template hipError_t fd_apply<float>(const FdPlan&, const float*, float*, bool, hipStream_t);
    template hipError_t fd_apply<double>(const FdPlan&, const double*, double*, bool, hipStream_t);
} // namespace ehp
