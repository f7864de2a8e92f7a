// ehp_kernels.hpp -- launchers of the projector kernels (all asynchronous on stream s).
#pragma once

#include "ehp_device.hpp"
#include "ehp_joseph_tile.hpp"
#include "ehp_siddon.hpp"
#include "ehp_voxel.hpp"

struct ehp_op;

namespace ehp
{
    template <class T>
    hipError_t launch_fill(T* p, int64_t n, T value, hipStream_t s);

    template <class T>
    hipError_t joseph_forward(const PoseGeom* geo, const Dims& dims, const T* vol, T* sino, hipStream_t s);
    template <class T>
    hipError_t joseph_adjoint(const PoseGeom* geo, const Dims& dims, const T* sino, T* vol, hipStream_t s);
    hipError_t joseph_info(const PoseGeom* geo, const Dims& dims, int32_t* axis, int64_t* nsteps,
                           unsigned long long* total, hipStream_t s);
    hipError_t joseph_ray_table(const PoseGeom* geo, const Dims& dims, RayRec* table, hipStream_t s);
    hipError_t joseph_pose_axes(const RayRec* table, const Dims& dims, int* pose_axes, hipStream_t s);
    size_t joseph_tile_lds_bytes(const Dims& dims, const TileCfg& tc);
    hipError_t joseph_adjoint_tiled(const PoseGeom* geo, const RayRec* table, const Dims& dims, const int* pose_axes,
                                    int axes_present, const TileCfg& tc, int batch, double slot_bound,
                                    unsigned int* d_maxbits, int* d_fixexp, const float* sino, float* vol,
                                    hipStream_t s);
    hipError_t fixed_point_exponent_launch(const float* y, int64_t n, double slot_bound, unsigned int* d_maxbits,
                                           int* d_fixexp, hipStream_t s);
    hipError_t siddon_ray_table(const PoseGeom* geo, const Dims& dims, SidRec* table, int* pose_axes, hipStream_t s);
    size_t siddon_tile_lds_bytes(const Dims& dims, const TileCfg& tc);
    hipError_t siddon_adjoint_tiled(const PoseGeom* geo, const SidRec* table, const Dims& dims, const int* pose_axes,
                                    int axes_present, const TileCfg& tc, int batch, double slot_bound,
                                    unsigned int* d_maxbits, int* d_fixexp, const float* sino, float* vol,
                                    hipStream_t s);
    size_t voxel_tile_lds_bytes(const TileCfg& tc);
    int voxel_max_lut();
    hipError_t voxel_forward_tiled(const PoseGeom* geo, const Dims& dims, const float* lut, int n_lut, float support,
                                   const int* pose_axis, const float* pose_margin, const TileCfg& tc, float lutscale,
                                   double slot_bound, unsigned int* d_maxbits, int* d_fixexp, const float* vol,
                                   float* sino, hipStream_t s);
    hipError_t joseph_view_copy(const Dims& dims, const PadLayout& L, const float* vol, float* view, hipStream_t s);
    hipError_t joseph_forward_planes(const PoseGeom* geo, const Dims& dims, const PadLayout* layouts,
                                     float* const* views, int axes_present, const float* vol, float* sino,
                                     hipStream_t s);

    template <class T>
    hipError_t siddon_forward(const PoseGeom* geo, const Dims& dims, const T* vol, T* sino, hipStream_t s);
    template <class T>
    hipError_t siddon_adjoint(const PoseGeom* geo, const Dims& dims, const T* sino, T* vol, hipStream_t s);
    hipError_t siddon_count(const PoseGeom* geo, const Dims& dims, unsigned long long* total, hipStream_t s);

    template <class T>
    hipError_t voxel_forward(const PoseGeom* geo, const Dims& dims, const T* lut, int n_lut, float support,
                             const T* vol, T* sino, hipStream_t s);
    template <class T>
    hipError_t voxel_adjoint(const PoseGeom* geo, const Dims& dims, const T* lut, int n_lut, float support,
                             const T* sino, T* vol, hipStream_t s);
    hipError_t voxel_count(const PoseGeom* geo, const Dims& dims, float support, unsigned long long* total,
                           hipStream_t s);

    // device-resident solvers (ehp_solvers.hip)
    int solver_cgls(ehp_op* op, int dtype, int64_t nd, int64_t nr, const void* b, void* x, int use_x0, int iterations,
                    double eps, double tol, bool graph, double* residuals, int* iters_done, hipStream_t s);
    int solver_apgd(ehp_op* op, int dtype, int64_t nd, int64_t nr, const void* b, void* x, int use_x0, int iterations,
                    double mu, double sigma, double epsilon, bool graph, double* grad_norms, int* iters_done,
                    double* mu_used, hipStream_t s);
    int solver_power_iterations(ehp_op* op, int dtype, int64_t nd, int64_t nr, int niters, double* L,
                                hipStream_t s);
} // namespace ehp
