// ehp_kernels.hpp -- launchers of the projector kernels (all asynchronous on stream s).
#pragma once

#include "ehp_device.hpp"
#include "ehp_joseph_tile.hpp"
#include "ehp_siddon.hpp"
#include "ehp_voxel.hpp"

#include <vector>

struct ehp_op;

namespace ehp
{
    // the views of one plane-major forward
    struct ViewSet {
        PadView v[3];
        int fb2;   // bit a: the fast in-plane axis of view a is b2
        int pairs; // bit a: view a is a row-pair view (3D)
    };

    // ---- batches of Joseph operators (FP32 fast kernels): one launch serves many operators.
    // A job is one operator's share of a launch; block b of the launch runs block
    // b - job.block0 of job block_job[b], exactly as that operator's own launch would. Volumes
    // and sinograms of the batch are packed back to back (offsets in elements).

    // what the batch needs of one operator (Joseph: table; Siddon: sidtable)
    struct JosephBatchItem {
        const PoseGeom* geo;
        const RayRec* table;
        const SidRec* sidtable;
        const int* pose_axes;
        float* views[3];
        PadLayout layouts[3];
        Dims dims;
        TileCfg tile;
        int axes_present;
        int batch;
        int64_t vol_off, sino_off;
    };

    struct CopyJob { // the view copy of one volume into one view
        float* view;
        Dims dims;
        PadLayout L;
        int64_t vol_off;
        int64_t block0;
        int gx, gy;
    };

    struct FwdJob { // the forward of one operator
        const PoseGeom* geo;
        ViewSet views;
        Dims dims;
        int tiles_u;
        int64_t sino_off;
        int64_t block0;
    };

    struct AdjJob { // the tiles of one operator along one driving axis
        const PoseGeom* geo;
        const RayRec* table;
        const int* pose_axes;
        Dims dims;
        TileCfg tc;
        int ntA, ntB;
        int batch;
        int item; // index of the operator in the batch (its fixed-point exponent)
        int64_t sino_off, vol_off;
        int64_t block0;
    };

    struct SegJob { // one operator's segment of a packed vector
        int64_t off, n;
        int64_t block0;
        int nblocks;
    };

    void joseph_forward_batch_jobs(const JosephBatchItem* items, int n, std::vector<CopyJob>& cjobs,
                                   std::vector<int>& cblocks, std::vector<FwdJob>& fjobs, std::vector<int>& fblocks);
    hipError_t joseph_forward_batch(int dim, const CopyJob* cjobs, const int* cblocks, int64_t ncblocks,
                                    const FwdJob* fjobs, const int* fblocks, int64_t nfblocks, const float* vols,
                                    float* sinos, hipStream_t s);
    void joseph_adjoint_batch_jobs(const JosephBatchItem* items, int n, std::vector<SegJob>& segs,
                                   std::vector<int>& sblocks, std::vector<AdjJob> ajobs[3],
                                   std::vector<int> ablocks[3], size_t lds[3]);
    hipError_t joseph_adjoint_batch(int dim, int nitems, const SegJob* segs, const int* sblocks, int64_t nsblocks,
                                    const double* bounds, const AdjJob* const ajobs[3], const int* const ablocks[3],
                                    const int64_t nablocks[3], const size_t lds[3], int64_t nvol_total,
                                    unsigned int* d_maxbits, int* d_fixexp, const float* sinos, float* vols,
                                    hipStream_t s);

    // the view copies of a batched forward (one launch) and the fixed-point exponents of a batched
    // back projection (absmax over every operator's segment, then one exponent per operator)
    int64_t view_copy_blocks(const Dims& dims, const PadLayout& L, int& gx, int& gy);
    hipError_t view_copy_batch(const CopyJob* cjobs, const int* cblocks, int64_t ncblocks, const float* vols,
                               hipStream_t s);
    hipError_t fixed_point_exponent_batch(int nitems, const SegJob* segs, const int* sblocks, int64_t nsblocks,
                                          const double* bounds, unsigned int* d_maxbits, int* d_fixexp,
                                          const float* y, hipStream_t s);
    void batch_segments(const int64_t* offs, const int64_t* sizes, int n, std::vector<SegJob>& segs,
                        std::vector<int>& sblocks);

    // ---- batches of Siddon operators (FP32 fast kernels), on the same scheme
    struct SidFwdJob { // the forward of one operator
        const SidRec* table;
        ViewSet views;
        Dims dims;
        int tiles_u;
        int64_t sino_off;
        int64_t block0;
    };

    struct SidAdjJob { // the tiles of one operator along one driving axis
        const PoseGeom* geo;
        const SidRec* table;
        const int* pose_axes;
        Dims dims;
        TileCfg tc;
        int ntA, ntB;
        int batch;
        int item;
        int64_t sino_off, vol_off;
        int64_t block0;
    };

    void siddon_forward_batch_jobs(const JosephBatchItem* items, int n, std::vector<CopyJob>& cjobs,
                                   std::vector<int>& cblocks, std::vector<SidFwdJob>& fjobs, std::vector<int>& fblocks);
    hipError_t siddon_forward_batch(int dim, const CopyJob* cjobs, const int* cblocks, int64_t ncblocks,
                                    const SidFwdJob* fjobs, const int* fblocks, int64_t nfblocks, const float* vols,
                                    float* sinos, hipStream_t s);
    void siddon_adjoint_batch_jobs(const JosephBatchItem* items, int n, std::vector<SidAdjJob> ajobs[3],
                                   std::vector<int> ablocks[3], size_t lds[3]);
    hipError_t siddon_adjoint_batch(int dim, const SidAdjJob* const ajobs[3], const int* const ablocks[3],
                                    const int64_t nablocks[3], const size_t lds[3], const int* d_fixexp,
                                    const float* sinos, float* vols, hipStream_t s);

    // ---- batches of voxel projectors (FP32 fast kernels): the tiled forward and the gather
    struct VoxBatchItem {
        const PoseGeom* geo;
        Dims dims;
        const float* lut;
        int n_lut;
        float support;
        int diff;
        const int* pose_axes;
        const float* pose_margin;
        TileCfg tile;
        int64_t vol_off, sino_off;
    };

    struct VoxFwdJob { // the tile-pose blocks of one operator's forward
        const PoseGeom* geo;
        Dims dims;
        const float* lut;
        int n_lut;
        float support;
        int diff;
        const int* pose_axis;
        const float* pose_margin;
        int TU, TV, ntu;
        int item;
        int64_t vol_off, sino_off;
        int64_t block0;
    };

    struct VoxAdjJob { // the voxel blocks of one operator's gather
        const PoseGeom* geo;
        Dims dims;
        const float* lut;
        int n_lut;
        float support;
        int diff;
        int64_t vol_off, sino_off;
        int64_t block0;
    };

    void voxel_forward_batch_jobs(const VoxBatchItem* items, int n, std::vector<VoxFwdJob>& jobs,
                                  std::vector<int>& blocks, size_t& lds);
    hipError_t voxel_forward_batch(int dim, const VoxFwdJob* jobs, const int* blocks, int64_t nblocks, size_t lds,
                                   const int* d_fixexp, const float* vols, float* sinos, hipStream_t s);
    void voxel_adjoint_batch_jobs(const VoxBatchItem* items, int n, std::vector<VoxAdjJob>& jobs,
                                  std::vector<int>& blocks);
    hipError_t voxel_adjoint_batch(int dim, const VoxAdjJob* jobs, const int* blocks, int64_t nblocks,
                                   const float* sinos, float* vols, hipStream_t s);

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
    // dynamic LDS of a tiled back projection with pose batches of `batch` (the largest by default)
    size_t joseph_tile_lds_bytes(const Dims& dims, const TileCfg& tc, int batch = kMaxPoseBatch);
    // ST: the sinogram's storage type; the volume is FP64 for FP64 sinograms, else FP32.
    // d_maxbits: kAbsmaxBlocks 64-bit words.
    template <class ST>
    hipError_t joseph_adjoint_tiled(const PoseGeom* geo, const RayRec* table, const Dims& dims, const int* pose_axes,
                                    int axes_present, const TileCfg& tc, int batch, double slot_bound,
                                    unsigned int* d_maxbits, int* d_fixexp, const ST* sino,
                                    typename Acc<ST>::type* vol, hipStream_t s);
    template <class T>
    hipError_t fixed_point_exponent_launch(const T* y, int64_t n, double slot_bound, unsigned int* d_maxbits,
                                           int* d_fixexp, hipStream_t s);
    // dst[i] = round-to-nearest-even bfloat16 of src[i]
    hipError_t launch_to_bf16(const float* src, bf16* dst, int64_t n, hipStream_t s);
    hipError_t siddon_ray_table(const PoseGeom* geo, const Dims& dims, SidRec* table, int* pose_axes, hipStream_t s);
    size_t siddon_tile_lds_bytes(const Dims& dims, const TileCfg& tc, int batch = kMaxPoseBatch);
    template <class T>
    hipError_t siddon_adjoint_tiled(const PoseGeom* geo, const SidRec* table, const Dims& dims, const int* pose_axes,
                                    int axes_present, const TileCfg& tc, int batch, double slot_bound,
                                    unsigned int* d_maxbits, int* d_fixexp, const T* sino, T* vol, hipStream_t s);
    size_t voxel_tile_lds_bytes(const TileCfg& tc, size_t lut_bytes = sizeof(float));
    int voxel_max_lut();
    template <class ST, class LT>
    hipError_t voxel_forward_tiled(const PoseGeom* geo, const Dims& dims, const LT* lut, int n_lut, float support,
                                   int diff, const int* pose_axis, const float* pose_margin, const TileCfg& tc,
                                   double slot_bound, unsigned int* d_maxbits, int* d_fixexp, const ST* vol, ST* sino,
                                   hipStream_t s);
    template <class T>
    hipError_t joseph_forward_planes(const PoseGeom* geo, const Dims& dims, const PadLayout* layouts,
                                     void* const* views, int axes_present, const T* vol, T* sino, hipStream_t s);
    // the volume copied into the interior of its plane-major view (float or bfloat16 elements)
    template <class E>
    hipError_t joseph_view_copy(const Dims& dims, const PadLayout& L, const E* vol, E* view, hipStream_t s);
    // Siddon forward over plane-major views (FP32, FP64), from the per-ray table
    template <class T>
    hipError_t siddon_forward_planes(const SidRec* table, const Dims& dims, const PadLayout* layouts,
                                     void* const* views, int axes_present, const T* vol, T* sino, hipStream_t s);

    template <class T>
    hipError_t siddon_forward(const PoseGeom* geo, const Dims& dims, const T* vol, T* sino, hipStream_t s);
    template <class T>
    hipError_t siddon_adjoint(const PoseGeom* geo, const Dims& dims, const T* sino, T* vol, hipStream_t s);
    hipError_t siddon_count(const PoseGeom* geo, const Dims& dims, unsigned long long* total, hipStream_t s);

    template <class T>
    hipError_t voxel_forward(const PoseGeom* geo, const Dims& dims, const T* lut, int n_lut, float support,
                             int diff, const T* vol, T* sino, hipStream_t s);
    // elsa's tap arithmetic (a division per tap): the one-thread-per-voxel kernels and FP64
    template <class T, class ST = T>
    hipError_t voxel_adjoint(const PoseGeom* geo, const Dims& dims, const T* lut, int n_lut, float support,
                             int diff, const ST* sino, ST* vol, hipStream_t s);
    // the default FP32 gather, with the tiled forward's tap arithmetic (ST: float or bfloat16)
    template <class ST>
    hipError_t voxel_adjoint_fast(const PoseGeom* geo, const Dims& dims, const float* lut, int n_lut, float support,
                                  int diff, const ST* sino, ST* vol, hipStream_t s);
    hipError_t voxel_count(const PoseGeom* geo, const Dims& dims, float support, unsigned long long* total,
                           hipStream_t s);

} // namespace ehp
