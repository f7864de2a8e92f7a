// ehp_joseph_tiled.hip -- Joseph back projection with LDS-privatised volume tiles (FP32).
//
// The matched adjoint of the Joseph forward (elsa/projectors/JosephsMethod.cpp:140-198,
// adjoint branch :180-186) without global atomics. The volume is cut into tiles that are
// long along one driving axis a (LA planes) and TB x TC voxels across. One workgroup owns a
// tile for all poses: it accumulates in LDS every sample of every ray whose driving axis is
// a and whose bilinear footprint touches the tile, then adds the tile to the volume once.
// One launch per driving axis that occurs in the trajectory ("pass"); each voxel is owned by
// exactly one tile per pass, so passes are plain read-modify-write.
//
// The per-ray state and the per-sample weights come from the same functions as the
// one-thread-per-ray kernels (ehp_device.hpp, ehp_joseph_tile.hpp), so the weights are
// bit-identical to the forward's; only the order of the additions differs.
#include "ehp_kernels.hpp"

namespace ehp
{
    namespace
    {
        constexpr int kViewBatch = 16; // poses whose candidate rays are flattened together
        constexpr int kThreads = 256;
    } // namespace

    template <int D>
    __global__ void __launch_bounds__(kThreads)
        joseph_adj_tiled_kernel(const PoseGeom* __restrict__ geo, const Dims dims, const int* __restrict__ pose_axes,
                                const int a, const TileCfg tc, const int ntA, const int ntB, const int ntC,
                                const int tiles_per_xcd, const float* __restrict__ sino, float* __restrict__ vol)
    {
        extern __shared__ float acc[];
        __shared__ int s_u0[kViewBatch], s_nu[kViewBatch], s_v0[kViewBatch], s_pose[kViewBatch];
        __shared__ int s_cnt[kViewBatch], s_pref[kViewBatch + 1];

        // consecutive workgroups land on different XCDs: give each XCD a contiguous block of tiles
        const int tile = (int) (blockIdx.x % 8u) * tiles_per_xcd + (int) (blockIdx.x / 8u);
        if (tile >= ntA * ntB * ntC)
            return;
        const int tid = threadIdx.x;
        const TileBox t = make_tile<D>(dims, tc, a, tile % ntA, (tile / ntA) % ntB, tile / (ntA * ntB));
        for (int i = tid; i < t.nslots; i += kThreads)
            acc[i] = 0.0f;
        __syncthreads();

        for (int pbase = 0; pbase < dims.nposes; pbase += kViewBatch) {
            if (tid < kViewBatch) {
                const int p = pbase + tid;
                int u0 = 0, u1 = -1, v0 = 0, v1 = -1;
                int cnt = 0;
                if (p < dims.nposes && (pose_axes[p] & (1 << a))) {
                    if (tile_bbox<D>(geo[p], t, dims, u0, u1, v0, v1))
                        cnt = (u1 - u0 + 1) * (v1 - v0 + 1);
                }
                s_u0[tid] = u0;
                s_nu[tid] = u1 - u0 + 1;
                s_v0[tid] = v0;
                s_pose[tid] = p;
                s_cnt[tid] = cnt;
            }
            __syncthreads();
            if (tid == 0) {
                s_pref[0] = 0;
                for (int j = 0; j < kViewBatch; ++j)
                    s_pref[j + 1] = s_pref[j] + s_cnt[j];
            }
            __syncthreads();
            const int total = s_pref[kViewBatch];
            for (int k = tid; k < total; k += kThreads) {
                int j = 0;
                while (k >= s_pref[j + 1])
                    ++j;
                const int loc = k - s_pref[j];
                const int u = s_u0[j] + loc % s_nu[j];
                const int v = s_v0[j] + loc / s_nu[j];
                const int p = s_pose[j];
                const float yv = sino[(int64_t) p * dims.pix_per_pose + (int64_t) v * dims.det[0] + u];
                if (yv == 0.0f)
                    continue;
                tile_ray_adjoint<D>(geo[p], dims, t, u, v, yv, [&](int slot, float val) { atomicAdd(&acc[slot], val); });
            }
            __syncthreads();
        }

        // add the owned voxels of the tile to the volume
        for (int i = tid; i < t.nslots; i += kThreads) {
            const int64_t lin = tile_slot_voxel<D>(t, dims, i);
            if (lin >= 0)
                vol[lin] += acc[i];
        }
    }

    // Which driving axes occur among the rays of each pose (bit a set when some ray that hits
    // the volume drives along a).
    template <int D>
    __global__ void __launch_bounds__(256)
        pose_axes_kernel(const PoseGeom* __restrict__ geo, const Dims dims, int* __restrict__ pose_axes)
    {
        const int pose = blockIdx.y;
        const int64_t pix = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
        if (pix >= dims.pix_per_pose)
            return;
        const int u = (int) (pix % dims.det[0]);
        const int v = (int) (pix / dims.det[0]);
        float o[3], d[3];
        make_ray<D>(geo[pose], u, v, o, d);
        JRay jr;
        joseph_setup<D>(o, d, dims.vol, jr);
        if (jr.nsteps > 0)
            atomicOr(pose_axes + pose, 1 << jr.axis);
    }

    hipError_t joseph_pose_axes(const PoseGeom* geo, const Dims& dims, int* pose_axes, hipStream_t s)
    {
        hipError_t e = launch_fill<int>(pose_axes, dims.nposes, 0, s);
        if (e != hipSuccess)
            return e;
        const dim3 grid((unsigned) ((dims.pix_per_pose + 255) / 256), (unsigned) dims.nposes);
        if (dims.dim == 2)
            hipLaunchKernelGGL((pose_axes_kernel<2>), grid, dim3(256), 0, s, geo, dims, pose_axes);
        else
            hipLaunchKernelGGL((pose_axes_kernel<3>), grid, dim3(256), 0, s, geo, dims, pose_axes);
        return hipGetLastError();
    }

    size_t joseph_tile_lds_bytes(const Dims& dims, const TileCfg& tc)
    {
        const int PC = dims.dim == 3 ? tc.TC + 2 : 1;
        return (size_t) tc.LA * PC * (tc.TB + 2) * sizeof(float);
    }

    hipError_t joseph_adjoint_tiled(const PoseGeom* geo, const Dims& dims, const int* pose_axes, int axes_present,
                                    const TileCfg& tc, const float* sino, float* vol, hipStream_t s)
    {
        hipError_t e = launch_fill<float>(vol, dims.nvox, 0.0f, s);
        if (e != hipSuccess)
            return e;
        const size_t lds = joseph_tile_lds_bytes(dims, tc);
        for (int a = 0; a < dims.dim; ++a) {
            if (!(axes_present & (1 << a)))
                continue;
            int ntA, ntB, ntC;
            tile_counts(dims, tc, a, ntA, ntB, ntC);
            const int per_xcd = (ntA * ntB * ntC + 7) / 8;
            const dim3 grid((unsigned) (per_xcd * 8));
            if (dims.dim == 2)
                hipLaunchKernelGGL((joseph_adj_tiled_kernel<2>), grid, dim3(kThreads), lds, s, geo, dims, pose_axes, a, tc,
                                   ntA, ntB, ntC, per_xcd, sino, vol);
            else
                hipLaunchKernelGGL((joseph_adj_tiled_kernel<3>), grid, dim3(kThreads), lds, s, geo, dims, pose_axes, a, tc,
                                   ntA, ntB, ntC, per_xcd, sino, vol);
            e = hipGetLastError();
            if (e != hipSuccess)
                return e;
        }
        return hipSuccess;
    }
} // namespace ehp
