// ehp_waves.hpp -- what the plane-walking forward projections (Joseph, Siddon) share: the
// pixels a wavefront and a block cover, and reductions over the 64 lanes of a wavefront.
#pragma once

#include "ehp_kernels.hpp"

namespace ehp
{
    constexpr int kFwdThreads = 256;
    // 3D: a wavefront covers WU x WV pixels (u x v, WV = 64 / WU), a block TU x TV (four
    // wavefronts side by side along u). Pixels of a compact patch have nearly the same entry and
    // exit planes (96 % of a Joseph wavefront's lane-steps carry a sample at 256^3 / 384^2 with
    // 8 x 8, against 77 % for a 64 x 1 row), while a wider patch reads fewer rows of a plane per
    // step. 8 x 8 for views with one element per voxel (the Joseph forward over them: 7.84 ms
    // at 256^3, 16 x 4: 7.95 ms); 16 x 4 for row-pair views, whose rows are twice as long (the
    // same forward: 5.7 ms, 8 x 8: 6.9 ms, 32 x 2: 6.0 ms). 2D: 64 pixels of the row per
    // wavefront, 256 per block.
    constexpr int kFwdWU = 8;
    constexpr int kFwdWUPairs = 16;

    template <int WU>
    struct FwdShape {
        static constexpr int WV = 64 / WU;
        static constexpr int TV = WV > 4 ? WV : 4;
        static constexpr int TU = kFwdThreads / TV;
        static_assert(WU * WV == 64 && TU % WU == 0 && TV % WV == 0, "forward wavefront shape");
    };

    // Pixels of a forward launch: blocks of TU x TV pixels of one pose, ordered detector-row band
    // first, then pose, then column: id = (band * nposes + pose) * tiles_u + tu.
    struct FwdGrid {
        int tiles_u;
        int64_t total; // blocks
        int64_t per_xcd;
    };

    // wu: the wavefront width WU of the launch (3D)
    inline FwdGrid fwd_grid(const Dims& dims, int wu = kFwdWU)
    {
        const int wv = 64 / wu;
        const int TU = dims.dim == 3 ? kFwdThreads / (wv > 4 ? wv : 4) : kFwdThreads;
        const int TV = kFwdThreads / TU;
        FwdGrid g;
        g.tiles_u = (dims.det[0] + TU - 1) / TU;
        const int bands = (dims.det[1] + TV - 1) / TV;
        g.total = (int64_t) g.tiles_u * bands * dims.nposes;
        g.per_xcd = (g.total + 7) / 8;
        return g;
    }

    // Pixel (u, v) and pose of this lane in block id of a forward launch (3D wavefronts WU wide).
    template <int D, int WU3 = kFwdWU>
    __device__ __forceinline__ void fwd_pixel(const Dims& dims, int tiles_u, int64_t id, int& u, int& v, int& pose)
    {
        constexpr int TU = (D == 3) ? FwdShape<WU3>::TU : kFwdThreads;
        constexpr int TV = kFwdThreads / TU;
        constexpr int WU = (D == 3) ? WU3 : 64, WV = 64 / WU;
        const int tu = (int) (id % tiles_u);
        const int64_t bp = id / tiles_u;
        pose = (int) (bp % dims.nposes);
        const int band = (int) (bp / dims.nposes);
        const int lane = (int) threadIdx.x % 64, w = (int) threadIdx.x / 64;
        u = tu * TU + (w % (TU / WU)) * WU + lane % WU;
        v = band * TV + (w / (TU / WU)) * WV + lane / WU;
    }

    // Reductions over the 64 lanes of a wavefront (every lane active), returned in a scalar
    // register.
    __device__ __forceinline__ int wave_min(int x)
    {
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
            x = min(x, __shfl_xor(x, o, 64));
        return __builtin_amdgcn_readfirstlane(x);
    }

    __device__ __forceinline__ int wave_max(int x)
    {
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
            x = max(x, __shfl_xor(x, o, 64));
        return __builtin_amdgcn_readfirstlane(x);
    }

    __device__ __forceinline__ int wave_or(int x)
    {
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
            x |= __shfl_xor(x, o, 64);
        return __builtin_amdgcn_readfirstlane(x);
    }
} // namespace ehp
