// ehp_waves.hpp -- what the plane-walking forward projections (Joseph, Siddon) share: the
// pixels a wavefront and a block cover, and reductions over the 64 lanes of a wavefront.
#pragma once

#include "ehp_kernels.hpp"

namespace ehp
{
    constexpr int kFwdThreads = 256;
    // 3D: a wavefront covers WU x WV pixels (u x v), a block TU x TV (four wavefronts side by
    // side along u). 8 x 8: neighbouring pixels of a square have nearly the same entry and exit
    // planes (96 % of a Joseph wavefront's lane-steps carry a sample at 256^3 / 384^2, against
    // 77 % for a 64 x 1 row), at about twice the cache lines per step of a row. 2D: 64 pixels
    // of the row per wavefront, 256 per block.
    constexpr int kFwdWU = 8;
    constexpr int kFwdWV = 64 / kFwdWU;
    constexpr int kFwdTV = kFwdWV > 4 ? kFwdWV : 4;
    constexpr int kFwdTU = kFwdThreads / kFwdTV;
    static_assert(kFwdWU * kFwdWV == 64 && kFwdTU % kFwdWU == 0 && kFwdTV % kFwdWV == 0, "forward wavefront shape");

    // Pixels of a forward launch: blocks of TU x TV pixels of one pose, ordered detector-row band
    // first, then pose, then column: id = (band * nposes + pose) * tiles_u + tu.
    struct FwdGrid {
        int tiles_u;
        int64_t total; // blocks
        int64_t per_xcd;
    };

    inline FwdGrid fwd_grid(const Dims& dims)
    {
        const int TU = dims.dim == 3 ? kFwdTU : kFwdThreads;
        const int TV = kFwdThreads / TU;
        FwdGrid g;
        g.tiles_u = (dims.det[0] + TU - 1) / TU;
        const int bands = (dims.det[1] + TV - 1) / TV;
        g.total = (int64_t) g.tiles_u * bands * dims.nposes;
        g.per_xcd = (g.total + 7) / 8;
        return g;
    }

    // Pixel (u, v) and pose of this lane in block id of a forward launch.
    template <int D>
    __device__ __forceinline__ void fwd_pixel(const Dims& dims, int tiles_u, int64_t id, int& u, int& v, int& pose)
    {
        constexpr int TU = (D == 3) ? kFwdTU : kFwdThreads;
        constexpr int TV = kFwdThreads / TU;
        constexpr int WU = (D == 3) ? kFwdWU : 64, WV = 64 / WU;
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
