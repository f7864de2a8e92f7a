// tile_logic_check.cpp -- runs the tile logic of the LDS-tiled Joseph kernels on the CPU.
//
// Built with g++ against elsa_amd/csrc/ehp_joseph_tile.hpp (the same functions the HIP
// kernels call), it walks every tile of every pass the way a workgroup does, with bounds
// checks on every LDS slot and volume index and a check that each voxel is written by exactly
// one tile per pass. Used by tests/test_tile_logic.py to compare against the oracle.
#include "ehp_joseph_tile.hpp"

#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace ehp;

namespace
{
    Dims make_dims(int dim, const int64_t* vol, const int64_t* det, int nposes)
    {
        Dims d{};
        d.dim = dim;
        for (int i = 0; i < 3; ++i)
            d.vol[i] = i < dim ? (int) vol[i] : 1;
        d.det[0] = (int) det[0];
        d.det[1] = dim == 3 ? (int) det[1] : 1;
        d.nposes = nposes;
        d.strides[0] = 1;
        for (int i = 1; i < 3; ++i)
            d.strides[i] = d.strides[i - 1] * d.vol[i - 1];
        d.nvox = (int64_t) d.vol[0] * d.vol[1] * d.vol[2];
        d.pix_per_pose = (int64_t) d.det[0] * d.det[1];
        d.nrays = d.pix_per_pose * nposes;
        return d;
    }

    template <int D>
    int run(const Dims& dims, const std::vector<PoseGeom>& geo, TileCfg tc, const float* sino, float* vol,
            int64_t* visits)
    {
        // driving axes per pose, as pose_axes_kernel computes them
        std::vector<int> axes(dims.nposes, 0);
        int present = 0;
        for (int p = 0; p < dims.nposes; ++p)
            for (int v = 0; v < dims.det[1]; ++v)
                for (int u = 0; u < dims.det[0]; ++u) {
                    float o[3], d[3];
                    make_ray<D>(geo[p], u, v, o, d);
                    JRay jr;
                    joseph_setup<D>(o, d, dims.vol, jr);
                    if (jr.nsteps > 0)
                        axes[p] |= 1 << jr.axis;
                }
        for (int p = 0; p < dims.nposes; ++p)
            present |= axes[p];
        for (int64_t i = 0; i < dims.nvox; ++i)
            vol[i] = 0.0f;
        std::vector<int> owner(dims.nvox);
        for (int a = 0; a < D; ++a) {
            if (!(present & (1 << a)))
                continue;
            std::fill(owner.begin(), owner.end(), 0);
            int ntA, ntB, ntC;
            tile_counts(dims, tc, a, ntA, ntB, ntC);
            for (int tile = 0; tile < ntA * ntB * ntC; ++tile) {
                const TileBox t = make_tile<D>(dims, tc, a, tile % ntA, (tile / ntA) % ntB, tile / (ntA * ntB));
                std::vector<float> acc(t.nslots, 0.0f);
                for (int p = 0; p < dims.nposes; ++p) {
                    if (!(axes[p] & (1 << a)))
                        continue;
                    int u0, u1, v0, v1;
                    if (!tile_bbox<D>(geo[p], t, dims, u0, u1, v0, v1))
                        continue;
                    if (u0 < 0 || u1 >= dims.det[0] || v0 < 0 || v1 >= dims.det[1]) {
                        std::fprintf(stderr, "bbox out of the detector\n");
                        return 2;
                    }
                    for (int v = v0; v <= v1; ++v)
                        for (int u = u0; u <= u1; ++u) {
                            const float yv = sino[(int64_t) p * dims.pix_per_pose + (int64_t) v * dims.det[0] + u];
                            if (yv == 0.0f)
                                continue;
                            bool bad = false;
                            tile_ray_adjoint<D>(geo[p], dims, t, u, v, yv, [&](int slot, float val) {
                                if (slot < 0 || slot >= t.nslots) {
                                    bad = true;
                                    return;
                                }
                                acc[slot] += val;
                            });
                            if (bad) {
                                std::fprintf(stderr, "LDS slot out of range (pass %d tile %d pose %d u %d v %d)\n", a,
                                             tile, p, u, v);
                                return 3;
                            }
                        }
                }
                for (int i = 0; i < t.nslots; ++i) {
                    const int64_t lin = tile_slot_voxel<D>(t, dims, i);
                    if (lin < 0)
                        continue;
                    if (lin >= dims.nvox) {
                        std::fprintf(stderr, "volume index out of range\n");
                        return 4;
                    }
                    owner[lin] += 1;
                    vol[lin] += acc[i];
                }
            }
            for (int64_t i = 0; i < dims.nvox; ++i) {
                if (owner[i] != 1) {
                    std::fprintf(stderr, "voxel %lld owned %d times in pass %d\n", (long long) i, owner[i], a);
                    return 5;
                }
            }
            if (visits)
                visits[a] = 1;
        }
        return 0;
    }
} // namespace

extern "C" int tile_adjoint(int dim, const int64_t* vol_dims, const int64_t* det_dims, int nposes, const float* P,
                            const float* Pinv, const float* C, int LA, int TB, int TC, const float* sino, float* vol,
                            int64_t* passes)
{
    const Dims dims = make_dims(dim, vol_dims, det_dims, nposes);
    const int nm = dim * (dim + 1);
    std::vector<PoseGeom> geo(nposes);
    for (int p = 0; p < nposes; ++p) {
        PoseGeom g{};
        for (int i = 0; i < nm; ++i) {
            g.P[i] = P[p * nm + i];
            g.Pinv[i] = Pinv[p * nm + i];
        }
        for (int i = 0; i < dim; ++i)
            g.C[i] = C[p * dim + i];
        geo[p] = g;
    }
    const TileCfg tc{LA, TB, TC};
    return dim == 2 ? run<2>(dims, geo, tc, sino, vol, passes) : run<3>(dims, geo, tc, sino, vol, passes);
}
