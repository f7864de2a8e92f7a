// voxel_tile_check.cpp -- runs the logic of the detector-tiled voxel forward projection on
// the CPU (g++ against elsa_amd/csrc/ehp_voxel.hpp, the functions the HIP kernel calls): the
// frustum enumeration of every (pose, tile), the clipped footprints and the fixed-point sums.
// Used by tests/test_tile_logic.py to compare against the oracle.
#include "ehp_voxel.hpp"

#include <cmath>
#include <cstdint>
#include <cstdio>
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

    // T = float (FP32 table and weights in the forward's reciprocal form) or double (FP64 table,
    // the weights of visit() and tap_weight, as the FP64 kernels form them)
    template <int D, class T>
    int run(const Dims& dims, const std::vector<PoseGeom>& geo, const T* lut, int n_lut, float support, int TU,
            int TV, const T* vol, T* sino, double* stats)
    {
        // host set-up as the library does it
        double smax_all = 0.0;
        std::vector<int> axis(dims.nposes);
        std::vector<float> margin(dims.nposes);
        for (int p = 0; p < dims.nposes; ++p) {
            const double smax = voxel_max_scaling(geo[p], dims);
            if (!(smax > 0.0))
                return 50;
            smax_all = smax > smax_all ? smax : smax_all;
            margin[p] = (float) (support * smax + 1.0);
            axis[p] = pose_axis<D>(geo[p], dims);
        }
        double lutmax = 0.0;
        for (int i = 0; i < n_lut; ++i)
            lutmax = std::fabs((double) lut[i]) > lutmax ? std::fabs((double) lut[i]) : lutmax;
        // the FP32 kernels' LDS pair table (lerp_table_entry, tap_weight_fast)
        std::vector<float> lut_f(n_lut), tab(2 * (size_t) lerp_table_size(n_lut));
        for (int i = 0; i < n_lut; ++i)
            lut_f[i] = (float) lut[i];
        for (int i = 0; i < lerp_table_size(n_lut); ++i)
            lerp_table_entry(lut_f.data(), n_lut, i, tab.data());
        const bool tail_zero = lut_f[n_lut - 1] == 0.0f; // the kernels' tap_weight_fast<D, true>
        double ls = 1.0;
        while (ls < lutmax)
            ls *= 2.0;
        const double bound = voxel_slot_bound(dims, support, smax_all, ls);
        double m = 0.0;
        for (int64_t i = 0; i < dims.nvox; ++i)
            m = std::fabs((double) vol[i]) > m ? std::fabs((double) vol[i]) : m;
        const int eg = fixed_point_exponent(m, bound);
        if (TV > dims.det[1])
            TV = dims.det[1];
        const int ntu = (dims.det[0] + TU - 1) / TU, ntv = (dims.det[1] + TV - 1) / TV;
        for (int p = 0; p < dims.nposes; ++p) {
            const PoseGeom& g = geo[p];
            // taps of the untiled footprints on the detector
            double taps_full = 0.0;
            for (int64_t lin = 0; lin < dims.nvox; ++lin) {
                const int x = (int) (lin % dims.vol[0]);
                const int y = (int) ((lin / dims.vol[0]) % dims.vol[1]);
                const int z = (int) (lin / ((int64_t) dims.vol[0] * dims.vol[1]));
                Footprint fp;
                float rs_unused;
                if (sizeof(T) == 8)
                    footprint<D>(g, dims, x, y, z, support, fp);
                else
                    footprint_fast<D>(g, dims, x, y, z, support, fp, rs_unused);
                const int nu = fp.hi[0] - fp.lo[0] + 1, nv = fp.hi[1] - fp.lo[1] + 1;
                if (nu > 0 && nv > 0)
                    taps_full += (double) nu * nv;
            }
            double taps_tiled = 0.0;
            const int A = axis[p];
            const int b1 = (D == 2) ? 1 - A : (A == 0 ? 1 : 0);
            const int b2 = (D == 2) ? 2 : (A == 2 ? 1 : 2);
            for (int tile = 0; tile < ntu * ntv; ++tile) {
                const int tu = tile % ntu, tv = tile / ntu;
                const int u0 = tu * TU, v0 = tv * TV;
                const int u1 = (u0 + TU < dims.det[0] ? u0 + TU : dims.det[0]) - 1;
                const int v1 = (v0 + TV < dims.det[1] ? v0 + TV : dims.det[1]) - 1;
                std::vector<__int128> acc((size_t) TU * TV, 0);
                float cdir[12];
                const int nc = frustum_corners<D>(g, u0, u1, v0, v1, margin[p], cdir);
                std::vector<char> seen(dims.nvox, 0);
                // the kernel's walk: the planes in chunks (5 here, so that a tile spans many chunks,
                // empty planes included), each chunk's candidates taken by kSimThreads threads as
                // k = t, t + kSimThreads, ... through the prefix of the planes' rectangle sizes
                constexpr int kChunk = 5, kSimThreads = 64;
                for (int pc0 = 0; pc0 < dims.vol[A]; pc0 += kChunk) {
                    const int npl = dims.vol[A] - pc0 < kChunk ? dims.vol[A] - pc0 : kChunk;
                    int rb[kChunk], rc[kChunk], rn[kChunk], pref[kChunk + 1];
                    float rr[kChunk];
                    pref[0] = 0;
                    for (int q = 0; q < npl; ++q) {
                        int blo, bhi, clo, chi, cnt = 0;
                        if (frustum_plane_range<D>(g, dims, A, pc0 + q, cdir, nc, blo, bhi, clo, chi))
                            cnt = (bhi - blo + 1) * (chi - clo + 1);
                        rn[q] = cnt > 0 ? bhi - blo + 1 : 1;
                        rb[q] = blo;
                        rc[q] = clo;
                        rr[q] = 1.0f / (float) rn[q];
                        pref[q + 1] = pref[q] + cnt;
                    }
                    const int total = pref[npl];
                    for (int t = 0; t < kSimThreads; ++t) {
                        int jp = 0, pj = 0, pn = pref[1];
                        for (int k = t; k < total; k += kSimThreads) {
                            while (k >= pn) {
                                ++jp;
                                pj = pn;
                                pn = pref[jp + 1];
                            }
                            int row, col;
                            rect_split(k - pj, rn[jp], rr[jp], row, col);
                            if (row < 0 || col < 0 || col >= rn[jp])
                                return 55;
                            int c[3] = {0, 0, 0};
                            c[A] = pc0 + jp;
                            c[b1] = rb[jp] + col;
                            if (D == 3)
                                c[b2] = rc[jp] + row;
                            const int64_t lin = c[0] + c[1] * dims.strides[1] + (int64_t) c[2] * dims.strides[2];
                            if (lin < 0 || lin >= dims.nvox || seen[lin])
                                return 51; // out of range or enumerated twice for one tile
                            seen[lin] = 1;
                            Footprint fp;
                            // the kernel's footprints: footprint() for FP64, footprint_fast for FP32
                            float rs = 0.0f;
                            if (sizeof(T) == 8)
                                footprint<D>(g, dims, c[0], c[1], c[2], support, fp);
                            else
                                footprint_fast<D>(g, dims, c[0], c[1], c[2], support, fp, rs);
                            const int lo0 = fp.lo[0] > u0 ? fp.lo[0] : u0, hi0 = fp.hi[0] < u1 ? fp.hi[0] : u1;
                            const int lo1 = fp.lo[1] > v0 ? fp.lo[1] : v0, hi1 = fp.hi[1] < v1 ? fp.hi[1] : v1;
                            if (lo0 > hi0 || lo1 > hi1)
                                continue;
                            taps_tiled += (double) (hi0 - lo0 + 1) * (hi1 - lo1 + 1);
                            const T xv = vol[lin];
                            if (xv == T(0))
                                continue;
                            const double xq = ldexp((double) xv, 62 - eg);
                            // the GPU's arithmetic: FP32, one reciprocal of the scaling per voxel and
                            // the table index as dist * (n - 1) / support; FP64, visit()'s divisions
                            const float kidx = (float) (n_lut - 1) / support;
                            for (int j = lo1; j <= hi1; ++j)
                                for (int i = lo0; i <= hi0; ++i) {
                                    double w;
                                    if (sizeof(T) == 8) {
                                        const float d0 = (fp.center[0] - (float) i) / fp.scaling;
                                        const float d1 = D == 3 ? (fp.center[1] - (float) j) / fp.scaling : 0.0f;
                                        const float dist = D == 2 ? sqrtf(d0 * d0) : sqrtf(d0 * d0 + d1 * d1);
                                        w = (double) tap_weight<D, T>(lut, n_lut, support, dist, d0, false);
                                    } else {
                                        const float d0 = (fp.center[0] - (float) i) * rs;
                                        const float d1 = D == 3 ? (fp.center[1] - (float) j) * rs : 0.0f;
                                        const float d2 = __builtin_fmaf(d0, d0, d1 * d1);
                                        w = (double) (tail_zero
                                                          ? tap_weight_fast<D, true>(tab.data(), n_lut, kidx, d2, d0, false)
                                                          : tap_weight_fast<D, false>(tab.data(), n_lut, kidx, d2, d0,
                                                                                      false));
                                    }
                                    if (w == 0.0)
                                        continue;
                                    // fixed_round's range: |x w| below 2^51 units
                                    if (!(std::fabs(xq * w) < 2251799813685248.0))
                                        return 52;
                                    acc[(size_t) (j - v0) * TU + (i - u0)] += (__int128) (long long) fixed_round(xq, w);
                                }
                        }
                    }
                }
                for (int v = v0; v <= v1; ++v)
                    for (int u = u0; u <= u1; ++u) {
                        const __int128 a = acc[(size_t) (v - v0) * TU + (u - u0)];
                        if ((double) (a < 0 ? -a : a) >= 9.2e18)
                            return 53;
                        sino[(int64_t) p * dims.pix_per_pose + (int64_t) v * dims.det[0] + u] =
                            (T) ((double) (int64_t) a * ldexp(1.0, eg - 62));
                    }
            }
            if (taps_tiled != taps_full) {
                std::fprintf(stderr, "pose %d: %.0f taps tiled, %.0f untiled\n", p, taps_tiled, taps_full);
                return 54;
            }
            stats[0] += taps_full;
        }
        return 0;
    }
} // namespace

extern "C" int voxel_tile_forward(int dim, const int64_t* vol_dims, const int64_t* det_dims, int nposes,
                                  const float* P, const float* ext, const float* C, const float* sdd,
                                  const float* Pinv, const float* lut, int n_lut, float support, int TU, int TV,
                                  const float* vol, float* sino, double* stats)
{
    const Dims dims = make_dims(dim, vol_dims, det_dims, nposes);
    const int nm = dim * (dim + 1);
    std::vector<PoseGeom> geo(nposes);
    for (int p = 0; p < nposes; ++p) {
        PoseGeom g{};
        for (int i = 0; i < nm; ++i) {
            g.P[i] = P[p * nm + i];
            g.Pinv[i] = Pinv[p * nm + i];
            g.ext[i] = ext[p * nm + i];
        }
        for (int i = 0; i < dim; ++i)
            g.C[i] = C[p * dim + i];
        g.sdd = sdd[p];
        geo[p] = g;
    }
    return dim == 2 ? run<2, float>(dims, geo, lut, n_lut, support, TU, TV, vol, sino, stats)
                    : run<3, float>(dims, geo, lut, n_lut, support, TU, TV, vol, sino, stats);
}

// The FP64 forward (FP64 table and data, FP32 geometry).
extern "C" int voxel_tile_forward64(int dim, const int64_t* vol_dims, const int64_t* det_dims, int nposes,
                                    const float* P, const float* ext, const float* C, const float* sdd,
                                    const float* Pinv, const double* lut, int n_lut, float support, int TU, int TV,
                                    const double* vol, double* sino, double* stats)
{
    const Dims dims = make_dims(dim, vol_dims, det_dims, nposes);
    const int nm = dim * (dim + 1);
    std::vector<PoseGeom> geo(nposes);
    for (int p = 0; p < nposes; ++p) {
        PoseGeom g{};
        for (int i = 0; i < nm; ++i) {
            g.P[i] = P[p * nm + i];
            g.Pinv[i] = Pinv[p * nm + i];
            g.ext[i] = ext[p * nm + i];
        }
        for (int i = 0; i < dim; ++i)
            g.C[i] = C[p * dim + i];
        g.sdd = sdd[p];
        geo[p] = g;
    }
    return dim == 2 ? run<2, double>(dims, geo, lut, n_lut, support, TU, TV, vol, sino, stats)
                    : run<3, double>(dims, geo, lut, n_lut, support, TU, TV, vol, sino, stats);
}
