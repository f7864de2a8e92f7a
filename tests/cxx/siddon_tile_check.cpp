// siddon_tile_check.cpp -- runs the Siddon traversal and the logic of the LDS-tiled Siddon
// back projection on the CPU (g++ against elsa_amd/csrc/ehp_siddon.hpp, the functions the HIP
// kernels call). Used by tests/test_tile_logic.py to compare against the oracle.
#include "ehp_siddon.hpp"

#include <cstdint>
#include <cstdio>
#include <map>
#include <vector>

using namespace ehp;

namespace
{
    // planar centres of a curved detector's pixels for the next call (nullptr: planar)
    const float* g_pix = nullptr;

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
        if (g_pix) {
            d.pix = g_pix;
            if (!curved_candidates(d, g_pix))
                d.dim = 0; // refused: the caller reports it
        }
        return d;
    }

    std::vector<PoseGeom> poses(int dim, int nposes, const float* P, const float* Pinv, const float* C)
    {
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
        return geo;
    }

    template <int D>
    SidRec ray_rec(const PoseGeom& g, const Dims& dims, int u, int v)
    {
        float o[3], d[3];
        make_ray_px<D>(g, dims, u, v, o, d);
        return siddon_setup<D>(o, d, dims.vol);
    }

    template <int D, int A>
    TileBox tile_of(const Dims& dims, const TileCfg& tc, int tile, int ntA, int ntB)
    {
        return make_tile<D, A>(dims, tc, tile % ntA, (tile / ntA) % ntB, tile / (ntA * ntB));
    }

    template <int D, int A>
    void tile_box(const TileBox& t, int* lo, int* hi, float* flo, float* fhi)
    {
        constexpr int b1 = TileAxes<D, A>::b1, b2 = TileAxes<D, A>::b2;
        lo[A] = t.A0;
        hi[A] = t.A1 - 1;
        lo[b1] = t.B0;
        hi[b1] = t.B1 - 1;
        if (D == 3) {
            lo[b2] = t.C0;
            hi[b2] = t.C1 - 1;
        }
        for (int i = 0; i < D; ++i) {
            flo[i] = (float) lo[i];
            fhi[i] = (float) (hi[i] + 1);
        }
    }

    // every ray: tiles of its pass reproduce the full traversal; candidates include it
    template <int D, int A>
    int check_cover_pass(const Dims& dims, const std::vector<PoseGeom>& geo, const TileCfg& tc, int stride,
                         double* stats)
    {
        int ntA, ntB, ntC;
        tile_counts(dims, tc, A, ntA, ntB, ntC);
        for (int p = 0; p < dims.nposes; ++p)
            for (int v = 0; v < dims.det[1]; ++v)
                for (int u = 0; u < dims.det[0]; u += stride) {
                    const SidRec r = ray_rec<D>(geo[p], dims, u, v);
                    if (!sid_hit(r) || sid_axis(r) != A)
                        continue;
                    std::map<int64_t, float> full;
                    siddon_traverse<D>(r, dims.vol, [&](int c0, int c1, int c2, float w) {
                        full[c0 + c1 * dims.strides[1] + (int64_t) c2 * dims.strides[2]] = w;
                    });
                    std::map<int64_t, float> tiled;
                    for (int tile = 0; tile < ntA * ntB * ntC; ++tile) {
                        const TileBox t = tile_of<D, A>(dims, tc, tile, ntA, ntB);
                        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
                        float flo[3] = {0, 0, 0}, fhi[3] = {0, 0, 0};
                        tile_box<D, A>(t, lo, hi, flo, fhi);
                        int n = 0;
                        bool bad = false;
                        siddon_box<D>(r, lo, hi, [&](int c0, int c1, int c2, float w) {
                            const int64_t lin = c0 + c1 * dims.strides[1] + (int64_t) c2 * dims.strides[2];
                            if (tiled.count(lin))
                                bad = true;
                            tiled[lin] = w;
                            if (w != 0.0f)
                                ++n;
                        });
                        if (bad)
                            return 31; // a voxel visited twice
                        if (n > 0) {
                            int u0, u1, v0, v1;
                            const bool any = box_bbox<D>(geo[p], flo, fhi, dims, u0, u1, v0, v1);
                            if (!any || u < u0 || u > u1 || v < v0 || v > v1)
                                return 32; // a crossing ray outside the candidate rectangle
                        }
                    }
                    if (tiled.size() != full.size())
                        return 33;
                    for (auto& kv : full) {
                        auto it = tiled.find(kv.first);
                        if (it == tiled.end())
                            return 34;
                        if (!(it->second == kv.second)) {
                            std::fprintf(stderr, "ray p %d u %d v %d voxel %lld: w %.9g vs %.9g\n", p, u, v,
                                         (long long) kv.first, it->second, kv.second);
                            return 35;
                        }
                    }
                    stats[0] += 1;
                    stats[1] += (double) full.size();
                }
        return 0;
    }

    // T = float or double (the FP32 and FP64 back projections: the same contributions,
    // round(y * length) in units of q by fixed_round)
    template <int D, int A, class T>
    int adj_pass(const Dims& dims, const std::vector<PoseGeom>& geo, const std::vector<SidRec>& table,
                 const std::vector<int>& axes, const TileCfg& tc, int eg, const T* sino, T* vol)
    {
        int ntA, ntB, ntC;
        tile_counts(dims, tc, A, ntA, ntB, ntC);
        const int nslots = tc.LA * tc.TB * (D == 3 ? tc.TC : 1);
        for (int tile = 0; tile < ntA * ntB * ntC; ++tile) {
            const TileBox t = tile_of<D, A>(dims, tc, tile, ntA, ntB);
            int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
            float flo[3] = {0, 0, 0}, fhi[3] = {0, 0, 0};
            tile_box<D, A>(t, lo, hi, flo, fhi);
            std::vector<__int128> iacc(nslots, 0);
            for (int p = 0; p < dims.nposes; ++p) {
                if (!(axes[p] & (1 << A)))
                    continue;
                int u0, u1, v0, v1;
                if (!box_bbox<D>(geo[p], flo, fhi, dims, u0, u1, v0, v1))
                    continue;
                for (int v = v0; v <= v1; ++v)
                    for (int u = u0; u <= u1; ++u) {
                        const int64_t ray = (int64_t) p * dims.pix_per_pose + (int64_t) v * dims.det[0] + u;
                        const T yv = sino[ray];
                        const SidRec& r = table[ray];
                        if (yv == T(0) || sid_axis(r) != A)
                            continue;
                        const double yq = ldexp((double) yv, 62 - eg);
                        bool bad = false;
                        siddon_box<D>(r, lo, hi, [&](int c0, int c1, int c2, float w) {
                            if (w == 0.0f)
                                return;
                            const int slot = sid_slot<D, A>(t, tc, c0, c1, c2);
                            if (slot < 0 || slot >= nslots) {
                                bad = true;
                                return;
                            }
                            // fixed_round's range: |y * length| below 2^51 units
                            if (!(fabs(yq * (double) w) < 2251799813685248.0)) {
                                bad = true;
                                return;
                            }
                            iacc[slot] += (__int128) (long long) fixed_round(yq, (double) w);
                        });
                        if (bad)
                            return 41;
                    }
            }
            const int TCe = D == 3 ? tc.TC : 1;
            for (int i = 0; i < nslots; ++i) {
                const int xb = i % tc.TB, rest = i / tc.TB, xc = rest % TCe, xa = rest / TCe;
                const int ca = t.A0 + xa, cb = t.B0 + xb, cc = t.C0 + xc;
                if (ca >= t.A1 || cb >= t.B1 || (D == 3 && cc >= t.C1)) {
                    if (iacc[i] != 0)
                        return 42; // a contribution to a slot outside the volume
                    continue;
                }
                int c[3] = {0, 0, 0};
                c[A] = ca;
                c[TileAxes<D, A>::b1] = cb;
                if (D == 3)
                    c[TileAxes<D, A>::b2] = cc;
                const int64_t lin = c[0] + c[1] * dims.strides[1] + (int64_t) c[2] * dims.strides[2];
                const double r = (double) (iacc[i] < 0 ? -iacc[i] : iacc[i]) / 9223372036854775808.0;
                if (r >= 1.0)
                    return 43;
                vol[lin] += (T) ((double) (int64_t) iacc[i] * ldexp(1.0, eg - 62));
            }
        }
        return 0;
    }

    template <int D, class T>
    int run_adjoint(const Dims& dims, const std::vector<PoseGeom>& geo, const TileCfg& tc, const T* sino, T* vol)
    {
        std::vector<SidRec> table(dims.nrays);
        std::vector<int> axes(dims.nposes, 0);
        for (int p = 0; p < dims.nposes; ++p)
            for (int v = 0; v < dims.det[1]; ++v)
                for (int u = 0; u < dims.det[0]; ++u) {
                    const SidRec r = ray_rec<D>(geo[p], dims, u, v);
                    table[(int64_t) p * dims.pix_per_pose + (int64_t) v * dims.det[0] + u] = r;
                    if (sid_hit(r))
                        axes[p] |= 1 << sid_axis(r);
                }
        const double bound = fixed_point_slot_bound(dims, geo.data());
        if (!(bound > 0.0))
            return 10;
        double m = 0.0;
        for (int64_t i = 0; i < dims.nrays; ++i)
            m = fabs((double) sino[i]) > m ? fabs((double) sino[i]) : m;
        const int eg = fixed_point_exponent(m, bound);
        for (int64_t i = 0; i < dims.nvox; ++i)
            vol[i] = T(0);
        int rc = adj_pass<D, 0, T>(dims, geo, table, axes, tc, eg, sino, vol);
        if (!rc)
            rc = adj_pass<D, 1, T>(dims, geo, table, axes, tc, eg, sino, vol);
        if (!rc && D == 3)
            rc = adj_pass<D, (D == 3 ? 2 : 1), T>(dims, geo, table, axes, tc, eg, sino, vol);
        return rc;
    }
} // namespace

extern "C" int siddon_forward(int dim, const int64_t* vol_dims, const int64_t* det_dims, int nposes, const float* P,
                              const float* Pinv, const float* C, const float* vol, float* sino)
{
    const Dims dims = make_dims(dim, vol_dims, det_dims, nposes);
    const std::vector<PoseGeom> geo = poses(dim, nposes, P, Pinv, C);
    for (int p = 0; p < nposes; ++p)
        for (int v = 0; v < dims.det[1]; ++v)
            for (int u = 0; u < dims.det[0]; ++u) {
                const SidRec r = dim == 2 ? ray_rec<2>(geo[p], dims, u, v) : ray_rec<3>(geo[p], dims, u, v);
                float acc = 0.0f;
                auto fn = [&](int c0, int c1, int c2, float w) {
                    acc += vol[c0 + c1 * dims.strides[1] + (int64_t) c2 * dims.strides[2]] * w;
                };
                if (dim == 2)
                    siddon_traverse<2>(r, dims.vol, fn);
                else
                    siddon_traverse<3>(r, dims.vol, fn);
                sino[(int64_t) p * dims.pix_per_pose + (int64_t) v * dims.det[0] + u] = acc;
            }
    return 0;
}

// stats: [0] rays checked, [1] crossings checked
extern "C" int siddon_tile_cover(int dim, const int64_t* vol_dims, const int64_t* det_dims, int nposes,
                                 const float* P, const float* Pinv, const float* C, int LA, int TB, int TC,
                                 int stride, double* stats)
{
    const Dims dims = make_dims(dim, vol_dims, det_dims, nposes);
    const std::vector<PoseGeom> geo = poses(dim, nposes, P, Pinv, C);
    const TileCfg tc{LA, TB, dim == 3 ? TC : 1};
    int rc;
    if (dim == 2) {
        rc = check_cover_pass<2, 0>(dims, geo, tc, stride, stats);
        if (!rc)
            rc = check_cover_pass<2, 1>(dims, geo, tc, stride, stats);
    } else {
        rc = check_cover_pass<3, 0>(dims, geo, tc, stride, stats);
        if (!rc)
            rc = check_cover_pass<3, 1>(dims, geo, tc, stride, stats);
        if (!rc)
            rc = check_cover_pass<3, 2>(dims, geo, tc, stride, stats);
    }
    return rc;
}

// The same for a curved detector whose pixels have the planar centres pix (dim - 1 per pixel,
// u fastest).
extern "C" int siddon_tile_cover_px(int dim, const int64_t* vol_dims, const int64_t* det_dims, int nposes,
                                    const float* P, const float* Pinv, const float* C, int LA, int TB, int TC,
                                    int stride, const float* pix, double* stats)
{
    g_pix = pix;
    const int rc = make_dims(dim, vol_dims, det_dims, nposes).dim == 0
                       ? 11
                       : siddon_tile_cover(dim, vol_dims, det_dims, nposes, P, Pinv, C, LA, TB, TC, stride, stats);
    g_pix = nullptr;
    return rc;
}

extern "C" int siddon_tile_adjoint(int dim, const int64_t* vol_dims, const int64_t* det_dims, int nposes,
                                   const float* P, const float* Pinv, const float* C, int LA, int TB, int TC,
                                   const float* sino, float* vol)
{
    const Dims dims = make_dims(dim, vol_dims, det_dims, nposes);
    const std::vector<PoseGeom> geo = poses(dim, nposes, P, Pinv, C);
    const TileCfg tc{LA, TB, dim == 3 ? TC : 1};
    return dim == 2 ? run_adjoint<2, float>(dims, geo, tc, sino, vol) : run_adjoint<3, float>(dims, geo, tc, sino, vol);
}

// The FP64 back projection (FP32 geometry and lengths, FP64 data).
extern "C" int siddon_tile_adjoint64(int dim, const int64_t* vol_dims, const int64_t* det_dims, int nposes,
                                     const float* P, const float* Pinv, const float* C, int LA, int TB, int TC,
                                     const double* sino, double* vol)
{
    const Dims dims = make_dims(dim, vol_dims, det_dims, nposes);
    const std::vector<PoseGeom> geo = poses(dim, nposes, P, Pinv, C);
    const TileCfg tc{LA, TB, dim == 3 ? TC : 1};
    return dim == 2 ? run_adjoint<2, double>(dims, geo, tc, sino, vol)
                    : run_adjoint<3, double>(dims, geo, tc, sino, vol);
}

// The GPU forward over zero-padded plane-major views (fan[a]: the fastest axis of the view of
// driving axis a), with its grouping: the pixels (u0 + l % wu, v0 + l / wu), l < 64, form a
// wavefront; its rays fall into groups by driving axis and step direction along it, and each
// group walks the planes of that axis together, every ray taking its own steps on the current
// plane (sid_walk_plane). Every read is checked against its view and against the plane. T: the
// element type of the views and of the sums (FP64: SiddonsMethodCUDA<double>, the FP32 lengths
// widened).
template <class T>
static int plane_forward(int dim, const int64_t* vol_dims, const int64_t* det_dims, int nposes, const float* P,
                         const float* Pinv, const float* C, const int* fan, const T* vol, T* sino, int wu)
{
    const Dims dims = make_dims(dim, vol_dims, det_dims, nposes);
    const std::vector<PoseGeom> geo = poses(dim, nposes, P, Pinv, C);
    if (dim == 2)
        wu = 64;
    if (wu < 1 || 64 % wu != 0)
        return 23;
    const int wv = 64 / wu;
    std::vector<std::vector<T>> views(3);
    PadLayout L[3];
    for (int a = 0; a < dim; ++a) {
        if (fan[a] == a || fan[a] < 0 || fan[a] > (dim == 3 ? 2 : 1))
            return 20;
        L[a] = pad_layout(dims, a, fan[a]);
        views[a].assign(L[a].size, T(0));
        for (int z = 0; z < dims.vol[2]; ++z)
            for (int y = 0; y < dims.vol[1]; ++y)
                for (int x = 0; x < dims.vol[0]; ++x)
                    views[a][pad_index(L[a], x, y, z)] = vol[x + dims.strides[1] * y + dims.strides[2] * z];
    }
    bool bad = false;
    for (int p = 0; p < nposes; ++p)
        for (int v0 = 0; v0 < dims.det[1]; v0 += wv)
            for (int u0 = 0; u0 < dims.det[0]; u0 += wu) {
                SidRec rr[64];
                int key[64];
                for (int l = 0; l < 64; ++l) {
                    key[l] = -1;
                    const int u = u0 + l % wu, v = v0 + l / wu;
                    if (u >= dims.det[0] || v >= dims.det[1])
                        continue;
                    rr[l] = dim == 2 ? ray_rec<2>(geo[p], dims, u, v) : ray_rec<3>(geo[p], dims, u, v);
                    if (sid_hit(rr[l])) {
                        const int a = sid_axis(rr[l]);
                        key[l] = 2 * a + (sid_step(rr[l], a) < 0 ? 1 : 0);
                    }
                }
                for (int k = 0; k < 2 * dim; ++k) {
                    const int a = k / 2, dir = (k & 1) ? -1 : 1;
                    int t0 = 0x7fffffff;
                    for (int l = 0; l < 64; ++l)
                        if (key[l] == k)
                            t0 = dir * rr[l].cur0[a] < t0 ? dir * rr[l].cur0[a] : t0;
                    if (t0 == 0x7fffffff)
                        continue;
                    SidWalk w[64];
                    T acc[64];
                    for (int l = 0; l < 64; ++l) {
                        acc[l] = T(0);
                        if (key[l] == k)
                            dim == 2 ? sid_walk_init<2>(rr[l], w[l]) : sid_walk_init<3>(rr[l], w[l]);
                    }
                    for (int pl = dir * t0; pl >= 0 && pl < dims.vol[a]; pl += dir) {
                        bool any = false;
                        for (int l = 0; l < 64; ++l)
                            any = any || (key[l] == k && w[l].alive);
                        if (!any)
                            break;
                        for (int l = 0; l < 64; ++l) {
                            if (key[l] != k)
                                continue;
                            auto fn = [&](int c0, int c1, int c2, float len) {
                                const int c[3] = {c0, c1, c2};
                                const int64_t o = pad_index(L[a], c0, c1, c2);
                                if (c[a] != pl || o < 0 || o >= L[a].size) {
                                    bad = true;
                                    return;
                                }
                                acc[l] += views[a][o] * (T) len;
                            };
                            if (dim == 2)
                                sid_walk_plane<2>(rr[l], dims.vol, a, pl, w[l], fn);
                            else
                                sid_walk_plane<3>(rr[l], dims.vol, a, pl, w[l], fn);
                        }
                    }
                    for (int l = 0; l < 64; ++l)
                        if (key[l] == k && w[l].alive)
                            return 24; // a ray left unfinished
                    for (int l = 0; l < 64; ++l)
                        if (key[l] == k)
                            sino[(int64_t) p * dims.pix_per_pose + (int64_t) (v0 + l / wu) * dims.det[0] + u0 +
                                 l % wu] = acc[l];
                }
                for (int l = 0; l < 64; ++l) {
                    const int u = u0 + l % wu, v = v0 + l / wu;
                    if (key[l] < 0 && u < dims.det[0] && v < dims.det[1])
                        sino[(int64_t) p * dims.pix_per_pose + (int64_t) v * dims.det[0] + u] = T(0);
                }
            }
    return bad ? 22 : 0;
}

extern "C" int siddon_plane_forward(int dim, const int64_t* vol_dims, const int64_t* det_dims, int nposes,
                                    const float* P, const float* Pinv, const float* C, const int* fan,
                                    const float* vol, float* sino, int wu)
{
    return plane_forward<float>(dim, vol_dims, det_dims, nposes, P, Pinv, C, fan, vol, sino, wu);
}

extern "C" int siddon_plane_forward64(int dim, const int64_t* vol_dims, const int64_t* det_dims, int nposes,
                                      const float* P, const float* Pinv, const float* C, const int* fan,
                                      const double* vol, double* sino, int wu)
{
    return plane_forward<double>(dim, vol_dims, det_dims, nposes, P, Pinv, C, fan, vol, sino, wu);
}
