// tile_logic_check.cpp -- runs the tile logic of the LDS-tiled Joseph kernels on the CPU.
//
// Built with g++ against elsa_amd/csrc/ehp_joseph_tile.hpp (the same functions the HIP
// kernels call), it builds the per-ray table, walks every tile of every pass the way a
// workgroup does -- pose batches, the low-face and test-free plane walks, contributions rounded
// to int32 in each ray's scale, 64-bit fixed-point sums -- with bounds checks on every LDS
// slot and volume index, a check that no sum leaves the int64 range, and a check that each
// voxel is written by exactly one tile per pass. It also runs the plane-major forward
// (plane_layout, plane_view, plane_ray_forward) over every layout a driving axis can get.
// Used by tests/test_tile_logic.py to compare against the oracle.
#define EHP_VIEW_CHECK 1
#include "ehp_joseph_tile.hpp"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace ehp
{
    bool g_view_oob = false;
}

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

    bool g_check_outside = false;

    struct Stats {
        double max_ratio = 0.0; // largest |slot sum| / 2^63 seen
    };

    // one candidate ray of a tile, LOW chosen as the kernel chooses it per tile; M: what add()
    // gets (TAP_FIXED: the FP32 back projection's 64-bit contributions, TAP_WEIGHT: the weights,
    // which the FP64 back projection scales)
    template <int D, int A, int M, class Y, class Add>
    void candidate(const RayRec& r, const TileBox& t, const TileOff& to, Y yq, Add add)
    {
        if (tile_low_face<D>(t))
            tile_candidate_fx<D, A, true, M>(r, t, to, yq, add);
        else
            tile_candidate_fx<D, A, false, M>(r, t, to, yq, add);
    }

    // T = float: the FP32 kernel's contributions (int32 in each ray's scale, shifted into units
    // of q); T = double: the FP64 kernel's ((y * len) * weight in FP64, rounded to a multiple of q)
    template <int D, int A, class T>
    int run_pass(const Dims& dims, const std::vector<PoseGeom>& geo, const std::vector<RayRec>& table,
                 const std::vector<int>& axes, TileCfg tc, int eg, const T* sino, T* vol, Stats& st)
    {
        std::vector<int> owner(dims.nvox, 0);
        int ntA, ntB, ntC;
        tile_counts(dims, tc, A, ntA, ntB, ntC);
        for (int tile = 0; tile < ntA * ntB * ntC; ++tile) {
            const TileBox t = make_tile<D, A>(dims, tc, tile % ntA, (tile / ntA) % ntB, tile / (ntA * ntB));
            const TileOff to = tile_offsets(t);
            std::vector<__int128> iacc(t.nslots, 0);
            for (int pbase = 0; pbase < dims.nposes; pbase += kPoseBatch) {
                for (int p = pbase; p < dims.nposes && p < pbase + kPoseBatch; ++p) {
                    if (!(axes[p] & (1 << A)))
                        continue;
                    int u0 = 0, u1 = -1, v0 = 0, v1 = -1;
                    const bool any = tile_bbox<D, A>(geo[p], t, dims, u0, u1, v0, v1);
                    if (g_check_outside) {
                        // no ray outside the candidate rectangle may reach a slot of the tile
                        for (int v = 0; v < dims.det[1]; ++v)
                            for (int u = 0; u < dims.det[0]; ++u) {
                                if (any && u >= u0 && u <= u1 && v >= v0 && v <= v1)
                                    continue;
                                const int64_t ray = (int64_t) p * dims.pix_per_pose + (int64_t) v * dims.det[0] + u;
                                bool hit = false;
                                candidate<D, A, TAP_WEIGHT>(table[ray], t, to, 1.0f, [&](uint32_t, float) { hit = true; });
                                if (hit) {
                                    std::fprintf(stderr, "ray outside the candidates (pass %d tile %d pose %d u %d v %d)\n",
                                                 A, tile, p, u, v);
                                    return 7;
                                }
                            }
                    }
                    if (!any)
                        continue;
                    if (u0 < 0 || u1 >= dims.det[0] || v0 < 0 || v1 >= dims.det[1]) {
                        std::fprintf(stderr, "bbox out of the detector\n");
                        return 2;
                    }
                    // the kernels' split of a rectangle entry into row and column
                    const int nu = u1 - u0 + 1, nv = v1 - v0 + 1;
                    const float rnu = 1.0f / (float) nu;
                    for (int k = 0; k < nu * nv; k += 1 + (nu * nv) / 4096) {
                        int row, col;
                        rect_split(k, nu, rnu, row, col);
                        if (row != k / nu || col != k % nu) {
                            std::fprintf(stderr, "rect_split(%d, %d) wrong\n", k, nu);
                            return 8;
                        }
                    }
                    for (int v = v0; v <= v1; ++v)
                        for (int u = u0; u <= u1; ++u) {
                            const int64_t ray = (int64_t) p * dims.pix_per_pose + (int64_t) v * dims.det[0] + u;
                            const T yv = sino[ray];
                            if (yv == T(0))
                                continue;
                            bool bad = false;
                            // the kernels' y * len in units of q: FP64 y times the FP32 length, or
                            // the FP32 product of the FP32 back projection; fixed_round needs
                            // |yq * w| < 2^51
                            const double yq = sizeof(T) == 8
                                                  ? ldexp((double) yv * (double) table[ray].len, 62 - eg)
                                                  : ldexp((double) ((float) yv * table[ray].len), 62 - eg);
                            if (!(fabs(yq) < 2251799813685248.0))
                                bad = true;
                            auto put = [&](uint32_t off, unsigned long long v) {
                                const int64_t slot = off / 8;
                                if (off % 8 != 0 || slot >= t.nslots) {
                                    bad = true;
                                    return;
                                }
                                iacc[slot] += (__int128) (long long) v;
                            };
                            if (sizeof(T) == 8) {
                                candidate<D, A, TAP_WEIGHT>(table[ray], t, to, yq, [&](uint32_t off, float w) {
                                    if (!(w >= 0.0f && w <= 1.0f))
                                        bad = true;
                                    put(off, fixed_round(yq, (double) w));
                                });
                            } else {
                                candidate<D, A, TAP_FIXED>(table[ray], t, to, yq, put);
                            }
                            if (bad) {
                                std::fprintf(stderr, "LDS slot or contribution out of range (pass %d tile %d pose %d u %d v %d)\n", A,
                                             tile, p, u, v);
                                return 3;
                            }
                        }
                }
            }
            for (int i = 0; i < t.nslots; ++i) {
                const double r = (double) (iacc[i] < 0 ? -iacc[i] : iacc[i]) / 9223372036854775808.0;
                st.max_ratio = r > st.max_ratio ? r : st.max_ratio;
                if (r >= 1.0) {
                    std::fprintf(stderr, "int64 slot sum overflow (pass %d tile %d slot %d)\n", A, tile, i);
                    return 6;
                }
            }
            for (int i = 0; i < t.nslots; ++i) {
                const int64_t lin = tile_slot_voxel<D, A>(t, dims, i);
                if (lin < 0)
                    continue;
                if (lin >= dims.nvox) {
                    std::fprintf(stderr, "volume index out of range\n");
                    return 4;
                }
                owner[lin] += 1;
                vol[lin] += (T) ((double) (int64_t) iacc[i] * ldexp(1.0, eg - 62));
            }
        }
        for (int64_t i = 0; i < dims.nvox; ++i) {
            if (owner[i] != 1) {
                std::fprintf(stderr, "voxel %lld owned %d times in pass %d\n", (long long) i, owner[i], A);
                return 5;
            }
        }
        return 0;
    }

    template <int D, class T>
    int run(const Dims& dims, const std::vector<PoseGeom>& geo, TileCfg tc, const T* sino, T* vol, double* info,
            const float* pix)
    {
        std::vector<RayRec> table(dims.nrays);
        std::vector<int> axes(dims.nposes, 0);
        for (int p = 0; p < dims.nposes; ++p)
            for (int v = 0; v < dims.det[1]; ++v)
                for (int u = 0; u < dims.det[0]; ++u) {
                    float o[3], d[3];
                    make_ray_px<D>(geo[p], dims, u, v, o, d);
                    JRay jr;
                    joseph_setup<D>(o, d, dims.vol, jr);
                    const RayRec r = make_ray_rec<D>(jr);
                    table[(int64_t) p * dims.pix_per_pose + (int64_t) v * dims.det[0] + u] = r;
                    if (r.nsteps > 0)
                        axes[p] |= 1 << (r.meta & 3);
                }
        int present = 0;
        for (int p = 0; p < dims.nposes; ++p)
            present |= axes[p];
        const double bound = fixed_point_slot_bound(dims, geo.data(), pix);
        info[0] = bound;
        if (!(bound > 0.0))
            return 10;
        T m = 0;
        for (int64_t i = 0; i < dims.nrays; ++i)
            m = fabs(sino[i]) > m ? (T) fabs(sino[i]) : m;
        const int eg = fixed_point_exponent(m, bound);
        info[1] = ldexp(1.0, eg - 62);
        for (int64_t i = 0; i < dims.nvox; ++i)
            vol[i] = 0;
        Stats st;
        for (int a = 0; a < D; ++a) {
            if (!(present & (1 << a)))
                continue;
            int rc = 0;
            if (a == 0)
                rc = run_pass<D, 0, T>(dims, geo, table, axes, tc, eg, sino, vol, st);
            else if (a == 1)
                rc = run_pass<D, 1, T>(dims, geo, table, axes, tc, eg, sino, vol, st);
            else
                rc = run_pass<D, (D == 3 ? 2 : 1), T>(dims, geo, table, axes, tc, eg, sino, vol, st);
            if (rc)
                return rc;
            info[3 + a] = 1;
        }
        info[2] = st.max_ratio;
        return 0;
    }
} // namespace

namespace
{
    template <class T>
    int tile_adjoint_t(int dim, const int64_t* vol_dims, const int64_t* det_dims, int nposes, const float* P,
                       const float* Pinv, const float* C, int LA, int TB, int TC, const T* sino, T* vol, double* info,
                       const float* pix = nullptr)
    {
        Dims dims = make_dims(dim, vol_dims, det_dims, nposes);
        if (pix) {
            dims.pix = pix;
            if (!curved_candidates(dims, pix))
                return 11;
            info[6] = dims.pix_uc;
            info[7] = dims.pix_margin;
        }
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
        return dim == 2 ? run<2, T>(dims, geo, tc, sino, vol, info, pix) : run<3, T>(dims, geo, tc, sino, vol, info, pix);
    }
} // namespace

// info: [0] slot bound, [1] q, [2] largest |slot sum| / 2^63, [3..5] passes run
extern "C" int tile_adjoint(int dim, const int64_t* vol_dims, const int64_t* det_dims, int nposes, const float* P,
                            const float* Pinv, const float* C, int LA, int TB, int TC, const float* sino, float* vol,
                            double* info)
{
    return tile_adjoint_t<float>(dim, vol_dims, det_dims, nposes, P, Pinv, C, LA, TB, TC, sino, vol, info);
}

// A curved detector: pix holds the planar centres of its pixels (dim - 1 per pixel, u fastest);
// check_outside also traces every ray outside each tile's candidate rectangle and fails when one
// reaches the tile. info[6], info[7]: the column and margin of the row search.
extern "C" int tile_adjoint_px(int dim, const int64_t* vol_dims, const int64_t* det_dims, int nposes, const float* P,
                               const float* Pinv, const float* C, int LA, int TB, int TC, const float* pix,
                               int check_outside, const float* sino, float* vol, double* info)
{
    g_check_outside = check_outside != 0;
    const int rc = tile_adjoint_t<float>(dim, vol_dims, det_dims, nposes, P, Pinv, C, LA, TB, TC, sino, vol, info, pix);
    g_check_outside = false;
    return rc;
}

// rect_split as the kernels call it (rnu = 1 / nu in FP32)
extern "C" void rect_split_host(int loc, int nu, int* row, int* col)
{
    rect_split(loc, nu, 1.0f / (float) nu, *row, *col);
}

// The same for the FP64 back projection (FP32 geometry and weights, FP64 data).
extern "C" int tile_adjoint64(int dim, const int64_t* vol_dims, const int64_t* det_dims, int nposes, const float* P,
                              const float* Pinv, const float* C, int LA, int TB, int TC, const double* sino,
                              double* vol, double* info)
{
    return tile_adjoint_t<double>(dim, vol_dims, det_dims, nposes, P, Pinv, C, LA, TB, TC, sino, vol, info);
}

// Forward projection through zero-padded plane-major views: fan[a] is the fastest axis of the
// view of driving axis a. The views are filled by the index map of the view copy kernels
// (pad_index), every offset checked against the view and written at most once; the padding
// stays zero.
static bool g_pairs = false; // row-pair views (FP32, 3D)

template <int D, int A>
static float fwd_one(const RayRec& r, const Dims& dims, const PadView& V, bool fb2, bool b16)
{
    if (D == 3 && g_pairs && b16)
        return fb2 ? plane_ray_forward<D, A, true, true, true>(r, dims, V)
                   : plane_ray_forward<D, A, false, true, true>(r, dims, V);
    if (D == 3 && g_pairs)
        return fb2 ? plane_ray_forward<D, A, true, false, true>(r, dims, V)
                   : plane_ray_forward<D, A, false, false, true>(r, dims, V);
    if (b16)
        return fb2 ? plane_ray_forward<D, A, true, true>(r, dims, V) : plane_ray_forward<D, A, false, true>(r, dims, V);
    return fb2 ? plane_ray_forward<D, A, true>(r, dims, V) : plane_ray_forward<D, A, false>(r, dims, V);
}

struct Span {
    int t0, niter, s_first, full0, full1;
};

// F64: an FP64 view (one element per voxel) and elsa's FP64 sums (plane_taps_add64)
template <int D, int A, bool NEG, bool F64>
static double span_one(const RayRec& r, const Dims& dims, const PadView& V, bool fb2, bool b16, const Span& q)
{
    if (F64)
        return fb2 ? plane_span_forward<D, A, NEG, true, false, false, true>(r, dims, V, q.t0, q.niter, q.s_first,
                                                                             q.full0, q.full1)
                   : plane_span_forward<D, A, NEG, false, false, false, true>(r, dims, V, q.t0, q.niter, q.s_first,
                                                                              q.full0, q.full1);
    if (D == 3 && g_pairs && b16)
        return fb2 ? plane_span_forward<D, A, NEG, true, true, true>(r, dims, V, q.t0, q.niter, q.s_first, q.full0,
                                                                     q.full1)
                   : plane_span_forward<D, A, NEG, false, true, true>(r, dims, V, q.t0, q.niter, q.s_first,
                                                                      q.full0, q.full1);
    if (D == 3 && g_pairs)
        return fb2 ? plane_span_forward<D, A, NEG, true, false, true>(r, dims, V, q.t0, q.niter, q.s_first, q.full0,
                                                                      q.full1)
                   : plane_span_forward<D, A, NEG, false, false, true>(r, dims, V, q.t0, q.niter, q.s_first,
                                                                       q.full0, q.full1);
    if (b16)
        return fb2 ? plane_span_forward<D, A, NEG, true, true>(r, dims, V, q.t0, q.niter, q.s_first, q.full0, q.full1)
                   : plane_span_forward<D, A, NEG, false, true>(r, dims, V, q.t0, q.niter, q.s_first, q.full0,
                                                                 q.full1);
    return fb2 ? plane_span_forward<D, A, NEG, true>(r, dims, V, q.t0, q.niter, q.s_first, q.full0, q.full1)
               : plane_span_forward<D, A, NEG, false>(r, dims, V, q.t0, q.niter, q.s_first, q.full0, q.full1);
}

template <int D, bool F64>
static double span_key(int key, const RayRec& r, const Dims& dims, const PadView* pv, const bool* fb2, bool b16,
                       const Span& q)
{
    switch (key) {
    case 0:
        return span_one<D, 0, false, F64>(r, dims, pv[0], fb2[0], b16, q);
    case 1:
        return span_one<D, 0, true, F64>(r, dims, pv[0], fb2[0], b16, q);
    case 2:
        return span_one<D, 1, false, F64>(r, dims, pv[1], fb2[1], b16, q);
    case 3:
        return span_one<D, 1, true, F64>(r, dims, pv[1], fb2[1], b16, q);
    case 4:
        return span_one<D, D - 1, false, F64>(r, dims, pv[D - 1], fb2[D - 1], b16, q);
    default:
        return span_one<D, D - 1, true, F64>(r, dims, pv[D - 1], fb2[D - 1], b16, q);
    }
}

// The GPU forward's grouping: the pixels (u0 + l % wu, v0 + l / wu), l < 64, form a wavefront;
// its rays fall into groups by driving axis and direction, each with one plane loop over the
// group's span (plane_ray_span, plane_span_forward). Pixels past the detector edge join no group.
static int64_t g_mixed_waves = 0; // wavefronts of the last plane_forward call with more than one group

template <int D, bool F64 = false>
static void wave_forward(const Dims& dims, const PoseGeom& g, int u0, int v0, int wu, const PadView* pv,
                         const bool* fb2, bool b16, double* out)
{
    RayRec rr[64];
    int key[64], tf[64], tl[64], s0[64];
    for (int l = 0; l < 64; ++l) {
        key[l] = -1;
        const int u = u0 + l % wu, v = v0 + l / wu;
        if (u >= dims.det[0] || v >= dims.det[1])
            continue;
        float o[3], d[3];
        JRay jr;
        make_ray<D>(g, u, v, o, d);
        joseph_setup<D>(o, d, dims.vol, jr);
        rr[l] = make_ray_rec<D>(jr);
        if (plane_ray_span(rr[l], dims.vol[jr.axis], tf[l], tl[l], s0[l]))
            key[l] = 2 * jr.axis + ((rr[l].meta & 4) ? 1 : 0);
        out[l] = 0.0;
    }
    int groups = 0;
    for (int k = 0; k < 2 * D; ++k) {
        int t0 = 0x7fffffff, t1 = -0x7fffffff, ta = -0x7fffffff, tb = 0x7fffffff;
        for (int l = 0; l < 64; ++l)
            if (key[l] == k) {
                t0 = tf[l] < t0 ? tf[l] : t0;
                t1 = tl[l] > t1 ? tl[l] : t1;
                ta = tf[l] > ta ? tf[l] : ta;
                tb = tl[l] < tb ? tl[l] : tb;
            }
        if (t0 > t1)
            continue;
        ++groups;
        const int full0 = ta - t0, full1 = tb - t0 + 1 > full0 ? tb - t0 + 1 : full0;
        for (int l = 0; l < 64; ++l)
            if (key[l] == k)
                out[l] = span_key<D, F64>(k, rr[l], dims, pv, fb2, b16,
                                          Span{t0, t1 - t0 + 1, s0[l] + t0 - tf[l], full0, full1});
    }
    g_mixed_waves += groups > 1;
}

extern "C" int64_t plane_forward_mixed_waves()
{
    return g_mixed_waves;
}

// flags: bit 0 bfloat16 views (the volume rounded to bfloat16), read by the bfloat16 variant of
// the kernel's ray loop; bit 1 the GPU's wavefront grouping (wave_forward) instead of one loop
// per ray, with wavefronts of wu = flags >> 8 pixels along u (3D; 0: 64) and 64 / wu along v;
// bit 2 row-pair views (3D: pad_store's two homes of every voxel, one read per sample)
extern "C" int plane_forward(int dim, const int64_t* vol_dims, const int64_t* det_dims, int nposes, const float* P,
                             const float* Pinv, const float* C, const int* fan, const float* vol, float* sino,
                             int flags)
{
    const bool b16 = flags & 1, waves = (flags & 2) != 0;
    g_pairs = (flags & 4) != 0;
    if (g_pairs && dim != 3)
        return 24;
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
    g_view_oob = false;
    std::vector<std::vector<float>> views(3);
    std::vector<std::vector<bf16>> views16(3);
    PadView pv[3];
    bool fb2[3] = {false, false, false};
    for (int a = 0; a < dim; ++a) {
        if (fan[a] == a || fan[a] < 0 || fan[a] > (dim == 3 ? 2 : 1))
            return 20;
        const PadLayout L = pad_layout(dims, a, fan[a], g_pairs);
        views[a].assign(L.size, 0.0f);
        std::vector<char> seen(L.size, 0);
        for (int z = 0; z < dims.vol[2]; ++z)
            for (int y = 0; y < dims.vol[1]; ++y)
                for (int x = 0; x < dims.vol[0]; ++x) {
                    const int64_t o = pad_index(L, x, y, z);
                    const int64_t o2 = L.pairs ? o - (int64_t) L.RS + 1 : o; // pad_store's second home
                    if (o < 0 || o >= L.size || seen[o] || o2 < 0 || o2 >= L.size || (L.pairs && seen[o2]))
                        return 21; // out of range or written twice
                    seen[o] = seen[o2] = 1;
                    const float v = vol[x + dims.strides[1] * y + dims.strides[2] * z];
                    pad_store<float>(L, views[a].data(), x, y, z, v);
                }
        if (b16) {
            views16[a].resize(L.size);
            for (int64_t i = 0; i < L.size; ++i)
                views16[a][i] = float_to_bf16(views[a][i]);
            pv[a] = PadView{views16[a].data(), L.PS * 2u, L.RS * 2u, L.size};
        } else {
            pv[a] = PadView{views[a].data(), L.PS * 4u, L.RS * 4u, L.size};
        }
        fb2[a] = dim == 3 && fan[a] == (a == 2 ? 1 : 2);
    }
    if (waves) {
        g_mixed_waves = 0;
        const int wu = dim == 2 || (flags >> 8) == 0 ? 64 : flags >> 8;
        if (wu < 1 || 64 % wu != 0)
            return 23;
        const int wv = 64 / wu;
        double lane[64];
        for (int p = 0; p < nposes; ++p)
            for (int v0 = 0; v0 < dims.det[1]; v0 += wv)
                for (int u0 = 0; u0 < dims.det[0]; u0 += wu) {
                    if (dim == 2)
                        wave_forward<2>(dims, geo[p], u0, v0, wu, pv, fb2, b16, lane);
                    else
                        wave_forward<3>(dims, geo[p], u0, v0, wu, pv, fb2, b16, lane);
                    for (int l = 0; l < 64; ++l) {
                        const int u = u0 + l % wu, v = v0 + l / wu;
                        if (u < dims.det[0] && v < dims.det[1])
                            sino[(int64_t) p * dims.pix_per_pose + (int64_t) v * dims.det[0] + u] = (float) lane[l];
                    }
                }
        return g_view_oob ? 22 : 0;
    }
    for (int p = 0; p < nposes; ++p)
        for (int v = 0; v < dims.det[1]; ++v)
            for (int u = 0; u < dims.det[0]; ++u) {
                float o[3], d[3];
                JRay jr;
                float val;
                if (dim == 2) {
                    make_ray<2>(geo[p], u, v, o, d);
                    joseph_setup<2>(o, d, dims.vol, jr);
                    const RayRec r = make_ray_rec<2>(jr);
                    val = jr.axis == 0 ? fwd_one<2, 0>(r, dims, pv[0], false, b16)
                                       : fwd_one<2, 1>(r, dims, pv[1], false, b16);
                } else {
                    make_ray<3>(geo[p], u, v, o, d);
                    joseph_setup<3>(o, d, dims.vol, jr);
                    const RayRec r = make_ray_rec<3>(jr);
                    val = jr.axis == 0   ? fwd_one<3, 0>(r, dims, pv[0], fb2[0], b16)
                          : jr.axis == 1 ? fwd_one<3, 1>(r, dims, pv[1], fb2[1], b16)
                                         : fwd_one<3, 2>(r, dims, pv[2], fb2[2], b16);
                }
                sino[(int64_t) p * dims.pix_per_pose + (int64_t) v * dims.det[0] + u] = val;
            }
    return g_view_oob ? 22 : 0;
}

// FP64 operators: the forward through FP64 views (one element per voxel, pad_store's layout),
// per ray (waves = 0) or with the GPU's wavefront grouping of wu x (64 / wu) pixels (waves = wu),
// with every 16-byte read checked against its view.
extern "C" int plane_forward64(int dim, const int64_t* vol_dims, const int64_t* det_dims, int nposes, const float* P,
                               const float* Pinv, const float* C, const int* fan, const double* vol, double* sino,
                               int waves)
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
    g_view_oob = false;
    std::vector<double> views[3];
    PadView pv[3];
    bool fb2[3] = {false, false, false};
    for (int a = 0; a < dim; ++a) {
        if (fan[a] == a || fan[a] < 0 || fan[a] > (dim == 3 ? 2 : 1))
            return 20;
        const PadLayout L = pad_layout(dims, a, fan[a], false);
        views[a].assign(L.size, 0.0);
        std::vector<char> seen(L.size, 0);
        for (int z = 0; z < dims.vol[2]; ++z)
            for (int y = 0; y < dims.vol[1]; ++y)
                for (int x = 0; x < dims.vol[0]; ++x) {
                    const int64_t o = pad_index(L, x, y, z);
                    if (o < 0 || o >= L.size || seen[o])
                        return 21;
                    seen[o] = 1;
                    pad_store<double>(L, views[a].data(), x, y, z, vol[x + dims.strides[1] * y + dims.strides[2] * z]);
                }
        pv[a] = PadView{views[a].data(), L.PS * 8u, L.RS * 8u, L.size};
        fb2[a] = dim == 3 && fan[a] == (a == 2 ? 1 : 2);
    }
    if (waves) {
        const int wu = dim == 2 ? 64 : waves;
        if (wu < 1 || 64 % wu != 0)
            return 23;
        const int wv = 64 / wu;
        double lane[64];
        for (int p = 0; p < nposes; ++p)
            for (int v0 = 0; v0 < dims.det[1]; v0 += wv)
                for (int u0 = 0; u0 < dims.det[0]; u0 += wu) {
                    if (dim == 2)
                        wave_forward<2, true>(dims, geo[p], u0, v0, wu, pv, fb2, false, lane);
                    else
                        wave_forward<3, true>(dims, geo[p], u0, v0, wu, pv, fb2, false, lane);
                    for (int l = 0; l < 64; ++l) {
                        const int u = u0 + l % wu, v = v0 + l / wu;
                        if (u < dims.det[0] && v < dims.det[1])
                            sino[(int64_t) p * dims.pix_per_pose + (int64_t) v * dims.det[0] + u] = lane[l];
                    }
                }
        return g_view_oob ? 22 : 0;
    }
    for (int p = 0; p < nposes; ++p)
        for (int v = 0; v < dims.det[1]; ++v)
            for (int u = 0; u < dims.det[0]; ++u) {
                float o[3], d[3];
                JRay jr;
                double val;
                if (dim == 2) {
                    make_ray<2>(geo[p], u, v, o, d);
                    joseph_setup<2>(o, d, dims.vol, jr);
                    const RayRec r = make_ray_rec<2>(jr);
                    val = jr.axis == 0 ? plane_ray_forward<2, 0, false, false, false, true>(r, dims, pv[0])
                                       : plane_ray_forward<2, 1, false, false, false, true>(r, dims, pv[1]);
                } else {
                    make_ray<3>(geo[p], u, v, o, d);
                    joseph_setup<3>(o, d, dims.vol, jr);
                    const RayRec r = make_ray_rec<3>(jr);
                    const int a = jr.axis;
                    const PadView& V = pv[a];
                    if (a == 0)
                        val = fb2[0] ? plane_ray_forward<3, 0, true, false, false, true>(r, dims, V)
                                     : plane_ray_forward<3, 0, false, false, false, true>(r, dims, V);
                    else if (a == 1)
                        val = fb2[1] ? plane_ray_forward<3, 1, true, false, false, true>(r, dims, V)
                                     : plane_ray_forward<3, 1, false, false, false, true>(r, dims, V);
                    else
                        val = fb2[2] ? plane_ray_forward<3, 2, true, false, false, true>(r, dims, V)
                                     : plane_ray_forward<3, 2, false, false, false, true>(r, dims, V);
                }
                sino[(int64_t) p * dims.pix_per_pose + (int64_t) v * dims.det[0] + u] = val;
            }
    return g_view_oob ? 22 : 0;
}
