// ehp_joseph_tile.hpp -- per-tile logic of the LDS-tiled Joseph kernels.
//
// Host- and device-callable so that the exact index arithmetic of the kernels can be run
// on the CPU against the oracle (tests/cxx/tile_logic_check.cpp) before it runs on a GPU.
#pragma once

#include "ehp_device.hpp"

namespace ehp
{
    // Tile of the LDS-privatised Joseph kernels: LA planes along the driving axis, TB x TC
    // voxels across (TC unused in 2D).
    struct TileCfg {
        int LA;
        int TB;
        int TC;
    };

    // One tile of one pass: planes [A0, A1) along the driving axis a, transverse voxel ranges
    // [B0, B1) along b1 and [C0, C1) along b2 (b1 < b2). The LDS block holds the planes of
    // the tile and the transverse ranges [B0 - 1, B0 + TB] x [C0 - 1, C0 + TC] (one slot of
    // halo on either side), b1 fastest: slot(pa, fc, fb) = (pa * PC + fc) * PB + fb.
    struct TileBox {
        int a, b1, b2;
        int A0, A1, B0, B1, C0, C1;
        int PB, PC;
        int nslots;
    };

    template <int D>
    EHP_HD inline TileBox make_tile(const Dims& dims, const TileCfg& tc, int a, int ia, int ib, int ic)
    {
        TileBox t;
        t.a = a;
        t.b1 = (D == 2) ? 1 - a : (a == 0 ? 1 : 0);
        t.b2 = (D == 2) ? 0 : (a == 2 ? 1 : 2);
        t.A0 = ia * tc.LA;
        t.A1 = dims.vol[a] < t.A0 + tc.LA ? dims.vol[a] : t.A0 + tc.LA;
        t.B0 = ib * tc.TB;
        t.B1 = dims.vol[t.b1] < t.B0 + tc.TB ? dims.vol[t.b1] : t.B0 + tc.TB;
        t.C0 = (D == 3) ? ic * tc.TC : 0;
        t.C1 = (D == 3) ? (dims.vol[t.b2] < t.C0 + tc.TC ? dims.vol[t.b2] : t.C0 + tc.TC) : 1;
        t.PB = tc.TB + 2;
        t.PC = (D == 3) ? tc.TC + 2 : 1;
        t.nslots = tc.LA * t.PC * t.PB;
        return t;
    }

    // Candidate pixels of one pose for a tile: the bounding rectangle of the projected corners
    // of [A0, A1] x [B0 - 1, B1 + 1] x [C0 - 1, C1 + 1], widened by one pixel. A point X lies on
    // the ray of pixel u when P X (homogeneous) maps to u + 0.5. Falls back to the whole
    // detector when a corner is not in front of the source.
    template <int D>
    EHP_HD inline bool tile_bbox(const PoseGeom& g, const TileBox& t, const Dims& dims, int& u0, int& u1, int& v0,
                                 int& v1)
    {
        float lo[3], hi[3];
        lo[t.a] = (float) t.A0;
        hi[t.a] = (float) t.A1;
        lo[t.b1] = (float) (t.B0 - 1);
        hi[t.b1] = (float) (t.B1 + 1);
        if (D == 3) {
            lo[t.b2] = (float) (t.C0 - 1);
            hi[t.b2] = (float) (t.C1 + 1);
        }
        float umin = 3.0e38f, umax = -3.0e38f, vmin = 3.0e38f, vmax = -3.0e38f;
        bool front = true;
        for (int c = 0; c < (1 << D); ++c) {
            float X[3];
            for (int i = 0; i < D; ++i)
                X[i] = ((c >> i) & 1) ? hi[i] : lo[i];
            float p[3];
            for (int k = 0; k < D; ++k) {
                float s = g.P[k + D * D];
                for (int j = 0; j < D; ++j)
                    s = __builtin_fmaf(g.P[k + j * D], X[j], s);
                p[k] = s;
            }
            const float den = p[D - 1];
            front = front && (den > 0.0f);
            const float pu = p[0] / den;
            umin = pu < umin ? pu : umin;
            umax = pu > umax ? pu : umax;
            if (D == 3) {
                const float pv = p[1] / den;
                vmin = pv < vmin ? pv : vmin;
                vmax = pv > vmax ? pv : vmax;
            }
        }
        if (!front) {
            u0 = 0;
            u1 = dims.det[0] - 1;
            v0 = 0;
            v1 = dims.det[1] - 1;
            return true;
        }
        const float W = (float) dims.det[0] + 4.0f, H = (float) dims.det[1] + 4.0f;
        umin = umin < -4.0f ? -4.0f : (umin > W ? W : umin);
        umax = umax < -4.0f ? -4.0f : (umax > W ? W : umax);
        u0 = (int) floorf(umin - 0.5f) - 1;
        u1 = (int) ceilf(umax - 0.5f) + 1;
        u0 = u0 < 0 ? 0 : u0;
        u1 = u1 > dims.det[0] - 1 ? dims.det[0] - 1 : u1;
        if (D == 3) {
            vmin = vmin < -4.0f ? -4.0f : (vmin > H ? H : vmin);
            vmax = vmax < -4.0f ? -4.0f : (vmax > H ? H : vmax);
            v0 = (int) floorf(vmin - 0.5f) - 1;
            v1 = (int) ceilf(vmax - 0.5f) + 1;
            v0 = v0 < 0 ? 0 : v0;
            v1 = v1 > dims.det[1] - 1 ? dims.det[1] - 1 : v1;
        } else {
            v0 = v1 = 0;
        }
        return u0 <= u1 && v0 <= v1;
    }

    // Narrow [slo, shi] to the samples whose position along one transverse axis may lie in
    // [L, H), with one sample of slack on either side (the exact test follows per sample).
    EHP_HD inline void clip_range(float p0, float st, float L, float H, int& slo, int& shi)
    {
        if (fabsf(st) < 1.0e-12f) {
            if (p0 < L - 1.0f || p0 >= H + 1.0f)
                shi = slo - 1;
            return;
        }
        float lo = (L - p0) / st, hi = (H - p0) / st;
        if (st < 0.0f) {
            const float t = lo;
            lo = hi;
            hi = t;
        }
        lo = lo < -2.0f ? -2.0f : (lo > 2.0e9f ? 2.0e9f : lo);
        hi = hi < -2.0f ? -2.0f : (hi > 2.0e9f ? 2.0e9f : hi);
        const int l = (int) ceilf(lo) - 1;
        const int h = (int) floorf(hi) + 1;
        slo = slo > l ? slo : l;
        shi = shi < h ? shi : h;
    }

    // Samples [slo, shi] of a ray (driving axis a) that can touch the tile's voxels.
    EHP_HD inline void tile_sample_range(const JRay& jr, const TileBox& t, int D, int& slo, int& shi)
    {
        const int P0 = (int) (jr.pos0[t.a] - 0.5f);
        if (jr.step[t.a] > 0.0f) {
            slo = t.A0 - P0;
            shi = t.A1 - 1 - P0;
        } else {
            slo = P0 - (t.A1 - 1);
            shi = P0 - t.A0;
        }
        slo = slo > 0 ? slo : 0;
        shi = shi < jr.nsteps - 1 ? shi : jr.nsteps - 1;
        clip_range(jr.pos0[t.b1], jr.step[t.b1], (float) t.B0 - 0.5f, (float) t.B1 + 0.5f, slo, shi);
        if (D == 3)
            clip_range(jr.pos0[t.b2], jr.step[t.b2], (float) t.C0 - 0.5f, (float) t.C1 + 0.5f, slo, shi);
    }

    // Back projection of one ray into the tile's LDS block: for every sample whose bilinear
    // footprint touches the tile, add y*len*weight to its four (2D: two) slots through add(slot,
    // value). Weights are formed as in elsa's doInterpolation (JosephsMethod.cpp:60-97): the
    // product of the transverse weights in axis order, the driving weight being 1.
    template <int D, class Add>
    EHP_HD inline void tile_ray_adjoint(const PoseGeom& g, const Dims& dims, const TileBox& t, int u, int v, float yv,
                                        Add add)
    {
        float o[3], d[3];
        make_ray<D>(g, u, v, o, d);
        JRay jr;
        joseph_setup<D>(o, d, dims.vol, jr);
        if (jr.axis != t.a || jr.nsteps == 0)
            return;
        int slo, shi;
        tile_sample_range(jr, t, D, slo, shi);
        const float yl = yv * jr.len;
        for (int s = slo; s <= shi; ++s) {
            JSample js;
            joseph_sample<D>(jr, s, js);
            const int f1 = js.fl[t.b1];
            if (f1 < t.B0 - 1 || f1 > t.B1 - 1)
                continue;
            const int pa = js.fl[t.a] - t.A0;
            if (D == 2) {
                const int base = pa * t.PB + (f1 - (t.B0 - 1));
                add(base, yl * js.w[t.b1]);
                add(base + 1, yl * js.cw[t.b1]);
            } else {
                const int f2 = js.fl[t.b2];
                if (f2 < t.C0 - 1 || f2 > t.C1 - 1)
                    continue;
                const int base = (pa * t.PC + (f2 - (t.C0 - 1))) * t.PB + (f1 - (t.B0 - 1));
                const float w1 = js.w[t.b1], c1 = js.cw[t.b1], w2 = js.w[t.b2], c2 = js.cw[t.b2];
                add(base, yl * (w1 * w2));
                add(base + 1, yl * (c1 * w2));
                add(base + t.PB, yl * (w1 * c2));
                add(base + t.PB + 1, yl * (c1 * c2));
            }
        }
    }

    // Linear volume index of LDS slot i if the tile owns that voxel, else -1.
    template <int D>
    EHP_HD inline int64_t tile_slot_voxel(const TileBox& t, const Dims& dims, int i)
    {
        const int fb = i % t.PB;
        const int rest = i / t.PB;
        const int fc = rest % t.PC;
        const int pa = rest / t.PC;
        const int xa = t.A0 + pa;
        const int xb = t.B0 - 1 + fb;
        const int xc = t.C0 - 1 + fc;
        if (xa >= t.A1 || xb < t.B0 || xb >= t.B1)
            return -1;
        if (D == 3 && (xc < t.C0 || xc >= t.C1))
            return -1;
        int c[3] = {0, 0, 0};
        c[t.a] = xa;
        c[t.b1] = xb;
        if (D == 3)
            c[t.b2] = xc;
        return c[0] * dims.strides[0] + c[1] * dims.strides[1] + c[2] * dims.strides[2];
    }

    inline void tile_counts(const Dims& dims, const TileCfg& tc, int a, int& ntA, int& ntB, int& ntC)
    {
        const int b1 = (dims.dim == 2) ? 1 - a : (a == 0 ? 1 : 0);
        const int b2 = (dims.dim == 2) ? -1 : (a == 2 ? 1 : 2);
        ntA = (dims.vol[a] + tc.LA - 1) / tc.LA;
        ntB = (dims.vol[b1] + tc.TB - 1) / tc.TB;
        ntC = dims.dim == 3 ? (dims.vol[b2] + tc.TC - 1) / tc.TC : 1;
    }
} // namespace ehp
