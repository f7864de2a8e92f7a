// ehp_joseph_tile.hpp -- per-tile logic of the LDS-tiled Joseph kernels.
//
// Host- and device-callable so that the exact index arithmetic of the kernels can be run
// on the CPU against the oracle (tests/cxx/tile_logic_check.cpp) before it runs on a GPU.
#pragma once

#include "ehp_device.hpp"

namespace ehp
{
    // a * b for operands below 2^24 (one full-rate 24-bit multiply on the GPU)
    EHP_HD inline uint32_t mul24(uint32_t a, uint32_t b)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return __umul24(a, b);
#else
        return a * b;
#endif
    }

    // a * b for |a|, |b| < 2^23 (one 24-bit multiply on the GPU)
    EHP_HD inline int smul24(int a, int b)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return __mul24(a, b);
#else
        return a * b;
#endif
    }

    // The two's-complement 64-bit integer nearest to a * b, for |a * b| < 2^51, in one FMA:
    // a * b + 1.5 * 2^52 is rounded once to an integer (its unit in the last place is 1), which
    // then sits in the low mantissa bits, offset by the constant's bit pattern; that pattern's
    // low word is 0, so removing it takes one 32-bit subtraction from the high word.
    constexpr double kRoundMagic = 6755399441055744.0; // 1.5 * 2^52, bits 0x4338000000000000
    EHP_HD inline unsigned long long fixed_round(double a, double b)
    {
        const double d = __builtin_fma(a, b, kRoundMagic);
        unsigned long long u;
        __builtin_memcpy(&u, &d, 8);
        return u - 0x4338000000000000ull;
    }

    // What the back projection's tap code hands to add(offset, value): the FP32 weight of the tap
    // (TAP_WEIGHT, for a caller that scales it itself), or the tap's 64-bit fixed-point
    // contribution (TAP_FIXED): round(yq * w1 * w2) with the exact bilinear weights 1 - c and c
    // of the FP32 fractions, formed in FP64.
    enum { TAP_WEIGHT = 1, TAP_FIXED = 2 };

    // Fixed-point exponent of a back projection whose largest |y| is not finite: the whole
    // output becomes NaN (integer sums cannot carry a NaN to just the voxels it touches).
    constexpr int kNonFinite = 1 << 20;

    // Tile of the LDS-privatised Joseph kernels: LA planes along the driving axis, TB x TC
    // voxels across (TC unused in 2D).
    struct TileCfg {
        int LA;
        int TB;
        int TC;
    };

    // Compact traversal state of one ray, built once per operator (joseph_ray_table): the
    // position of sample 0 and the step along the two transverse axes b1 < b2 of the ray's own
    // driving axis, the number of samples, the driving axis, the sign of its step, the plane
    // of sample 0 and the step length. Sample s lies on plane P0 + s*sign at transverse
    // position fma(s, step, pos0) -- the same arithmetic as joseph_sample.
    struct alignas(16) RayRec {
        float p1, p2, s1, s2;
        int nsteps;
        int meta; // driving axis in bits 0-1, bit 2 set when the step along it is negative
        int P0;
        float len;
    };
    static_assert(sizeof(RayRec) == 32, "RayRec must be 32 bytes");

    template <int D>
    EHP_HD inline RayRec make_ray_rec(const JRay& jr)
    {
        const int a = jr.axis;
        RayRec r;
        float p[3] = {jr.pos0[0], jr.pos0[1], jr.pos0[2]};
        float st[3] = {jr.step[0], jr.step[1], jr.step[2]};
        // transverse axes of a: (1, 2), (0, 2), (0, 1) in 3D; 1 - a in 2D
        r.p1 = (a == 0) ? p[1] : p[0];
        r.s1 = (a == 0) ? st[1] : st[0];
        r.p2 = (D == 3) ? ((a == 2) ? p[1] : p[D - 1]) : 0.0f;
        r.s2 = (D == 3) ? ((a == 2) ? st[1] : st[D - 1]) : 0.0f;
        r.nsteps = jr.nsteps;
        r.meta = a | ((pick<D>(st, a) < 0.0f) ? 4 : 0);
        r.P0 = (int) (pick<D>(p, a) - 0.5f);
        r.len = jr.len;
        return r;
    }

    // One tile of one pass: planes [A0, A1) along the driving axis a, transverse voxel ranges
    // [B0, B1) along b1 and [C0, C1) along b2 (b1 < b2). The LDS block holds the planes of
    // the tile and the transverse ranges [B0 - 1, B0 + TB] x [C0 - 1, C0 + TC] (one slot of
    // halo on either side), b1 fastest: slot(pa, fc, fb) = (pa * PC + fc) * PB + fb.
    struct TileBox {
        int A0, A1, B0, B1, C0, C1;
        int PB, PC;
        int nslots;
        float B0m1f, C0m1f; // B0 - 1 and C0 - 1 (the first slot's voxel along b1 and b2)
    };

    // The transverse axes b1 < b2 of driving axis A (compile-time, so that per-axis arrays are
    // indexed statically and stay in registers).
    template <int D, int A>
    struct TileAxes {
        static constexpr int b1 = (D == 2) ? 1 - A : (A == 0 ? 1 : 0);
        static constexpr int b2 = (D == 2) ? 0 : (A == 2 ? 1 : 2);
    };

    template <int D, int A>
    EHP_HD inline TileBox make_tile(const Dims& dims, const TileCfg& tc, int ia, int ib, int ic)
    {
        constexpr int b1 = TileAxes<D, A>::b1, b2 = TileAxes<D, A>::b2;
        TileBox t;
        t.A0 = ia * tc.LA;
        t.A1 = dims.vol[A] < t.A0 + tc.LA ? dims.vol[A] : t.A0 + tc.LA;
        t.B0 = ib * tc.TB;
        t.B1 = dims.vol[b1] < t.B0 + tc.TB ? dims.vol[b1] : t.B0 + tc.TB;
        t.C0 = (D == 3) ? ic * tc.TC : 0;
        t.C1 = (D == 3) ? (dims.vol[b2] < t.C0 + tc.TC ? dims.vol[b2] : t.C0 + tc.TC) : 1;
        t.PB = tc.TB + 2;
        t.PC = (D == 3) ? tc.TC + 2 : 1;
        t.nslots = tc.LA * t.PC * t.PB;
        t.B0m1f = (float) (t.B0 - 1);
        t.C0m1f = (float) (t.C0 - 1);
        return t;
    }

    // Number of leading entries k of the increasing sequence f(0..n-1) with f(k) < x (LE:
    // f(k) <= x), by binary lifting: pos + step only moves while the predicate holds, over the
    // bits of the answer from the highest.
    template <bool LE, class F>
    EHP_HD inline int count_below(int n, float x, F f)
    {
        int pos = 0;
        for (int step = n > 0 ? 1 << (31 - __builtin_clz((unsigned) n)) : 0; step > 0; step >>= 1) {
            if (pos + step <= n) {
                const float y = f(pos + step - 1);
                if (LE ? y <= x : y < x)
                    pos += step;
            }
        }
        return pos;
    }

    // Row and column of entry loc of a row-major rectangle nu wide: loc / nu through the
    // reciprocal rnu = 1 / nu, then one correction step. The float quotient has a relative error
    // below 2^-22, so for rows below 2^21 it is within one of the true row; exact for loc < 2^31
    // and nu < 2^23 (the tiled operators keep detectors within these sizes).
    EHP_HD inline void rect_split(int loc, int nu, float rnu, int& row, int& col)
    {
        int r = (int) (((float) loc + 0.5f) * rnu);
        // loc - r * nu lies in [-nu, 2 nu); modulo 2^32 (r * nu itself may pass 2^31)
        int c = (int) ((uint32_t) loc - mul24((uint32_t) r, (uint32_t) nu));
        if (c < 0) {
            r -= 1;
            c += nu;
        } else if (c >= nu) {
            r += 1;
            c -= nu;
        }
        row = r;
        col = c;
    }

    // Detectors the tiled back projections take: rect_split's limits, and a batch of
    // kMaxPoseBatch whole-detector rectangles still counted in an int.
    EHP_HD inline bool tiled_detector_fits(const Dims& dims)
    {
        return dims.det[0] < (1 << 23) && dims.det[1] < (1 << 21) && dims.pix_per_pose < (int64_t) (1 << 26);
    }

    // Candidate pixels of a curved detector whose rays can pass through a region projecting
    // into [umin, umax] x [vmin, vmax] on the planar detector (the ray of pixel (u, v) runs
    // through its planar centre (pu(u), pv(u, v))). Requires what the operator checks when it
    // is created: pu the same on every row and increasing along it, pv increasing down every
    // column, and dims.pix_uc / pix_margin set. Columns: pu within a quarter pixel of [umin,
    // umax]; rows: those where pv at some candidate column can be within a quarter pixel of
    // [vmin, vmax] -- on row v, pv over the columns [u0, u1] lies within pix_margin of its
    // values at u0, u1 and clamp(pix_uc, u0, u1), whose maximum (minimum) increases with v.
    template <int D>
    EHP_HD inline bool curved_rect(const Dims& dims, float umin, float umax, float vmin, float vmax, int& u0,
                                   int& u1, int& v0, int& v1)
    {
        const float* px = dims.pix;
        const int W = dims.det[0];
        const float alo = umin - 0.25f, ahi = umax + 0.25f;
        u0 = count_below<false>(W, alo, [&](int u) { return px[(int64_t) u * (D - 1)]; });
        u1 = count_below<true>(W, ahi, [&](int u) { return px[(int64_t) u * (D - 1)]; }) - 1;
        v0 = v1 = 0;
        if (u0 > u1)
            return false;
        if (D == 3) {
            const int c2 = dims.pix_uc < u0 ? u0 : (dims.pix_uc > u1 ? u1 : dims.pix_uc);
            const float m = 0.25f + dims.pix_margin;
            const float blo = vmin - m, bhi = vmax + m;
            const int64_t W2 = 2 * (int64_t) W;
            const float* c0 = px + 2 * (int64_t) u0 + 1;
            const float* c1 = px + 2 * (int64_t) u1 + 1;
            const float* cc = px + 2 * (int64_t) c2 + 1;
            v0 = count_below<false>(dims.det[1], blo, [&](int v) {
                const int64_t o = v * W2;
                return fmaxf(fmaxf(c0[o], c1[o]), cc[o]);
            });
            v1 = count_below<true>(dims.det[1], bhi, [&](int v) {
                     const int64_t o = v * W2;
                     return fminf(fminf(c0[o], c1[o]), cc[o]);
                 }) - 1;
        }
        return u0 <= u1 && v0 <= v1;
    }

    // Candidate pixels of one pose for the box [lo, hi]: the bounding rectangle of the box's
    // projected corners (a point X lies on the ray of pixel u when P X, homogeneous, maps to
    // u + 0.5, or to its planar centre on a curved detector), widened by a quarter pixel for
    // rounding. Falls back to the whole detector when a corner is not in front of the source.
    template <int D>
    EHP_HD inline bool box_bbox(const PoseGeom& g, const float* lo, const float* hi, const Dims& dims, int& u0,
                                int& u1, int& v0, int& v1)
    {
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
            // approximate reciprocal: a few ulp, covered by the quarter-pixel margin
            const float rden = ehp_rcp(den);
            const float pu = p[0] * rden;
            umin = pu < umin ? pu : umin;
            umax = pu > umax ? pu : umax;
            if (D == 3) {
                const float pv = p[1] * rden;
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
        if (dims.pix)
            return curved_rect<D>(dims, umin, umax, vmin, vmax, u0, u1, v0, v1);
        const float W = (float) dims.det[0] + 4.0f, H = (float) dims.det[1] + 4.0f;
        umin = umin < -4.0f ? -4.0f : (umin > W ? W : umin);
        umax = umax < -4.0f ? -4.0f : (umax > W ? W : umax);
        // pixels u with u + 0.5 in [umin - 0.25, umax + 0.25]
        u0 = (int) ceilf(umin - 0.75f);
        u1 = (int) floorf(umax - 0.25f);
        u0 = u0 < 0 ? 0 : u0;
        u1 = u1 > dims.det[0] - 1 ? dims.det[0] - 1 : u1;
        if (D == 3) {
            vmin = vmin < -4.0f ? -4.0f : (vmin > H ? H : vmin);
            vmax = vmax < -4.0f ? -4.0f : (vmax > H ? H : vmax);
            v0 = (int) ceilf(vmin - 0.75f);
            v1 = (int) floorf(vmax - 0.25f);
            v0 = v0 < 0 ? 0 : v0;
            v1 = v1 > dims.det[1] - 1 ? dims.det[1] - 1 : v1;
        } else {
            v0 = v1 = 0;
        }
        return u0 <= u1 && v0 <= v1;
    }

    // Candidate pixels of one pose for a Joseph tile. A ray touches the tile's slots only with
    // a sample point in [A0 + 0.5, A1 - 0.5] x [B0 - 0.5, B1 + 0.5) x [C0 - 0.5, C1 + 0.5)
    // (sample points lie on plane centres; floor(q - 0.5) in [B0 - 1, B1 - 1] is
    // q in [B0 - 0.5, B1 + 0.5)), and a ray through that box projects inside box_bbox of it.
    template <int D, int A>
    EHP_HD inline bool tile_bbox(const PoseGeom& g, const TileBox& t, const Dims& dims, int& u0, int& u1, int& v0,
                                 int& v1)
    {
        constexpr int b1 = TileAxes<D, A>::b1, b2 = TileAxes<D, A>::b2;
        float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {0.0f, 0.0f, 0.0f};
        lo[A] = (float) t.A0 + 0.5f;
        hi[A] = (float) t.A1 - 0.5f;
        lo[b1] = (float) t.B0 - 0.5f;
        hi[b1] = (float) t.B1 + 0.5f;
        if (D == 3) {
            lo[b2] = (float) t.C0 - 0.5f;
            hi[b2] = (float) t.C1 + 0.5f;
        }
        return box_bbox<D>(g, lo, hi, dims, u0, u1, v0, v1);
    }

    // ---------------------------------------------------------------------------------
    // back projection of one ray into a tile (the matched adjoint of the forward's samples)

    // floor(x) as an int in one instruction (v_cvt_flr_i32_f32); |x| < 2^31
    EHP_HD inline int cvt_flr(float x)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        int r;
        asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
        return r;
#else
        return (int) floorf(x);
#endif
    }

    // x - floor(x) (v_fract_f32); exact for finite x >= 0
    EHP_HD inline float fract_pos(float x)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return __builtin_amdgcn_fractf(x);
#else
        return x - floorf(x);
#endif
    }

    // a + b + c in one instruction (v_add3_u32), c uniform; keeps the compiler from splitting the
    // slot offset back into its parts
    EHP_HD inline uint32_t add3_u32(uint32_t a, uint32_t b, uint32_t c)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        uint32_t r;
        asm("v_add3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
        return r;
#else
        return a + b + c;
#endif
    }

    // a * b + c for |a|, |b| < 2^23 in one instruction (v_mad_i32_i24), b uniform, c in a
    // vector register (one scalar operand per instruction)
    EHP_HD inline int mad_i24(int a, int b, int c)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        int r;
        asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
        return r;
#else
        return a * b + c;
#endif
    }

    // Whether p holds on every active lane of the wavefront (one lane on the CPU).
    EHP_HD inline bool wave_all(bool p)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return __all(p);
#else
        return p;
#endif
    }

    // Offsets of a tile's LDS slots (bytes; slot (pa, i2, i1) at base + pa * plane + i2 * pb8 + 8 i1) in
    // the form the per-sample code adds them: a sample with floor indices g1, g2 touches the
    // tile when g1 is in [B0 - 1, B1 - 1] and g2 in [C0 - 1, C1 - 1], that is when
    // o1 = 8 g1 + nb1 <= max1 and o2 = g2 pb8 + nb2 <= max2 as unsigned numbers.
    struct TileOff {
        uint32_t nb1, max1;
        int nb2;
        uint32_t max2;
        int pb8;
        uint32_t plane;
        uint32_t base; // byte address of slot 0 (the kernels: in LDS; 0 on the CPU)
    };

    EHP_HD inline TileOff tile_offsets(const TileBox& t, uint32_t base = 0)
    {
        TileOff o;
        o.base = base;
        o.pb8 = t.PB * 8;
        o.nb1 = (uint32_t) (-8 * (t.B0 - 1));
        o.max1 = (uint32_t) (8 * (t.B1 - t.B0));
        o.nb2 = -(t.C0 - 1) * o.pb8;
#if defined(__HIP_DEVICE_COMPILE__)
        // mad_i24 takes nb2 from a vector register: copy it there once, not in every sample
        asm("v_mov_b32 %0, %1" : "=v"(o.nb2) : "s"(o.nb2));
#endif
        o.max2 = (uint32_t) ((t.C1 - t.C0) * o.pb8);
        o.plane = (uint32_t) (t.PC * o.pb8);
        return o;
    }

    // Whether elsa's weight form must be evaluated as written in this tile: with B0 >= 1 (and
    // C0 >= 1) every sample the tile keeps has q - 0.5 >= 0, where q - 0.5 is exact and
    // c = (q - floor(q - 0.5)) - 0.5 equals fract(q - 0.5) bit for bit; samples of the tiles on
    // the volume's low faces may have q < 0.5.
    template <int D>
    EHP_HD inline bool tile_low_face(const TileBox& t)
    {
        return t.B0 == 0 || (D == 3 && t.C0 == 0);
    }

    // Back projection of the sample at fs = s (as a float) of ray r lying on the tile plane at
    // byte offset poff: if its bilinear footprint touches the tile, hand add(offset, value) one
    // value for each of its four (2D: two) slots. M = TAP_FIXED: yq is y * len in units of the
    // fixed-point unit q (a power-of-two scaling of elsa's FP32 product y * len) and the value is
    // round(yq * w2 * w1) as a 64-bit integer, with the weights 1 - c and c of the FP32
    // fractions c taken exactly (FP64; the products are exact, only the final rounding is not).
    // M = TAP_WEIGHT: the value is the tap's FP32 weight as elsa forms it (3D: the product of
    // the two transverse weights, the driving axis' weight being 1), for a caller that scales it
    // in FP64. The position and fractions are elsa's (doInterpolation, JosephsMethod.cpp:60-97)
    // with the forward's position rule fma(s, step, pos0), so positions and floors equal the
    // forward's bit for bit (the forward interpolates in lerp form: the exact weights up to FP32
    // rounding). LOW: evaluate c as elsa writes it (tiles on the low faces, tile_low_face);
    // otherwise as fract(q - 0.5), the same value there.
    template <int D, bool LOW, int M, class Y, class Add>
    EHP_HD inline void tile_sample_fx(const RayRec& r, float fs, uint32_t poff, const TileOff& to, Y yq, Add add)
    {
        const float q1 = __builtin_fmaf(fs, r.s1, r.p1);
        const float t1 = q1 - 0.5f;
        const int g1 = cvt_flr(t1);
        const uint32_t o1 = ((uint32_t) g1 << 3) + to.nb1;
        if (o1 > to.max1)
            return;
        if (D == 2) {
            const float c1 = LOW ? (q1 - (float) g1) - 0.5f : fract_pos(t1);
            const uint32_t off = o1 + poff;
            if constexpr (M == TAP_FIXED) {
                const double c1d = c1;
                add(off, fixed_round(yq, 1.0 - c1d));
                add(off + 8u, fixed_round(yq, c1d));
            } else {
                add(off, 1.0f - c1);
                add(off + 8u, c1);
            }
        } else {
            const float q2 = __builtin_fmaf(fs, r.s2, r.p2);
            const float t2 = q2 - 0.5f;
            const int g2 = cvt_flr(t2);
            const uint32_t o2 = (uint32_t) mad_i24(g2, to.pb8, to.nb2);
            if (o2 > to.max2)
                return;
            const float c1 = LOW ? (q1 - (float) g1) - 0.5f : fract_pos(t1);
            const float c2 = LOW ? (q2 - (float) g2) - 0.5f : fract_pos(t2);
            const uint32_t off = add3_u32(o1, o2, poff);
            const uint32_t off2 = off + (uint32_t) to.pb8;
            if constexpr (M == TAP_FIXED) {
                const double c1d = c1, c2d = c2;
                const double w1d = 1.0 - c1d;
                const double a = yq * (1.0 - c2d), b = yq * c2d;
                add(off, fixed_round(a, w1d));
                add(off + 8u, fixed_round(a, c1d));
                add(off2, fixed_round(b, w1d));
                add(off2 + 8u, fixed_round(b, c1d));
            } else {
                const float w1 = 1.0f - c1, w2 = 1.0f - c2;
                add(off, w1 * w2);
                add(off + 8u, c1 * w2);
                add(off2, w1 * c2);
                add(off2 + 8u, c1 * c2);
            }
        }
    }

    // Plane walk of one ray through the la planes of a tile: s0 is the ray's sample index on the
    // tile's first plane and sg the step of the sample index from plane to plane. Each sample in
    // [0, nsteps) goes to tile_sample_fx. FULL: the caller knows that all la samples lie in
    // [0, nsteps) (no test per plane). Every ray visits the planes in the same order, so the
    // lanes of a wavefront address one LDS plane at each step.
    template <int D, bool LOW, bool FULL, int M, class Y, class Add>
    EHP_HD inline void tile_ray_fx(const RayRec& r, int s0, int sg, int la, const TileOff& to, Y yq, Add add)
    {
        float fs = (float) s0;
        const float fsg = (float) sg;
        int s = s0;
        uint32_t poff = to.base;
        for (int pa = 0; pa < la; ++pa, fs += fsg, s += sg, poff += to.plane)
            if (FULL || (uint32_t) s < (uint32_t) r.nsteps)
                tile_sample_fx<D, LOW, M>(r, fs, poff, to, yq, add);
    }

    // Back projection of one candidate ray into a tile (planes [A0, A1) along A): rays of
    // another driving axis, without samples or with yq = 0 add nothing (tile_sample_fx; with
    // TAP_WEIGHT, yq is any value that is 0 exactly when y * len is). Takes the test-free plane
    // walk when every active lane of the wavefront has samples on all the tile's planes.
    template <int D, int A, bool LOW, int M, class Y, class Add>
    EHP_HD inline void tile_candidate_fx(const RayRec& r, const TileBox& t, const TileOff& to, Y yq, Add add)
    {
        if ((r.meta & 3) != A || r.nsteps == 0 || yq == Y(0))
            return;
        const bool neg = (r.meta & 4) != 0;
        const int sg = neg ? -1 : 1;
        const int s0 = neg ? r.P0 - t.A0 : t.A0 - r.P0; // the ray's sample on the tile's first plane
        const int la = t.A1 - t.A0;
        const int s1 = s0 + (la - 1) * sg;
        const int lo = s0 < s1 ? s0 : s1, hi = s0 < s1 ? s1 : s0;
        if (wave_all(lo >= 0 && hi < r.nsteps))
            tile_ray_fx<D, LOW, true, M>(r, s0, sg, la, to, yq, add);
        else
            tile_ray_fx<D, LOW, false, M>(r, s0, sg, la, to, yq, add);
    }

    // Linear volume index of LDS slot i if the tile owns that voxel, else -1.
    template <int D, int A>
    EHP_HD inline int64_t tile_slot_voxel(const TileBox& t, const Dims& dims, int i)
    {
        constexpr int b1 = TileAxes<D, A>::b1, b2 = TileAxes<D, A>::b2;
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
        c[A] = xa;
        c[b1] = xb;
        if (D == 3)
            c[b2] = xc;
        return c[0] * dims.strides[0] + c[1] * dims.strides[1] + c[2] * dims.strides[2];
    }

    // Poses whose candidate rays one workgroup gathers between two barriers (at most
    // kMaxPoseBatch).
    // Pose j of candidate k of a pose batch: the largest j with pref[j] <= k (pref is the
    // prefix sum of the batch's per-pose candidate counts, pref[0] = 0, k < pref[batch]).
    EHP_HD inline int batch_pose_slot(const int* pref, int batch, int k)
    {
        int lo = 0, hi = batch - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (pref[mid] <= k)
                lo = mid;
            else
                hi = mid - 1;
        }
        return lo;
    }

    constexpr int kMaxPoseBatch = 32;

    // Candidate rectangle of one pose of a batch in the tiled back projections: rnu = 1 / width,
    // nu = width, base = ray index of its first pixel.
    struct CandRect {
        float rnu;
        int nu;
        long long base;
    };
    // Dynamic LDS of a tiled back projection after the tile's 8-byte slots: the rectangles,
    // candidate counts and their prefix sums of a batch of `batch` poses, sized by the launch's
    // batch (at 8 x 17 x 17 tiles and 8 poses that keeps a workgroup under a seventh of the
    // CU's 160 KB, where 32 poses' worth would allow only six workgroups per CU).
    EHP_HD constexpr size_t batch_lds_bytes(int batch)
    {
        return ((size_t) batch * sizeof(CandRect) + (size_t) (2 * batch + 1) * sizeof(int) + 7) & ~(size_t) 7;
    }

    struct BatchLds {
        CandRect* rect;
        int *cnt, *pref;
    };

    EHP_HD inline BatchLds batch_lds(long long* lds, int nslots, int batch)
    {
        BatchLds b;
        b.rect = reinterpret_cast<CandRect*>(lds + nslots);
        b.cnt = reinterpret_cast<int*>(b.rect + batch);
        b.pref = b.cnt + batch;
        return b;
    }
    // partial maxima of the fixed-point scale (one per block of the max-|y| launch)
    constexpr int kAbsmaxBlocks = 2048;
    constexpr int kPoseBatch = 8;

    // Upper bound of sum |len * weight| over the contributions one LDS slot can receive from
    // all poses. Per pose a slot receives one sample of each ray whose crossing point with the
    // slot's plane lies within the slot's 2x2 (2D: 2) interpolation footprint; with
    // neighbouring rays at least h apart there are at most (2/h + 2)^(dim-1) of them, each with
    // weight <= 1. h is bounded below by half the smallest angle between neighbouring pixel
    // rays (sampled at the detector corners and centre) times the smallest distance from the
    // source to the volume. On a curved detector (pix: its planar centres, host memory, with
    // the structure curved_candidates checks) neighbouring rays run through planar points at
    // least the smallest planar spacing apart, and the corners are those of the planar centres'
    // extent. Returns a negative value when no useful bound exists (source inside or touching
    // the volume); the caller then uses the one-thread-per-ray kernels.
    inline double fixed_point_slot_bound(const Dims& dims, const PoseGeom* poses, const float* pix = nullptr)
    {
        const int D = dims.dim;
        const int W = dims.det[0], H = dims.det[1];
        double su = 1.0, sv = 1.0, ulo = 0.5, uhi = W - 0.5, vlo = 0.5, vhi = H - 0.5;
        if (pix) {
            const int s = D - 1;
            ulo = pix[0];
            uhi = pix[(int64_t) (W - 1) * s];
            for (int u = 0; u + 1 < W; ++u) {
                const double d = (double) pix[(int64_t) (u + 1) * s] - (double) pix[(int64_t) u * s];
                su = (u == 0 || d < su) ? d : su;
            }
            if (D == 3) {
                auto pv = [&](int u, int v) { return (double) pix[((int64_t) v * W + u) * 2 + 1]; };
                vlo = 1.0e300;
                vhi = -1.0e300;
                bool first = true;
                for (int u = 0; u < W; ++u) {
                    vlo = pv(u, 0) < vlo ? pv(u, 0) : vlo;
                    vhi = pv(u, H - 1) > vhi ? pv(u, H - 1) : vhi;
                    for (int v = 0; v + 1 < H; ++v) {
                        const double d = pv(u, v + 1) - pv(u, v);
                        sv = (first || d < sv) ? d : sv;
                        first = false;
                    }
                }
            }
        }
        double theta = 1.0e300, dmin = 1.0e300;
        for (int p = 0; p < dims.nposes; ++p) {
            const PoseGeom& g = poses[p];
            auto dir = [&](double u, double v, double* n) {
                const double h[3] = {u, D == 3 ? v : 1.0, 1.0};
                double y[3] = {0, 0, 0}, n2 = 0;
                for (int i = 0; i < D; ++i) {
                    for (int j = 0; j < D; ++j)
                        y[i] += (double) g.Pinv[i + j * (D + 1)] * h[j];
                    n2 += y[i] * y[i];
                }
                const double nn = n2 > 0 ? 1.0 / sqrt(n2) : 0.0;
                for (int i = 0; i < D; ++i)
                    n[i] = y[i] * nn;
            };
            const double us[3] = {ulo, uhi, 0.5 * (ulo + uhi)};
            const double vs[3] = {vlo, vhi, 0.5 * (vlo + vhi)};
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < (D == 3 ? 3 : 1); ++b) {
                    double n0[3], n1[3], n2[3];
                    dir(us[a], vs[b], n0);
                    dir(us[a] + su, vs[b], n1);
                    double d1 = 0, d2 = 0;
                    for (int i = 0; i < D; ++i)
                        d1 += (n1[i] - n0[i]) * (n1[i] - n0[i]);
                    theta = sqrt(d1) < theta ? sqrt(d1) : theta;
                    if (D == 3) {
                        dir(us[a], vs[b] + sv, n2);
                        for (int i = 0; i < D; ++i)
                            d2 += (n2[i] - n0[i]) * (n2[i] - n0[i]);
                        theta = sqrt(d2) < theta ? sqrt(d2) : theta;
                    }
                }
            double dist2 = 0;
            for (int i = 0; i < D; ++i) {
                const double c = g.C[i];
                const double lo = 0.0, hi = dims.vol[i];
                const double e = c < lo ? lo - c : (c > hi ? c - hi : 0.0);
                dist2 += e * e;
            }
            const double dd = sqrt(dist2);
            dmin = dd < dmin ? dd : dmin;
        }
        const double h = 0.5 * theta * dmin;
        if (!(h > 1.0e-3))
            return -1.0;
        const double per_pose = D == 3 ? (2.0 / h + 2.0) * (2.0 / h + 2.0) : (2.0 / h + 2.0);
        const double len_max = D == 3 ? 1.7320509 : 1.4142137; // |step| with one component 1
        // at least 4096 (> sqrt(3) 2^11): a single term, at most max|y| len_max, then stays below
        // 2^51 units, the range of fixed_round
        const double b = dims.nposes * per_pose * len_max;
        return b > 4096.0 ? b : 4096.0;
    }

    // Checks that a curved detector's planar centres pix (host memory, dim - 1 per pixel, u
    // fastest) have the structure curved_rect relies on -- all finite, planar u the same on
    // every row and increasing along it, planar v increasing down every column -- and sets
    // dims.pix_uc to the column whose planar v spans the least (the middle of a curved
    // detector) and dims.pix_margin to the least distance that keeps, on every row, pv over any
    // column range within reach of its values at the range's ends and at clamp(pix_uc, .):
    // within [0, uc] (and within [uc, W - 1]) a row that rises by at most r and falls by at
    // most f (later minus earlier entries) stays within min(r, f) of its two ends. Returns
    // false without that structure; the caller then keeps the one-thread-per-ray kernels.
    inline bool curved_candidates(Dims& dims, const float* pix)
    {
        const int D = dims.dim, W = dims.det[0], H = dims.det[1];
        const int s = D - 1;
        for (int u = 0; u < W; ++u) {
            const float a = pix[(int64_t) u * s];
            if (!(a - a == 0.0f) || (u > 0 && !(a > pix[(int64_t) (u - 1) * s])))
                return false;
            for (int v = 1; v < (D == 3 ? H : 1); ++v)
                if (!(pix[((int64_t) v * W + u) * 2] == a))
                    return false;
        }
        dims.pix_uc = 0;
        dims.pix_margin = 0.0f;
        if (D == 2)
            return true;
        auto pv = [&](int u, int v) { return (double) pix[((int64_t) v * W + u) * 2 + 1]; };
        double least = 0.0;
        int uc = 0;
        for (int u = 0; u < W; ++u) {
            for (int v = 0; v < H; ++v)
                if (!(pv(u, v) - pv(u, v) == 0.0) || (v > 0 && !(pv(u, v) > pv(u, v - 1))))
                    return false;
            const double span = pv(u, H - 1) - pv(u, 0);
            if (u == 0 || span < least) {
                least = span;
                uc = u;
            }
        }
        double margin = 0.0;
        for (int v = 0; v < H; ++v)
            for (int part = 0; part < 2; ++part) {
                const int a = part == 0 ? 0 : uc, b = part == 0 ? uc : W - 1;
                double lo = pv(a, v), hi = lo, rise = 0.0, fall = 0.0;
                for (int u = a + 1; u <= b; ++u) {
                    const double x = pv(u, v);
                    rise = x - lo > rise ? x - lo : rise;
                    fall = hi - x > fall ? hi - x : fall;
                    lo = x < lo ? x : lo;
                    hi = x > hi ? x : hi;
                }
                const double m = rise < fall ? rise : fall;
                margin = m > margin ? m : margin;
            }
        float mf = (float) margin;
        if ((double) mf < margin)
            mf = nextafterf(mf, 3.0e38f);
        dims.pix_uc = uc;
        dims.pix_margin = mf;
        return true;
    }

    // Fixed-point format of a back projection: slot sums are 64-bit integers in units of
    // q = 2^(eg - 62), where 2^eg > max|y| * bound, so no sum reaches 2^62 in magnitude.
    // Returns eg; 0 when max|y| is 0 (nothing to add), kNonFinite when it is not finite.

    EHP_HD inline int fixed_point_exponent(double maxabs, double bound)
    {
        const double X = maxabs * bound;
        if (X == 0.0)
            return 0;
        if (!(X <= 1.0e300))
            return kNonFinite;
        int e;
        frexp(X, &e); // X = f 2^e, f in [0.5, 1)
        return e < -900 ? -900 : e;
    }

    EHP_HD inline int fixed_point_exponent(float maxabs, double bound)
    {
        return fixed_point_exponent((double) maxabs, bound);
    }

    // Per-ray scale of the contributions: with |yl| < 2^er (er raised to at least eg - 32),
    // ys = yl * 2^(30 - er), so every contribution ys * w (w <= 1) truncates to an int32 of
    // magnitude at most 2^30, and shifting it left by sh = er - eg + 32 (0..32) puts it in units
    // of q. The truncation error of a contribution is below 2^(er - 30), at most 2^-29 |yl|
    // (2^-30 of the bound 2^er): far below an FP32 rounding of the sums it enters.
    EHP_HD inline void ray_scale(float yl, int eg, float& ys, int& sh)
    {
        int er;
        frexpf(yl, &er);
        er = er < eg - 32 ? eg - 32 : er;
        ys = ldexpf(yl, 30 - er);
        sh = er - eg + 32;
    }

    // The 64-bit fixed-point value of one contribution.
    // Truncation (one conversion instruction) rather than rounding to nearest (two).
    EHP_HD inline unsigned long long fixed_contrib(float t, int sh)
    {
        return (unsigned long long) (long long) (int) t << sh;
    }


    inline void tile_counts(const Dims& dims, const TileCfg& tc, int a, int& ntA, int& ntB, int& ntC)
    {
        const int b1 = (dims.dim == 2) ? 1 - a : (a == 0 ? 1 : 0);
        const int b2 = (dims.dim == 2) ? -1 : (a == 2 ? 1 : 2);
        ntA = (dims.vol[a] + tc.LA - 1) / tc.LA;
        ntB = (dims.vol[b1] + tc.TB - 1) / tc.TB;
        ntC = dims.dim == 3 ? (dims.vol[b2] + tc.TC - 1) / tc.TC : 1;
    }

    // ---------------------------------------------------------------------------------
    // plane-major forward projection

    // Zero-padded plane-major view of the volume for the rays of driving axis A: A slowest,
    // then the in-plane axis m, then the in-plane axis f fastest, with two zero voxels on
    // either side of each in-plane axis (2D: of f only). Voxel (c_A, c_m, c_f) lives at
    // c_A * PS + (c_m + mpad) * RS + (c_f + 2). A tap outside the volume then reads a zero
    // instead of needing a bounds test, which gives the same sum as elsa's zero weight for
    // finite volumes.
    //
    // Row-pair views (pairs = 1, the 3D Joseph forward): row m of a plane holds, for each f, the
    // two elements (m, f) and (m + 1, f) side by side, so that the 2 x 2 footprint of a sample,
    // (m, f), (m + 1, f), (m, f + 1), (m + 1, f + 1), is 16 contiguous bytes (8 in bfloat16): one
    // load instead of two. Each voxel is stored twice (slot 0 of its own row, slot 1 of the row
    // before).
    struct PadLayout {
        int A, m, f;
        int mpad;      // 2 in 3D, 0 in 2D
        int pairs;     // 1: row-pair view
        uint32_t RS;   // row stride in elements ((n_f + 4, rounded up to even) times 1 + pairs)
        uint32_t PS;   // plane stride
        int64_t size;  // elements in the view
    };

    inline PadLayout pad_layout(const Dims& dims, int A, int f, bool pairs = false)
    {
        PadLayout L{};
        L.A = A;
        L.f = f;
        L.m = 3 - A - f; // in 2D the third axis has one voxel
        L.mpad = dims.dim == 3 ? 2 : 0;
        L.pairs = pairs && dims.dim == 3 ? 1 : 0;
        // rows of an even number of elements, so that with bfloat16 elements every row and plane
        // starts on a 4-byte boundary
        L.RS = (uint32_t) ((dims.vol[f] + 5) & ~1) << L.pairs;
        L.PS = L.RS * (uint32_t) (dims.vol[L.m] + 2 * L.mpad);
        L.size = (int64_t) L.PS * dims.vol[A];
        return L;
    }

    // Offset in the view of voxel (c0, c1, c2) of the input.
    EHP_HD inline int64_t pad_index(const PadLayout& L, int c0, int c1, int c2)
    {
        const int c[3] = {c0, c1, c2};
        const int ca = L.A == 0 ? c[0] : (L.A == 1 ? c[1] : c[2]);
        const int cm = L.m == 0 ? c[0] : (L.m == 1 ? c[1] : c[2]);
        const int cf = L.f == 0 ? c[0] : (L.f == 1 ? c[1] : c[2]);
        return (int64_t) ca * L.PS + (int64_t) (cm + L.mpad) * L.RS + ((int64_t) (cf + 2) << L.pairs);
    }

    // Store voxel (c0, c1, c2) in its view: at pad_index, and in a row-pair view also in slot 1
    // of the row before (which exists: rows start at m = -mpad = -2).
    template <class E>
    EHP_HD inline void pad_store(const PadLayout& L, E* out, int c0, int c1, int c2, E v)
    {
        const int64_t i = pad_index(L, c0, c1, c2);
        out[i] = v;
        if (L.pairs)
            out[i - (int64_t) L.RS + 1] = v;
    }

    // Bytes allocated past the last element of a bfloat16 view (ld_view2_bf16 reads up to 4).
    constexpr int kViewSlack = 8;

    // What the kernels read of a view; strides in bytes (a view is below 4 GiB). Views hold
    // FP32, bfloat16 or FP64 elements.
    struct PadView {
        const void* base;
        uint32_t PS4, RS4;
        int64_t size; // elements in the view (bounds checks of the CPU build)
    };

#if defined(EHP_VIEW_CHECK)
    // CPU check builds: set when a read leaves its view
    extern bool g_view_oob;
#endif

    // Two neighbouring floats of a view (one 8-byte load; 4-byte alignment suffices).
    struct __attribute__((aligned(4))) Pair {
        float a, b;
    };

    // The two elements at byte offset off from pb, a pointer into view V: pb is the same for a
    // whole wavefront (a plane or row start, held in scalar registers), off a 32-bit offset.
    EHP_HD inline Pair ld_view2(const PadView& V, const char* pb, uint32_t off)
    {
#if defined(EHP_VIEW_CHECK)
        const int64_t o = (int64_t) (pb - (const char*) V.base) + off;
        if ((o & 3) || o < 0 || (o >> 2) + 1 >= V.size) {
            g_view_oob = true;
            return Pair{0.0f, 0.0f};
        }
#endif
        Pair p;
        __builtin_memcpy(&p, pb + off, sizeof(Pair));
        return p;
    }

    // Two neighbouring bfloat16 elements of a view, widened to FP32. pb is 4-byte aligned (rows
    // have an even number of elements), the pair half the time not: the GPU reads the aligned
    // 8 bytes around it and takes the pair out with one byte-align, instead of a misaligned
    // 4-byte load. Those 8 bytes can reach 4 bytes past the view's last element, so bfloat16
    // views are allocated with kViewSlack bytes to spare.
    EHP_HD inline Pair ld_view2_bf16(const PadView& V, const char* pb, uint32_t off)
    {
#if defined(EHP_VIEW_CHECK)
        const int64_t o = (int64_t) (pb - (const char*) V.base) + off;
        if ((o & 1) || o < 0 || (o >> 1) + 1 >= V.size || ((pb - (const char*) V.base) & 3)) {
            g_view_oob = true;
            return Pair{0.0f, 0.0f};
        }
#endif
#if defined(__HIP_DEVICE_COMPILE__)
        uint32_t w[2];
        __builtin_memcpy(w, __builtin_assume_aligned(pb + (off & ~3u), 4), sizeof(w));
        const uint32_t pair = __builtin_amdgcn_alignbyte(w[1], w[0], off);
        return Pair{__uint_as_float(pair << 16), __uint_as_float(pair & 0xffff0000u)};
#else
        uint16_t h[2];
        __builtin_memcpy(h, pb + off, sizeof(h));
        return Pair{bf16_to_float(bf16{h[0]}), bf16_to_float(bf16{h[1]})};
#endif
    }

    template <bool V16>
    EHP_HD inline Pair ld_pair(const PadView& V, const char* pb, uint32_t off)
    {
        return V16 ? ld_view2_bf16(V, pb, off) : ld_view2(V, pb, off);
    }

    // The 2 x 2 footprint of a row-pair view at byte offset off from pb: (m, f), (m + 1, f),
    // (m, f + 1), (m + 1, f + 1). One 16-byte load; 8-byte alignment suffices.
    struct __attribute__((aligned(8))) Quad {
        float a, b, c, d;
    };

    // The same footprint in a bfloat16 row-pair view: one 8-byte load (4-byte alignment suffices),
    // widened to FP32.
    EHP_HD inline Quad ld_view4_bf16(const PadView& V, const char* pb, uint32_t off)
    {
#if defined(EHP_VIEW_CHECK)
        const int64_t o = (int64_t) (pb - (const char*) V.base) + off;
        if ((o & 3) || o < 0 || (o >> 1) + 3 >= V.size) {
            g_view_oob = true;
            return Quad{0.0f, 0.0f, 0.0f, 0.0f};
        }
#endif
        uint32_t w[2];
        __builtin_memcpy(w, pb + off, sizeof(w));
        // element (m, f) is the low half of w[0], (m + 1, f) its high half; w[1] holds f + 1
#if defined(__HIP_DEVICE_COMPILE__)
        return Quad{__uint_as_float(w[0] << 16), __uint_as_float(w[0] & 0xffff0000u), __uint_as_float(w[1] << 16),
                    __uint_as_float(w[1] & 0xffff0000u)};
#else
        return Quad{bf16_to_float(bf16{(uint16_t) w[0]}), bf16_to_float(bf16{(uint16_t) (w[0] >> 16)}),
                    bf16_to_float(bf16{(uint16_t) w[1]}), bf16_to_float(bf16{(uint16_t) (w[1] >> 16)})};
#endif
    }

    EHP_HD inline Quad ld_view4(const PadView& V, const char* pb, uint32_t off)
    {
#if defined(EHP_VIEW_CHECK)
        const int64_t o = (int64_t) (pb - (const char*) V.base) + off;
        if ((o & 7) || o < 0 || (o >> 2) + 3 >= V.size) {
            g_view_oob = true;
            return Quad{0.0f, 0.0f, 0.0f, 0.0f};
        }
#endif
        Quad q;
        __builtin_memcpy(&q, pb + off, sizeof(Quad));
        return q;
    }

    // Two neighbouring doubles of an FP64 view (one 16-byte load; 8-byte alignment suffices).
    struct __attribute__((aligned(8))) Pair64 {
        double a, b;
    };

    EHP_HD inline Pair64 ld_view2_f64(const PadView& V, const char* pb, uint32_t off)
    {
#if defined(EHP_VIEW_CHECK)
        const int64_t o = (int64_t) (pb - (const char*) V.base) + off;
        if ((o & 7) || o < 0 || (o >> 3) + 1 >= V.size) {
            g_view_oob = true;
            return Pair64{0.0, 0.0};
        }
#endif
        Pair64 p;
        __builtin_memcpy(&p, pb + off, sizeof(Pair64));
        return p;
    }

    // floor value g clamped to [-2, n]: taps outside the volume land in the zero padding
    EHP_HD inline int pad_clamp(float g, int n)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        const float c = __builtin_amdgcn_fmed3f(g, -2.0f, (float) n);
#else
        const float c = g < -2.0f ? -2.0f : (g > (float) n ? (float) n : g);
#endif
        return (int) c;
    }

    // Samples [s0, s1) of a ray whose planes lie inside the volume of na planes along its
    // driving axis (a ray entering exactly through the far face has samples outside, which
    // contribute 0). Empty when s0 >= s1.
    EHP_HD inline void plane_sample_range(const RayRec& r, int na, int& s0, int& s1)
    {
        const bool neg = (r.meta & 4) != 0;
        s0 = neg ? r.P0 - na + 1 : -r.P0;
        s1 = neg ? r.P0 + 1 : na - r.P0;
        s0 = s0 > 0 ? s0 : 0;
        s1 = s1 < r.nsteps ? s1 : r.nsteps;
    }

    // The taps of one sample and its fractional positions: v00 = (F1, F2), v10 = (C1, F2),
    // v01 = (F1, C2), v11 = (C1, C2) along the transverse axes b1, b2 of its driving axis, C = F + 1,
    // and c1, c2 = position - 0.5 - F along b1, b2 (2D: v00, v10 and c1 only).
    struct SampleTaps {
        float v00, v10, v01, v11;
        float c1, c2;
    };

    // floor(t) clamped to the padded range [-2, n] of a view (t finite), as floor(med3(t, -2, n)):
    // the same integer, since -2 and n are integers and floor is monotone (two instructions)
    EHP_HD inline int pad_floor(float t, int n)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return cvt_flr(__builtin_amdgcn_fmed3f(t, -2.0f, (float) n));
#else
        const float c = t < -2.0f ? -2.0f : (t > (float) n ? (float) n : t);
        return (int) floorf(c);
#endif
    }

    // The fraction t - floor(t) of a sample position t = q - 0.5 in one instruction (v_fract_f32):
    // exact for t >= 0; for t < 0 (q below 0.5, at the volume's low faces) within one rounding.
    EHP_HD inline float pad_fract(float t)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return __builtin_amdgcn_fractf(t);
#else
        return t - floorf(t);
#endif
    }

    template <int D, int A, bool FB2, bool V16, bool PR = false>
    EHP_HD inline SampleTaps plane_sample_taps(const RayRec& r, const Dims& dims, const PadView& V, uint32_t poff,
                                               float fs)
    {
        constexpr uint32_t ES = V16 ? 1u : 2u; // log2 of the element size
        constexpr int b1 = TileAxes<D, A>::b1, b2 = TileAxes<D, A>::b2;
        const int n1 = dims.vol[b1], n2 = (D == 3) ? dims.vol[b2] : 1;
        const char* pb = (const char*) V.base + poff; // the plane (the same for the whole wavefront)
        const float q1 = __builtin_fmaf(fs, r.s1, r.p1);
        const float t1 = q1 - 0.5f;
        const int k1 = pad_floor(t1, n1);
        SampleTaps T;
        T.c1 = pad_fract(t1);
        if (D == 2) {
            // offset of element k1 + 2 of the plane's only row
            const Pair fk = ld_pair<V16>(V, pb, ((uint32_t) k1 << ES) + (2u << ES));
            T.v00 = fk.a;
            T.v10 = fk.b;
            T.v01 = T.v11 = T.c2 = 0.0f;
        } else {
            const float q2 = __builtin_fmaf(fs, r.s2, r.p2);
            const float t2 = q2 - 0.5f;
            const int k2 = pad_floor(t2, n2);
            T.c2 = pad_fract(t2);
            // the F and F + 1 taps along the fast axis are neighbours: two 8-byte loads, one per
            // row of the slow axis; (slow + 2) * RS + (fast + 2) elements in, the 2s being the padding
            const int ks = FB2 ? k1 : k2, kf = FB2 ? k2 : k1;
            Pair r0, r1;
            if (PR) {
                // row-pair view: element (ks, kf) holds rows ks and ks + 1 (8 bytes per element in
                // FP32, 4 in bfloat16)
                constexpr uint32_t EP = ES + 1; // log2 of the pair element's size
                const uint32_t o = (uint32_t) mad_i24(ks, (int) V.RS4, (int) (((uint32_t) kf << EP) +
                                                                              (2u * V.RS4 + (2u << EP))));
                const Quad q = V16 ? ld_view4_bf16(V, pb, o) : ld_view4(V, pb, o);
                r0 = Pair{q.a, q.c};
                r1 = Pair{q.b, q.d};
            } else {
                const uint32_t o =
                    (uint32_t) (smul24(ks, (int) V.RS4) + (int) (2u * V.RS4 + (2u << ES))) + ((uint32_t) kf << ES);
                r0 = ld_pair<V16>(V, pb, o);
                r1 = ld_pair<V16>(V, pb + V.RS4, o);
            }
            // FB2: rows are b1 = F1, F1 + 1, pairs run along b2; else rows are b2, pairs along b1
            T.v00 = r0.a;
            T.v10 = FB2 ? r1.a : r0.b;
            T.v01 = FB2 ? r0.b : r1.a;
            T.v11 = r1.b;
        }
        return T;
    }

    // acc plus the sample's bilinear (2D: linear) interpolation times len: the interpolation in
    // FP32 lerp form (along b1 with c1, then in 3D along b2 with c2), the sum in FP64 (one
    // conversion and one FMA per sample). An FP32 sum would round every step the same way where
    // the volume is constant (equal increments), an error growing with the number of samples.
    template <int D>
    EHP_HD inline double plane_taps_add(const SampleTaps& T, double len, double acc)
    {
        const float a0 = __builtin_fmaf(T.c1, T.v10 - T.v00, T.v00);
        if (D == 2)
            return __builtin_fma((double) a0, len, acc);
        const float a1 = __builtin_fmaf(T.c1, T.v11 - T.v01, T.v01);
        return __builtin_fma((double) __builtin_fmaf(T.c2, a1 - a0, a0), len, acc);
    }

    // The taps of one sample of an FP64 view (one element per voxel) and elsa's fractions
    // c = (q - F) - 0.5 (doInterpolation's form, also for q < 0.5 at the low faces).
    struct SampleTaps64 {
        double v00, v10, v01, v11;
        float c1, c2;
    };

    template <int D, int A, bool FB2>
    EHP_HD inline SampleTaps64 plane_sample_taps64(const RayRec& r, const Dims& dims, const PadView& V,
                                                   uint32_t poff, float fs)
    {
        constexpr int b1 = TileAxes<D, A>::b1, b2 = TileAxes<D, A>::b2;
        const int n1 = dims.vol[b1], n2 = (D == 3) ? dims.vol[b2] : 1;
        const char* pb = (const char*) V.base + poff;
        const float q1 = __builtin_fmaf(fs, r.s1, r.p1);
        const float g1 = floorf(q1 - 0.5f);
        const int k1 = pad_clamp(g1, n1);
        SampleTaps64 T;
        T.c1 = (q1 - g1) - 0.5f;
        if (D == 2) {
            const Pair64 fk = ld_view2_f64(V, pb, ((uint32_t) k1 << 3) + 16u);
            T.v00 = fk.a;
            T.v10 = fk.b;
            T.v01 = T.v11 = 0.0;
            T.c2 = 0.0f;
        } else {
            const float q2 = __builtin_fmaf(fs, r.s2, r.p2);
            const float g2 = floorf(q2 - 0.5f);
            const int k2 = pad_clamp(g2, n2);
            T.c2 = (q2 - g2) - 0.5f;
            const int ks = FB2 ? k1 : k2, kf = FB2 ? k2 : k1;
            const uint32_t o = (uint32_t) (smul24(ks, (int) V.RS4) + (int) (2u * V.RS4 + 16u)) + ((uint32_t) kf << 3);
            const Pair64 r0 = ld_view2_f64(V, pb, o);
            const Pair64 r1 = ld_view2_f64(V, pb + V.RS4, o);
            T.v00 = r0.a;
            T.v10 = FB2 ? r1.a : r0.b;
            T.v01 = FB2 ? r0.b : r1.a;
            T.v11 = r1.b;
        }
        return T;
    }

    // acc plus one FP64 sample as elsa sums it (JosephsMethod.cpp:60-97 with data_t = double):
    // (value * len) * w per tap in doInterpolation's tap order, w the FP32 product of the
    // transverse weights (the driving axis' weight is 1). 2D taps (F0, C1), (C0, F1); 3D (F0,
    // F1, F2), (F0, C1, C2), (C0, F1, C2), (C0, C1, F2), the driving axis' F = C being its plane.
    template <int D, int A>
    EHP_HD inline double plane_taps_add64(const SampleTaps64& T, double len, double acc)
    {
        const float w1 = 1.0f - T.c1;
        auto tap = [&](double v, float w) { acc = acc + (v * len) * (double) w; };
        if (D == 2) {
            if (A == 0) { // (plane, C1), (plane, F1)
                tap(T.v10, T.c1);
                tap(T.v00, w1);
            } else { // (F0, plane), (C0, plane)
                tap(T.v00, w1);
                tap(T.v10, T.c1);
            }
            return acc;
        }
        const float w2 = 1.0f - T.c2;
        if (A == 0) { // b1 = 1, b2 = 2: (F1, F2), (C1, C2), (F1, C2), (C1, F2)
            tap(T.v00, w1 * w2);
            tap(T.v11, T.c1 * T.c2);
            tap(T.v01, w1 * T.c2);
            tap(T.v10, T.c1 * w2);
        } else if (A == 1) { // b1 = 0, b2 = 2: (F0, F2), (F0, C2), (C0, C2), (C0, F2)
            tap(T.v00, w1 * w2);
            tap(T.v01, w1 * T.c2);
            tap(T.v11, T.c1 * T.c2);
            tap(T.v10, T.c1 * w2);
        } else { // b1 = 0, b2 = 1: (F0, F1), (F0, C1), (C0, F1), (C0, C1)
            tap(T.v00, w1 * w2);
            tap(T.v01, w1 * T.c2);
            tap(T.v10, T.c1 * w2);
            tap(T.v11, T.c1 * T.c2);
        }
        return acc;
    }

    // Sample s (fs = s) of a ray (driving axis A) on the plane at byte offset poff of its padded
    // view, added to acc; FB2 says that the fast in-plane axis of the view is b2 (else b1). The
    // interpolation is elsa's doInterpolation (JosephsMethod.cpp:60-97: bilinear in the
    // transverse axes with floor F = floor(q - 0.5) and fraction q - 0.5 - F, the driving axis on
    // its plane, taps outside the volume 0). FP32 and bfloat16 views (F64 = false): in lerp form,
    // which needs 7 operations per 3D sample where elsa's weight products need 16, summed in
    // FP64: elsa's sums to FP32 rounding (not bit for bit). FP64 views (F64): elsa's weights and
    // sums, plane_taps_add64. C = F + 1 is read even where elsa's ceil gives C = F: its weight is
    // 0 then. Positions and floors are those of joseph_sample and the tiled back projection, bit
    // for bit.
    template <int D, int A, bool FB2, bool V16, bool PR, bool F64>
    EHP_HD inline double plane_sample_add(const RayRec& r, const Dims& dims, const PadView& V, uint32_t poff,
                                          float fs, double len, double acc)
    {
        if constexpr (F64)
            return plane_taps_add64<D, A>(plane_sample_taps64<D, A, FB2>(r, dims, V, poff, fs), len, acc);
        else
            return plane_taps_add<D>(plane_sample_taps<D, A, FB2, V16, PR>(r, dims, V, poff, fs), len, acc);
    }

    // Forward projection of one ray (driving axis A) through its padded view: the sum of its
    // samples in the volume (plane_sample_add), in traversal order, in FP64 (the FP32 and
    // bfloat16 operators round it once).
    template <int D, int A, bool FB2, bool V16 = false, bool PR = false, bool F64 = false>
    EHP_HD inline double plane_ray_forward(const RayRec& r, const Dims& dims, const PadView& V)
    {
        int s0, s1;
        plane_sample_range(r, dims.vol[A], s0, s1);
        double acc = 0.0;
        if (s0 >= s1)
            return 0.0;
        const double len = r.len;
        const bool neg = (r.meta & 4) != 0;
        uint32_t poff = (uint32_t) (neg ? r.P0 - s0 : r.P0 + s0) * V.PS4;
        const uint32_t dp = neg ? 0u - V.PS4 : V.PS4;
        for (int s = s0; s < s1; ++s, poff += dp)
            acc = plane_sample_add<D, A, FB2, V16, PR, F64>(r, dims, V, poff, (float) s, len, acc);
        return acc;
    }

    // A ray's samples in traversal coordinates, t = plane when it steps towards higher planes and
    // t = -plane when it steps towards lower ones (NEG), so that t rises by one per sample:
    // samples s0 .. s1 - 1 lie at t = tfirst .. tlast. False for a ray with no sample in the volume.
    EHP_HD inline bool plane_ray_span(const RayRec& r, int na, int& tfirst, int& tlast, int& s0)
    {
        int s1;
        plane_sample_range(r, na, s0, s1);
        if (s0 >= s1)
            return false;
        const bool neg = (r.meta & 4) != 0;
        tfirst = neg ? s0 - r.P0 : r.P0 + s0;
        tlast = tfirst + (s1 - s0 - 1);
        return true;
    }

    // plane_ray_forward with the plane loop shared by a group of rays of one driving axis and
    // direction (a wavefront, on the GPU): iteration it visits the plane at traversal coordinate
    // t0 + it, the same plane for every ray of the group, so their reads fall on neighbouring
    // addresses; the ray adds its sample s = s_first + it when s0 <= s < s1. t0 and niter span
    // the group's [min tfirst, max tlast] (plane_ray_span), s_first = s0 + t0 - tfirst. In
    // iterations [full0, full1) every ray of the group has a sample (full0 = max tfirst - t0,
    // full1 = max(full0, min tlast - t0 + 1)), so they need no test. Same samples, positions
    // and order as plane_ray_forward, so the same sum bit for bit.
    template <int D, int A, bool NEG, bool FB2, bool V16 = false, bool PR = false, bool F64 = false>
    EHP_HD inline double plane_span_forward(const RayRec& r, const Dims& dims, const PadView& V, int t0, int niter,
                                            int s_first, int full0, int full1)
    {
        int s0, s1;
        plane_sample_range(r, dims.vol[A], s0, s1);
        const uint32_t n = (uint32_t) (s1 - s0);
        uint32_t j = (uint32_t) (s_first - s0); // the ray's sample index minus s0 at this iteration
        float fs = (float) s_first;            // its sample index (exact: |s| < 2^24)
        uint32_t poff = (uint32_t) (NEG ? -t0 : t0) * V.PS4;
        const uint32_t dp = NEG ? 0u - V.PS4 : V.PS4;
        const double len = r.len;
        double acc = 0.0;
        int it = 0;
        for (; it < full0; ++it, ++j, fs += 1.0f, poff += dp)
            if (j < n)
                acc = plane_sample_add<D, A, FB2, V16, PR, F64>(r, dims, V, poff, fs, len, acc);
        // two samples per step, both fetched before either is added, so that each wavefront
        // has four loads in flight
        for (; it + 1 < full1; it += 2, fs += 2.0f, poff += 2u * dp) {
            if constexpr (F64) {
                const SampleTaps64 Ta = plane_sample_taps64<D, A, FB2>(r, dims, V, poff, fs);
                const SampleTaps64 Tb = plane_sample_taps64<D, A, FB2>(r, dims, V, poff + dp, fs + 1.0f);
                acc = plane_taps_add64<D, A>(Ta, len, acc);
                acc = plane_taps_add64<D, A>(Tb, len, acc);
            } else {
                const SampleTaps Ta = plane_sample_taps<D, A, FB2, V16, PR>(r, dims, V, poff, fs);
                const SampleTaps Tb = plane_sample_taps<D, A, FB2, V16, PR>(r, dims, V, poff + dp, fs + 1.0f);
                acc = plane_taps_add<D>(Ta, len, acc);
                acc = plane_taps_add<D>(Tb, len, acc);
            }
        }
        for (; it < full1; ++it, fs += 1.0f, poff += dp)
            acc = plane_sample_add<D, A, FB2, V16, PR, F64>(r, dims, V, poff, fs, len, acc);
        j += (uint32_t) (full1 - full0);
        for (; it < niter; ++it, ++j, fs += 1.0f, poff += dp)
            if (j < n)
                acc = plane_sample_add<D, A, FB2, V16, PR, F64>(r, dims, V, poff, fs, len, acc);
        return acc;
    }
} // namespace ehp
