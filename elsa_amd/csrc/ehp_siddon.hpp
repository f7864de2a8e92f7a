// ehp_siddon.hpp -- Siddon traversal shared by the GPU kernels and the CPU checks.
//
// The traversal of elsa's TraverseAABB (elsa/projectors/TraverseAABB.cpp:6-184): the voxel
// containing the entry point (a coordinate on a voxel face counts to the lower voxel when the
// ray runs downwards, :136-147), the parameters T[i] of the next face crossing per axis, and
// per step the axes whose T is smallest -- ties step several axes at once (:59-69) -- with the
// segment length tExit - tEntry, tExit starting at 0 (TraverseAABB.h). The crossing
// parameters use the direct rule T[i](k) = fma(k, tDelta[i], T0[i]) after k steps along axis i
// (the oracle's MODE_DIRECT; elsa adds tDelta repeatedly). The state at any point of a ray
// then follows from its step counts, so a tile can restart a ray exactly where it enters.
#pragma once

#include "ehp_joseph_tile.hpp"

namespace ehp
{
    constexpr float kSidMax = 3.4028234663852886e+38f; // numeric_limits<float>::max()

    // Traversal state of one ray after set-up, 32 bytes (one record of the per-operator table).
    struct alignas(16) SidRec {
        float T0[3];     // T of the first face crossing per axis (kSidMax: axis never steps)
        float dt[3];     // tDelta per axis (kSidMax: axis never steps)
        short cur0[3];   // first voxel
        unsigned short meta; // bits 2i..2i+1: step of axis i (0: none, 1: +1, 2: -1); bit 6: the
                             // ray traverses the volume; bits 7-8: driving axis (largest |d|)
    };
    static_assert(sizeof(SidRec) == 32, "SidRec must be 32 bytes");

    EHP_HD inline int sid_step(const SidRec& r, int i)
    {
        const int s = (r.meta >> (2 * i)) & 3;
        return s == 1 ? 1 : (s == 2 ? -1 : 0);
    }
    EHP_HD inline bool sid_hit(const SidRec& r)
    {
        return (r.meta >> 6) & 1;
    }
    EHP_HD inline int sid_axis(const SidRec& r)
    {
        return (r.meta >> 7) & 3;
    }

    // T of axis i after k steps
    EHP_HD inline float sid_T(const SidRec& r, int i, int k)
    {
        return r.dt[i] == kSidMax ? kSidMax : __builtin_fmaf((float) k, r.dt[i], r.T0[i]);
    }

    // Set-up of TraverseAABB for the ray o + t d (d normalised) against [0, vol).
    template <int D>
    EHP_HD inline SidRec siddon_setup(const float* o, const float* d, const int* vol)
    {
        SidRec r;
        for (int i = 0; i < 3; ++i) {
            r.T0[i] = kSidMax;
            r.dt[i] = kSidMax;
            r.cur0[i] = 0;
        }
        int a = 0;
        float m = fabsf(d[0]);
        if (fabsf(d[1]) > m) {
            m = fabsf(d[1]);
            a = 1;
        }
        if (D == 3 && fabsf(d[D - 1]) > m)
            a = 2;
        unsigned meta = (unsigned) a << 7;
        float tmin, tmax;
        bool inbox = intersect<D>(o, d, vol, tmin, tmax);
        if (inbox) {
            float entry[3], cur[3];
            int stepdir[3];
#pragma unroll
            for (int i = 0; i < D; ++i) {
                const float hi = (float) vol[i];
                entry[i] = o[i] + tmin * d[i];
                entry[i] = entry[i] < 0.0f ? 0.0f : entry[i];
                entry[i] = entry[i] > hi ? hi : entry[i];
                stepdir[i] = (int) sgnf(d[i]);
            }
#pragma unroll
            for (int i = 0; i < D; ++i) {
                float lower = floorf(entry[i]);
                if (lower == entry[i] && stepdir[i] < 0)
                    lower = lower - 1.0f;
                cur[i] = lower;
                if (cur[i] >= (float) vol[i] || cur[i] < 0.0f)
                    inbox = false;
            }
            const float EPS = 1.1920928955078125e-07f; // numeric_limits<float>::epsilon()
#pragma unroll
            for (int i = 0; i < D; ++i) {
                const bool moves = fabsf(d[i]) > EPS;
                r.dt[i] = moves ? (float) stepdir[i] / d[i] : kSidMax;
                r.T0[i] = moves ? ((d[i] > 0.0f ? cur[i] + 1.0f : cur[i]) - entry[i]) / d[i] : kSidMax;
                r.cur0[i] = inbox ? (short) cur[i] : (short) 0;
                const int st = moves ? stepdir[i] : 0;
                meta |= (unsigned) (st == 1 ? 1 : (st == -1 ? 2 : 0)) << (2 * i);
            }
        }
        if (inbox)
            meta |= 1u << 6;
        r.meta = (unsigned short) meta;
        return r;
    }

    // The whole traversal: fn(c0, c1, c2, w) per crossed voxel (c2 = 0 in 2D), in ray order.
    // An axis whose |d| <= epsilon never steps (its T stays kSidMax), as in elsa.
    template <int D, class Fn>
    EHP_HD inline void siddon_traverse(const SidRec& r, const int* vol, Fn fn)
    {
        if (!sid_hit(r))
            return;
        int cur[3] = {r.cur0[0], r.cur0[1], D == 3 ? r.cur0[2] : 0};
        int k[3] = {0, 0, 0};
        float Tn[3] = {r.T0[0], r.T0[1], D == 3 ? r.T0[2] : kSidMax};
        float tExit = 0.0f;
        // a ray crosses at most vol[0] + vol[1] + vol[2] voxels; the cap only matters for
        // degenerate (non-finite) rays
        int guard = vol[0] + vol[1] + (D == 3 ? vol[2] : 0) + 2;
        bool inbox = true;
        while (inbox && guard-- > 0) {
            const float tEntry = tExit;
            float tm = Tn[0];
#pragma unroll
            for (int i = 1; i < D; ++i)
                tm = fminf(tm, Tn[i]);
            tExit = tm;
            bool mask[3];
#pragma unroll
            for (int i = 0; i < D; ++i) {
                bool mi = true;
#pragma unroll
                for (int j = 0; j < D; ++j)
                    if (j != i)
                        mi = mi && (Tn[i] <= Tn[j]);
                mask[i] = mi;
            }
            fn(cur[0], cur[1], cur[2], tExit - tEntry);
#pragma unroll
            for (int i = 0; i < D; ++i) {
                if (mask[i]) {
                    cur[i] += sid_step(r, i);
                    k[i] += 1;
                    Tn[i] = sid_T(r, i, k[i]);
                }
            }
#pragma unroll
            for (int i = 0; i < D; ++i)
                inbox = inbox && cur[i] < vol[i] && cur[i] >= 0;
        }
    }

    // A Siddon traversal that can pause between the planes of one axis: the state of
    // siddon_traverse after some of its steps (voxel, step direction and count per axis as a
    // float -- exact, as counts stay below 2^24 -- next crossing per axis, the last crossing).
    struct SidWalk {
        int cur[3];
        int st[3];
        float fk[3];
        float Tn[3];
        float tExit;
        bool alive;
    };

    template <int D>
    EHP_HD inline void sid_walk_init(const SidRec& r, SidWalk& w)
    {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            w.cur[i] = i < D ? r.cur0[i] : 0;
            w.st[i] = i < D ? sid_step(r, i) : 0;
            w.fk[i] = 0.0f;
            w.Tn[i] = i < D ? r.T0[i] : kSidMax;
        }
        w.tExit = 0.0f;
        w.alive = sid_hit(r);
    }

    // Steps of one plane: a ray crossing plane p of its driving axis A (the largest |d|) in it
    // moves less than one voxel along every other axis, so it visits at most D voxels there.
    constexpr int kSidPlaneSteps = 8;

    // The steps of siddon_traverse whose voxel lies on plane p of axis A, fn(c0, c1, c2, w) for
    // each in ray order, leaving the walk at the first voxel beyond that plane (or ended). The
    // same arithmetic as siddon_traverse step by step, so walking plane after plane gives its
    // crossings bit for bit. (Only an axis that moves can have the smallest T, as the driving
    // axis moves, so the update needs no test for a non-moving one.) At most kSidPlaneSteps steps
    // per call, which also stops a degenerate ray.
    template <int D, class Fn>
    EHP_HD inline void sid_walk_plane(const SidRec& r, const int* vol, int A, int p, SidWalk& w, Fn fn)
    {
        for (int g = 0; g < kSidPlaneSteps && w.alive && w.cur[A] == p; ++g) {
            const float tEntry = w.tExit;
            float tm = w.Tn[0];
#pragma unroll
            for (int i = 1; i < D; ++i)
                tm = fminf(tm, w.Tn[i]);
            w.tExit = tm;
            bool mask[3];
#pragma unroll
            for (int i = 0; i < D; ++i) {
                bool mi = true;
#pragma unroll
                for (int j = 0; j < D; ++j)
                    if (j != i)
                        mi = mi && (w.Tn[i] <= w.Tn[j]);
                mask[i] = mi;
            }
            fn(w.cur[0], w.cur[1], w.cur[2], w.tExit - tEntry);
            bool in = true;
#pragma unroll
            for (int i = 0; i < D; ++i) {
                if (mask[i]) {
                    w.cur[i] += w.st[i];
                    w.fk[i] += 1.0f;
                    w.Tn[i] = __builtin_fmaf(w.fk[i], r.dt[i], r.T0[i]);
                }
                in = in && (uint32_t) w.cur[i] < (uint32_t) vol[i];
            }
            w.alive = in;
        }
    }

    // The part of the traversal inside the voxel box lo[i] <= c[i] <= hi[i]: fn(c0, c1, c2, w)
    // for exactly the crossings of siddon_traverse whose voxel lies in the box, with the same
    // w. The ray is restarted at the last face crossing before the box: for each axis the step
    // counts keeping it inside the box form a range [a_i, b_i]; the crossing that brings the
    // last axis into its range happens at t_lo = max_i T_i(a_i - 1), and at t_lo axis i has
    // taken every step whose T_i(k) <= t_lo.
    template <int D, class Fn>
    EHP_HD inline void siddon_box(const SidRec& r, const int* lo, const int* hi, Fn fn)
    {
        if (!sid_hit(r))
            return;
        int a[3], st[3];
        float t_lo = 0.0f;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            st[i] = sid_step(r, i);
            const int c0 = r.cur0[i];
            int ai, bi;
            if (st[i] == 0) {
                if (c0 < lo[i] || c0 > hi[i])
                    return;
                ai = 0;
                bi = 0x7fffffff;
            } else if (st[i] > 0) {
                ai = lo[i] - c0;
                bi = hi[i] - c0;
            } else {
                ai = c0 - hi[i];
                bi = c0 - lo[i];
            }
            if (bi < 0)
                return;
            ai = ai > 0 ? ai : 0;
            a[i] = ai;
            if (ai > 0) {
                const float te = sid_T(r, i, ai - 1);
                t_lo = te > t_lo ? te : t_lo;
            }
        }
        int cur[3] = {0, 0, 0};
        float fk[3] = {0.0f, 0.0f, 0.0f}; // step counts (exact below 2^24)
        float Tn[3] = {kSidMax, kSidMax, kSidMax};
#pragma unroll
        for (int i = 0; i < D; ++i) {
            int j = 0;
            if (st[i] != 0 && t_lo > 0.0f) {
                // smallest j with T_i(j) > t_lo: estimate (an approximate reciprocal will do),
                // then step to it
                const float est = (t_lo - r.T0[i]) * ehp_rcp(r.dt[i]);
                j = est < 0.0f ? 0 : (est > 1.0e7f ? 10000000 : (int) est);
                j = j > a[i] ? a[i] : j;
                while (sid_T(r, i, j) <= t_lo)
                    ++j;
                while (j > 0 && sid_T(r, i, j - 1) > t_lo)
                    --j;
            }
            fk[i] = (float) j;
            cur[i] = r.cur0[i] + st[i] * j;
            Tn[i] = sid_T(r, i, j);
        }
        float tEntry = t_lo;
        int guard = 0;
#pragma unroll
        for (int i = 0; i < D; ++i)
            guard += hi[i] - lo[i] + 1;
        while (guard-- > 0) {
            bool in = true;
#pragma unroll
            for (int i = 0; i < D; ++i)
                in = in && (uint32_t) (cur[i] - lo[i]) <= (uint32_t) (hi[i] - lo[i]);
            if (!in)
                return;
            float tm = Tn[0];
#pragma unroll
            for (int i = 1; i < D; ++i)
                tm = fminf(tm, Tn[i]);
            bool mask[3];
#pragma unroll
            for (int i = 0; i < D; ++i) {
                bool mi = true;
#pragma unroll
                for (int j = 0; j < D; ++j)
                    if (j != i)
                        mi = mi && (Tn[i] <= Tn[j]);
                mask[i] = mi;
            }
            fn(cur[0], cur[1], cur[2], tm - tEntry);
            tEntry = tm;
            // only an axis that moves can have the smallest T (the driving axis moves)
#pragma unroll
            for (int i = 0; i < D; ++i) {
                if (mask[i]) {
                    cur[i] += st[i];
                    fk[i] += 1.0f;
                    Tn[i] = __builtin_fmaf(fk[i], r.dt[i], r.T0[i]);
                }
            }
        }
    }

    // Slot of voxel c in a tile of the Siddon back projection (no halo: the tile's own voxels,
    // b1 fastest, then b2, then the driving axis).
    template <int D, int A>
    EHP_HD inline int sid_slot(const TileBox& t, const TileCfg& tc, int c0, int c1, int c2)
    {
        constexpr int b1 = TileAxes<D, A>::b1, b2 = TileAxes<D, A>::b2;
        const int c[3] = {c0, c1, c2};
        const int xa = (A == 0 ? c[0] : (A == 1 ? c[1] : c[2])) - t.A0;
        const int xb = (b1 == 0 ? c[0] : (b1 == 1 ? c[1] : c[2])) - t.B0;
        const int xc = D == 3 ? (b2 == 0 ? c[0] : (b2 == 1 ? c[1] : c[2])) - t.C0 : 0;
        return (int) (mul24((uint32_t) xa, (uint32_t) (tc.TB * (D == 3 ? tc.TC : 1))) +
                      mul24((uint32_t) xc, (uint32_t) tc.TB)) +
               xb;
    }
} // namespace ehp
