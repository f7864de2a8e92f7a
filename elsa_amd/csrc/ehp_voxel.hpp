// ehp_voxel.hpp -- footprint arithmetic of the voxel-driven projectors, shared by the GPU
// kernels and the CPU checks.
//
// elsa/projectors/VoxelComputation.h:40-107 (projectAndScaleVoxelOnDetector,
// visitDetector): the voxel centre projected to the detector (minus 0.5), the scaling
// sdd / |ext x|, and the detector pixels within radius * scaling of the centre -- the full
// rectangle [ceil(c - r s), floor(c + r s)] (elsa's CPU loop visits it correctly only for
// square footprints, VoxelComputation.h:90-105) -- each weighted with the projected basis
// function at the distance |c - pixel| / scaling, read from a table by Luts.hpp:60-78.
#pragma once

#include "ehp_joseph_tile.hpp"

namespace ehp
{
    // Largest table of the LDS-table kernels (tiled forward, FP32 gather back projection)
    constexpr int kVoxMaxLut = 1024;

    // Entries of the FP32 kernels' pair table of an n-entry table: the entries and zeros past
    // sqrt(2) (n - 1), the largest index a tap of a footprint rectangle reaches (|d0|, |d1| at most
    // the radius, up to rounding), plus the entry after it.
    EHP_HD inline int lerp_table_size(int n)
    {
        return (int) (1.4143f * (float) (n - 1)) + 3;
    }
    constexpr int kVoxLutSpan = (int) (1.4143f * (float) (kVoxMaxLut - 1)) + 3;

    template <class T>
    EHP_HD inline T lut_eval(const T* __restrict__ lut, int n, float support, float distance)
    {
        const float index = (distance / support) * (float) (n - 1);
        if (index < 0.0f || index > (float) (n - 1))
            return T(0);
        const int a = (int) floorf(index);
        const int b = (int) ceilf(index);
        const T fa = lut[a];
        const T fb = lut[b];
        const T t = (T) (index - (float) a);
        return t * fb + (T(1) - t) * fa;
    }

    // Footprint of voxel (x, y, z) in one pose (VoxelComputation.h:40-107, full rectangle).
    struct Footprint {
        float center[2];
        float scaling;
        int lo[2];
        int hi[2];
    };

    // P (x + 0.5) and ext (x + 0.5) of voxel (x, y, z), homogeneous (VoxelComputation.h:40-60)
    template <int D>
    EHP_HD inline void voxel_project(const PoseGeom& g, int x, int y, int z, float* pc, float* ce)
    {
        float hv[4];
        hv[0] = (float) x + 0.5f;
        hv[1] = (float) y + 0.5f;
        hv[2] = (float) z + 0.5f;
        hv[D] = 1.0f;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            float acc = g.P[i] * hv[0];
            float acc2 = g.ext[i] * hv[0];
#pragma unroll
            for (int j = 1; j <= D; ++j) {
                acc = acc + g.P[i + j * D] * hv[j];
                acc2 = acc2 + g.ext[i + j * D] * hv[j];
            }
            pc[i] = acc;
            ce[i] = acc2;
        }
    }

    // The pixel rectangle [ceil(c - r s), floor(c + r s)] of a footprint, clipped to the detector
    template <int D>
    EHP_HD inline void footprint_rect(const Dims& dims, float radius, Footprint& fp)
    {
        const float rod = fp.scaling * radius;
#pragma unroll
        for (int i = 0; i < D - 1; ++i) {
            const int l = (int) ceilf(fp.center[i] - rod);
            const int h = (int) floorf(fp.center[i] + rod);
            fp.lo[i] = l < 0 ? 0 : l;
            fp.hi[i] = h > dims.det[i] - 1 ? dims.det[i] - 1 : h;
        }
        if (D == 2) {
            fp.lo[1] = 0;
            fp.hi[1] = 0;
        }
    }

    template <int D>
    EHP_HD inline void footprint(const PoseGeom& g, const Dims& dims, int x, int y, int z, float radius,
                                     Footprint& fp)
    {
        float pc[3], ce[3];
        voxel_project<D>(g, x, y, z, pc, ce);
#pragma unroll
        for (int i = 0; i < D - 1; ++i)
            fp.center[i] = pc[i] / pc[D - 1] - 0.5f;
        float n2 = ce[0] * ce[0];
#pragma unroll
        for (int i = 1; i < D; ++i)
            n2 = n2 + ce[i] * ce[i];
        fp.scaling = g.sdd / sqrtf(n2);
        footprint_rect<D>(dims, radius, fp);
    }

    // Reciprocal and reciprocal square root in one instruction each (v_rcp_f32, v_rsq_f32, within
    // 1 ulp); correctly rounded on the CPU.
    EHP_HD inline float fast_rcp(float x)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return __builtin_amdgcn_rcpf(x);
#else
        return 1.0f / x;
#endif
    }

    EHP_HD inline float fast_rsq(float x)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return __builtin_amdgcn_rsqf(x);
#else
        return 1.0f / sqrtf(x);
#endif
    }

    // The FP32 kernels' footprint (tiled forward, gather back projection): footprint()'s
    // projection with the hardware's reciprocal and reciprocal square root in place of three
    // correctly rounded divisions and a square root (about 45 VALU instructions fewer per
    // footprint). The homogeneous division is the quotient q = p rcp(w) with one correction,
    // q + (p - q w) rcp(w) (two FMAs): within about half an ulp of p / w, as the centre, a
    // coordinate of hundreds of pixels from which the taps' pixel indices are subtracted, needs.
    // The scaling sdd / |ext x| = sdd rsq(|ext x|^2) and rs = 1 / scaling come within an ulp or
    // two (relative errors of 1e-7 in the tap distances). Forward and back projection share it, so
    // they stay each other's adjoint.
    template <int D>
    EHP_HD inline void footprint_fast(const PoseGeom& g, const Dims& dims, int x, int y, int z, float radius,
                                      Footprint& fp, float& rs)
    {
        float pc[3], ce[3];
        voxel_project<D>(g, x, y, z, pc, ce);
        const float w = pc[D - 1];
        const float rw = fast_rcp(w);
#pragma unroll
        for (int i = 0; i < D - 1; ++i) {
            const float q = pc[i] * rw;
            fp.center[i] = __builtin_fmaf(__builtin_fmaf(-q, w, pc[i]), rw, q) - 0.5f;
        }
        float n2 = ce[0] * ce[0];
#pragma unroll
        for (int i = 1; i < D; ++i)
            n2 = n2 + ce[i] * ce[i];
        fp.scaling = g.sdd * fast_rsq(n2);
        rs = fast_rcp(fp.scaling);
        footprint_rect<D>(dims, radius, fp);
    }

    // fn(pixel, |d|, d_u) for every pixel of the footprint, d = (centre - pixel) / scaling
    template <int D, class T, class Fn>
    EHP_HD inline void visit(const Footprint& fp, const Dims& dims, Fn fn)
    {
        for (int j = fp.lo[1]; j <= fp.hi[1]; ++j)
            for (int i = fp.lo[0]; i <= fp.hi[0]; ++i) {
                const float d0 = (fp.center[0] - (float) i) / fp.scaling;
                float dist;
                if (D == 2) {
                    dist = sqrtf(d0 * d0);
                } else {
                    const float d1 = (fp.center[1] - (float) j) / fp.scaling;
                    dist = sqrtf(d0 * d0 + d1 * d1);
                }
                fn((int64_t) j * dims.det[0] + i, dist, d0);
            }
    }

    // Weight of one footprint tap: the table at |d| (VoxelComputation.h:20-24), or for the
    // phase-contrast projectors (diff) times sgn(d_u) in 2D and d_u in 3D (VoxelComputation.h:
    // 26-38, ProjectVoxelsCUDA.cu:73-76, :160-165).
    template <int D, class T>
    EHP_HD inline T tap_weight(const T* __restrict__ lut, int n, float support, float dist, float d0, bool diff)
    {
        const T w = lut_eval<T>(lut, n, support, dist);
        if (!diff)
            return w;
        if (D == 2)
            return d0 > 0.0f ? w : (d0 < 0.0f ? -w : T(0));
        return w * (T) d0;
    }

    // Square root for the FP32 taps: the hardware's v_sqrt_f32 (within 1 ulp; the correctly
    // rounded sqrtf adds a dozen instructions per tap on this part); sqrtf on the CPU.
    EHP_HD inline float tap_sqrt(float x)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return __builtin_amdgcn_sqrtf(x);
#else
        return sqrtf(x);
#endif
    }

    // The FP32 kernels' table (tiled forward, gather back projection): entry i < lerp_table_size(n)
    // of a table of n values as the pair (lut[i], lut[i + 1] - lut[i]) in tab[2 i], tab[2 i + 1],
    // with lut[i] = 0 for i >= n, so that one aligned 8-byte read gives an interpolation's two
    // operands.
    EHP_HD inline void lerp_table_entry(const float* __restrict__ lut, int n, int i, float* __restrict__ tab)
    {
        const float cur = i < n ? lut[i] : 0.0f;
        const float next = i + 1 < n ? lut[i + 1] : 0.0f;
        tab[2 * i] = cur;
        tab[2 * i + 1] = next - cur;
    }

    // The FP32 kernels' tap weight from the squared distance d2 = |d|^2 (d0 * d0 + d1 * d1 with one
    // rounding): the table index as |d| * k, k = (n - 1) / support (one multiplication instead of
    // a division per tap), and with a = floor(index), t = index - a the interpolation
    // lut[a] + t * (lut[a + 1] - lut[a]) in one FMA over the pair table (lerp_table_entry; lut_eval
    // forms t * lut[b] + (1 - t) * lut[a], b = ceil(index): the same up to FP32 rounding, and
    // exactly lut[a] at an integer index), 0 outside [0, n - 1]. Branch-free. TZ: the table's last
    // entry is 0 (blob and B-spline tables vanish at the support), so that past n - 1 the zero
    // pairs of the padded table give 0 by themselves and the clamp and range test drop out.
    template <int D, bool TZ = false>
    EHP_HD inline float tap_weight_fast(const float* __restrict__ tab, int n, float k, float d2, float d0, bool diff)
    {
        // index >= 0, and below sqrt(2) (n - 1) for a tap of the footprint rectangle (|d0|, |d1| at
        // most the radius in table units), so its integer part fits an int. In 2D |d| is |d0|,
        // which the correctly rounded sqrt(d0 * d0) equals (d0 * d0 normal).
        const float index = (D == 2 ? fabsf(d0) : tap_sqrt(d2)) * k;
        const float last = (float) (n - 1);
        const int ia = (int) index;
        const int a = TZ ? ia : (ia < n - 1 ? ia : n - 1); // floor(index) where index <= n - 1
        const float t = fract_pos(index);                  // index - a there (exact)
        float w = __builtin_fmaf(t, tab[2 * a + 1], tab[2 * a]);
        if (!TZ)
            w = index <= last ? w : 0.0f;
        if (!diff)
            return w;
        if (D == 2)
            return d0 > 0.0f ? w : (d0 < 0.0f ? -w : 0.0f);
        return w * d0;
    }

    // Row stride (slots) of a detector tile's LDS sums in the tiled voxel forward: odd in 3D, so
    // that the 8-byte slots of neighbouring detector rows fall on different LDS banks.
    EHP_HD inline int vox_acc_stride(int TU, int TV)
    {
        return TV > 1 ? (TU | 1) : TU;
    }

    // What one detector pixel of the tiled voxel forward can receive, in units of max|x|: at most
    // (2 r s + 2)^(dim-1) voxels per plane (s the largest footprint scaling), each with a
    // weight of at most ls (a power of two at least the largest |weight|), over the longest
    // axis of planes; at least 4096 ls, so that one term x * w stays below 2^51 units (the range
    // of fixed_round).
    inline double voxel_slot_bound(const Dims& dims, double support, double smax, double ls)
    {
        int maxdim = 0;
        for (int i = 0; i < dims.dim; ++i)
            maxdim = dims.vol[i] > maxdim ? dims.vol[i] : maxdim;
        const double per_plane = 2.0 * support * smax + 2.0;
        const double b = (double) maxdim * (dims.dim == 3 ? per_plane * per_plane : per_plane) * ls;
        return b > 4096.0 * ls ? b : 4096.0 * ls;
    }

    // Largest footprint scaling of a pose over the volume: sdd over the smallest distance from
    // the source to the volume box (scaling is sdd / |ext x| and |ext x| is the distance from
    // the source to the voxel centre x).
    inline double voxel_max_scaling(const PoseGeom& g, const Dims& dims)
    {
        double dist2 = 0.0;
        for (int i = 0; i < dims.dim; ++i) {
            const double c = g.C[i];
            const double e = c < 0.0 ? -c : (c > dims.vol[i] ? c - dims.vol[i] : 0.0);
            dist2 += e * e;
        }
        return dist2 > 0.0 ? (double) g.sdd / sqrt(dist2) : -1.0;
    }

    // Driving axis of a pose for the tiled forward: the largest component of the central ray.
    template <int D>
    EHP_HD inline int pose_axis(const PoseGeom& g, const Dims& dims)
    {
        float o[3], d[3];
        make_ray<D>(g, dims.det[0] / 2, D == 3 ? dims.det[1] / 2 : 0, o, d);
        int a = 0;
        float m = fabsf(d[0]);
        for (int i = 1; i < D; ++i)
            if (fabsf(d[i]) > m) {
                m = fabsf(d[i]);
                a = i;
            }
        return a;
    }

    // Voxels of plane `pa` (along axis A) whose footprint can reach the detector rectangle
    // [u0, u1] x [v0, v1]: the footprint centre c = (P x)_hom - 0.5 of such a voxel lies within
    // margin m of the rectangle, i.e. P x lies in [u0 + 0.5 - m, u1 + 0.5 + m] x ..., the
    // frustum through the source spanned by the rays of those four corners. Its cross-section
    // with the plane of voxel centres pa + 0.5 is the quadrilateral of the corner rays' hits;
    // the voxel ranges cover its bounding rectangle with one voxel to spare. Returns false
    // when no voxel qualifies; falls back to the whole plane when a corner ray does not reach
    // the plane in front of the source.
    template <int D>
    EHP_HD inline bool frustum_plane_range(const PoseGeom& g, const Dims& dims, int A, int pa, const float* cdir,
                                           int ncorners, int& blo, int& bhi, int& clo, int& chi)
    {
        const int b1 = (D == 2) ? 1 - A : (A == 0 ? 1 : 0);
        const int b2 = (D == 2) ? -1 : (A == 2 ? 1 : 2);
        const float X = (float) pa + 0.5f;
        const float CA = pick<D>(g.C, A);
        float bmin = 3.0e38f, bmax = -3.0e38f, cmin = 3.0e38f, cmax = -3.0e38f;
        bool ok = true;
        for (int k = 0; k < ncorners; ++k) {
            const float* dk = cdir + 3 * k;
            const float da = pick<D>(dk, A);
            const float t = (X - CA) / da;
            ok = ok && (da != 0.0f) && (t > 0.0f) && (t < 3.0e38f);
            const float pb = pick<D>(g.C, b1) + t * pick<D>(dk, b1);
            bmin = pb < bmin ? pb : bmin;
            bmax = pb > bmax ? pb : bmax;
            if (D == 3) {
                const float pc = pick<D>(g.C, b2) + t * pick<D>(dk, b2);
                cmin = pc < cmin ? pc : cmin;
                cmax = pc > cmax ? pc : cmax;
            }
        }
        const int nb = dims.vol[b1], nc = D == 3 ? dims.vol[b2] : 1;
        if (!ok) {
            blo = 0;
            bhi = nb - 1;
            clo = 0;
            chi = nc - 1;
            return true;
        }
        // voxel centres b + 0.5 in [bmin, bmax], one voxel to spare
        bmin = bmin < -2.0f ? -2.0f : (bmin > (float) nb + 2.0f ? (float) nb + 2.0f : bmin);
        bmax = bmax < -2.0f ? -2.0f : (bmax > (float) nb + 2.0f ? (float) nb + 2.0f : bmax);
        blo = (int) floorf(bmin - 0.5f) - 1;
        bhi = (int) ceilf(bmax - 0.5f) + 1;
        blo = blo < 0 ? 0 : blo;
        bhi = bhi > nb - 1 ? nb - 1 : bhi;
        if (D == 3) {
            cmin = cmin < -2.0f ? -2.0f : (cmin > (float) nc + 2.0f ? (float) nc + 2.0f : cmin);
            cmax = cmax < -2.0f ? -2.0f : (cmax > (float) nc + 2.0f ? (float) nc + 2.0f : cmax);
            clo = (int) floorf(cmin - 0.5f) - 1;
            chi = (int) ceilf(cmax - 0.5f) + 1;
            clo = clo < 0 ? 0 : clo;
            chi = chi > nc - 1 ? nc - 1 : chi;
        } else {
            clo = chi = 0;
        }
        return blo <= bhi && clo <= chi;
    }

    // Directions of the corner rays of the detector rectangle [u0, u1] x [v0, v1] widened by
    // margin m (unnormalised: Pinv applied to the homogeneous detector point).
    template <int D>
    EHP_HD inline int frustum_corners(const PoseGeom& g, int u0, int u1, int v0, int v1, float m, float* cdir)
    {
        const float us[2] = {(float) u0 + 0.5f - m, (float) u1 + 0.5f + m};
        const float vs[2] = {(float) v0 + 0.5f - m, (float) v1 + 0.5f + m};
        int n = 0;
        for (int jv = 0; jv < (D == 3 ? 2 : 1); ++jv)
            for (int ju = 0; ju < 2; ++ju) {
                const float h[3] = {us[ju], D == 3 ? vs[jv] : 1.0f, 1.0f};
                for (int i = 0; i < 3; ++i)
                    cdir[3 * n + i] = 0.0f;
                for (int i = 0; i < D; ++i) {
                    float acc = 0.0f;
                    for (int j = 0; j < D; ++j)
                        acc += g.Pinv[i + j * (D + 1)] * h[j];
                    cdir[3 * n + i] = acc;
                }
                ++n;
            }
        return n;
    }
} // namespace ehp
