// ehp_device.hpp -- device-side geometry shared by the Joseph, Siddon and voxel kernels.
//
// Everything here is evaluated in FP32 (elsa's real_t) with the operation order of elsa's
// CPU code and without contraction (the library is built with -ffp-contract=off), so the
// per-ray traversal state -- driving axis, first sample and sample count -- is the same on
// the GPU as in the CPU restatement. Where a fused multiply-add is meant it is written.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define EHP_HD __host__ __device__
#else
// plain C++ build of the geometry (CPU checks of the kernel index logic)
#include <math.h>
#define EHP_HD
#endif

namespace ehp
{
    // Approximate reciprocal (one rounding error of a few ulp at most) for quantities that are
    // only bounds; exact quotients keep the division operator.
    EHP_HD inline float ehp_rcp(float x)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return __builtin_amdgcn_rcpf(x);
#else
        return 1.0f / x;
#endif
    }
} // namespace ehp

namespace ehp
{
    // bfloat16 storage: the upper half of an IEEE single, kept as raw bits. Arithmetic on it is
    // FP32: up() widens exactly, down() rounds to nearest even (NaN stays a quiet NaN).
    struct bf16 {
        uint16_t bits;
    };

    EHP_HD inline float bf16_to_float(bf16 v)
    {
        const uint32_t u = (uint32_t) v.bits << 16;
        float f;
        __builtin_memcpy(&f, &u, 4);
        return f;
    }

    EHP_HD inline bf16 float_to_bf16(float f)
    {
        uint32_t u;
        __builtin_memcpy(&u, &f, 4);
        if ((u & 0x7fffffffu) > 0x7f800000u)
            return bf16{(uint16_t) ((u >> 16) | 0x40u)};
        u += 0x7fffu + ((u >> 16) & 1u);
        return bf16{(uint16_t) (u >> 16)};
    }

    // arithmetic type of a storage type
    template <class T>
    struct Acc {
        using type = T;
    };
    template <>
    struct Acc<bf16> {
        using type = float;
    };

    template <class T>
    EHP_HD inline typename Acc<T>::type up(T v)
    {
        return v;
    }
    EHP_HD inline float up(bf16 v)
    {
        return bf16_to_float(v);
    }

    template <class T>
    EHP_HD inline T down(typename Acc<T>::type v)
    {
        return v;
    }
    template <>
    EHP_HD inline bf16 down<bf16>(float v)
    {
        return float_to_bf16(v);
    }

    // Per-pose geometry as kernels read it. One 64-float record per pose so a wavefront can
    // read it with scalar loads. Column-major blocks as in elsa (Eigen).
    struct PoseGeom {
        float Pinv[12]; // (dim+1) x dim
        float C[4];     // camera centre
        float P[12];    // dim x (dim+1)
        float ext[12];  // dim x (dim+1)
        float sdd;
        float pad[23];
    };
    static_assert(sizeof(PoseGeom) == 64 * sizeof(float), "PoseGeom must be 256 bytes");

    struct Dims {
        int dim;
        int vol[3];    // voxels per axis (1 for unused)
        int det[2];    // detector pixels (1 for unused)
        int nposes;
        int64_t strides[3];
        int64_t nvox;
        int64_t pix_per_pose;
        int64_t nrays;
        // planar detector coordinates of the pixel centres (dim - 1 per pixel, u fastest) of a
        // curved detector; nullptr for a planar detector (centre of pixel u at u + 0.5)
        const float* pix;
        // curved detectors in the tiled back projections (3D): over any column range [a, b] the
        // planar v of a row stays within pix_margin of its values at columns a, b and
        // clamp(pix_uc, a, b)
        int pix_uc;
        float pix_margin;
    };

    EHP_HD inline float minNum(float x, float y)
    {
        return (y != y) ? x : (x < y ? x : y);
    }
    EHP_HD inline float maxNum(float x, float y)
    {
        return (y != y) ? x : (x > y ? x : y);
    }
    EHP_HD inline float sgnf(float v)
    {
        return v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f);
    }

    // Ray through the centre of pixel (u, v): origin C, direction normalize(Pinv h).
    // elsa/core/Descriptors/DetectorDescriptor.cpp:90-128.
    template <int D>
    EHP_HD inline void make_ray_at(const PoseGeom& g, float hu, float hv, float* o, float* d)
    {
        float h[3];
        h[0] = hu;
        if (D == 3) {
            h[1] = hv;
            h[2] = 1.0f;
        } else {
            h[1] = 1.0f;
        }
        float y[3];
#pragma unroll
        for (int i = 0; i < D; ++i) {
            float acc = g.Pinv[i] * h[0];
#pragma unroll
            for (int j = 1; j < D; ++j)
                acc = acc + g.Pinv[i + j * (D + 1)] * h[j];
            y[i] = acc;
        }
        float n2 = y[0] * y[0];
#pragma unroll
        for (int i = 1; i < D; ++i)
            n2 = n2 + y[i] * y[i];
        const float n = sqrtf(n2);
#pragma unroll
        for (int i = 0; i < D; ++i) {
            d[i] = n2 > 0.0f ? y[i] / n : y[i];
            o[i] = g.C[i];
        }
    }

    template <int D>
    EHP_HD inline void make_ray(const PoseGeom& g, int u, int v, float* o, float* d)
    {
        make_ray_at<D>(g, (float) u + 0.5f, (float) v + 0.5f, o, d);
    }

    // The ray of pixel (u, v) of a detector that may be curved: through its planar coordinate
    // (elsa/core/Descriptors/CurvedDetectorDescriptor.cpp:97-119) when dims.pix is set.
    template <int D>
    EHP_HD inline void make_ray_px(const PoseGeom& g, const Dims& dims, int u, int v, float* o, float* d)
    {
        if (dims.pix) {
            const float* c = dims.pix + ((int64_t) v * dims.det[0] + u) * (D - 1);
            make_ray_at<D>(g, c[0], D == 3 ? c[D == 3 ? 1 : 0] : 0.0f, o, d);
        } else {
            make_ray<D>(g, u, v, o, d);
        }
    }

    // Slab test against [0, vol) with elsa's NaN handling and hit rule.
    // elsa/projectors/Intersection.cpp:7-46.
    template <int D>
    EHP_HD inline bool intersect(const float* o, const float* d, const int* vol, float& tmin_out,
                                     float& tmax_out)
    {
        float invDir = 1.0f / d[0];
        float t1 = (0.0f - o[0]) * invDir;
        float t2 = ((float) vol[0] - o[0]) * invDir;
        float tmin = invDir >= 0.0f ? t1 : t2;
        float tmax = invDir >= 0.0f ? t2 : t1;
#pragma unroll
        for (int i = 1; i < D; ++i) {
            invDir = 1.0f / d[i];
            t1 = (0.0f - o[i]) * invDir;
            t2 = ((float) vol[i] - o[i]) * invDir;
            tmin = maxNum(tmin, invDir >= 0.0f ? t1 : t2);
            tmax = minNum(tmax, invDir >= 0.0f ? t2 : t1);
        }
        if (tmax == 0.0f && tmin == 0.0f)
            return false;
        if (tmax >= maxNum(tmin, 0.0f)) {
            tmin_out = tmin;
            tmax_out = tmax;
            return true;
        }
        return false;
    }

    // Traversal state of elsa's DrivingDirectionTraversal after construction
    // (elsa/projectors/DrivingDirectionTraversal.cpp:6-57): driving axis, first sample
    // position (snapped to a voxel-centre plane), step (unit along the driving axis),
    // number of samples and the length of one step.
    struct JRay {
        int axis;
        int nsteps;
        float pos0[3];
        float step[3];
        float len;
    };

    // v[a] for a runtime a without dynamic indexing (keeps small arrays in registers).
    template <int D>
    EHP_HD inline float pick(const float* v, int a)
    {
        return (D == 3 && a == 2) ? v[D - 1] : (a == 1 ? v[1] : v[0]);
    }

    template <int D>
    EHP_HD inline void joseph_setup(const float* o, const float* d, const int* vol, JRay& jr)
    {
        // driving axis: first index of the largest |d| (Eigen's maxCoeff keeps the first)
        int a = 0;
        float m = fabsf(d[0]);
        if (fabsf(d[1]) > m) {
            m = fabsf(d[1]);
            a = 1;
        }
        if (D == 3 && fabsf(d[D - 1]) > m) {
            m = fabsf(d[D - 1]);
            a = 2;
        }
        jr.axis = a;
        jr.nsteps = 0;
        jr.len = 0.0f;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            jr.pos0[i] = 0.0f;
            jr.step[i] = 0.0f;
        }
        float tmin, tmax;
        if (!intersect<D>(o, d, vol, tmin, tmax))
            return;
        float cur[3], ex[3];
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const float hi = (float) vol[i];
            cur[i] = o[i] + tmin * d[i];
            cur[i] = cur[i] < 0.0f ? 0.0f : cur[i];
            cur[i] = cur[i] > hi ? hi : cur[i];
            ex[i] = o[i] + tmax * d[i];
            ex[i] = ex[i] < 0.0f ? 0.0f : ex[i];
            ex[i] = ex[i] > hi ? hi : ex[i];
        }
        const float ad = m; // |d[a]|
        float step[3];
#pragma unroll
        for (int i = 0; i < D; ++i)
            step[i] = d[i] / ad;
        const float cur_a = pick<D>(cur, a);
        const float dist = cur_a - floorf(cur_a);
        const float s = sgnf(pick<D>(step, a));
#pragma unroll
        for (int i = 0; i < D; ++i)
            cur[i] = cur[i] + step[i] * (0.5f - dist) * s;
        const float snapped = floorf(pick<D>(cur, a)) + 0.5f;
#pragma unroll
        for (int i = 0; i < D; ++i)
            cur[i] = (i == a) ? snapped : cur[i];
        jr.nsteps = (int) ceilf(fabsf(pick<D>(ex, a) - snapped));
        const float n2 = (D == 3) ? step[0] * step[0] + (step[1] * step[1] + step[D - 1] * step[D - 1])
                                  : step[0] * step[0] + step[1] * step[1];
        jr.len = sqrtf(n2);
#pragma unroll
        for (int i = 0; i < D; ++i) {
            jr.pos0[i] = cur[i];
            jr.step[i] = step[i];
        }
    }

    // Bilinear weights in the plane normal to the driving axis for sample s (position
    // pos0 + s*step, one FMA per transverse axis), as getLinearInterpolationWeights
    // (JosephsMethod.cpp:37-52) evaluates them. A >= 0 fixes the driving axis at compile
    // time; A = -1 reads it from the ray.
    struct JSample {
        int fl[3];
        int cl[3];
        float w[3];
        float cw[3];
    };

    template <int D, int A = -1>
    EHP_HD inline void joseph_sample(const JRay& jr, int s, JSample& js)
    {
        const int ax = A >= 0 ? A : jr.axis;
        const float fs = (float) s;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const float p = (i == ax) ? jr.pos0[i] + fs * jr.step[i] : __builtin_fmaf(fs, jr.step[i], jr.pos0[i]);
            const float f = floorf(p - 0.5f);
            js.fl[i] = (int) f;
            js.cl[i] = (int) ceilf(p - 0.5f);
            const float cw = (p - f) - 0.5f;
            js.cw[i] = (i == ax) ? 1.0f : cw;
            js.w[i] = (i == ax) ? 1.0f : 1.0f - cw;
        }
    }

    // The taps of one sample in elsa's order (doInterpolation, JosephsMethod.cpp:60-97):
    // 2D (F0,C1)(C0,F1); 3D (F0,F1,F2)(F0,C1,C2)(C0,F1,C2)(C0,C1,F2). A tap outside the
    // volume gets weight 0 and a clamped index.
    template <int D, class Fn>
    EHP_HD inline void joseph_taps(const JSample& js, const int* vol, const int64_t* strides, Fn fn)
    {
        auto tap = [&](int c0, int c1, int c2, float wt) {
            const int c[3] = {c0, c1, c2};
            bool in = true;
            int64_t lin = 0;
#pragma unroll
            for (int i = 0; i < D; ++i) {
                in = in && (c[i] >= 0) && (c[i] < vol[i]);
                const int cc = c[i] < 0 ? 0 : (c[i] > vol[i] - 1 ? vol[i] - 1 : c[i]);
                lin += (int64_t) cc * strides[i];
            }
            fn(lin, in ? wt : 0.0f);
        };
        if (D == 2) {
            tap(js.fl[0], js.cl[1], 0, js.w[0] * js.cw[1]);
            tap(js.cl[0], js.fl[1], 0, js.cw[0] * js.w[1]);
        } else {
            tap(js.fl[0], js.fl[1], js.fl[2], js.w[0] * js.w[1] * js.w[2]);
            tap(js.fl[0], js.cl[1], js.cl[2], js.w[0] * js.cw[1] * js.cw[2]);
            tap(js.cl[0], js.fl[1], js.cl[2], js.cw[0] * js.w[1] * js.cw[2]);
            tap(js.cl[0], js.cl[1], js.fl[2], js.cw[0] * js.cw[1] * js.w[2]);
        }
    }

#if defined(__HIPCC__)
    // Hardware float/double atomic add (global_atomic_add_f32 / _f64 on gfx950).
    __device__ inline void atomic_add(float* p, float v)
    {
        unsafeAtomicAdd(p, v);
    }
    __device__ inline void atomic_add(double* p, double v)
    {
        unsafeAtomicAdd(p, v);
    }
#endif
} // namespace ehp
