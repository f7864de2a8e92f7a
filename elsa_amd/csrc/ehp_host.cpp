// ehp_host.cpp -- host-only parts of the C-ABI: elsa's geometry, the circular trajectory
// and the basis-function tables. Nothing here touches the GPU.
//
//   geometry     elsa/core/Geometry.cpp:12-109 (constructors), :225-270 (buildMatrices)
//   trajectory   elsa/generators/CircleTrajectoryGenerator.cpp:9-24,
//                BaseCircleTrajectoryGenerator.cpp:12-74, TrajectoryGenerator.cpp:9-47
//   blob table   elsa/projectors/Blobs.h:67-98, :249-317 with Bessel.cpp:7-80
//   B-spline     elsa/projectors/BSplines.h:22-38, :72-81, :132-208
//   derivative tables (phase contrast)
//                elsa/projectors/Blobs.h:129-207, :249-317; BSplines.h:40-45, :83-92, :132-208
//   tables       elsa/projectors/Luts.hpp:14-29
//
// FP32 throughout (elsa's real_t), products and sums in elsa's order, no contraction
// (-ffp-contract=off). Products with the structural zeros and ones of S, [I|0] and K^-1
// are exact and left out.
#include "elsa_hip.h"

#include <cmath>
#include <cstring>
#include <string>

namespace
{
    // r x c matrix, column-major, small.
    struct M {
        int r, c;
        float a[16];
        float& at(int i, int j) { return a[i + j * r]; }
        float at(int i, int j) const { return a[i + j * r]; }
    };

    M matmul(const M& A, const M& B)
    {
        M C{A.r, B.c, {}};
        for (int i = 0; i < A.r; ++i)
            for (int j = 0; j < B.c; ++j) {
                float s = A.at(i, 0) * B.at(0, j);
                for (int k = 1; k < A.c; ++k)
                    s = s + A.at(i, k) * B.at(k, j);
                C.at(i, j) = s;
            }
        return C;
    }

    // x = A^-1 b for a 2x2 or 3x3 system, evaluated in double by cofactors.
    void solve_double(int d, const double* A /*col-major*/, const double* b, double* x)
    {
        auto a = [&](int i, int j) { return A[i + j * d]; };
        if (d == 2) {
            const double det = a(0, 0) * a(1, 1) - a(0, 1) * a(1, 0);
            x[0] = (a(1, 1) * b[0] - a(0, 1) * b[1]) / det;
            x[1] = (a(0, 0) * b[1] - a(1, 0) * b[0]) / det;
            return;
        }
        const double c00 = a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1);
        const double c01 = a(1, 2) * a(2, 0) - a(1, 0) * a(2, 2);
        const double c02 = a(1, 0) * a(2, 1) - a(1, 1) * a(2, 0);
        const double det = a(0, 0) * c00 + a(0, 1) * c01 + a(0, 2) * c02;
        const double inv[3][3] = {
            {c00, a(0, 2) * a(2, 1) - a(0, 1) * a(2, 2), a(0, 1) * a(1, 2) - a(0, 2) * a(1, 1)},
            {c01, a(0, 0) * a(2, 2) - a(0, 2) * a(2, 0), a(0, 2) * a(1, 0) - a(0, 0) * a(1, 2)},
            {c02, a(0, 1) * a(2, 0) - a(0, 0) * a(2, 1), a(0, 0) * a(1, 1) - a(0, 1) * a(1, 0)}};
        for (int i = 0; i < 3; ++i)
            x[i] = (inv[i][0] * b[0] + inv[i][1] * b[1] + inv[i][2] * b[2]) / det;
    }

    // buildMatrices: ext = [R|t] S, P = K [I|0] [R|t;0 1] S, C = -P33^-1 p4,
    // Pinv = S^-1 [R^T | -R^T t; 0 1] [I;0] K^-1. K is upper triangular with K(d-1,d-1) = 1.
    void build(int d, const M& R, const float* t, const M& K, const float* spacing, float* P, float* Pinv,
               float* C, float* ext)
    {
        // extFull = [R|t;0 1] * S: column j < d is R(:,j)*s_j, column d is t
        M E{d, d + 1, {}};
        for (int i = 0; i < d; ++i) {
            for (int j = 0; j < d; ++j)
                E.at(i, j) = R.at(i, j) * spacing[j];
            E.at(i, d) = t[i];
        }
        // P = [K|0] * extFull: P(i,j) = sum_k K(i,k) E(k,j) over the non-zero K(i,k)
        M Pm{d, d + 1, {}};
        for (int j = 0; j <= d; ++j) {
            for (int i = 0; i < d - 1; ++i)
                Pm.at(i, j) = K.at(i, i) * E.at(i, j) + K.at(i, d - 1) * E.at(d - 1, j);
            Pm.at(d - 1, j) = E.at(d - 1, j);
        }
        // camera centre in double, rounded once
        double A[9], b[3], x[3];
        for (int i = 0; i < d; ++i) {
            for (int j = 0; j < d; ++j)
                A[i + j * d] = (double) Pm.at(i, j);
            b[i] = (double) Pm.at(i, d);
        }
        solve_double(d, A, b, x);
        // S^-1 [R^T | -R^T t; 0 1] [I; 0] keeps only S^-1 R^T (the translation column is
        // dropped and the last row is zero), so Pinv does not depend on t.
        M Q{d + 1, d, {}};
        for (int i = 0; i < d; ++i) {
            const float sinv = 1.0f / spacing[i];
            for (int j = 0; j < d; ++j)
                Q.at(i, j) = sinv * R.at(j, i);
        }
        // times K^-1 = [1/K00, 0, -K02/K00; 0, 1/K11, -K12/K11; 0, 0, 1] (d = 3) or
        //              [1/K00, -K01/K00; 0, 1] (d = 2). Row d of the product is zero.
        float kdiag[2], kofs[2];
        for (int i = 0; i < d - 1; ++i) {
            kdiag[i] = 1.0f / K.at(i, i);
            kofs[i] = -K.at(i, d - 1) / K.at(i, i);
        }
        M Pi{d + 1, d, {}};
        for (int i = 0; i < d; ++i) {
            for (int j = 0; j < d - 1; ++j)
                Pi.at(i, j) = Q.at(i, j) * kdiag[j];
            float s = Q.at(i, 0) * kofs[0];
            for (int k = 1; k < d - 1; ++k)
                s = s + Q.at(i, k) * kofs[k];
            Pi.at(i, d - 1) = s + Q.at(i, d - 1);
        }
        const int n = d * (d + 1);
        for (int i = 0; i < n; ++i) {
            P[i] = Pm.a[i];
            ext[i] = E.a[i];
            Pinv[i] = Pi.a[i];
        }
        for (int i = 0; i < d; ++i)
            C[i] = (float) -x[i];
    }

    // Bessel I_0, I_1, I_2 (polynomial approximations of Abramowitz & Stegun 9.8.1-9.8.4 as
    // elsa evaluates them, Bessel.cpp:7-80).
    double i0(double x)
    {
        const double ax = std::fabs(x);
        if (ax < 3.75) {
            const double y = std::pow(x / 3.75, 2);
            const double c[7] = {1.0, 3.5156229, 3.0899424, 1.2067492, 0.2659732, 0.360768e-1, 0.45813e-2};
            double p = c[6];
            for (int k = 5; k >= 0; --k)
                p = c[k] + y * p;
            return p;
        }
        const double y = 3.75 / ax;
        const double c[9] = {0.39894228,  0.1328592e-1, 0.225319e-2,  -0.157565e-2, 0.916281e-2,
                             -0.2057706e-1, 0.2635537e-1, -0.1647633e-1, 0.392377e-2};
        double p = c[8];
        for (int k = 7; k >= 0; --k)
            p = c[k] + y * p;
        return (std::exp(ax) / std::sqrt(ax)) * p;
    }

    double i1(double x)
    {
        const double ax = std::fabs(x);
        double r;
        if (ax < 3.75) {
            const double y = std::pow(x / 3.75, 2);
            const double c[7] = {0.5, 0.87890594, 0.51498869, 0.15084934, 0.2658733e-1, 0.301532e-2, 0.32411e-3};
            double p = c[6];
            for (int k = 5; k >= 0; --k)
                p = c[k] + y * p;
            r = ax * p;
        } else {
            const double y = 3.75 / ax;
            // innermost term is 0.1787654e-1 - y * 0.420059e-2 (elsa subtracts there)
            double p = 0.1787654e-1 - y * 0.420059e-2;
            const double c[7] = {0.39894228, -0.3988024e-1, -0.362018e-2, 0.163801e-2,
                                 -0.1031555e-1, 0.2282967e-1, -0.2895312e-1};
            for (int k = 6; k >= 0; --k)
                p = c[k] + y * p;
            r = p * (std::exp(ax) / std::sqrt(ax));
        }
        return x < 0.0 ? -r : r;
    }

    double i2(double x) { return x == 0 ? 0 : i0(x) - (2.0 / x) * i1(x); }

    // Closed-form line integral of the Kaiser-Bessel blob of order 0..2 (Blobs.h:67-98),
    // evaluated in the table's precision T.
    template <class T>
    T blob_line_integral(T s, T a, T alpha, int64_t m)
    {
        const T q = s / a;
        const T w = 1.f - std::pow(q, 2.f);
        if (!(w > 1.0e-10))
            return T(0);
        const T arg = alpha * std::sqrt(w);
        const T pre = 2.f * a / alpha;
        switch (m) {
        case 0:
            return pre * std::sinh(arg) / static_cast<T>(i0(alpha));
        case 1:
            return pre * std::sqrt(w) * (std::cosh(arg) - std::sinh(arg) / arg) / static_cast<T>(i1(alpha));
        default:
            return pre * w * ((3.f / (arg * arg) + 1.f) * std::sinh(arg) - (3.f / arg) * std::cosh(arg))
                   / static_cast<T>(i2(alpha));
        }
    }

    // Line integral of the blob's derivative (Blobs.h:129-153), or that divided by s (the
    // normalized gradient, :183-207), orders 1 and 2. As in elsa the factors -2.0 s / a and
    // the Bessel values are double and sinh / cosh stay in T.
    template <class T>
    T blob_derivative_line_integral(T s, T a, T alpha, int64_t m, bool normalized)
    {
        const T q = s / a;
        const T w = static_cast<T>(1.0 - static_cast<T>(std::pow(q, 2)));
        if (!(w > 1.0e-10))
            return T(0);
        const T arg = alpha * std::sqrt(w);
        const double lead = normalized ? -2.0 : -2.0 * s;
        if (m == 1)
            return static_cast<T>((lead / a) * std::sinh(arg) / i1(alpha));
        return static_cast<T>((lead / alpha / a) * (std::cosh(arg) * arg - std::sinh(arg)) / i2(alpha));
    }

    int64_t fact(int64_t n) { return n <= 1 ? 1 : n * fact(n - 1); }

    int64_t choose(int64_t n, int64_t k)
    {
        if (k < 0 || k > n)
            return 0;
        int64_t r = 1;
        for (int64_t i = 1; i <= k; ++i)
            r = r * (n - k + i) / i;
        return r;
    }

    // Centred cardinal B-spline of order m via the truncated-power sum (BSplines.h:22-38);
    // the power terms are double, the running sum is T.
    template <class T>
    T bspline(T x, int64_t m)
    {
        const T xs = x + static_cast<T>(m + 1) / T(2);
        T sum = 0;
        for (int n = 0; n <= m + 1; ++n) {
            const T dx = xs - static_cast<T>(n);
            if (!(dx > 0))
                continue; // Heaviside(dx) with H(0) = 0
            const double term = std::pow(static_cast<double>(dx), static_cast<int>(m))
                                * static_cast<double>(static_cast<T>(choose(m + 1, n)))
                                * (std::pow(-1.0, n) / static_cast<double>(static_cast<T>(fact(m))));
            sum += static_cast<T>(term);
        }
        return sum;
    }

    // Derivative of the projected B-spline (BSplines.h:40-45 with :83-92 at dim - 1): the
    // difference of two double B-splines of order m - 1 at x +- 0.5, in T, times b(0) (double)
    // per dimension beyond the second; normalized divides by x, x = 0 taken as 1e-10
    // (BSplines.h:182-188).
    template <class T>
    T bspline_derivative(T x, int64_t m, int64_t dim, bool normalized)
    {
        if (normalized && x == T(0))
            x = static_cast<T>(1e-10);
        T v = static_cast<T>(bspline<double>(static_cast<double>(x) + 0.5, m - 1) -
                             bspline<double>(static_cast<double>(x) - 0.5, m - 1));
        for (int64_t i = 1; i < dim - 1; ++i)
            v = static_cast<T>(v * bspline<double>(0.0, m));
        return normalized ? v / x : v;
    }

    template <class T, class F>
    void table(T support, int n, T* out, F f)
    {
        const T step = support / static_cast<T>(n - 1);
        for (int i = 0; i < n; ++i)
            out[i] = f(step * static_cast<T>(i));
    }
} // namespace

extern "C" {

int ehp_geometry_pose(int dim, float s2c, float c2d, const float* vol_spacing, const float* vol_origin,
                      const float* det_spacing, const float* det_origin, float gamma, float beta,
                      float alpha, const float* principal_offset, const float* rotation_offset, float* P,
                      float* Pinv, float* C, float* ext, float* sdd)
{
    if (dim != 2 && dim != 3)
        return EHP_ERR_ARGUMENT;
    if (!vol_spacing || !vol_origin || !det_spacing || !det_origin || !P || !Pinv || !C || !ext || !sdd)
        return EHP_ERR_ARGUMENT;
    float zeros[3] = {0, 0, 0};
    const float* ppo = principal_offset ? principal_offset : zeros;
    const float* cor = rotation_offset ? rotation_offset : zeros;
    M R{dim, dim, {}};
    if (dim == 2) {
        const float c = std::cos(gamma), s = std::sin(gamma);
        R.at(0, 0) = c;
        R.at(0, 1) = -s;
        R.at(1, 0) = s;
        R.at(1, 1) = c;
    } else {
        // YZY: R = Ry(gamma) Rz(beta) Ry(alpha)
        const float cg = std::cos(gamma), sg = std::sin(gamma);
        const float cb = std::cos(beta), sb = std::sin(beta);
        const float ca = std::cos(alpha), sa = std::sin(alpha);
        const M G{3, 3, {cg, 0, -sg, 0, 1, 0, sg, 0, cg}};
        const M B{3, 3, {cb, sb, 0, -sb, cb, 0, 0, 0, 1}};
        const M A{3, 3, {ca, 0, -sa, 0, 1, 0, sa, 0, ca}};
        R = matmul(matmul(G, B), A);
    }
    float t[3];
    for (int i = 0; i < dim; ++i) {
        float s = R.at(i, 0) * (-cor[0] - vol_origin[0]);
        for (int k = 1; k < dim; ++k)
            s = s + R.at(i, k) * (-cor[k] - vol_origin[k]);
        t[i] = s;
    }
    t[dim - 1] += s2c;
    M K{dim, dim, {}};
    for (int i = 0; i < dim; ++i)
        K.at(i, i) = 1.0f;
    for (int i = 0; i < dim - 1; ++i) {
        K.at(i, i) = (s2c + c2d) / det_spacing[i];
        K.at(i, dim - 1) = det_origin[i] / det_spacing[i] + ppo[i];
    }
    build(dim, R, t, K, vol_spacing, P, Pinv, C, ext);
    *sdd = s2c + c2d;
    return EHP_OK;
}

int ehp_circle_trajectory(int dim, int64_t nposes, const int64_t* vol_dims, float arc_degrees, float s2c,
                          float c2d, const int64_t* det_dims, int64_t* det_out, float* P, float* Pinv, float* C,
                          float* ext, float* sdd)
{
    if ((dim != 2 && dim != 3) || nposes <= 0 || !vol_dims || !P || !Pinv || !C || !ext || !sdd)
        return EHP_ERR_ARGUMENT;
    float vsp[3] = {1, 1, 1}, vorg[3], dsp[2] = {1, 1}, dorg[2];
    int64_t det[2] = {1, 1};
    for (int i = 0; i < dim; ++i) {
        if (vol_dims[i] <= 0)
            return EHP_ERR_ARGUMENT;
        vorg[i] = 0.5f * ((float) vol_dims[i] * vsp[i]);
    }
    for (int i = 0; i < dim - 1; ++i) {
        // default detector: floor(N * sqrt(2)) evaluated in float (TrajectoryGenerator.cpp:136-144)
        det[i] = det_dims ? det_dims[i] : (int64_t) ((float) vol_dims[i] * std::sqrt(2.f));
        if (det[i] <= 0)
            return EHP_ERR_ARGUMENT;
        dorg[i] = 0.5f * ((float) det[i] * dsp[i]);
        if (det_out)
            det_out[i] = det[i];
    }
    const int nm = dim * (dim + 1);
    // Eigen LinSpaced(n, 0, arc): i * ((arc - 0) / (n - 1)), the last one exactly arc
    const float stepdeg = nposes > 1 ? arc_degrees / (float) (nposes - 1) : 0.0f;
    const float pi_f = 3.14159265358979323846f;
    for (int64_t p = 0; p < nposes; ++p) {
        const float deg = (nposes == 1) ? arc_degrees : (p == nposes - 1 ? arc_degrees : (float) p * stepdeg);
        const float rad = deg * pi_f / 180.0f;
        const int rc = ehp_geometry_pose(dim, s2c, c2d, vsp, vorg, dsp, dorg, rad, 0.0f, 0.0f, nullptr, nullptr,
                                         P + p * nm, Pinv + p * nm, C + p * dim, ext + p * nm, sdd + p);
        if (rc != EHP_OK)
            return rc;
    }
    return EHP_OK;
}

int ehp_blob_lut(int dtype, double radius, double alpha, int64_t order, int n, void* out, double* support)
{
    if (!out || n < 2 || order < 0 || order > 2 || !(radius > 0) || !(alpha > 0))
        return EHP_ERR_ARGUMENT;
    if (dtype == EHP_F32) {
        const float a = (float) radius, al = (float) alpha;
        table<float>(a, n, (float*) out, [&](float s) { return blob_line_integral<float>(s, a, al, order); });
    } else if (dtype == EHP_F64) {
        table<double>(radius, n, (double*) out,
                      [&](double s) { return blob_line_integral<double>(s, radius, alpha, order); });
    } else {
        return EHP_ERR_ARGUMENT;
    }
    if (support)
        *support = dtype == EHP_F32 ? (double) (float) radius : radius;
    return EHP_OK;
}

int ehp_bspline_lut(int dtype, int64_t dim, int64_t order, int n, void* out, double* support)
{
    if (!out || n < 2 || order < 0 || dim < 2)
        return EHP_ERR_ARGUMENT;
    // ProjectedBSpline: support (order + 1) / 2, value b(x) * b(0)^(dim - 2)
    if (dtype == EHP_F32) {
        const float r = (float) ((order + 1) * 0.5);
        table<float>(r, n, (float*) out, [&](float x) {
            float v = bspline<float>(x, order);
            for (int64_t i = 1; i < dim - 1; ++i)
                v *= bspline<float>(0.f, order);
            return v;
        });
        if (support)
            *support = r;
    } else if (dtype == EHP_F64) {
        const double r = (order + 1) * 0.5;
        table<double>(r, n, (double*) out, [&](double x) {
            double v = bspline<double>(x, order);
            for (int64_t i = 1; i < dim - 1; ++i)
                v *= bspline<double>(0.0, order);
            return v;
        });
        if (support)
            *support = r;
    } else {
        return EHP_ERR_ARGUMENT;
    }
    return EHP_OK;
}

int ehp_curved_detector_coords(int dim, const int64_t* det_dims, float angle, float s2d, float* out)
{
    // CurvedDetectorDescriptor::setup / mapCurvedCoordToPlanarCoord
    // (elsa/core/Descriptors/CurvedDetectorDescriptor.cpp:36-95), in FP32 as elsa's real_t
    if ((dim != 2 && dim != 3) || !det_dims || !out || !(angle > 0.0f) || !(s2d > 0.0f))
        return EHP_ERR_ARGUMENT;
    const int64_t len = det_dims[0], height = dim == 3 ? det_dims[1] : 1;
    if (len <= 0 || height <= 0)
        return EHP_ERR_ARGUMENT;
    const float radius = (float) len / angle;
    const float delta = angle / (float) len;
    const float cx = (float) len / 2.0f, cy = (float) height / 2.0f;
    const float ccx = cx, ccy = s2d - radius; // centre of the circle
    int64_t k = 0;
    for (int64_t h = 0; h < height; ++h)
        for (int64_t l = 0; l < len; ++l) {
            const float pu = (float) l + 0.5f, pv = (float) h + 0.5f;
            const float ray_angle = (pu - cx) * delta;
            const float dx = std::sin(ray_angle) * radius;
            const float depth = std::cos(ray_angle) * radius;
            const float x = ccx + dx, z = ccy + depth;
            const float factor = z == 0.0f ? 1.0f : s2d / z;
            out[k++] = (x - cx) * factor + cx;
            if (dim == 3)
                out[k++] = (pv - cy) * factor + cy;
        }
    return EHP_OK;
}

int ehp_blob_derivative_lut(int dtype, double radius, double alpha, int64_t order, int n, int normalized, void* out,
                            double* support)
{
    if (!out || n < 2 || order < 0 || order > 2 || !(radius > 0) || !(alpha > 0))
        return EHP_ERR_ARGUMENT;
    const bool nz = normalized != 0;
    if (dtype == EHP_F32) {
        const float a = (float) radius, al = (float) alpha;
        table<float>(a, n, (float*) out, [&](float s) {
            return order > 0 ? blob_derivative_line_integral<float>(s, a, al, order, nz) : 0.0f;
        });
    } else if (dtype == EHP_F64) {
        table<double>(radius, n, (double*) out, [&](double s) {
            return order > 0 ? blob_derivative_line_integral<double>(s, radius, alpha, order, nz) : 0.0;
        });
    } else {
        return EHP_ERR_ARGUMENT;
    }
    if (support)
        *support = dtype == EHP_F32 ? (double) (float) radius : radius;
    return EHP_OK;
}

int ehp_bspline_derivative_lut(int dtype, int64_t dim, int64_t order, int n, int normalized, void* out,
                               double* support)
{
    if (!out || n < 2 || order < 0 || dim < 2)
        return EHP_ERR_ARGUMENT;
    const bool nz = normalized != 0;
    if (dtype == EHP_F32) {
        const float r = (float) ((order + 1) * 0.5);
        table<float>(r, n, (float*) out,
                     [&](float x) { return order > 0 ? bspline_derivative<float>(x, order, dim, nz) : 0.0f; });
        if (support)
            *support = r;
    } else if (dtype == EHP_F64) {
        const double r = (order + 1) * 0.5;
        table<double>(r, n, (double*) out,
                      [&](double x) { return order > 0 ? bspline_derivative<double>(x, order, dim, nz) : 0.0; });
        if (support)
            *support = r;
    } else {
        return EHP_ERR_ARGUMENT;
    }
    return EHP_OK;
}

} // extern "C"
