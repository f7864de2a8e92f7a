// ehp_solvers.hip -- elsa's CGLS, APGD (FISTA) and power iterations, resident on the device.
//
// The iterations call the projector through the C-ABI (ehp_forward / ehp_adjoint) and keep
// every scalar of the method -- step lengths, momentum, norms, the convergence state -- in
// device memory, so an iteration needs no host round trip and can be captured once into a
// hipGraph and replayed. Norms and dot products are FP64 two-stage reductions in a fixed
// order (bitwise reproducible); the vector updates run in the operator's data type.
//
// elsa returns early when a method has converged (CGLS.cpp:54-58, APGD.cpp:83-87). Here the
// check sets a device flag; once it is set the updates of x are skipped, so the result is the
// iterate at which elsa would have returned, while the remaining replays do no harm.
#include "elsa_hip.h"
#include "ehp_kernels.hpp"

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

struct ehp_op; // defined in ehp_capi.hip
int ehp_solver_fail(int code, const char* msg);

namespace ehp
{
    namespace
    {
        constexpr int kRedBlocks = 1024;
        constexpr int kThreads = 256;

        // scalar slots
        enum {
            S_KOLD = 0, // CGLS: ||r||^2 of the previous iteration
            S_K,        // CGLS: ||r||^2
            S_DELTA,    // CGLS: ||q||^2
            S_CN,       // CGLS: ||c||^2 (damped problems)
            S_ALPHA,    // CGLS step length
            S_BETA,     // CGLS direction update
            S_DONE,     // converged (1) or not (0)
            S_ITERS,    // iterations taken
            S_SN,       // CGLS: ||s||^2; APGD: ||grad||^2
            S_T,        // APGD: t of the previous iteration (in the data type)
            S_COEF,     // APGD: (t_prev - 1) / t
            S_TMP,      // scratch (power iterations)
            S_COUNT
        };

        template <class T>
        __global__ void __launch_bounds__(kThreads)
            dot_partial_kernel(const T* __restrict__ a, const T* __restrict__ b, int64_t n,
                               double* __restrict__ partial)
        {
            double acc = 0.0;
            const int64_t stride = (int64_t) gridDim.x * kThreads;
            for (int64_t i = (int64_t) blockIdx.x * kThreads + threadIdx.x; i < n; i += stride)
                acc += (double) a[i] * (double) b[i];
            __shared__ double red[kThreads / 64];
            for (int off = 32; off > 0; off >>= 1)
                acc += __shfl_down(acc, off, 64);
            if ((threadIdx.x & 63) == 0)
                red[threadIdx.x >> 6] = acc;
            __syncthreads();
            if (threadIdx.x == 0) {
                double s = 0.0;
                for (int w = 0; w < kThreads / 64; ++w)
                    s += red[w];
                partial[blockIdx.x] = s;
            }
        }

        // out = sum of partial[0..np) in a fixed order; hist, when given, gets a copy at
        // index iters (only while not converged)
        __global__ void __launch_bounds__(kThreads)
            dot_final_kernel(const double* __restrict__ partial, int np, double* __restrict__ sc, int slot,
                             double* __restrict__ hist)
        {
            __shared__ double red[kThreads];
            double acc = 0.0;
            for (int i = threadIdx.x; i < np; i += kThreads)
                acc += partial[i];
            red[threadIdx.x] = acc;
            __syncthreads();
            for (int off = kThreads / 2; off > 0; off >>= 1) {
                if ((int) threadIdx.x < off)
                    red[threadIdx.x] += red[threadIdx.x + off];
                __syncthreads();
            }
            if (threadIdx.x == 0) {
                sc[slot] = red[0];
                if (hist && sc[S_DONE] == 0.0)
                    hist[(int) sc[S_ITERS]] = red[0];
            }
        }

        template <class T>
        __global__ void __launch_bounds__(kThreads) copy_kernel(const T* __restrict__ a, T* __restrict__ b, int64_t n)
        {
            const int64_t i = (int64_t) blockIdx.x * kThreads + threadIdx.x;
            if (i < n)
                b[i] = a[i];
        }

        // s = b - s (elsa: s *= -1; s += b)
        template <class T>
        __global__ void __launch_bounds__(kThreads) rsub_kernel(const T* __restrict__ b, T* __restrict__ s, int64_t n)
        {
            const int64_t i = (int64_t) blockIdx.x * kThreads + threadIdx.x;
            if (i < n)
                s[i] = b[i] - s[i];
        }

        // y = a * x + y with a host scalar
        template <class T>
        __global__ void __launch_bounds__(kThreads) axpy_kernel(T a, const T* __restrict__ x, T* __restrict__ y, int64_t n)
        {
            const int64_t i = (int64_t) blockIdx.x * kThreads + threadIdx.x;
            if (i < n)
                y[i] = a * x[i] + y[i];
        }

        // CGLS: x += alpha c over nd elements and s -= alpha q over nr elements
        template <class T>
        __global__ void __launch_bounds__(kThreads)
            cgls_update_kernel(T* __restrict__ x, const T* __restrict__ c, int64_t nd, T* __restrict__ s,
                               const T* __restrict__ q, int64_t nr, const double* __restrict__ sc)
        {
            if (sc[S_DONE] != 0.0)
                return;
            const T alpha = (T) sc[S_ALPHA];
            const int64_t i = (int64_t) blockIdx.x * kThreads + threadIdx.x;
            if (i < nd)
                x[i] += alpha * c[i];
            if (i < nr)
                s[i] -= alpha * q[i];
        }

        // c = r + beta c
        template <class T>
        __global__ void __launch_bounds__(kThreads)
            cgls_direction_kernel(const T* __restrict__ r, T* __restrict__ c, int64_t n, const double* __restrict__ sc)
        {
            const T beta = (T) sc[S_BETA];
            const int64_t i = (int64_t) blockIdx.x * kThreads + threadIdx.x;
            if (i < n)
                c[i] = r[i] + beta * c[i];
        }

        // start of a CGLS iteration: stop when kold < tol (CGLS.cpp:54-58), else count it
        __global__ void cgls_check_kernel(double* sc, double tol)
        {
            if (sc[S_DONE] == 0.0 && sc[S_KOLD] < tol)
                sc[S_DONE] = 1.0;
            if (sc[S_DONE] == 0.0)
                sc[S_ITERS] += 1.0;
        }

        __global__ void cgls_alpha_kernel(double* sc, double damp)
        {
            const double d = damp == 0.0 ? sc[S_DELTA] : sc[S_DELTA] + damp * sc[S_CN];
            sc[S_ALPHA] = sc[S_KOLD] / d;
        }

        __global__ void cgls_beta_kernel(double* sc)
        {
            if (sc[S_DONE] != 0.0) {
                sc[S_BETA] = 0.0;
                return;
            }
            sc[S_BETA] = sc[S_K] / sc[S_KOLD];
            sc[S_KOLD] = sc[S_K];
        }

        __global__ void set_scalar_kernel(double* sc, int slot, double v)
        {
            sc[slot] = v;
        }

        // APGD: x = prox(y - mu grad), prox the identity (sigma < 0) or L1 soft thresholding with
        // threshold mu * sigma (ProximalL1.cpp:32-43)
        template <class T>
        __global__ void __launch_bounds__(kThreads)
            apgd_prox_kernel(const T* __restrict__ y, const T* __restrict__ grad, T mu, T sigma, T* __restrict__ x,
                             int64_t n, const double* __restrict__ sc)
        {
            if (sc[S_DONE] != 0.0)
                return;
            const int64_t i = (int64_t) blockIdx.x * kThreads + threadIdx.x;
            if (i >= n)
                return;
            const T z = y[i] - mu * grad[i];
            if (sigma < T(0)) {
                x[i] = z;
            } else {
                T tmp = (z < T(0) ? -z : z) - mu * sigma;
                tmp = T(0.5) * (tmp + (tmp < T(0) ? -tmp : tmp));
                const T sg = z > T(0) ? T(1) : (z < T(0) ? T(-1) : T(0));
                x[i] = sg * tmp;
            }
        }

        // APGD momentum scalars in the data type, as elsa computes them (APGD.cpp:74-78)
        template <class T>
        __global__ void apgd_t_kernel(double* sc)
        {
            if (sc[S_DONE] != 0.0)
                return;
            const T tprev = (T) sc[S_T];
            const T t = (T(1) + sqrt(T(1) + T(4) * tprev * tprev)) / T(2);
            sc[S_COEF] = (double) ((tprev - T(1)) / t);
            sc[S_T] = (double) t;
        }

        // y = x + coef (x - xprev); xprev = x
        template <class T>
        __global__ void __launch_bounds__(kThreads)
            apgd_momentum_kernel(const T* __restrict__ x, T* __restrict__ xprev, T* __restrict__ y, int64_t n,
                                 const double* __restrict__ sc)
        {
            if (sc[S_DONE] != 0.0)
                return;
            const int64_t i = (int64_t) blockIdx.x * kThreads + threadIdx.x;
            if (i >= n)
                return;
            const T coef = (T) sc[S_COEF];
            const T xi = x[i];
            y[i] = xi + coef * (xi - xprev[i]);
            xprev[i] = xi;
        }

        // APGD convergence: ||grad||^2 <= epsilon ends the method after this iteration
        // (APGD.cpp:83-87)
        __global__ void apgd_check_kernel(double* sc, double eps)
        {
            if (sc[S_DONE] != 0.0)
                return;
            sc[S_ITERS] += 1.0;
            if (sc[S_SN] <= eps)
                sc[S_DONE] = 1.0;
        }

        // u = u / sqrt(sc[slot])
        template <class T>
        __global__ void __launch_bounds__(kThreads) normalize_kernel(T* __restrict__ u, int64_t n, const double* __restrict__ sc,
                                                                   int slot)
        {
            const int64_t i = (int64_t) blockIdx.x * kThreads + threadIdx.x;
            if (i < n)
                u[i] = (T) ((double) u[i] / sqrt(sc[slot]));
        }

        inline unsigned blocks(int64_t n)
        {
            return (unsigned) ((n + kThreads - 1) / kThreads);
        }

        template <class T>
        hipError_t dot(const T* a, const T* b, int64_t n, double* partial, double* sc, int slot, double* hist,
                       hipStream_t s)
        {
            int64_t nb = (n + kThreads - 1) / kThreads;
            nb = nb > kRedBlocks ? kRedBlocks : (nb < 1 ? 1 : nb);
            hipLaunchKernelGGL((dot_partial_kernel<T>), dim3((unsigned) nb), dim3(kThreads), 0, s, a, b, n, partial);
            hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kThreads), 0, s, partial, (int) nb, sc, slot, hist);
            return hipGetLastError();
        }

        // Device buffers of one solve, freed on scope exit.
        struct Scratch {
            std::vector<void*> ptrs;
            ~Scratch()
            {
                for (void* p : ptrs)
                    (void) hipFree(p);
            }
            template <class P>
            hipError_t alloc(P** p, size_t bytes)
            {
                hipError_t e = hipMalloc((void**) p, bytes);
                if (e == hipSuccess)
                    ptrs.push_back((void*) *p);
                return e;
            }
        };

        // A stream of the solver's own (default priority) ordered after the caller's stream.
        struct SolveStream {
            hipStream_t s = nullptr;
            hipError_t init(hipStream_t caller)
            {
                hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
                if (e != hipSuccess)
                    return e;
                hipEvent_t ev;
                e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
                if (e != hipSuccess)
                    return e;
                e = hipEventRecord(ev, caller);
                if (e == hipSuccess)
                    e = hipStreamWaitEvent(s, ev, 0);
                (void) hipEventDestroy(ev);
                return e;
            }
            ~SolveStream()
            {
                if (s) {
                    (void) hipStreamSynchronize(s);
                    (void) hipStreamDestroy(s);
                }
            }
        };

        // Runs body() once eagerly per iteration, or captures it into a graph and replays it.
        template <class Body>
        hipError_t iterate(hipStream_t s, int iterations, bool graph, Body body)
        {
            if (iterations <= 0)
                return hipSuccess;
            if (!graph) {
                for (int i = 0; i < iterations; ++i) {
                    const int rc = body();
                    if (rc != EHP_OK)
                        return hipErrorLaunchFailure;
                }
                return hipSuccess;
            }
            hipGraph_t g = nullptr;
            hipGraphExec_t ge = nullptr;
            hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess)
                return e;
            const int rc = body();
            e = hipStreamEndCapture(s, &g);
            if (rc != EHP_OK) {
                if (g)
                    (void) hipGraphDestroy(g);
                return hipErrorLaunchFailure;
            }
            if (e != hipSuccess)
                return e;
            e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            for (int i = 0; e == hipSuccess && i < iterations; ++i)
                e = hipGraphLaunch(ge, s);
            if (ge)
                (void) hipGraphExecDestroy(ge);
            (void) hipGraphDestroy(g);
            return e;
        }

        template <class T>
        int cgls(ehp_op* op, int64_t nd, int64_t nr, const T* b, T* x, int use_x0, int iterations, double eps,
                 double tol, bool graph, double* residuals, int* iters_done, hipStream_t caller)
        {
            SolveStream ss;
            hipError_t e = ss.init(caller);
            if (e != hipSuccess)
                return ehp_solver_fail(EHP_ERR_HIP, hipGetErrorString(e));
            hipStream_t s = ss.s;
            Scratch m;
            T *r, *c, *sv, *q;
            double *partial, *sc, *hist;
            const int nh = iterations + 1;
            if (m.alloc(&r, nd * sizeof(T)) != hipSuccess || m.alloc(&c, nd * sizeof(T)) != hipSuccess ||
                m.alloc(&sv, nr * sizeof(T)) != hipSuccess || m.alloc(&q, nr * sizeof(T)) != hipSuccess ||
                m.alloc(&partial, kRedBlocks * sizeof(double)) != hipSuccess ||
                m.alloc(&sc, S_COUNT * sizeof(double)) != hipSuccess || m.alloc(&hist, nh * sizeof(double)) != hipSuccess)
                return ehp_solver_fail(EHP_ERR_ALLOC, "hipMalloc of the CGLS workspace failed");
            const double damp = eps * eps;
            int rc = EHP_OK;
            auto ok = [&](hipError_t err) {
                if (err != hipSuccess && rc == EHP_OK)
                    rc = ehp_solver_fail(EHP_ERR_HIP, hipGetErrorString(err));
                return rc == EHP_OK;
            };
            ok(launch_fill<double>(sc, S_COUNT, 0.0, s));
            ok(launch_fill<double>(hist, nh, -1.0, s));
            // s = b - A x (or b), r = A^T s - damp x (CGLS.cpp:30-45)
            if (use_x0) {
                if (rc == EHP_OK)
                    rc = ehp_forward(op, x, sv, s);
                hipLaunchKernelGGL((rsub_kernel<T>), dim3(blocks(nr)), dim3(kThreads), 0, s, b, sv, nr);
            } else {
                ok(launch_fill<T>(x, nd, T(0), s));
                hipLaunchKernelGGL((copy_kernel<T>), dim3(blocks(nr)), dim3(kThreads), 0, s, b, sv, nr);
            }
            if (rc == EHP_OK)
                rc = ehp_adjoint(op, sv, r, s);
            if (use_x0 && damp != 0.0)
                hipLaunchKernelGGL((axpy_kernel<T>), dim3(blocks(nd)), dim3(kThreads), 0, s, (T) -damp, x, r, nd);
            hipLaunchKernelGGL((copy_kernel<T>), dim3(blocks(nd)), dim3(kThreads), 0, s, r, c, nd);
            ok(dot<T>(r, r, nd, partial, sc, S_KOLD, nullptr, s));
            ok(dot<T>(sv, sv, nr, partial, sc, S_SN, hist, s)); // hist[0] = ||s||^2 at the start
            if (rc != EHP_OK)
                return rc;

            const int64_t nmax = nd > nr ? nd : nr;
            auto body = [&]() -> int {
                hipLaunchKernelGGL(cgls_check_kernel, dim3(1), dim3(1), 0, s, sc, tol);
                int r2 = ehp_forward(op, c, q, s); // q = A c
                if (r2 != EHP_OK)
                    return r2;
                if (dot<T>(q, q, nr, partial, sc, S_DELTA, nullptr, s) != hipSuccess)
                    return EHP_ERR_HIP;
                if (damp != 0.0 && dot<T>(c, c, nd, partial, sc, S_CN, nullptr, s) != hipSuccess)
                    return EHP_ERR_HIP;
                hipLaunchKernelGGL(cgls_alpha_kernel, dim3(1), dim3(1), 0, s, sc, damp);
                hipLaunchKernelGGL((cgls_update_kernel<T>), dim3(blocks(nmax)), dim3(kThreads), 0, s, x, c, nd, sv, q,
                                   nr, sc);
                r2 = ehp_adjoint(op, sv, r, s); // r = A^T s
                if (r2 != EHP_OK)
                    return r2;
                if (damp != 0.0)
                    hipLaunchKernelGGL((axpy_kernel<T>), dim3(blocks(nd)), dim3(kThreads), 0, s, (T) -damp, x, r, nd);
                if (dot<T>(r, r, nd, partial, sc, S_K, nullptr, s) != hipSuccess)
                    return EHP_ERR_HIP;
                hipLaunchKernelGGL(cgls_beta_kernel, dim3(1), dim3(1), 0, s, sc);
                hipLaunchKernelGGL((cgls_direction_kernel<T>), dim3(blocks(nd)), dim3(kThreads), 0, s, r, c, nd, sc);
                // ||s||^2 of this iteration, recorded at index iters while not converged
                if (dot<T>(sv, sv, nr, partial, sc, S_SN, hist, s) != hipSuccess)
                    return EHP_ERR_HIP;
                return hipGetLastError() == hipSuccess ? EHP_OK : EHP_ERR_HIP;
            };
            e = iterate(s, iterations, graph, body);
            if (e != hipSuccess)
                return rc != EHP_OK ? rc : ehp_solver_fail(EHP_ERR_HIP, hipGetErrorString(e));
            // the last check: elsa returns before iteration `iterations` if kold < tol
            double hsc[S_COUNT];
            std::vector<double> hh(nh);
            e = hipMemcpyAsync(hsc, sc, sizeof(hsc), hipMemcpyDeviceToHost, s);
            if (e == hipSuccess)
                e = hipMemcpyAsync(hh.data(), hist, nh * sizeof(double), hipMemcpyDeviceToHost, s);
            if (e == hipSuccess)
                e = hipStreamSynchronize(s);
            if (e != hipSuccess)
                return ehp_solver_fail(EHP_ERR_HIP, hipGetErrorString(e));
            if (iters_done)
                *iters_done = (int) hsc[S_ITERS];
            if (residuals)
                for (int i = 0; i < nh; ++i)
                    residuals[i] = hh[i] >= 0.0 ? std::sqrt(hh[i]) : -1.0;
            return EHP_OK;
        }

        template <class T>
        int power_iterations(ehp_op* op, int64_t nd, int64_t nr, int niters, double* L, hipStream_t s, Scratch& m,
                             double* partial, double* sc)
        {
            T *u, *t;
            if (m.alloc(&u, nd * sizeof(T)) != hipSuccess || m.alloc(&t, nr * sizeof(T)) != hipSuccess)
                return ehp_solver_fail(EHP_ERR_ALLOC, "hipMalloc of the power-iteration workspace failed");
            if (launch_fill<T>(u, nd, T(1), s) != hipSuccess)
                return ehp_solver_fail(EHP_ERR_HIP, "fill");
            // u = A^T A u; u /= ||u|| (PowerIterations.cpp:18-23)
            for (int i = 0; i < niters; ++i) {
                int rc = ehp_forward(op, u, t, s);
                if (rc == EHP_OK)
                    rc = ehp_adjoint(op, t, u, s);
                if (rc != EHP_OK)
                    return rc;
                if (dot<T>(u, u, nd, partial, sc, S_TMP, nullptr, s) != hipSuccess)
                    return ehp_solver_fail(EHP_ERR_HIP, "dot");
                hipLaunchKernelGGL((normalize_kernel<T>), dim3(blocks(nd)), dim3(kThreads), 0, s, u, nd, sc, S_TMP);
            }
            // u . (A^T A u) / ||u||  (PowerIterations.cpp:25)
            T* v;
            if (m.alloc(&v, nd * sizeof(T)) != hipSuccess)
                return ehp_solver_fail(EHP_ERR_ALLOC, "hipMalloc of the power-iteration workspace failed");
            int rc = ehp_forward(op, u, t, s);
            if (rc == EHP_OK)
                rc = ehp_adjoint(op, t, v, s);
            if (rc != EHP_OK)
                return rc;
            // (S_KOLD is free here and takes ||u||^2)
            if (dot<T>(u, v, nd, partial, sc, S_TMP, nullptr, s) != hipSuccess ||
                dot<T>(u, u, nd, partial, sc, S_KOLD, nullptr, s) != hipSuccess)
                return ehp_solver_fail(EHP_ERR_HIP, "dot");
            double h[S_COUNT];
            hipError_t e = hipMemcpyAsync(h, sc, sizeof(h), hipMemcpyDeviceToHost, s);
            if (e == hipSuccess)
                e = hipStreamSynchronize(s);
            if (e != hipSuccess)
                return ehp_solver_fail(EHP_ERR_HIP, hipGetErrorString(e));
            *L = h[S_TMP] / std::sqrt(h[S_KOLD]);
            return EHP_OK;
        }

        template <class T>
        int apgd(ehp_op* op, int64_t nd, int64_t nr, const T* b, T* x, int use_x0, int iterations, double mu_in,
                 double sigma, double epsilon, bool graph, double* grad_norms, int* iters_done, double* mu_used,
                 hipStream_t caller)
        {
            SolveStream ss;
            hipError_t e = ss.init(caller);
            if (e != hipSuccess)
                return ehp_solver_fail(EHP_ERR_HIP, hipGetErrorString(e));
            hipStream_t s = ss.s;
            Scratch m;
            double *partial, *sc, *hist;
            const int nh = iterations + 1;
            if (m.alloc(&partial, kRedBlocks * sizeof(double)) != hipSuccess ||
                m.alloc(&sc, S_COUNT * sizeof(double)) != hipSuccess || m.alloc(&hist, nh * sizeof(double)) != hipSuccess)
                return ehp_solver_fail(EHP_ERR_ALLOC, "hipMalloc of the APGD workspace failed");
            if (launch_fill<double>(sc, S_COUNT, 0.0, s) != hipSuccess ||
                launch_fill<double>(hist, nh, -1.0, s) != hipSuccess)
                return ehp_solver_fail(EHP_ERR_HIP, "fill");
            // mu = 0.45 / L with L from power iterations on A^T A (APGD.cpp:19-27)
            double mu = mu_in;
            if (!(mu > 0.0)) {
                double L = 0.0;
                const int rc = power_iterations<T>(op, nd, nr, 5, &L, s, m, partial, sc);
                if (rc != EHP_OK)
                    return rc;
                mu = (double) (T(0.45) / (T) L);
            }
            if (mu_used)
                *mu_used = mu;
            T *xprev, *y, *grad, *atb, *ay;
            if (m.alloc(&xprev, nd * sizeof(T)) != hipSuccess || m.alloc(&y, nd * sizeof(T)) != hipSuccess ||
                m.alloc(&grad, nd * sizeof(T)) != hipSuccess || m.alloc(&atb, nd * sizeof(T)) != hipSuccess ||
                m.alloc(&ay, nr * sizeof(T)) != hipSuccess)
                return ehp_solver_fail(EHP_ERR_ALLOC, "hipMalloc of the APGD workspace failed");
            int rc = EHP_OK;
            if (!use_x0 && launch_fill<T>(x, nd, T(0), s) != hipSuccess)
                return ehp_solver_fail(EHP_ERR_HIP, "fill");
            hipLaunchKernelGGL((copy_kernel<T>), dim3(blocks(nd)), dim3(kThreads), 0, s, x, xprev, nd);
            hipLaunchKernelGGL((copy_kernel<T>), dim3(blocks(nd)), dim3(kThreads), 0, s, x, y, nd);
            hipLaunchKernelGGL(set_scalar_kernel, dim3(1), dim3(1), 0, s, sc, (int) S_T, 1.0);
            rc = ehp_adjoint(op, b, atb, s); // A^T b
            // grad = A^T (A y) - A^T b (APGD.cpp:55-60)
            auto gradient = [&]() -> int {
                int r2 = ehp_forward(op, y, ay, s);
                if (r2 == EHP_OK)
                    r2 = ehp_adjoint(op, ay, grad, s);
                if (r2 == EHP_OK)
                    hipLaunchKernelGGL((axpy_kernel<T>), dim3(blocks(nd)), dim3(kThreads), 0, s, T(-1), atb, grad, nd);
                return r2;
            };
            if (rc == EHP_OK)
                rc = gradient();
            if (rc != EHP_OK)
                return rc;
            const T tmu = (T) mu, tsig = sigma >= 0.0 ? (T) sigma : T(-1);
            auto body = [&]() -> int {
                hipLaunchKernelGGL((apgd_prox_kernel<T>), dim3(blocks(nd)), dim3(kThreads), 0, s, y, grad, tmu, tsig, x,
                                   nd, sc);
                hipLaunchKernelGGL((apgd_t_kernel<T>), dim3(1), dim3(1), 0, s, sc);
                hipLaunchKernelGGL((apgd_momentum_kernel<T>), dim3(blocks(nd)), dim3(kThreads), 0, s, x, xprev, y, nd, sc);
                if (dot<T>(grad, grad, nd, partial, sc, S_SN, hist, s) != hipSuccess)
                    return EHP_ERR_HIP;
                hipLaunchKernelGGL(apgd_check_kernel, dim3(1), dim3(1), 0, s, sc, epsilon);
                const int r2 = gradient();
                if (r2 != EHP_OK)
                    return r2;
                return hipGetLastError() == hipSuccess ? EHP_OK : EHP_ERR_HIP;
            };
            e = iterate(s, iterations, graph, body);
            if (e != hipSuccess)
                return ehp_solver_fail(EHP_ERR_HIP, hipGetErrorString(e));
            double hsc[S_COUNT];
            std::vector<double> hh(nh);
            e = hipMemcpyAsync(hsc, sc, sizeof(hsc), hipMemcpyDeviceToHost, s);
            if (e == hipSuccess)
                e = hipMemcpyAsync(hh.data(), hist, nh * sizeof(double), hipMemcpyDeviceToHost, s);
            if (e == hipSuccess)
                e = hipStreamSynchronize(s);
            if (e != hipSuccess)
                return ehp_solver_fail(EHP_ERR_HIP, hipGetErrorString(e));
            if (iters_done)
                *iters_done = (int) hsc[S_ITERS];
            if (grad_norms)
                for (int i = 0; i < nh; ++i)
                    grad_norms[i] = hh[i];
            return EHP_OK;
        }
    } // namespace

    // dispatch helpers for the C-ABI (ehp_capi.hip owns the operator)
    int solver_cgls(ehp_op* op, int dtype, int64_t nd, int64_t nr, const void* b, void* x, int use_x0, int iterations,
                    double eps, double tol, bool graph, double* residuals, int* iters_done, hipStream_t s)
    {
        if (dtype == EHP_F64)
            return cgls<double>(op, nd, nr, (const double*) b, (double*) x, use_x0, iterations, eps, tol, graph,
                                residuals, iters_done, s);
        return cgls<float>(op, nd, nr, (const float*) b, (float*) x, use_x0, iterations, eps, tol, graph, residuals,
                           iters_done, s);
    }

    int solver_apgd(ehp_op* op, int dtype, int64_t nd, int64_t nr, const void* b, void* x, int use_x0, int iterations,
                    double mu, double sigma, double epsilon, bool graph, double* grad_norms, int* iters_done,
                    double* mu_used, hipStream_t s)
    {
        if (dtype == EHP_F64)
            return apgd<double>(op, nd, nr, (const double*) b, (double*) x, use_x0, iterations, mu, sigma, epsilon,
                                graph, grad_norms, iters_done, mu_used, s);
        return apgd<float>(op, nd, nr, (const float*) b, (float*) x, use_x0, iterations, mu, sigma, epsilon, graph,
                           grad_norms, iters_done, mu_used, s);
    }

    int solver_power_iterations(ehp_op* op, int dtype, int64_t nd, int64_t nr, int niters, double* L, hipStream_t caller)
    {
        SolveStream ss;
        hipError_t e = ss.init(caller);
        if (e != hipSuccess)
            return ehp_solver_fail(EHP_ERR_HIP, hipGetErrorString(e));
        Scratch m;
        double *partial, *sc;
        if (m.alloc(&partial, kRedBlocks * sizeof(double)) != hipSuccess ||
            m.alloc(&sc, S_COUNT * sizeof(double)) != hipSuccess)
            return ehp_solver_fail(EHP_ERR_ALLOC, "hipMalloc of the power-iteration workspace failed");
        if (dtype == EHP_F64)
            return power_iterations<double>(op, nd, nr, niters, L, ss.s, m, partial, sc);
        return power_iterations<float>(op, nd, nr, niters, L, ss.s, m, partial, sc);
    }
} // namespace ehp
