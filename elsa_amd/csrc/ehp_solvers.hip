// ehp_solvers.hip -- elsa's CGLS, APGD (FISTA) and power iterations, resident on the device.
//
// The iterations call the projector through the C-ABI (an operator's ehp_forward /
// ehp_adjoint, or a batch's) and keep every scalar of the method -- step lengths, momentum,
// norms, the convergence state -- in device memory, so an iteration needs no host round trip
// and can be captured once into a hipGraph and replayed. Norms and dot products are FP64
// two-stage reductions in a fixed order (bitwise reproducible); the vector updates run in the
// operator's data type.
//
// A solve runs on nseg independent problems at once (a batch; a single operator is a batch of
// one): vectors are the problems' own vectors packed back to back, every scalar exists once
// per problem, and each problem follows its own iteration of the method exactly as if it were
// solved alone -- its own step lengths, its own convergence test. Element-wise kernels give
// each problem ceil(len / 256) blocks; reductions give it up to 1024 partial sums.
//
// elsa returns early when a method has converged (CGLS.cpp:54-58, APGD.cpp:83-87). Here the
// check sets a per-problem device flag; once it is set the updates of that problem's x are
// skipped, so its result is the iterate at which elsa would have returned, while the remaining
// replays do no harm.
#include "elsa_hip.h"
#include "ehp_solvers.hpp"

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

int ehp_solver_fail(int code, const char* msg);

namespace ehp
{
    namespace
    {
        constexpr int kRedBlocks = 1024; // partial sums per problem at most
        constexpr int kThreads = 256;

        // scalar slots (per problem)
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
            S_MU,       // APGD step length (in the data type)
            S_COUNT
        };

        // Blocks of one vector space (domain or range) of the packed problems, on the device.
        struct SegSpace {
            const int64_t* off; // nseg + 1 element offsets
            const int* eseg;    // element-wise blocks: problem of each block
            const int64_t* eb0; // nseg + 1: first element-wise block of each problem
            const int* rseg;    // reduction blocks: problem of each block
            const int64_t* rb0; // nseg + 1: first reduction block of each problem
            int64_t neb, nrb;   // block counts
        };

        // Element i (packed index) of this thread in an element-wise launch over S; false past the
        // end of its problem.
        __device__ inline bool seg_elem(const SegSpace& S, int& seg, int64_t& i)
        {
            seg = S.eseg[blockIdx.x];
            i = S.off[seg] + ((int64_t) blockIdx.x - S.eb0[seg]) * kThreads + threadIdx.x;
            return i < S.off[seg + 1];
        }

        __device__ inline double& scal(double* sc, int seg, int slot)
        {
            return sc[seg * S_COUNT + slot];
        }

        template <class T>
        __global__ void __launch_bounds__(kThreads)
            dot_partial_kernel(const T* __restrict__ a, const T* __restrict__ b, const SegSpace S,
                               double* __restrict__ partial)
        {
            const int seg = S.rseg[blockIdx.x];
            const int64_t lb = (int64_t) blockIdx.x - S.rb0[seg];
            const int64_t nb = S.rb0[seg + 1] - S.rb0[seg];
            const int64_t end = S.off[seg + 1];
            double acc = 0.0;
            for (int64_t i = S.off[seg] + lb * kThreads + threadIdx.x; i < end; i += nb * kThreads)
                acc += (double) up(a[i]) * (double) up(b[i]);
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

        // per problem (one block each): slot = sum of its partial sums in a fixed order; hist, when
        // given, gets a copy at index iters of the problem's row (only while not converged)
        __global__ void __launch_bounds__(kThreads)
            dot_final_kernel(const double* __restrict__ partial, const SegSpace S, double* __restrict__ sc, int slot,
                             double* __restrict__ hist, int nh)
        {
            const int seg = blockIdx.x;
            __shared__ double red[kThreads];
            double acc = 0.0;
            for (int64_t i = S.rb0[seg] + threadIdx.x; i < S.rb0[seg + 1]; i += kThreads)
                acc += partial[i];
            red[threadIdx.x] = acc;
            __syncthreads();
            for (int off = kThreads / 2; off > 0; off >>= 1) {
                if ((int) threadIdx.x < off)
                    red[threadIdx.x] += red[threadIdx.x + off];
                __syncthreads();
            }
            if (threadIdx.x == 0) {
                scal(sc, seg, slot) = red[0];
                if (hist && scal(sc, seg, S_DONE) == 0.0)
                    hist[(int64_t) seg * nh + (int) scal(sc, seg, S_ITERS)] = red[0];
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
                s[i] = down<T>(up(b[i]) - up(s[i]));
        }

        // y = a * x + y with a host scalar (in the arithmetic type)
        template <class T>
        __global__ void __launch_bounds__(kThreads)
            axpy_kernel(typename Acc<T>::type a, const T* __restrict__ x, T* __restrict__ y, int64_t n)
        {
            const int64_t i = (int64_t) blockIdx.x * kThreads + threadIdx.x;
            if (i < n)
                y[i] = down<T>(a * up(x[i]) + up(y[i]));
        }

        // CGLS: y += sign * alpha c per problem (x += alpha c over the domain, s -= alpha q over the
        // range)
        template <class T>
        __global__ void __launch_bounds__(kThreads)
            cgls_update_kernel(T* __restrict__ y, const T* __restrict__ c, const SegSpace S, double* __restrict__ sc,
                               int negate)
        {
            int seg;
            int64_t i;
            if (!seg_elem(S, seg, i) || scal(sc, seg, S_DONE) != 0.0)
                return;
            using A = typename Acc<T>::type;
            const A alpha = (A) scal(sc, seg, S_ALPHA);
            if (negate)
                y[i] = down<T>(up(y[i]) - alpha * up(c[i]));
            else
                y[i] = down<T>(up(y[i]) + alpha * up(c[i]));
        }

        // c = r + beta c
        template <class T>
        __global__ void __launch_bounds__(kThreads)
            cgls_direction_kernel(const T* __restrict__ r, T* __restrict__ c, const SegSpace S, double* __restrict__ sc)
        {
            int seg;
            int64_t i;
            if (!seg_elem(S, seg, i))
                return;
            using A = typename Acc<T>::type;
            const A beta = (A) scal(sc, seg, S_BETA);
            c[i] = down<T>(up(r[i]) + beta * up(c[i]));
        }

        // start of a CGLS iteration: stop when kold < tol (CGLS.cpp:54-58), else count it
        __global__ void cgls_check_kernel(double* sc, int nseg, double tol)
        {
            const int seg = blockIdx.x * blockDim.x + threadIdx.x;
            if (seg >= nseg)
                return;
            if (scal(sc, seg, S_DONE) == 0.0 && scal(sc, seg, S_KOLD) < tol)
                scal(sc, seg, S_DONE) = 1.0;
            if (scal(sc, seg, S_DONE) == 0.0)
                scal(sc, seg, S_ITERS) += 1.0;
        }

        __global__ void cgls_alpha_kernel(double* sc, int nseg, double damp)
        {
            const int seg = blockIdx.x * blockDim.x + threadIdx.x;
            if (seg >= nseg)
                return;
            const double d = damp == 0.0 ? scal(sc, seg, S_DELTA) : scal(sc, seg, S_DELTA) + damp * scal(sc, seg, S_CN);
            scal(sc, seg, S_ALPHA) = scal(sc, seg, S_KOLD) / d;
        }

        __global__ void cgls_beta_kernel(double* sc, int nseg)
        {
            const int seg = blockIdx.x * blockDim.x + threadIdx.x;
            if (seg >= nseg)
                return;
            if (scal(sc, seg, S_DONE) != 0.0) {
                scal(sc, seg, S_BETA) = 0.0;
                return;
            }
            scal(sc, seg, S_BETA) = scal(sc, seg, S_K) / scal(sc, seg, S_KOLD);
            scal(sc, seg, S_KOLD) = scal(sc, seg, S_K);
        }

        __global__ void set_scalar_kernel(double* sc, int nseg, int slot, double v)
        {
            const int seg = blockIdx.x * blockDim.x + threadIdx.x;
            if (seg < nseg)
                scal(sc, seg, slot) = v;
        }

        // APGD: x = prox(y - mu grad), prox the identity (sigma < 0) or L1 soft thresholding with
        // threshold mu * sigma (ProximalL1.cpp:32-43)
        template <class T>
        __global__ void __launch_bounds__(kThreads)
            apgd_prox_kernel(const T* __restrict__ y, const T* __restrict__ grad, T sigma, T* __restrict__ x,
                             const SegSpace S, double* __restrict__ sc)
        {
            int seg;
            int64_t i;
            if (!seg_elem(S, seg, i) || scal(sc, seg, S_DONE) != 0.0)
                return;
            const T mu = (T) scal(sc, seg, S_MU);
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
        __global__ void apgd_t_kernel(double* sc, int nseg)
        {
            const int seg = blockIdx.x * blockDim.x + threadIdx.x;
            if (seg >= nseg || scal(sc, seg, S_DONE) != 0.0)
                return;
            const T tprev = (T) scal(sc, seg, S_T);
            const T t = (T(1) + sqrt(T(1) + T(4) * tprev * tprev)) / T(2);
            scal(sc, seg, S_COEF) = (double) ((tprev - T(1)) / t);
            scal(sc, seg, S_T) = (double) t;
        }

        // y = x + coef (x - xprev); xprev = x
        template <class T>
        __global__ void __launch_bounds__(kThreads)
            apgd_momentum_kernel(const T* __restrict__ x, T* __restrict__ xprev, T* __restrict__ y, const SegSpace S,
                                 double* __restrict__ sc)
        {
            int seg;
            int64_t i;
            if (!seg_elem(S, seg, i) || scal(sc, seg, S_DONE) != 0.0)
                return;
            const T coef = (T) scal(sc, seg, S_COEF);
            const T xi = x[i];
            y[i] = xi + coef * (xi - xprev[i]);
            xprev[i] = xi;
        }

        // APGD convergence: ||grad||^2 <= epsilon ends the method after this iteration
        // (APGD.cpp:83-87)
        __global__ void apgd_check_kernel(double* sc, int nseg, double eps)
        {
            const int seg = blockIdx.x * blockDim.x + threadIdx.x;
            if (seg >= nseg || scal(sc, seg, S_DONE) != 0.0)
                return;
            scal(sc, seg, S_ITERS) += 1.0;
            if (scal(sc, seg, S_SN) <= eps)
                scal(sc, seg, S_DONE) = 1.0;
        }

        // u = u / sqrt(||u||^2) per problem
        template <class T>
        __global__ void __launch_bounds__(kThreads)
            normalize_kernel(T* __restrict__ u, const SegSpace S, double* __restrict__ sc, int slot)
        {
            int seg;
            int64_t i;
            if (seg_elem(S, seg, i))
                u[i] = down<T>((typename Acc<T>::type) ((double) up(u[i]) / sqrt(scal(sc, seg, slot))));
        }

        inline unsigned blocks(int64_t n)
        {
            return (unsigned) ((n + kThreads - 1) / kThreads);
        }

        inline unsigned seg_blocks(int nseg)
        {
            return (unsigned) ((nseg + 63) / 64);
        }

        template <class T>
        hipError_t dot(const T* a, const T* b, const SegSpace& S, int nseg, double* partial, double* sc, int slot,
                       double* hist, int nh, hipStream_t s)
        {
            hipLaunchKernelGGL((dot_partial_kernel<T>), dim3((unsigned) S.nrb), dim3(kThreads), 0, s, a, b, S, partial);
            hipLaunchKernelGGL(dot_final_kernel, dim3((unsigned) nseg), dim3(kThreads), 0, s, partial, S, sc, slot,
                               hist, nh);
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

        // The block tables of one vector space with segment offsets off[0..nseg], uploaded with s
        // (synchronously: the host tables are local).
        hipError_t make_space(const int64_t* off, int nseg, Scratch& m, SegSpace& S, hipStream_t s)
        {
            std::vector<int64_t> eb0(nseg + 1), rb0(nseg + 1);
            std::vector<int> eseg, rseg;
            for (int j = 0; j < nseg; ++j) {
                const int64_t len = off[j + 1] - off[j];
                const int64_t ne = (len + kThreads - 1) / kThreads;
                int64_t nr = (len + 8 * kThreads - 1) / (8 * kThreads);
                nr = nr < 1 ? 1 : (nr > kRedBlocks ? kRedBlocks : nr);
                eb0[j] = (int64_t) eseg.size();
                rb0[j] = (int64_t) rseg.size();
                eseg.insert(eseg.end(), (size_t) ne, j);
                rseg.insert(rseg.end(), (size_t) nr, j);
            }
            eb0[nseg] = (int64_t) eseg.size();
            rb0[nseg] = (int64_t) rseg.size();
            if (eseg.size() > 0x7fffffffu)
                return hipErrorInvalidValue;
            int64_t *d_off, *d_eb0, *d_rb0;
            int *d_eseg, *d_rseg;
            const size_t n64 = (size_t) (nseg + 1) * sizeof(int64_t);
            hipError_t e = m.alloc(&d_off, n64);
            if (e == hipSuccess)
                e = m.alloc(&d_eb0, n64);
            if (e == hipSuccess)
                e = m.alloc(&d_rb0, n64);
            if (e == hipSuccess)
                e = m.alloc(&d_eseg, (eseg.size() + 1) * sizeof(int));
            if (e == hipSuccess)
                e = m.alloc(&d_rseg, (rseg.size() + 1) * sizeof(int));
            if (e == hipSuccess)
                e = hipMemcpyAsync(d_off, off, n64, hipMemcpyHostToDevice, s);
            if (e == hipSuccess)
                e = hipMemcpyAsync(d_eb0, eb0.data(), n64, hipMemcpyHostToDevice, s);
            if (e == hipSuccess)
                e = hipMemcpyAsync(d_rb0, rb0.data(), n64, hipMemcpyHostToDevice, s);
            if (e == hipSuccess && !eseg.empty())
                e = hipMemcpyAsync(d_eseg, eseg.data(), eseg.size() * sizeof(int), hipMemcpyHostToDevice, s);
            if (e == hipSuccess)
                e = hipMemcpyAsync(d_rseg, rseg.data(), rseg.size() * sizeof(int), hipMemcpyHostToDevice, s);
            if (e == hipSuccess)
                e = hipStreamSynchronize(s);
            if (e != hipSuccess)
                return e;
            S.off = d_off;
            S.eb0 = d_eb0;
            S.rb0 = d_rb0;
            S.eseg = d_eseg;
            S.rseg = d_rseg;
            S.neb = (int64_t) eseg.size();
            S.nrb = (int64_t) rseg.size();
            return hipSuccess;
        }

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

        // common set-up of a solve: its stream, workspace and block tables
        struct Solve {
            Scratch m;       // declared first: freed after the stream has drained
            SolveStream ss;
            SegSpace D{}, R{};
            double *partial = nullptr, *sc = nullptr;
            int nseg = 0;
            int64_t nd = 0, nr = 0;

            int init(const SolveOps& A, hipStream_t caller, const char* what)
            {
                nseg = A.nseg;
                nd = A.doff[nseg];
                nr = A.roff[nseg];
                hipError_t e = ss.init(caller);
                if (e == hipSuccess)
                    e = make_space(A.doff, nseg, m, D, ss.s);
                if (e == hipSuccess)
                    e = make_space(A.roff, nseg, m, R, ss.s);
                if (e != hipSuccess)
                    return ehp_solver_fail(EHP_ERR_HIP, hipGetErrorString(e));
                const int64_t np = D.nrb > R.nrb ? D.nrb : R.nrb;
                if (m.alloc(&partial, (size_t) np * sizeof(double)) != hipSuccess ||
                    m.alloc(&sc, (size_t) nseg * S_COUNT * sizeof(double)) != hipSuccess)
                    return ehp_solver_fail(EHP_ERR_ALLOC, what);
                if (launch_fill<double>(sc, (int64_t) nseg * S_COUNT, 0.0, ss.s) != hipSuccess)
                    return ehp_solver_fail(EHP_ERR_HIP, "fill");
                return EHP_OK;
            }
        };

        // device scalars and history rows back to the host
        int fetch(hipStream_t s, const double* sc, int nseg, const double* hist, int nh, std::vector<double>& hsc,
                  std::vector<double>& hh)
        {
            hsc.resize((size_t) nseg * S_COUNT);
            hh.resize((size_t) nseg * nh);
            hipError_t e = hipMemcpyAsync(hsc.data(), sc, hsc.size() * sizeof(double), hipMemcpyDeviceToHost, s);
            if (e == hipSuccess && hist)
                e = hipMemcpyAsync(hh.data(), hist, hh.size() * sizeof(double), hipMemcpyDeviceToHost, s);
            if (e == hipSuccess)
                e = hipStreamSynchronize(s);
            if (e != hipSuccess)
                return ehp_solver_fail(EHP_ERR_HIP, hipGetErrorString(e));
            return EHP_OK;
        }

        template <class T>
        int cgls(const SolveOps& A, const T* b, T* x, int use_x0, int iterations, double eps, double tol, bool graph,
                 double* residuals, int* iters_done, hipStream_t caller)
        {
            Solve sv;
            int rc = sv.init(A, caller, "hipMalloc of the CGLS workspace failed");
            if (rc != EHP_OK)
                return rc;
            hipStream_t s = sv.ss.s;
            const int nseg = sv.nseg;
            const int64_t nd = sv.nd, nr = sv.nr;
            const SegSpace &D = sv.D, &R = sv.R;
            double *partial = sv.partial, *sc = sv.sc, *hist;
            T *r, *c, *sres, *q;
            const int nh = iterations + 1;
            Scratch& m = sv.m;
            if (m.alloc(&r, nd * sizeof(T)) != hipSuccess || m.alloc(&c, nd * sizeof(T)) != hipSuccess ||
                m.alloc(&sres, nr * sizeof(T)) != hipSuccess || m.alloc(&q, nr * sizeof(T)) != hipSuccess ||
                m.alloc(&hist, (size_t) nseg * nh * sizeof(double)) != hipSuccess)
                return ehp_solver_fail(EHP_ERR_ALLOC, "hipMalloc of the CGLS workspace failed");
            const double damp = eps * eps;
            auto ok = [&](hipError_t err) {
                if (err != hipSuccess && rc == EHP_OK)
                    rc = ehp_solver_fail(EHP_ERR_HIP, hipGetErrorString(err));
                return rc == EHP_OK;
            };
            ok(launch_fill<double>(hist, (int64_t) nseg * nh, -1.0, s));
            // s = b - A x (or b), r = A^T s - damp x (CGLS.cpp:30-45)
            if (use_x0) {
                if (rc == EHP_OK)
                    rc = A.fwd(x, sres, s);
                hipLaunchKernelGGL((rsub_kernel<T>), dim3(blocks(nr)), dim3(kThreads), 0, s, b, sres, nr);
            } else {
                ok(launch_fill<T>(x, nd, down<T>(0), s));
                hipLaunchKernelGGL((copy_kernel<T>), dim3(blocks(nr)), dim3(kThreads), 0, s, b, sres, nr);
            }
            if (rc == EHP_OK)
                rc = A.adj(sres, r, s);
            if (use_x0 && damp != 0.0)
                hipLaunchKernelGGL((axpy_kernel<T>), dim3(blocks(nd)), dim3(kThreads), 0, s, (typename Acc<T>::type) -damp, x,
                                   r, nd);
            hipLaunchKernelGGL((copy_kernel<T>), dim3(blocks(nd)), dim3(kThreads), 0, s, r, c, nd);
            ok(dot<T>(r, r, D, nseg, partial, sc, S_KOLD, nullptr, nh, s));
            ok(dot<T>(sres, sres, R, nseg, partial, sc, S_SN, hist, nh, s)); // hist[0] = ||s||^2 at the start
            if (rc != EHP_OK)
                return rc;

            auto body = [&]() -> int {
                hipLaunchKernelGGL(cgls_check_kernel, dim3(seg_blocks(nseg)), dim3(64), 0, s, sc, nseg, tol);
                int r2 = A.fwd(c, q, s); // q = A c
                if (r2 != EHP_OK)
                    return r2;
                if (dot<T>(q, q, R, nseg, partial, sc, S_DELTA, nullptr, nh, s) != hipSuccess)
                    return EHP_ERR_HIP;
                if (damp != 0.0 && dot<T>(c, c, D, nseg, partial, sc, S_CN, nullptr, nh, s) != hipSuccess)
                    return EHP_ERR_HIP;
                hipLaunchKernelGGL(cgls_alpha_kernel, dim3(seg_blocks(nseg)), dim3(64), 0, s, sc, nseg, damp);
                hipLaunchKernelGGL((cgls_update_kernel<T>), dim3((unsigned) D.neb), dim3(kThreads), 0, s, x, c, D, sc, 0);
                hipLaunchKernelGGL((cgls_update_kernel<T>), dim3((unsigned) R.neb), dim3(kThreads), 0, s, sres, q, R, sc,
                                   1);
                r2 = A.adj(sres, r, s); // r = A^T s
                if (r2 != EHP_OK)
                    return r2;
                if (damp != 0.0)
                    hipLaunchKernelGGL((axpy_kernel<T>), dim3(blocks(nd)), dim3(kThreads), 0, s, (typename Acc<T>::type) -damp, x,
                                   r, nd);
                if (dot<T>(r, r, D, nseg, partial, sc, S_K, nullptr, nh, s) != hipSuccess)
                    return EHP_ERR_HIP;
                hipLaunchKernelGGL(cgls_beta_kernel, dim3(seg_blocks(nseg)), dim3(64), 0, s, sc, nseg);
                hipLaunchKernelGGL((cgls_direction_kernel<T>), dim3((unsigned) D.neb), dim3(kThreads), 0, s, r, c, D, sc);
                // ||s||^2 of this iteration, recorded at index iters while not converged
                if (dot<T>(sres, sres, R, nseg, partial, sc, S_SN, hist, nh, s) != hipSuccess)
                    return EHP_ERR_HIP;
                return hipGetLastError() == hipSuccess ? EHP_OK : EHP_ERR_HIP;
            };
            hipError_t e = iterate(s, iterations, graph, body);
            if (e != hipSuccess)
                return rc != EHP_OK ? rc : ehp_solver_fail(EHP_ERR_HIP, hipGetErrorString(e));
            std::vector<double> hsc, hh;
            rc = fetch(s, sc, nseg, hist, nh, hsc, hh);
            if (rc != EHP_OK)
                return rc;
            for (int j = 0; j < nseg; ++j) {
                if (iters_done)
                    iters_done[j] = (int) hsc[(size_t) j * S_COUNT + S_ITERS];
                if (residuals)
                    for (int i = 0; i < nh; ++i) {
                        const double h = hh[(size_t) j * nh + i];
                        residuals[(size_t) j * nh + i] = h >= 0.0 ? std::sqrt(h) : -1.0;
                    }
            }
            return EHP_OK;
        }

        // L[j] of every problem: 5 (niters) power iterations on A_j^T A_j from the ones vector
        template <class T>
        int power_iterations(const SolveOps& A, Solve& sv, int niters, double* L)
        {
            hipStream_t s = sv.ss.s;
            const int nseg = sv.nseg;
            T *u, *t, *v;
            if (sv.m.alloc(&u, sv.nd * sizeof(T)) != hipSuccess || sv.m.alloc(&t, sv.nr * sizeof(T)) != hipSuccess ||
                sv.m.alloc(&v, sv.nd * sizeof(T)) != hipSuccess)
                return ehp_solver_fail(EHP_ERR_ALLOC, "hipMalloc of the power-iteration workspace failed");
            if (launch_fill<T>(u, sv.nd, down<T>(1), s) != hipSuccess)
                return ehp_solver_fail(EHP_ERR_HIP, "fill");
            // u = A^T A u; u /= ||u|| (PowerIterations.cpp:18-23)
            for (int i = 0; i < niters; ++i) {
                int rc = A.fwd(u, t, s);
                if (rc == EHP_OK)
                    rc = A.adj(t, u, s);
                if (rc != EHP_OK)
                    return rc;
                if (dot<T>(u, u, sv.D, nseg, sv.partial, sv.sc, S_TMP, nullptr, 0, s) != hipSuccess)
                    return ehp_solver_fail(EHP_ERR_HIP, "dot");
                hipLaunchKernelGGL((normalize_kernel<T>), dim3((unsigned) sv.D.neb), dim3(kThreads), 0, s, u, sv.D, sv.sc,
                                   (int) S_TMP);
            }
            // u . (A^T A u) / ||u||  (PowerIterations.cpp:25)
            int rc = A.fwd(u, t, s);
            if (rc == EHP_OK)
                rc = A.adj(t, v, s);
            if (rc != EHP_OK)
                return rc;
            // (S_KOLD is free here and takes ||u||^2)
            if (dot<T>(u, v, sv.D, nseg, sv.partial, sv.sc, S_TMP, nullptr, 0, s) != hipSuccess ||
                dot<T>(u, u, sv.D, nseg, sv.partial, sv.sc, S_KOLD, nullptr, 0, s) != hipSuccess)
                return ehp_solver_fail(EHP_ERR_HIP, "dot");
            std::vector<double> hsc, hh;
            rc = fetch(s, sv.sc, nseg, nullptr, 0, hsc, hh);
            if (rc != EHP_OK)
                return rc;
            for (int j = 0; j < nseg; ++j)
                L[j] = hsc[(size_t) j * S_COUNT + S_TMP] / std::sqrt(hsc[(size_t) j * S_COUNT + S_KOLD]);
            return EHP_OK;
        }

        template <class T>
        int apgd(const SolveOps& A, const T* b, T* x, int use_x0, int iterations, const double* mu_in, double sigma,
                 double epsilon, bool graph, double* grad_norms, int* iters_done, double* mu_used, hipStream_t caller)
        {
            Solve sv;
            int rc = sv.init(A, caller, "hipMalloc of the APGD workspace failed");
            if (rc != EHP_OK)
                return rc;
            hipStream_t s = sv.ss.s;
            const int nseg = sv.nseg;
            const int64_t nd = sv.nd, nr = sv.nr;
            const SegSpace& D = sv.D;
            double *partial = sv.partial, *sc = sv.sc, *hist;
            const int nh = iterations + 1;
            Scratch& m = sv.m;
            if (m.alloc(&hist, (size_t) nseg * nh * sizeof(double)) != hipSuccess)
                return ehp_solver_fail(EHP_ERR_ALLOC, "hipMalloc of the APGD workspace failed");
            if (launch_fill<double>(hist, (int64_t) nseg * nh, -1.0, s) != hipSuccess)
                return ehp_solver_fail(EHP_ERR_HIP, "fill");
            // mu = 0.45 / L with L from power iterations on A^T A (APGD.cpp:19-27), per problem
            std::vector<double> mu(nseg), L(nseg, 0.0);
            bool need_L = false;
            for (int j = 0; j < nseg; ++j) {
                mu[j] = mu_in ? mu_in[j] : 0.0;
                need_L = need_L || !(mu[j] > 0.0);
            }
            if (need_L) {
                rc = power_iterations<T>(A, sv, 5, L.data());
                if (rc != EHP_OK)
                    return rc;
            }
            for (int j = 0; j < nseg; ++j) {
                if (!(mu[j] > 0.0))
                    mu[j] = (double) (T(0.45) / (T) L[j]);
                if (mu_used)
                    mu_used[j] = mu[j];
                hipLaunchKernelGGL(set_scalar_kernel, dim3(1), dim3(1), 0, s, sc + (size_t) j * S_COUNT, 1, (int) S_MU,
                                   mu[j]);
            }
            T *xprev, *y, *grad, *atb, *ay;
            if (m.alloc(&xprev, nd * sizeof(T)) != hipSuccess || m.alloc(&y, nd * sizeof(T)) != hipSuccess ||
                m.alloc(&grad, nd * sizeof(T)) != hipSuccess || m.alloc(&atb, nd * sizeof(T)) != hipSuccess ||
                m.alloc(&ay, nr * sizeof(T)) != hipSuccess)
                return ehp_solver_fail(EHP_ERR_ALLOC, "hipMalloc of the APGD workspace failed");
            if (!use_x0 && launch_fill<T>(x, nd, down<T>(0), s) != hipSuccess)
                return ehp_solver_fail(EHP_ERR_HIP, "fill");
            hipLaunchKernelGGL((copy_kernel<T>), dim3(blocks(nd)), dim3(kThreads), 0, s, x, xprev, nd);
            hipLaunchKernelGGL((copy_kernel<T>), dim3(blocks(nd)), dim3(kThreads), 0, s, x, y, nd);
            hipLaunchKernelGGL(set_scalar_kernel, dim3(seg_blocks(nseg)), dim3(64), 0, s, sc, nseg, (int) S_T, 1.0);
            rc = A.adj(b, atb, s); // A^T b
            // grad = A^T (A y) - A^T b (APGD.cpp:55-60)
            auto gradient = [&]() -> int {
                int r2 = A.fwd(y, ay, s);
                if (r2 == EHP_OK)
                    r2 = A.adj(ay, grad, s);
                if (r2 == EHP_OK)
                    hipLaunchKernelGGL((axpy_kernel<T>), dim3(blocks(nd)), dim3(kThreads), 0, s, (typename Acc<T>::type) -1, atb,
                                       grad, nd);
                return r2;
            };
            if (rc == EHP_OK)
                rc = gradient();
            if (rc != EHP_OK)
                return rc;
            const T tsig = sigma >= 0.0 ? (T) sigma : T(-1);
            auto body = [&]() -> int {
                hipLaunchKernelGGL((apgd_prox_kernel<T>), dim3((unsigned) D.neb), dim3(kThreads), 0, s, y, grad, tsig,
                                   x, D, sc);
                hipLaunchKernelGGL((apgd_t_kernel<T>), dim3(seg_blocks(nseg)), dim3(64), 0, s, sc, nseg);
                hipLaunchKernelGGL((apgd_momentum_kernel<T>), dim3((unsigned) D.neb), dim3(kThreads), 0, s, x, xprev, y,
                                   D, sc);
                if (dot<T>(grad, grad, D, nseg, partial, sc, S_SN, hist, nh, s) != hipSuccess)
                    return EHP_ERR_HIP;
                hipLaunchKernelGGL(apgd_check_kernel, dim3(seg_blocks(nseg)), dim3(64), 0, s, sc, nseg, epsilon);
                const int r2 = gradient();
                if (r2 != EHP_OK)
                    return r2;
                return hipGetLastError() == hipSuccess ? EHP_OK : EHP_ERR_HIP;
            };
            hipError_t e = iterate(s, iterations, graph, body);
            if (e != hipSuccess)
                return ehp_solver_fail(EHP_ERR_HIP, hipGetErrorString(e));
            std::vector<double> hsc, hh;
            rc = fetch(s, sc, nseg, hist, nh, hsc, hh);
            if (rc != EHP_OK)
                return rc;
            for (int j = 0; j < nseg; ++j) {
                if (iters_done)
                    iters_done[j] = (int) hsc[(size_t) j * S_COUNT + S_ITERS];
                if (grad_norms)
                    for (int i = 0; i < nh; ++i)
                        grad_norms[(size_t) j * nh + i] = hh[(size_t) j * nh + i];
            }
            return EHP_OK;
        }
        // ---- ADMML2 (ADMML2.cpp:46-97) with the identity or forward finite differences as
        // regularisation operator A: the x-update is CGNE (CGNE.cpp) on K = [op; lambda A],
        // K^T w = op^T b + lambda A^T (z - u)

        // rhs = atb + lambda (z - u), as BlockLinearOperator's adjoint sums the blocks
        template <class T>
        __global__ void __launch_bounds__(kThreads)
            admm_rhs_kernel(const T* __restrict__ atb, const T* __restrict__ z, const T* __restrict__ u, T lambda,
                            T* __restrict__ rhs, int64_t n)
        {
            const int64_t i = (int64_t) blockIdx.x * kThreads + threadIdx.x;
            if (i < n) {
                const T zu = z[i] - u[i];
                rhs[i] = atb[i] + lambda * zu;
            }
        }

        // K^T K v = op^T op v + lambda (lambda v): ktk holds op^T op v on entry
        template <class T>
        __global__ void __launch_bounds__(kThreads)
            admm_normal_kernel(T* __restrict__ ktk, const T* __restrict__ v, T lambda, int64_t n)
        {
            const int64_t i = (int64_t) blockIdx.x * kThreads + threadIdx.x;
            if (i < n) {
                const T lv = lambda * v[i];
                ktk[i] = ktk[i] + lambda * lv;
            }
        }

        // CGNE start: r = rhs - K^T K x (elsa: r = A x; r *= -1; r += b), c = r
        template <class T>
        __global__ void __launch_bounds__(kThreads)
            cgne_start_kernel(const T* __restrict__ rhs, const T* __restrict__ ktkx, T* __restrict__ r,
                              T* __restrict__ c, int64_t n)
        {
            const int64_t i = (int64_t) blockIdx.x * kThreads + threadIdx.x;
            if (i < n) {
                const T ri = -ktkx[i] + rhs[i];
                r[i] = ri;
                c[i] = ri;
            }
        }

        // z = prox(x + u) with threshold tau * sigma (ProximalL1.cpp:32-43; sigma < 0: identity),
        // then u += x; u -= z
        template <class T>
        __global__ void __launch_bounds__(kThreads)
            admm_zu_kernel(const T* __restrict__ x, T* __restrict__ z, T* __restrict__ u, T tau, T sigma, int64_t n)
        {
            const int64_t i = (int64_t) blockIdx.x * kThreads + threadIdx.x;
            if (i >= n)
                return;
            const T v = x[i] + u[i];
            T zi;
            if (sigma < T(0)) {
                zi = v;
            } else {
                T tmp = (v < T(0) ? -v : v) - tau * sigma;
                tmp = T(0.5) * (tmp + (tmp < T(0) ? -tmp : tmp));
                const T sg = v > T(0) ? T(1) : (v < T(0) ? T(-1) : T(0));
                zi = sg * tmp;
            }
            z[i] = zi;
            const T ux = u[i] + x[i];
            u[i] = ux - zi;
        }

        // out = a - b
        template <class T>
        __global__ void __launch_bounds__(kThreads)
            admm_sub_kernel(const T* __restrict__ a, const T* __restrict__ b, T* __restrict__ out, int64_t n)
        {
            const int64_t i = (int64_t) blockIdx.x * kThreads + threadIdx.x;
            if (i < n)
                out[i] = a[i] - b[i];
        }

        // v = lambda v (elsa's scalar-times-operator: the scalar applied to the operator's result)
        template <class T>
        __global__ void __launch_bounds__(kThreads) admm_scale_kernel(T* __restrict__ v, T lambda, int64_t n)
        {
            const int64_t i = (int64_t) blockIdx.x * kThreads + threadIdx.x;
            if (i < n)
                v[i] = lambda * v[i];
        }

        // out = a + lambda v (a block operator's adjoint: the blocks' adjoints summed in order)
        template <class T>
        __global__ void __launch_bounds__(kThreads)
            admm_axpy_kernel(const T* __restrict__ a, T lambda, const T* __restrict__ v, T* __restrict__ out,
                             int64_t n)
        {
            const int64_t i = (int64_t) blockIdx.x * kThreads + threadIdx.x;
            if (i < n) {
                const T lv = lambda * v[i];
                out[i] = a[i] + lv;
            }
        }

        // r = b - q formed in FP64 and rounded to bfloat16, and partial sums of its square (the
        // residual norm of the refinement, FP64, before rounding)
        __global__ void __launch_bounds__(kThreads)
            refine_residual_kernel(const float* __restrict__ b, const float* __restrict__ q, bf16* __restrict__ r,
                                   const SegSpace S, double* __restrict__ partial)
        {
            const int seg = S.rseg[blockIdx.x];
            const int64_t lb = (int64_t) blockIdx.x - S.rb0[seg];
            const int64_t nb = S.rb0[seg + 1] - S.rb0[seg];
            const int64_t end = S.off[seg + 1];
            double acc = 0.0;
            for (int64_t i = S.off[seg] + lb * kThreads + threadIdx.x; i < end; i += nb * kThreads) {
                const double ri = (double) b[i] - (double) q[i];
                r[i] = float_to_bf16((float) ri);
                acc += ri * ri;
            }
            __shared__ double red[kThreads / 64];
            for (int off = 32; off > 0; off >>= 1)
                acc += __shfl_down(acc, off, 64);
            if ((threadIdx.x & 63) == 0)
                red[threadIdx.x >> 6] = acc;
            __syncthreads();
            if (threadIdx.x == 0) {
                double sum = 0.0;
                for (int w = 0; w < kThreads / 64; ++w)
                    sum += red[w];
                partial[blockIdx.x] = sum;
            }
        }

        // x += d (FP32 += bfloat16)
        __global__ void __launch_bounds__(kThreads)
            refine_update_kernel(float* __restrict__ x, const bf16* __restrict__ d, int64_t n)
        {
            const int64_t i = (int64_t) blockIdx.x * kThreads + threadIdx.x;
            if (i < n)
                x[i] += bf16_to_float(d[i]);
        }
    } // namespace

    int solve_cgls_refined(const SolveOps& Ahi, const SolveOps& Alo, const float* b, float* x, int use_x0, int outer,
                           int inner, bool graph, double* residuals, hipStream_t caller)
    {
        Solve sv;
        int rc = sv.init(Ahi, caller, "hipMalloc of the refinement workspace failed");
        if (rc != EHP_OK)
            return rc;
        hipStream_t s = sv.ss.s;
        float* q;
        bf16 *r, *d;
        if (sv.m.alloc(&q, sv.nr * sizeof(float)) != hipSuccess || sv.m.alloc(&r, sv.nr * sizeof(bf16)) != hipSuccess ||
            sv.m.alloc(&d, sv.nd * sizeof(bf16)) != hipSuccess)
            return ehp_solver_fail(EHP_ERR_ALLOC, "hipMalloc of the refinement workspace failed");
        if (!use_x0 && launch_fill<float>(x, sv.nd, 0.0f, s) != hipSuccess)
            return ehp_solver_fail(EHP_ERR_HIP, "fill");
        for (int k = 0; k <= outer; ++k) {
            // r = b - A_hi x in FP64, ||r||^2 in FP64
            rc = Ahi.fwd(x, q, s);
            if (rc != EHP_OK)
                return rc;
            hipLaunchKernelGGL(refine_residual_kernel, dim3((unsigned) sv.R.nrb), dim3(kThreads), 0, s, b, q, r, sv.R,
                               sv.partial);
            hipLaunchKernelGGL(dot_final_kernel, dim3((unsigned) sv.nseg), dim3(kThreads), 0, s, sv.partial, sv.R,
                               sv.sc, (int) S_SN, (double*) nullptr, 0);
            double rr = 0.0;
            hipError_t e = hipMemcpyAsync(&rr, sv.sc + S_SN, sizeof(double), hipMemcpyDeviceToHost, s);
            if (e == hipSuccess)
                e = hipStreamSynchronize(s);
            if (e != hipSuccess)
                return ehp_solver_fail(EHP_ERR_HIP, hipGetErrorString(e));
            if (residuals)
                residuals[k] = std::sqrt(rr);
            if (k == outer)
                break;
            // d ~ argmin ||A_lo d - r|| by CGLS on bfloat16 storage, then x += d
            rc = solve_cgls(Alo, EHP_BF16, r, d, 0, inner, 0.0, 0.0, graph, nullptr, nullptr, s);
            if (rc != EHP_OK)
                return rc;
            hipLaunchKernelGGL(refine_update_kernel, dim3(blocks(sv.nd)), dim3(kThreads), 0, s, x, d, sv.nd);
            e = hipGetLastError();
            if (e != hipSuccess)
                return ehp_solver_fail(EHP_ERR_HIP, hipGetErrorString(e));
        }
        return EHP_OK;
    }

    namespace
    {
        // frees the finite-difference plan when the solve (declared after it) has drained
        struct FdGuard {
            FdPlan plan;
            ~FdGuard()
            {
                fd_plan_destroy(plan);
            }
        };

        template <class T>
        int admml2(const SolveOps& A, const T* b, T* x, int use_x0, int iterations, double tau, double sigma,
                   int ninner, bool fd, bool graph, double* residuals, double* tau_used, hipStream_t caller)
        {
            FdGuard reg;
            const int nseg0 = A.nseg;
            if (fd) {
                if ((int) A.shapes.size() != 4 * nseg0)
                    return ehp_solver_fail(EHP_ERR_ARGUMENT, "finite differences need the volume shapes");
                const int rc0 = fd_plan_create(A.shapes.data(), A.doff, nseg0, reg.plan);
                if (rc0 != EHP_OK)
                    return ehp_solver_fail(rc0, "finite-difference plan");
            }
            // elsa: tau = 0.9 * (1 / powerIterations(A^T A)) (ADMML2.cpp:30-43): 0.9 for the
            // identity; for finite differences from 5 power iterations on the device
            T ttau;
            if (tau > 0.0) {
                ttau = (T) tau;
            } else if (!fd) {
                ttau = (T) 0.9;
            } else {
                if (nseg0 > 1)
                    return ehp_solver_fail(EHP_ERR_ARGUMENT,
                                           "ADMML2 with finite differences on several problems needs tau > 0");
                SolveOps D;
                D.nseg = nseg0;
                D.doff = A.doff;
                D.roff = reg.plan.reg_off.data();
                const FdPlan* pl = &reg.plan;
                D.fwd = [pl](const void* in, void* out, hipStream_t s2) {
                    return fd_apply<T>(*pl, (const T*) in, (T*) out, false, s2) == hipSuccess ? EHP_OK : EHP_ERR_HIP;
                };
                D.adj = [pl](const void* in, void* out, hipStream_t s2) {
                    return fd_apply<T>(*pl, (const T*) in, (T*) out, true, s2) == hipSuccess ? EHP_OK : EHP_ERR_HIP;
                };
                double L = 0.0;
                const int rc0 =
                    solve_power_iterations(D, sizeof(T) == 8 ? EHP_F64 : EHP_F32, 5, &L, caller);
                if (rc0 != EHP_OK)
                    return rc0;
                ttau = (T) (0.9 * (double) (T(1) / (T) L));
            }
            if (tau_used)
                *tau_used = (double) ttau;
            Solve sv;
            int rc = sv.init(A, caller, "hipMalloc of the ADMML2 workspace failed");
            if (rc != EHP_OK)
                return rc;
            hipStream_t s = sv.ss.s;
            const int nseg = sv.nseg;
            const int64_t nd = sv.nd, nr = sv.nr;
            const int64_t nz = fd ? reg.plan.nreg : nd; // the regulariser's range: z, u
            const SegSpace &D = sv.D, &R = sv.R;
            Scratch& m = sv.m;
            T *z, *u, *atb, *rhs, *r, *c, *ac, *q, *t = nullptr, *w = nullptr;
            const int nh = iterations + 1;
            if (m.alloc(&z, nz * sizeof(T)) != hipSuccess || m.alloc(&u, nz * sizeof(T)) != hipSuccess ||
                m.alloc(&atb, nd * sizeof(T)) != hipSuccess || m.alloc(&rhs, nd * sizeof(T)) != hipSuccess ||
                m.alloc(&r, nd * sizeof(T)) != hipSuccess || m.alloc(&c, nd * sizeof(T)) != hipSuccess ||
                m.alloc(&ac, nd * sizeof(T)) != hipSuccess || m.alloc(&q, nr * sizeof(T)) != hipSuccess ||
                (fd && (m.alloc(&t, nz * sizeof(T)) != hipSuccess || m.alloc(&w, nd * sizeof(T)) != hipSuccess)))
                return ehp_solver_fail(EHP_ERR_ALLOC, "hipMalloc of the ADMML2 workspace failed");
            // elsa: lambda = 1 / sqrt(tau) ("sqrttau", ADMML2.cpp:70), in the data type
            const T lambda = T(1) / std::sqrt(ttau), tsig = sigma >= 0.0 ? (T) sigma : T(-1);
            bool okf = true;
            if (!use_x0)
                okf = launch_fill<T>(x, nd, T(0), s) == hipSuccess;
            okf = okf && launch_fill<T>(z, nz, T(0), s) == hipSuccess && launch_fill<T>(u, nz, T(0), s) == hipSuccess;
            if (!okf)
                return ehp_solver_fail(EHP_ERR_HIP, "fill");
            rc = A.adj(b, atb, s); // op^T b, the same every iteration
            if (rc != EHP_OK)
                return rc;
            auto regop = [&](const T* in, T* out, bool adjoint) -> int {
                return fd_apply<T>(reg.plan, in, out, adjoint, s) == hipSuccess ? EHP_OK : EHP_ERR_HIP;
            };
            // K^T K v = op^T op v + lambda A^T (lambda A v) into out
            auto normal = [&](const T* v, T* out) -> int {
                int r2 = A.fwd(v, q, s);
                if (r2 == EHP_OK)
                    r2 = A.adj(q, out, s);
                if (r2 != EHP_OK)
                    return r2;
                if (!fd) {
                    hipLaunchKernelGGL((admm_normal_kernel<T>), dim3(blocks(nd)), dim3(kThreads), 0, s, out, v, lambda,
                                       nd);
                    return EHP_OK;
                }
                r2 = regop(v, t, false);
                if (r2 != EHP_OK)
                    return r2;
                hipLaunchKernelGGL((admm_scale_kernel<T>), dim3(blocks(nz)), dim3(kThreads), 0, s, t, lambda, nz);
                r2 = regop(t, w, true);
                if (r2 == EHP_OK)
                    hipLaunchKernelGGL((admm_axpy_kernel<T>), dim3(blocks(nd)), dim3(kThreads), 0, s, out, lambda, w,
                                       out, nd);
                return r2;
            };
            auto body = [&]() -> int {
                // x = CGNE(K, [b; z - u], ninner, x) (RegularizedInversion.cpp, CGNE.cpp)
                int r2 = EHP_OK;
                if (!fd) {
                    hipLaunchKernelGGL((admm_rhs_kernel<T>), dim3(blocks(nd)), dim3(kThreads), 0, s, atb, z, u, lambda,
                                       rhs, nd);
                } else {
                    hipLaunchKernelGGL((admm_sub_kernel<T>), dim3(blocks(nz)), dim3(kThreads), 0, s, z, u, t, nz);
                    r2 = regop(t, w, true);
                    if (r2 != EHP_OK)
                        return r2;
                    hipLaunchKernelGGL((admm_axpy_kernel<T>), dim3(blocks(nd)), dim3(kThreads), 0, s, atb, lambda, w,
                                       rhs, nd);
                }
                r2 = normal(x, ac);
                if (r2 != EHP_OK)
                    return r2;
                hipLaunchKernelGGL((cgne_start_kernel<T>), dim3(blocks(nd)), dim3(kThreads), 0, s, rhs, ac, r, c, nd);
                if (dot<T>(r, r, D, nseg, sv.partial, sv.sc, S_KOLD, nullptr, nh, s) != hipSuccess)
                    return EHP_ERR_HIP;
                for (int it = 0; it < ninner; ++it) {
                    r2 = normal(c, ac);
                    if (r2 != EHP_OK)
                        return r2;
                    if (dot<T>(c, ac, D, nseg, sv.partial, sv.sc, S_DELTA, nullptr, nh, s) != hipSuccess)
                        return EHP_ERR_HIP;
                    hipLaunchKernelGGL(cgls_alpha_kernel, dim3(seg_blocks(nseg)), dim3(64), 0, s, sv.sc, nseg, 0.0);
                    hipLaunchKernelGGL((cgls_update_kernel<T>), dim3((unsigned) D.neb), dim3(kThreads), 0, s, x, c, D,
                                       sv.sc, 0);
                    hipLaunchKernelGGL((cgls_update_kernel<T>), dim3((unsigned) D.neb), dim3(kThreads), 0, s, r, ac, D,
                                       sv.sc, 1);
                    if (dot<T>(r, r, D, nseg, sv.partial, sv.sc, S_K, nullptr, nh, s) != hipSuccess)
                        return EHP_ERR_HIP;
                    hipLaunchKernelGGL(cgls_beta_kernel, dim3(seg_blocks(nseg)), dim3(64), 0, s, sv.sc, nseg);
                    hipLaunchKernelGGL((cgls_direction_kernel<T>), dim3((unsigned) D.neb), dim3(kThreads), 0, s, r, c, D,
                                       sv.sc);
                }
                // z = prox(A x + u, tau); u += A x; u -= z (ADMML2.cpp:74-85)
                const T* ax = x;
                if (fd) {
                    r2 = regop(x, t, false);
                    if (r2 != EHP_OK)
                        return r2;
                    ax = t;
                }
                hipLaunchKernelGGL((admm_zu_kernel<T>), dim3(blocks(nz)), dim3(kThreads), 0, s, ax, z, u, ttau, tsig, nz);
                return hipGetLastError() == hipSuccess ? EHP_OK : EHP_ERR_HIP;
            };
            if (!residuals) {
                hipError_t e = iterate(s, iterations, graph, body);
                if (e == hipSuccess)
                    e = hipStreamSynchronize(s);
                return e == hipSuccess ? EHP_OK : ehp_solver_fail(EHP_ERR_HIP, hipGetErrorString(e));
            }
            // with a residual history the iterations run eagerly: ||b - op x|| (FP64) at the
            // start and after each iteration
            std::vector<double> hsc, hh;
            for (int k = 0; k <= iterations; ++k) {
                if (k > 0) {
                    rc = body();
                    if (rc != EHP_OK)
                        return rc == EHP_ERR_HIP ? ehp_solver_fail(EHP_ERR_HIP, "ADMML2 iteration") : rc;
                }
                rc = A.fwd(x, q, s);
                if (rc != EHP_OK)
                    return rc;
                hipLaunchKernelGGL((rsub_kernel<T>), dim3(blocks(nr)), dim3(kThreads), 0, s, b, q, nr);
                if (dot<T>(q, q, R, nseg, sv.partial, sv.sc, S_SN, nullptr, nh, s) != hipSuccess)
                    return ehp_solver_fail(EHP_ERR_HIP, "dot");
                rc = fetch(s, sv.sc, nseg, nullptr, 0, hsc, hh);
                if (rc != EHP_OK)
                    return rc;
                for (int j = 0; j < nseg; ++j)
                    residuals[(size_t) j * nh + k] = std::sqrt(hsc[(size_t) j * S_COUNT + S_SN]);
            }
            return EHP_OK;
        }
    } // namespace

    int solve_admml2(const SolveOps& A, int dtype, const void* b, void* x, int use_x0, int iterations, double tau,
                     double sigma, int ninner, bool fd, bool graph, double* residuals, double* tau_used, hipStream_t s)
    {
        if (dtype == EHP_F64)
            return admml2<double>(A, (const double*) b, (double*) x, use_x0, iterations, tau, sigma, ninner, fd,
                                  graph, residuals, tau_used, s);
        if (dtype == EHP_F32)
            return admml2<float>(A, (const float*) b, (float*) x, use_x0, iterations, tau, sigma, ninner, fd, graph,
                                 residuals, tau_used, s);
        return ehp_solver_fail(EHP_ERR_UNSUPPORTED, "ADMML2 needs an FP32 or FP64 operator");
    }

    int solve_cgls(const SolveOps& A, int dtype, const void* b, void* x, int use_x0, int iterations, double eps,
                   double tol, bool graph, double* residuals, int* iters_done, hipStream_t s)
    {
        if (dtype == EHP_BF16)
            return cgls<bf16>(A, (const bf16*) b, (bf16*) x, use_x0, iterations, eps, tol, graph, residuals, iters_done,
                              s);
        if (dtype == EHP_F64)
            return cgls<double>(A, (const double*) b, (double*) x, use_x0, iterations, eps, tol, graph, residuals,
                                iters_done, s);
        return cgls<float>(A, (const float*) b, (float*) x, use_x0, iterations, eps, tol, graph, residuals, iters_done,
                           s);
    }

    int solve_apgd(const SolveOps& A, int dtype, const void* b, void* x, int use_x0, int iterations, const double* mu,
                   double sigma, double epsilon, bool graph, double* grad_norms, int* iters_done, double* mu_used,
                   hipStream_t s)
    {
        if (dtype == EHP_BF16)
            return ehp_solver_fail(EHP_ERR_UNSUPPORTED, "APGD on bfloat16 storage is not implemented");
        if (dtype == EHP_F64)
            return apgd<double>(A, (const double*) b, (double*) x, use_x0, iterations, mu, sigma, epsilon, graph,
                                grad_norms, iters_done, mu_used, s);
        return apgd<float>(A, (const float*) b, (float*) x, use_x0, iterations, mu, sigma, epsilon, graph, grad_norms,
                           iters_done, mu_used, s);
    }

    int solve_power_iterations(const SolveOps& A, int dtype, int niters, double* L, hipStream_t caller)
    {
        Solve sv;
        const int rc = sv.init(A, caller, "hipMalloc of the power-iteration workspace failed");
        if (rc != EHP_OK)
            return rc;
        if (dtype == EHP_F64)
            return power_iterations<double>(A, sv, niters, L);
        if (dtype == EHP_BF16)
            return power_iterations<bf16>(A, sv, niters, L);
        return power_iterations<float>(A, sv, niters, L);
    }
} // namespace ehp
