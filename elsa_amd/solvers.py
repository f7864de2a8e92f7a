"""elsa's CGLS and APGD (FISTA) solvers, run on the GPU by libelsahip.so.

  CGLS(A, b, eps=0, tol=1e-4).solve(iterations, x0=None)
      elsa/solvers/CGLS.h:45-46, CGLS.cpp:19-91
  APGD(A, b, prox, mu=None, epsilon=eps(data_t)).solve(iterations, x0=None)
      elsa/solvers/APGD.h:45-47, APGD.cpp:14-102 (FISTA), with prox ProximalL1(sigma)
      (ProximalL1.cpp:32-43) or ProximalIdentity()
  powerIterations(A, niters=5)
      elsa/solvers/PowerIterations.cpp:10-27 applied to A^T A

A is one of the projector classes of elsa_amd.projectors, or an elsa_amd.batch.ProjectorBatch:
then every problem of the batch is solved independently at once (per-problem step lengths and
convergence), b and x are packed vectors, mu may be a sequence of per-problem steps, and
`iterations_done`, `residuals`, `mu_used` (and powerIterations' result) get one row or entry
per problem. The whole iteration runs on the device: step lengths, momentum and norms never leave it (FP64 reductions), and with
graph=True one iteration is captured into a hipGraph and replayed. b is a GPU tensor or a
numpy array; solve() returns the same kind. After solve(), `iterations_done` and
`residuals` (CGLS: ||b - Ax|| at the start and per iteration; APGD: ||grad||^2 per
iteration) hold what the run recorded.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def _torch():
    import torch

    return torch


class ProximalL1:
    """Soft thresholding sign(v) * max(|v| - t * sigma, 0) (elsa: ProximalL1(sigma))."""

    def __init__(self, sigma=1.0):
        self.sigma = float(sigma)


class ProximalIdentity:
    """prox(v) = v (elsa: ProximalIdentity)."""

    sigma = -1.0


def _is_batch(A):
    from .batch import ProjectorBatch

    return isinstance(A, ProjectorBatch)


def _prepare(A, b, x0):
    from .projectors import BF16, _torch_dtype

    torch = _torch()
    tdt = _torch_dtype(A._dtype)
    to_numpy = isinstance(b, np.ndarray)
    if to_numpy:
        if A._dtype == BF16:
            bt = torch.from_numpy(np.ascontiguousarray(b, dtype=np.float32)).cuda().to(tdt)
        else:
            bt = torch.from_numpy(np.ascontiguousarray(b, dtype=A._dtype)).cuda()
    else:
        if not (b.is_cuda and b.dtype == tdt):
            raise ValueError("b must be a GPU tensor of the operator's dtype")
        bt = b.contiguous()
    if bt.numel() != A._n_sino:
        raise ValueError(f"b has {bt.numel()} elements, expected {A._n_sino}")
    shape = A.getDomainDescriptor().shape
    if x0 is None:
        xt = torch.empty(shape, dtype=tdt, device=bt.device)
    elif isinstance(x0, np.ndarray):
        xt = torch.from_numpy(np.array(x0, dtype=np.float32 if A._dtype == BF16 else A._dtype)
                              .reshape(shape)).cuda().to(tdt)
    else:
        xt = x0.detach().to(dtype=tdt).reshape(shape).contiguous().clone()
    if xt.numel() != A._n_vol:
        raise ValueError(f"x0 has {xt.numel()} elements, expected {A._n_vol}")
    stream = ctypes.c_void_p(torch.cuda.current_stream(bt.device).cuda_stream)
    return bt, xt, to_numpy, stream


class CGLS:
    """Conjugate gradients on the normal equations of min ||Ax - b||^2 + eps^2 ||x||^2."""

    def __init__(self, A, b, eps=0.0, tol=1e-4, graph=False):
        self.A, self.b, self.eps, self.tol, self.graph = A, b, float(eps), float(tol), bool(graph)
        self.iterations_done = 0
        self.residuals = None

    def solve(self, iterations, x0=None):
        bt, xt, to_numpy, stream = _prepare(self.A, self.b, x0)
        batch = _is_batch(self.A)
        nseg = len(self.A) if batch else 1
        res = np.empty((nseg, iterations + 1), np.float64)
        n = np.zeros(nseg, np.int32)
        flags = _lib.EHP_SOLVE_GRAPH if self.graph else 0
        fn = self.A._L.ehp_batch_cgls if batch else self.A._L.ehp_cgls
        _lib.check(fn(self.A.handle, ctypes.c_void_p(bt.data_ptr()), ctypes.c_void_p(xt.data_ptr()),
                      int(x0 is not None), int(iterations), self.eps, self.tol, flags,
                      res.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                      n.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), stream), "CGLS")
        if batch:
            self.iterations_done, self.residuals = n, res
        else:
            self.iterations_done, self.residuals = int(n[0]), res[0]
        return (xt.float() if xt.dtype == _torch().bfloat16 else xt).cpu().numpy() if to_numpy else xt


class APGD:
    """Accelerated proximal gradient (FISTA) on 0.5 ||Ax - b||^2 + g(x)."""

    def __init__(self, A, b, prox=None, mu=None, epsilon=None, graph=False):
        self.A, self.b, self.graph = A, b, bool(graph)
        self.prox = prox if prox is not None else ProximalIdentity()
        if _is_batch(A):
            mu = np.zeros(len(A)) if mu is None else np.broadcast_to(np.asarray(mu, np.float64), (len(A),))
            self.mu = np.ascontiguousarray(mu, np.float64)
        else:
            self.mu = float(mu) if mu is not None else 0.0
        self.epsilon = float(np.finfo(A._dtype if A._dtype != "bfloat16" else np.float32).eps) if epsilon is None \
            else float(epsilon)
        self.iterations_done = 0
        self.residuals = None
        self.mu_used = None

    def solve(self, iterations, x0=None):
        bt, xt, to_numpy, stream = _prepare(self.A, self.b, x0)
        batch = _is_batch(self.A)
        nseg = len(self.A) if batch else 1
        res = np.empty((nseg, iterations + 1), np.float64)
        n = np.zeros(nseg, np.int32)
        mu_used = np.zeros(nseg, np.float64)
        flags = _lib.EHP_SOLVE_GRAPH if self.graph else 0
        dp = ctypes.POINTER(ctypes.c_double)
        args = (ctypes.c_void_p(bt.data_ptr()), ctypes.c_void_p(xt.data_ptr()), int(x0 is not None), int(iterations))
        tail = (float(self.prox.sigma), self.epsilon, flags, res.ctypes.data_as(dp),
                n.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), mu_used.ctypes.data_as(dp), stream)
        if batch:
            _lib.check(self.A._L.ehp_batch_apgd(self.A.handle, *args, self.mu.ctypes.data_as(dp), *tail), "APGD")
            self.iterations_done, self.residuals, self.mu_used = n, res, mu_used
        else:
            _lib.check(self.A._L.ehp_apgd(self.A.handle, *args, self.mu, *tail), "APGD")
            self.iterations_done, self.residuals, self.mu_used = int(n[0]), res[0], float(mu_used[0])
        return (xt.float() if xt.dtype == _torch().bfloat16 else xt).cpu().numpy() if to_numpy else xt


class ADMML2:
    """ADMM with an L2 data term, 0.5 ||op x - b||^2 + g(A x), for elsa's ADMML2 (ADMML2.h:62-64,
    ADMML2.cpp:20-97) with regularisation operator A the identity (regularizer=None) or forward
    finite differences on every axis (regularizer="finite_differences", elsa's
    FiniteDifferences(domain); with ProximalL1 the anisotropic TV problem): ninneriters CGNE
    iterations per x-update, prox ProximalL1(sigma) or ProximalIdentity(), tau None for elsa's
    default 0.9 / powerIterations(A^T A) (0.9 with the identity; a batch with finite differences
    needs tau). After solve(), `tau_used` holds tau and `residuals` ||b - op x|| at the start and
    after each iteration when record=True (the iterations then run without graph capture)."""

    def __init__(self, op, b, prox=None, tau=None, ninneriters=5, graph=False, record=False, regularizer=None):
        self.A, self.b, self.graph, self.record = op, b, bool(graph), bool(record)
        self.prox = prox if prox is not None else ProximalIdentity()
        self.tau = float(tau) if tau is not None else 0.0
        self.ninner = int(ninneriters)
        if regularizer not in (None, "identity", "finite_differences"):
            raise ValueError(f"unknown regularizer {regularizer!r} (None, 'identity' or 'finite_differences')")
        self.fd = regularizer == "finite_differences"
        self.residuals = None
        self.tau_used = None

    def solve(self, iterations, x0=None):
        bt, xt, to_numpy, stream = _prepare(self.A, self.b, x0)
        batch = _is_batch(self.A)
        nseg = len(self.A) if batch else 1
        res = np.empty((nseg, iterations + 1), np.float64) if self.record else None
        tau = ctypes.c_double()
        flags = (_lib.EHP_SOLVE_GRAPH if self.graph else 0) | (_lib.EHP_ADMM_FINITE_DIFFERENCES if self.fd else 0)
        fn = self.A._L.ehp_batch_admml2 if batch else self.A._L.ehp_admml2
        dp = ctypes.POINTER(ctypes.c_double)
        _lib.check(fn(self.A.handle, ctypes.c_void_p(bt.data_ptr()), ctypes.c_void_p(xt.data_ptr()),
                      int(x0 is not None), int(iterations), self.tau, float(self.prox.sigma), self.ninner, flags,
                      res.ctypes.data_as(dp) if res is not None else None, ctypes.byref(tau), stream), "ADMML2")
        self.tau_used = tau.value
        if res is not None:
            self.residuals = res if batch else res[0]
        return xt.cpu().numpy() if to_numpy else xt


class CGLSRefined:
    """CGLS with iterative refinement in mixed precision (no counterpart in elsa): x (FP32)
    takes `outer` steps x += d, each d from `inner` CGLS iterations on A_lo, a bfloat16
    operator, applied to r = b - A_hi x, the residual of the FP32 operator A_hi of the same
    geometry formed in FP64. b and x are FP32. After solve(), `residuals` holds ||b - A_hi x||
    before the first step and after each."""

    def __init__(self, A_hi, A_lo, b, inner=10, graph=False):
        self.A_hi, self.A_lo, self.b, self.inner, self.graph = A_hi, A_lo, b, int(inner), bool(graph)
        self.residuals = None

    def solve(self, outer, x0=None):
        bt, xt, to_numpy, stream = _prepare(self.A_hi, self.b, x0)
        res = np.empty(outer + 1, np.float64)
        flags = _lib.EHP_SOLVE_GRAPH if self.graph else 0
        _lib.check(self.A_hi._L.ehp_cgls_refined(self.A_hi.handle, self.A_lo.handle, ctypes.c_void_p(bt.data_ptr()),
                                                 ctypes.c_void_p(xt.data_ptr()), int(x0 is not None), int(outer),
                                                 self.inner, flags, res.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                 stream), "CGLSRefined")
        self.residuals = res
        return xt.cpu().numpy() if to_numpy else xt


def powerIterations(A, niters=5):
    """Largest eigenvalue estimate of A^T A (elsa: powerIterations(adjoint(A) * A, niters));
    for a batch, an array with one estimate per problem."""
    torch = _torch()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if _is_batch(A):
        L = np.zeros(len(A), np.float64)
        _lib.check(A._L.ehp_batch_power_iterations(A.handle, int(niters), L.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                   stream), "powerIterations")
        return L
    L = ctypes.c_double()
    _lib.check(A._L.ehp_power_iterations(A.handle, int(niters), ctypes.byref(L), stream), "powerIterations")
    return L.value
