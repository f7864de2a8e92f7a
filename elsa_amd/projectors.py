"""GPU X-ray projectors with elsa's class names and LinearOperator surface.

  JosephsMethodCUDA(domainDescriptor, rangeDescriptor, fast=True)
      elsa/projectors_cuda/JosephsMethodCUDA.h:78-79, pyelsa bind_projectors_cuda.cpp:30-44
  SiddonsMethodCUDA(domainDescriptor, rangeDescriptor)
      elsa/projectors_cuda/SiddonsMethodCUDA.h:71-72
  BlobVoxelProjectorCUDA(domainDescriptor, rangeDescriptor, radius=2, alpha=10.83, order=2)
      elsa/projectors_cuda/VoxelProjectorCUDA.h:39-43
  BSplineVoxelProjectorCUDA(domainDescriptor, rangeDescriptor, order=2)
      elsa/projectors_cuda/VoxelProjectorCUDA.h:82-84
  PhaseContrastBlobVoxelProjectorCUDA(domainDescriptor, rangeDescriptor, radius=2, alpha=10.83,
                                      order=2)
      elsa/projectors_cuda/VoxelProjectorCUDA.h:109-150
  PhaseContrastBSplineVoxelProjectorCUDA(domainDescriptor, rangeDescriptor, order=2)
      elsa/projectors_cuda/VoxelProjectorCUDA.h:152-191

apply(x[, out]) / applyAdjoint(y[, out]) follow elsa::LinearOperator
(elsa/core/LinearOperator.h:75-111): the result is overwritten. Data are torch tensors on the
GPU (the projection runs on torch's current stream) or numpy arrays (copied to the GPU and
back). Volume arrays have shape ([z,] y, x) and sinograms (poses, [v,] u): C order of elsa's
x-fastest layout. Every projection runs a HIP kernel of libelsahip.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .descriptors import DetectorDescriptor, VolumeDescriptor

# bfloat16 storage (FP32 arithmetic; Joseph and the voxel projectors): dtype="bfloat16" or torch.bfloat16. Data are
# torch.bfloat16 GPU tensors, or numpy arrays rounded to bfloat16 on the way in and returned as
# float32.
BF16 = "bfloat16"
_DT = {np.dtype(np.float32): _lib.EHP_F32, np.dtype(np.float64): _lib.EHP_F64, BF16: _lib.EHP_BF16}


def _dtype_key(dtype):
    if isinstance(dtype, str) and dtype.lower() in ("bfloat16", "bf16"):
        return BF16
    if str(dtype) == "torch.bfloat16":
        return BF16
    return np.dtype(dtype)


def _torch_dtype(key):
    torch = _torch()
    return {BF16: torch.bfloat16, np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64}[key]


def _torch():
    import torch

    return torch


class LinearOperator:
    """apply / applyAdjoint over a domain and a range descriptor."""

    def __init__(self, domainDescriptor, rangeDescriptor):
        self._domain = domainDescriptor
        self._range = rangeDescriptor

    def getDomainDescriptor(self):
        return self._domain

    def getRangeDescriptor(self):
        return self._range

    def apply(self, x, out=None):
        return self._run(x, out, adjoint=False)

    def applyAdjoint(self, y, out=None):
        return self._run(y, out, adjoint=True)

    def _run(self, x, out, adjoint):
        raise NotImplementedError


class _HipProjector(LinearOperator):
    """Holds one ehp_op handle; the device geometry lives as long as the object."""

    _dtype = np.dtype(np.float32)

    def __init__(self, domainDescriptor: VolumeDescriptor, rangeDescriptor: DetectorDescriptor, dtype=np.float32):
        super().__init__(domainDescriptor, rangeDescriptor)
        dim = domainDescriptor.getNumberOfDimensions()
        if dim not in (2, 3):
            raise ValueError(f"{type(self).__name__}: only supporting 2d/3d operations")
        if dim != rangeDescriptor.getNumberOfDimensions():
            raise ValueError(f"{type(self).__name__}: domain and range dimension need to match")
        if rangeDescriptor.getNumberOfGeometryPoses() == 0:
            raise ValueError(f"{type(self).__name__}: geometry list was empty")
        self._dtype = _dtype_key(dtype)
        if self._dtype not in _DT:
            raise ValueError(f"unsupported dtype {dtype}")
        self._op = ctypes.c_void_p()
        self._L = _lib.load()
        P, Pinv, C, ext, sdd = rangeDescriptor.pose_arrays()
        vol = domainDescriptor.getNumberOfCoefficientsPerDimension()
        det = rangeDescriptor.getNumberOfCoefficientsPerDimension()[: dim - 1].copy()
        geo = ctypes.c_void_p()
        fp = ctypes.POINTER(ctypes.c_float)
        i64p = ctypes.POINTER(ctypes.c_int64)
        _lib.check(self._L.ehp_geometry_create(dim, rangeDescriptor.getNumberOfGeometryPoses(), vol.ctypes.data_as(i64p),
                                               det.ctypes.data_as(i64p), P.ctypes.data_as(fp), Pinv.ctypes.data_as(fp),
                                               C.ctypes.data_as(fp), ext.ctypes.data_as(fp), sdd.ctypes.data_as(fp),
                                               ctypes.byref(geo)), type(self).__name__)
        try:
            coords = getattr(rangeDescriptor, "pixel_coords", None)
            if coords is not None:  # a curved detector
                _lib.check(self._L.ehp_geometry_set_pixel_coords(geo, coords().ctypes.data_as(fp)), type(self).__name__)
            self._create(geo)
        finally:
            self._L.ehp_geometry_destroy(geo)
        nv, ns = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(self._L.ehp_op_sizes(self._op, ctypes.byref(nv), ctypes.byref(ns)))
        self._n_vol, self._n_sino = nv.value, ns.value

    def _create(self, geo):
        raise NotImplementedError

    def __del__(self):
        op = getattr(self, "_op", None)
        if op is not None and op.value:
            self._L.ehp_destroy(op)
            self._op = None

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._op

    def _ctor_kwargs(self):
        return {}

    def clone(self):
        """A new operator on the same descriptors with the same settings (elsa:
        LinearOperator::clone / cloneImpl): its own device state, the same results bit for bit;
        the tile and poses per step of its tiled kernels are copied."""
        dt = BF16 if self._dtype == BF16 else self._dtype
        c = type(self)(self._domain, self._range, dtype=dt, **self._ctor_kwargs())
        for knob in ("tile", "pose_batch"):
            try:
                cur = getattr(self, knob)()
            except ValueError:
                continue  # no tiled kernel of that kind
            getattr(c, knob)(*cur) if knob == "tile" else c.pose_batch(cur)
        return c

    def tile(self, la: int = 0, tb: int = 0, tc: int = 0):
        """Set (la > 0) and return the tile of the LDS-tiled kernel (FP32): (LA, TB, TC) of the
        Joseph / Siddon back projection, (width, height, -) in pixels of the voxel forward."""
        o = [ctypes.c_int(), ctypes.c_int(), ctypes.c_int()]
        _lib.check(self._L.ehp_tile(self._op, la, tb, tc, *(ctypes.byref(x) for x in o)), "tile")
        return tuple(x.value for x in o)

    def pose_batch(self, n: int = 0) -> int:
        """Set (n > 0) and return the poses per step of the tiled back projection (FP32 Joseph
        and Siddon)."""
        o = ctypes.c_int()
        _lib.check(self._L.ehp_pose_batch(self._op, int(n), ctypes.byref(o)), "pose_batch")
        return o.value

    def sample_count(self) -> int:
        """Exact number of samples (Joseph), voxel crossings (Siddon) or footprint taps
        (voxel projectors) of one pass."""
        s = ctypes.c_uint64()
        _lib.check(self._L.ehp_sample_count(self._op, ctypes.byref(s)))
        return int(s.value)

    def _run(self, x, out, adjoint):
        torch = _torch()
        n_in, n_out = (self._n_sino, self._n_vol) if adjoint else (self._n_vol, self._n_sino)
        shape_out = self._domain.shape if adjoint else self._range.shape
        to_numpy = isinstance(x, np.ndarray)
        tdt = _torch_dtype(self._dtype)
        if to_numpy:
            if self._dtype == BF16:
                xt = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda().to(torch.bfloat16)
            else:
                xt = torch.from_numpy(np.ascontiguousarray(x, dtype=self._dtype)).cuda()
        else:
            if not x.is_cuda:
                raise ValueError("input tensor must live on the GPU")
            if x.dtype != tdt:
                raise ValueError(f"input dtype {x.dtype} does not match the operator's {tdt}")
            xt = x.contiguous()
        if xt.numel() != n_in:
            raise ValueError(f"input has {xt.numel()} elements, expected {n_in}")
        if out is None or to_numpy:
            ot = torch.empty(shape_out, dtype=tdt, device=xt.device)
        else:
            ot = out
            if not (ot.is_cuda and ot.dtype == tdt and ot.is_contiguous() and ot.numel() == n_out):
                raise ValueError("output tensor must be a contiguous GPU tensor of the range size and dtype")
            if ot.data_ptr() == xt.data_ptr():
                raise ValueError("input and output must not alias")
        stream = ctypes.c_void_p(torch.cuda.current_stream(xt.device).cuda_stream)
        f = self._L.ehp_adjoint if adjoint else self._L.ehp_forward
        _lib.check(f(self._op, ctypes.c_void_p(xt.data_ptr()), ctypes.c_void_p(ot.data_ptr()), stream),
                   "applyAdjoint" if adjoint else "apply")
        if to_numpy:
            return ot.float().cpu().numpy() if self._dtype == BF16 else ot.cpu().numpy()
        return ot


class JosephsMethodCUDA(_HipProjector):
    """Joseph's method with the discretization of elsa's CPU JosephsMethod
    (JosephsMethod.cpp:140-198) and its matched adjoint. ``fast`` is accepted for
    interface compatibility; the adjoint is always the matched one.

    ``kernels="fast"`` (FP32) runs the forward over plane-major volume views and the back
    projection over LDS-privatised tiles with 64-bit fixed-point sums; ``"simple"`` runs one
    thread per ray for both (the back projection with float atomics). FP64 with
    ``kernels="fast"`` runs the forward over FP64 plane-major views (elsa's weights and sums, bit
    for bit) and the LDS-tiled back projection, each FP64 contribution rounded to a multiple of a
    per-call unit and summed in 64-bit integers (bitwise reproducible); ``"simple"`` one thread
    per ray, with an FP64-atomic back projection."""

    def __init__(self, domainDescriptor, rangeDescriptor, fast: bool = True, dtype=np.float32, kernels: str = "fast"):
        self.fast = bool(fast)
        if kernels not in ("fast", "simple"):
            raise ValueError("kernels must be 'fast' or 'simple'")
        self.kernels = kernels
        super().__init__(domainDescriptor, rangeDescriptor, dtype)

    def _ctor_kwargs(self):
        return {"fast": self.fast, "kernels": self.kernels}

    def _create(self, geo):
        flags = _lib.EHP_SIMPLE_KERNELS if self.kernels == "simple" else 0
        _lib.check(self._L.ehp_joseph_create(geo, _DT[self._dtype], flags, ctypes.byref(self._op)),
                   "JosephsMethodCUDA")

    def ray_info(self):
        """Driving axis and sample count of every ray as the device computes them."""
        axis = np.empty(self._n_sino, np.int32)
        nsteps = np.empty(self._n_sino, np.int64)
        _lib.check(self._L.ehp_joseph_ray_info(self._op, axis.ctypes.data_as(ctypes.c_void_p),
                                               nsteps.ctypes.data_as(ctypes.c_void_p)))
        return axis, nsteps


class SiddonsMethodCUDA(_HipProjector):
    """Siddon's method (traversal of TraverseAABB.cpp:6-184, crossing parameters T0 + k tDelta).

    ``kernels="fast"`` runs the forward over plane-major volume views (each ray's crossings in
    traversal order, the one-thread-per-ray sums bit for bit) and the back projection over
    LDS-privatised tiles with 64-bit fixed-point sums, in FP32 and in FP64 (each FP64
    contribution rounded once to the fixed-point unit, bitwise reproducible); ``"simple"`` runs
    one thread per ray with float (FP64) atomics."""

    def __init__(self, domainDescriptor, rangeDescriptor, dtype=np.float32, kernels: str = "fast"):
        if kernels not in ("fast", "simple"):
            raise ValueError("kernels must be 'fast' or 'simple'")
        self.kernels = kernels
        super().__init__(domainDescriptor, rangeDescriptor, dtype)

    def _ctor_kwargs(self):
        return {"kernels": self.kernels}

    def _create(self, geo):
        flags = _lib.EHP_SIMPLE_KERNELS if self.kernels == "simple" else 0
        _lib.check(self._L.ehp_siddon_create(geo, _DT[self._dtype], flags, ctypes.byref(self._op)),
                   "SiddonsMethodCUDA")


def blob_lut(radius=2.0, alpha=10.83, order=2, n=100, dtype=np.float32):
    """Table of elsa's ProjectedBlob<data_t, n> (Blobs.h:249-317) and its support."""
    out = np.empty(n, dtype)
    sup = ctypes.c_double()
    _lib.check(_lib.load().ehp_blob_lut(_DT[np.dtype(dtype)], radius, alpha, order, n,
                                        out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(sup)), "blob_lut")
    return out, sup.value


def bspline_lut(dim, order=2, n=100, dtype=np.float32):
    """Table of elsa's ProjectedBSpline<data_t, n> (BSplines.h:132-208) and its support."""
    out = np.empty(n, dtype)
    sup = ctypes.c_double()
    _lib.check(_lib.load().ehp_bspline_lut(_DT[np.dtype(dtype)], dim, order, n,
                                           out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(sup)), "bspline_lut")
    return out, sup.value


def blob_derivative_lut(radius=2.0, alpha=10.83, order=2, n=100, dtype=np.float32, normalized=False):
    """ProjectedBlob::get_derivative_lut() or, normalized, get_normalized_gradient_lut()
    (Blobs.h:249-317) and the support."""
    out = np.empty(n, dtype)
    sup = ctypes.c_double()
    _lib.check(_lib.load().ehp_blob_derivative_lut(_DT[np.dtype(dtype)], radius, alpha, order, n, int(normalized),
                                                   out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(sup)),
               "blob_derivative_lut")
    return out, sup.value


def bspline_derivative_lut(dim, order=2, n=100, dtype=np.float32, normalized=False):
    """ProjectedBSpline::get_derivative_lut() or, normalized, get_normalized_gradient_lut()
    (BSplines.h:132-208) and the support."""
    out = np.empty(n, dtype)
    sup = ctypes.c_double()
    _lib.check(_lib.load().ehp_bspline_derivative_lut(_DT[np.dtype(dtype)], dim, order, n, int(normalized),
                                                      out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(sup)),
               "bspline_derivative_lut")
    return out, sup.value


def _table_dtype(dtype):
    """The basis-function table's dtype: the operator's, FP32 for bfloat16 storage."""
    key = _dtype_key(dtype)
    return np.float32 if key == BF16 else key


class _VoxelProjector(_HipProjector):
    """``kernels="fast"`` runs the forward projection over LDS-privatised detector tiles with
    64-bit fixed-point sums (FP64: with the simple kernels' weights, each term rounded once to the
    fixed-point unit); ``"simple"`` runs one thread per voxel and pose with float atomics. The
    back projection is a gather in both (FP64: the simple kernels' form). bfloat16 operators run
    the fast kernels only."""

    def _set_kernels(self, kernels):
        if kernels not in ("fast", "simple"):
            raise ValueError("kernels must be 'fast' or 'simple'")
        self.kernels = kernels

    _differential = False

    def _ctor_kwargs(self):
        kw = {"order": self.order, "kernels": self.kernels}
        if hasattr(self, "radius"):
            kw.update(radius=self.radius, alpha=self.alpha)
        return kw

    def _create(self, geo):
        flags = _lib.EHP_SIMPLE_KERNELS if self.kernels == "simple" else 0
        if self._differential:
            flags |= _lib.EHP_DIFFERENTIAL
        _lib.check(self._L.ehp_voxel_create(geo, _DT[self._dtype], self.lut.ctypes.data_as(ctypes.c_void_p),
                                            int(self.lut.shape[0]), float(self.support), flags,
                                            ctypes.byref(self._op)),
                   type(self).__name__)


class BlobVoxelProjectorCUDA(_VoxelProjector):
    """Voxel-driven projector with Kaiser-Bessel blobs (elsa BlobVoxelProjectorCUDA)."""

    def __init__(self, domainDescriptor, rangeDescriptor, radius=2.0, alpha=10.83, order=2, dtype=np.float32,
                 kernels: str = "fast"):
        self.radius, self.alpha, self.order = radius, alpha, int(order)
        self.lut, self.support = blob_lut(radius, alpha, order, 100, _table_dtype(dtype))
        self._set_kernels(kernels)
        super().__init__(domainDescriptor, rangeDescriptor, dtype)


class BSplineVoxelProjectorCUDA(_VoxelProjector):
    """Voxel-driven projector with B-splines (elsa BSplineVoxelProjectorCUDA)."""

    def __init__(self, domainDescriptor, rangeDescriptor, order=2, dtype=np.float32, kernels: str = "fast"):
        self.order = int(order)
        self.lut, self.support = bspline_lut(domainDescriptor.getNumberOfDimensions(), order, 100, _table_dtype(dtype))
        self._set_kernels(kernels)
        super().__init__(domainDescriptor, rangeDescriptor, dtype)


class PhaseContrastBlobVoxelProjectorCUDA(_VoxelProjector):
    """Differential phase-contrast projector with Kaiser-Bessel blobs (elsa
    PhaseContrastBlobVoxelProjectorCUDA, VoxelProjectorCUDA.cpp:75-114): the blob's derivative
    table weighted by sgn(d_u) in 2D, its normalized-gradient table weighted by d_u in 3D."""

    _differential = True

    def __init__(self, domainDescriptor, rangeDescriptor, radius=2.0, alpha=10.83, order=2, dtype=np.float32,
                 kernels: str = "fast"):
        self.radius, self.alpha, self.order = radius, alpha, int(order)
        dim = domainDescriptor.getNumberOfDimensions()
        self.lut, self.support = blob_derivative_lut(radius, alpha, order, 100, _table_dtype(dtype),
                                                     normalized=dim == 3)
        self._set_kernels(kernels)
        super().__init__(domainDescriptor, rangeDescriptor, dtype)


class PhaseContrastBSplineVoxelProjectorCUDA(_VoxelProjector):
    """Differential phase-contrast projector with B-splines (elsa
    PhaseContrastBSplineVoxelProjectorCUDA, VoxelProjectorCUDA.cpp:196-235)."""

    _differential = True

    def __init__(self, domainDescriptor, rangeDescriptor, order=2, dtype=np.float32, kernels: str = "fast"):
        self.order = int(order)
        dim = domainDescriptor.getNumberOfDimensions()
        self.lut, self.support = bspline_derivative_lut(dim, order, 100, _table_dtype(dtype), normalized=dim == 3)
        self._set_kernels(kernels)
        super().__init__(domainDescriptor, rangeDescriptor, dtype)


# pyelsa exposes float and double variants with f / d suffixes and un-suffixed float aliases
def _variant(cls, dtype, name):
    def init(self, *args, **kwargs):
        kwargs.setdefault("dtype", dtype)
        cls.__init__(self, *args, **kwargs)

    return type(name, (cls,), {"__init__": init, "__doc__": cls.__doc__})


JosephsMethodCUDAf = _variant(JosephsMethodCUDA, np.float32, "JosephsMethodCUDAf")
JosephsMethodCUDAd = _variant(JosephsMethodCUDA, np.float64, "JosephsMethodCUDAd")
SiddonsMethodCUDAf = _variant(SiddonsMethodCUDA, np.float32, "SiddonsMethodCUDAf")
SiddonsMethodCUDAd = _variant(SiddonsMethodCUDA, np.float64, "SiddonsMethodCUDAd")
BlobVoxelProjectorCUDAf = _variant(BlobVoxelProjectorCUDA, np.float32, "BlobVoxelProjectorCUDAf")
BlobVoxelProjectorCUDAd = _variant(BlobVoxelProjectorCUDA, np.float64, "BlobVoxelProjectorCUDAd")
BSplineVoxelProjectorCUDAf = _variant(BSplineVoxelProjectorCUDA, np.float32, "BSplineVoxelProjectorCUDAf")
BSplineVoxelProjectorCUDAd = _variant(BSplineVoxelProjectorCUDA, np.float64, "BSplineVoxelProjectorCUDAd")
PhaseContrastBlobVoxelProjectorCUDAf = _variant(PhaseContrastBlobVoxelProjectorCUDA, np.float32,
                                                "PhaseContrastBlobVoxelProjectorCUDAf")
PhaseContrastBlobVoxelProjectorCUDAd = _variant(PhaseContrastBlobVoxelProjectorCUDA, np.float64,
                                                "PhaseContrastBlobVoxelProjectorCUDAd")
PhaseContrastBSplineVoxelProjectorCUDAf = _variant(PhaseContrastBSplineVoxelProjectorCUDA, np.float32,
                                                   "PhaseContrastBSplineVoxelProjectorCUDAf")
PhaseContrastBSplineVoxelProjectorCUDAd = _variant(PhaseContrastBSplineVoxelProjectorCUDA, np.float64,
                                                   "PhaseContrastBSplineVoxelProjectorCUDAd")
