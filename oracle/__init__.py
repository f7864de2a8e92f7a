"""CPU oracle for the elsa X-ray projectors -- TEST INFRASTRUCTURE ONLY.

A ctypes front end over ``oracle/liboracle.so`` (built from ``oracle/elsa_oracle.cpp`` by
``oracle/Makefile``). It restates elsa's CPU projectors and geometry in FP32 (elsa's own
precision) and FP64. Only ``tests/``, the smoke check and the ``cpu_baseline`` leg of
``bench.py`` use it; the product package ``elsa_amd`` never imports it.

Every function names the reference routine it follows (paths relative to the elsa tree).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ELSA_ORACLE_LIB names another build of the same source (the AddressSanitizer /
# UndefinedBehaviorSanitizer build of `make -C oracle san`, run under LD_PRELOAD of the runtimes)
_LIB_PATH = os.environ.get("ELSA_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None

JOSEPH, SIDDON = 0, 1
MODE_ELSA, MODE_DIRECT = 0, 1
# Joseph only: MODE_DIRECT's sample points, weights formed in the data type (with FP64 data, the
# exact bilinear interpolation at the FP32 sample points)
MODE_EXACT_WEIGHTS = 2

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile liboracle.so with the committed Makefile (g++, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = ctypes.CDLL(_LIB_PATH)
    c_int, c_i64, c_f, c_d = ctypes.c_int, ctypes.c_int64, ctypes.c_float, ctypes.c_double
    L.orc_xray_f32.argtypes = [c_int, c_int, c_int, c_int, _i64p, _i64p, c_i64, _f32p, _f32p, _f32p, _f32p, c_int]
    L.orc_xray_f64.argtypes = [c_int, c_int, c_int, c_int, _i64p, _i64p, c_i64, _f64p, _f64p, _f64p, _f64p, c_int]
    L.orc_xray_f32_acc64.argtypes = [c_int, c_int, c_int, c_int, _i64p, _i64p, c_i64, _f32p, _f32p, _f64p, _f64p,
                                     c_int]
    L.orc_xray_px_f32.argtypes = [c_int, c_int, c_int, c_int, _i64p, _i64p, c_i64, _f32p, _f32p, _f32p, _f32p,
                                  _f32p, c_int]
    L.orc_xray_px_f64.argtypes = [c_int, c_int, c_int, c_int, _i64p, _i64p, c_i64, _f64p, _f64p, _f64p, _f64p,
                                  _f64p, c_int]
    L.orc_voxel_f32.argtypes = [c_int, c_int, _i64p, _i64p, c_i64, _f32p, _f32p, _f32p, _f32p, c_int, c_f,
                                c_int, _f32p, _f32p, c_int]
    L.orc_voxel_f64.argtypes = [c_int, c_int, _i64p, _i64p, c_i64, _f64p, _f64p, _f64p, _f64p, c_int, c_d,
                                c_int, _f64p, _f64p, c_int]
    L.orc_voxel_f32_acc64.argtypes = [c_int, c_int, _i64p, _i64p, c_i64, _f32p, _f32p, _f32p, _f64p, c_int, c_f,
                                      c_int, _f64p, _f64p, c_int]
    vp = ctypes.c_void_p
    L.orc_joseph_rays_f32.argtypes = [c_int, _i64p, _i64p, c_i64, _f32p, _f32p, vp, vp, vp, vp, vp]
    L.orc_joseph_rays_f64.argtypes = [c_int, _i64p, _i64p, c_i64, _f64p, _f64p, vp, vp, vp, vp, vp]
    L.orc_make_ray_f32.argtypes = [c_int, _f32p, _f32p, c_i64, c_i64, _f32p, _f32p]
    L.orc_intersect_f32.argtypes = [c_int, _f32p, _f32p, _i64p, ctypes.POINTER(c_f), ctypes.POINTER(c_f)]
    L.orc_intersect_f32.restype = c_int
    L.orc_geometry2d_f32.argtypes = [c_f, c_f, c_f, _f32p, _f32p, c_f, c_f, c_f, _f32p, _f32p, _f32p, _f32p,
                                     _f32p, ctypes.POINTER(c_f)]
    L.orc_geometry2d_f64.argtypes = [c_d, c_d, c_d, _f64p, _f64p, c_d, c_d, c_d, _f64p, _f64p, _f64p, _f64p,
                                     _f64p, ctypes.POINTER(c_d)]
    L.orc_geometry3d_f32.argtypes = [c_f, c_f, _f32p, _f32p, _f32p, _f32p, c_f, c_f, c_f, _f32p, _f32p, _f32p,
                                     _f32p, _f32p, _f32p, ctypes.POINTER(c_f)]
    L.orc_geometry3d_f64.argtypes = [c_d, c_d, _f64p, _f64p, _f64p, _f64p, c_d, c_d, c_d, _f64p, _f64p, _f64p,
                                     _f64p, _f64p, _f64p, ctypes.POINTER(c_d)]
    L.orc_bessi.argtypes = [c_i64, c_d]
    L.orc_bessi.restype = c_d
    L.orc_blob_evaluate_f64.argtypes = [c_d, c_d, c_d, c_i64]
    L.orc_blob_evaluate_f64.restype = c_d
    L.orc_blob_projected_f64.argtypes = [c_d, c_d, c_d, c_i64]
    L.orc_blob_projected_f64.restype = c_d
    L.orc_blob_lut_f32.argtypes = [c_f, c_f, c_i64, c_int, _f32p]
    L.orc_blob_lut_f64.argtypes = [c_d, c_d, c_i64, c_int, _f64p]
    L.orc_bspline_lut_f32.argtypes = [c_i64, c_i64, c_int, _f32p]
    L.orc_bspline_lut_f64.argtypes = [c_i64, c_i64, c_int, _f64p]
    L.orc_blob_deriv_lut_f32.argtypes = [c_f, c_f, c_i64, c_int, c_int, _f32p]
    L.orc_blob_deriv_lut_f64.argtypes = [c_d, c_d, c_i64, c_int, c_int, _f64p]
    L.orc_bspline_deriv_lut_f32.argtypes = [c_i64, c_i64, c_int, c_int, _f32p]
    L.orc_bspline_deriv_lut_f64.argtypes = [c_i64, c_i64, c_int, c_int, _f64p]
    L.orc_lut_eval_f64.argtypes = [_f64p, c_int, c_d, c_d]
    L.orc_lut_eval_f64.restype = c_d
    L.orc_lut_eval_f32.argtypes = [_f32p, c_int, c_f, c_f]
    L.orc_lut_eval_f32.restype = c_f
    _lib = L
    return L


# --------------------------------------------------------------------------------------
# Geometry (elsa/core/Geometry.cpp) and the circular trajectory
# (elsa/generators/CircleTrajectoryGenerator.cpp:9-24, BaseCircleTrajectoryGenerator.cpp:12-74,
# TrajectoryGenerator.cpp:9-47).


@dataclass
class Geom:
    """Per-view matrices of a trajectory, column-major blocks as Eigen stores them."""

    dim: int
    vol: tuple          # voxels per axis, x fastest
    det: tuple          # detector pixels (u[, v])
    P: np.ndarray       # nviews x dim*(dim+1)
    Pinv: np.ndarray    # nviews x (dim+1)*dim
    C: np.ndarray       # nviews x dim
    ext: np.ndarray     # nviews x dim*(dim+1)
    sdd: np.ndarray     # nviews

    @property
    def nviews(self) -> int:
        return int(self.sdd.shape[0])

    def astype(self, dt) -> "Geom":
        return Geom(self.dim, self.vol, self.det, *(np.ascontiguousarray(a, dtype=dt) for a in
                                                     (self.P, self.Pinv, self.C, self.ext, self.sdd)))

    @property
    def sino_shape(self):
        """numpy shape of the sinogram (slowest first): (views[, v], u)."""
        return (self.nviews,) + tuple(reversed(self.det))

    @property
    def vol_shape(self):
        """numpy shape of the volume (slowest first): ([z,] y, x)."""
        return tuple(reversed(self.vol))


def geometry_pose(dim, s2c, c2d, vol, det, angle_rad=0.0, beta=0.0, alpha=0.0, vol_spacing=None,
                  det_spacing=None, ppo=None, cor=None, dtype=np.float32):
    """One pose of elsa's Geometry (Geometry.cpp:12-109); origins at the data centres."""
    L = lib()
    dt = np.dtype(dtype)
    sfx = "f32" if dt == np.float32 else "f64"
    ct = ctypes.c_float if dt == np.float32 else ctypes.c_double
    vsp = np.ones(dim, dt) if vol_spacing is None else np.asarray(vol_spacing, dt)
    vorg = (dt.type(0.5) * (np.asarray(vol, dt) * vsp)).astype(dt)
    ssp = np.ones(dim - 1, dt) if det_spacing is None else np.asarray(det_spacing, dt)
    sorg = (dt.type(0.5) * (np.asarray(det, dt) * ssp)).astype(dt)
    cor = np.zeros(dim, dt) if cor is None else np.asarray(cor, dt)
    P = np.zeros(dim * (dim + 1), dt)
    Pinv = np.zeros(dim * (dim + 1), dt)
    C = np.zeros(dim, dt)
    ext = np.zeros(dim * (dim + 1), dt)
    sdd = ct()
    if dim == 2:
        ppo = 0.0 if ppo is None else float(np.asarray(ppo).reshape(-1)[0])
        getattr(L, f"orc_geometry2d_{sfx}")(s2c, c2d, angle_rad, vsp, vorg, ssp[0], sorg[0], ppo, cor, P, Pinv, C,
                                            ext, ctypes.byref(sdd))
    else:
        ppo = np.zeros(2, dt) if ppo is None else np.asarray(ppo, dt)
        getattr(L, f"orc_geometry3d_{sfx}")(s2c, c2d, vsp, vorg, ssp, sorg, angle_rad, beta, alpha, ppo, cor, P, Pinv,
                                            C, ext, ctypes.byref(sdd))
    return P, Pinv, C, ext, dt.type(sdd.value)


def linspaced_f32(n, high):
    """Eigen 3.4 LinSpaced(n, 0, high) in float: low + i*step, last element exactly high."""
    if n == 1:
        return np.array([high], np.float32)
    step = np.float32(np.float32(high) / np.float32(n - 1))
    out = np.arange(n, dtype=np.float32) * step
    out[-1] = np.float32(high)
    return out


def circle_geometry(dim, vol, nviews, arc_deg, s2c, c2d, det=None, dtype=np.float32, angles_deg=None):
    """Circular trajectory of elsa's CircleTrajectoryGenerator; detector defaults to floor(N*sqrt(2))."""
    vol = tuple(int(v) for v in vol)
    if det is None:
        det = tuple(int(np.float32(v) * np.float32(np.sqrt(np.float32(2.0)))) for v in vol[:dim - 1])
    det = tuple(int(d) for d in det)
    if angles_deg is None:
        thetas = linspaced_f32(nviews, arc_deg)
    else:
        thetas = np.asarray(angles_deg, np.float32)
    dt = np.dtype(dtype)
    Ps, Pis, Cs, Es, Ss = [], [], [], [], []
    pi_f = np.float32(np.pi)
    for th in thetas:
        if dt == np.float32:
            ang = np.float32(np.float32(th * pi_f) / np.float32(180))
        else:
            ang = float(th) * np.pi / 180.0
        P, Pi, C, E, sdd = geometry_pose(dim, s2c, c2d, vol, det, angle_rad=ang, dtype=dt)
        Ps.append(P); Pis.append(Pi); Cs.append(C); Es.append(E); Ss.append(sdd)
    return Geom(dim, vol, det, np.stack(Ps), np.stack(Pis), np.stack(Cs), np.stack(Es), np.array(Ss, dt))


def geom_from_poses(dim, vol, det, poses, dtype=np.float32):
    """Stack explicit poses (tuples returned by geometry_pose) into a Geom."""
    dt = np.dtype(dtype)
    cols = list(zip(*poses))
    return Geom(dim, tuple(vol), tuple(det), *(np.ascontiguousarray(np.stack(c) if i < 4 else np.array(c), dtype=dt)
                                               for i, c in enumerate(cols)))


# --------------------------------------------------------------------------------------
# Projectors


def _dims(g: Geom):
    vol = np.array(list(g.vol) + [1] * (3 - g.dim), np.int64)
    det = np.array(list(g.det) + [1] * (2 - len(g.det)), np.int64)
    return vol, det


def xray(kind, g: Geom, data, adjoint=False, mode=MODE_ELSA, nthreads=0, acc64=False, pix=None):
    """Joseph (kind=JOSEPH; JosephsMethod.cpp:140-198) or Siddon (kind=SIDDON;
    SiddonsMethod.cpp:76-124) forward (adjoint=False) or back projection. acc64 with an FP32
    geometry keeps the FP32 geometry and weights but takes FP64 data and sums (the exact sum
    of the FP32 terms). pix: a curved detector's planar coordinates of the pixel centres
    (CurvedDetectorDescriptor::getPlanarCoords, dim - 1 per pixel, u fastest)."""
    L = lib()
    dt = np.float64 if g.Pinv.dtype == np.float64 else np.float32
    if pix is not None:
        if acc64:
            raise ValueError("curved detectors: acc64 is not provided")
        f = L.orc_xray_px_f64 if dt == np.float64 else L.orc_xray_px_f32
        vol, det = _dims(g)
        data = np.ascontiguousarray(data, dtype=dt)
        out = np.empty(g.vol_shape if adjoint else g.sino_shape, dt)
        rc = f(kind, g.dim, int(adjoint), mode, vol, det, g.nviews, np.ascontiguousarray(g.Pinv),
               np.ascontiguousarray(g.C), np.ascontiguousarray(pix, dtype=dt).reshape(-1), data.reshape(-1),
               out.reshape(-1), nthreads)
        assert rc == 0
        return out
    f = L.orc_xray_f64 if dt == np.float64 else L.orc_xray_f32
    if acc64 and dt == np.float32:
        f, dt = L.orc_xray_f32_acc64, np.float64
    vol, det = _dims(g)
    data = np.ascontiguousarray(data, dtype=dt)
    out = np.empty(g.vol_shape if adjoint else g.sino_shape, dt)
    rc = f(kind, g.dim, int(adjoint), mode, vol, det, g.nviews, np.ascontiguousarray(g.Pinv),
           np.ascontiguousarray(g.C), data.reshape(-1), out.reshape(-1), nthreads)
    assert rc == 0
    return out


def joseph(g, data, adjoint=False, mode=MODE_ELSA, nthreads=0, acc64=False, pix=None):
    return xray(JOSEPH, g, data, adjoint, mode, nthreads, acc64, pix)


def siddon(g, data, adjoint=False, nthreads=0, mode=MODE_ELSA, acc64=False, pix=None):
    return xray(SIDDON, g, data, adjoint, mode, nthreads, acc64, pix)


def curved_planar_coords(det, angle, s2d, dtype=np.float32):
    """CurvedDetectorDescriptor::setup / mapCurvedCoordToPlanarCoord
    (CurvedDetectorDescriptor.cpp:36-95): the planar coordinate of every pixel centre of a
    curved detector of det = (len[, height]) pixels spanning `angle` radians at source-detector
    distance s2d, evaluated in dtype (elsa: real_t = float). Shape (pixels, dim - 1)."""
    dt = np.dtype(dtype).type
    det = tuple(int(d) for d in det)
    length, height = det[0], (det[1] if len(det) > 1 else 1)
    radius = dt(length) / dt(angle)
    delta = dt(angle) / dt(length)
    cx, cy = dt(length) / dt(2), dt(height) / dt(2)
    ccy = dt(s2d) - radius
    out = []
    for h in range(height):
        for l in range(length):
            pu, pv = dt(l) + dt(0.5), dt(h) + dt(0.5)
            ra = (pu - cx) * delta
            x = cx + np.sin(ra, dtype=dt) * radius
            z = ccy + np.cos(ra, dtype=dt) * radius
            factor = dt(1) if z == 0 else dt(s2d) / z
            row = [(x - cx) * factor + cx]
            if len(det) > 1:
                row.append((pv - cy) * factor + cy)
            out.append(row)
    return np.array(out, dtype=dt)


def voxel(g: Geom, data, lut, support, adjoint=False, nthreads=0, differential=False, acc64=False):
    """Blob / B-spline voxel projector (VoxelComputation.h:40-185, full-rectangle footprint);
    differential=True gives the phase-contrast weight (VoxelComputation.h:26-38, used by
    PhaseContrastProjector.cpp with the derivative LUT in 2D, the normalized-gradient LUT in 3D).
    acc64 with an FP32 geometry: FP32 footprints with an FP64 table, data and sums (elsa's
    VoxelProjectorCUDA<double>, real_t geometry)."""
    L = lib()
    dt = np.float64 if g.P.dtype == np.float64 else np.float32
    f = L.orc_voxel_f64 if dt == np.float64 else L.orc_voxel_f32
    if acc64 and dt == np.float32:
        vol, det = _dims(g)
        data = np.ascontiguousarray(data, dtype=np.float64)
        lut = np.ascontiguousarray(lut, dtype=np.float64)
        out = np.empty(g.vol_shape if adjoint else g.sino_shape, np.float64)
        rc = L.orc_voxel_f32_acc64(g.dim, int(adjoint), vol, det, g.nviews, np.ascontiguousarray(g.P),
                                   np.ascontiguousarray(g.ext), np.ascontiguousarray(g.sdd), lut, lut.shape[0],
                                   support, int(differential), data.reshape(-1), out.reshape(-1), nthreads)
        assert rc == 0
        return out
    vol, det = _dims(g)
    data = np.ascontiguousarray(data, dtype=dt)
    lut = np.ascontiguousarray(lut, dtype=dt)
    out = np.empty(g.vol_shape if adjoint else g.sino_shape, dt)
    rc = f(g.dim, int(adjoint), vol, det, g.nviews, np.ascontiguousarray(g.P), np.ascontiguousarray(g.ext),
           np.ascontiguousarray(g.sdd), lut, lut.shape[0], support, int(differential), data.reshape(-1),
           out.reshape(-1), nthreads)
    assert rc == 0
    return out


def joseph_rays(g: Geom):
    """Per-ray DrivingDirectionTraversal state: (axis, nsteps, pos0, step, len)."""
    L = lib()
    dt = np.float64 if g.Pinv.dtype == np.float64 else np.float32
    f = L.orc_joseph_rays_f64 if dt == np.float64 else L.orc_joseph_rays_f32
    vol, det = _dims(g)
    n = g.nviews * int(np.prod(g.det))
    axis = np.empty(n, np.int32)
    nsteps = np.empty(n, np.int64)
    pos0 = np.empty((n, 3), dt)
    step = np.empty((n, 3), dt)
    ln = np.empty(n, dt)
    f(g.dim, vol, det, g.nviews, np.ascontiguousarray(g.Pinv), np.ascontiguousarray(g.C),
      axis.ctypes.data, nsteps.ctypes.data, pos0.ctypes.data, step.ctypes.data, ln.ctypes.data)
    return axis, nsteps, pos0, step, ln


def joseph_sample_count(g: Geom) -> int:
    return int(joseph_rays(g)[1].sum())


def make_ray(g: Geom, view, u, v=0):
    L = lib()
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    L.orc_make_ray_f32(g.dim, np.ascontiguousarray(g.Pinv[view], np.float32),
                       np.ascontiguousarray(g.C[view], np.float32), u, v, o, d)
    return o[:g.dim], d[:g.dim]


def intersect(o, d, vol):
    """elsa intersectRay (Intersection.cpp:35-46); returns (tmin, tmax) or None."""
    L = lib()
    o = np.ascontiguousarray(np.pad(np.asarray(o, np.float32), (0, 3 - len(o))))
    dd = np.ascontiguousarray(np.pad(np.asarray(d, np.float32), (0, 3 - len(d))))
    tmin, tmax = ctypes.c_float(), ctypes.c_float()
    hit = L.orc_intersect_f32(len(vol), o, dd, np.array(list(vol) + [1] * (3 - len(vol)), np.int64),
                              ctypes.byref(tmin), ctypes.byref(tmax))
    return (tmin.value, tmax.value) if hit else None


# --------------------------------------------------------------------------------------
# Basis functions (Blobs.h, BSplines.h, Luts.hpp, Bessel.cpp)


def bessi(m, x):
    return lib().orc_bessi(m, x)


def blob_evaluate(r, a=2.0, alpha=10.83, m=2):
    return lib().orc_blob_evaluate_f64(r, a, alpha, m)


def blob_projected(s, a=2.0, alpha=10.83, m=2):
    return lib().orc_blob_projected_f64(s, a, alpha, m)


def blob_lut(radius=2.0, alpha=10.83, order=2, N=100, dtype=np.float32):
    out = np.empty(N, dtype)
    if np.dtype(dtype) == np.float32:
        lib().orc_blob_lut_f32(radius, alpha, order, N, out)
    else:
        lib().orc_blob_lut_f64(radius, alpha, order, N, out)
    return out


def bspline_lut(dim, order=2, N=100, dtype=np.float32):
    out = np.empty(N, dtype)
    if np.dtype(dtype) == np.float32:
        lib().orc_bspline_lut_f32(dim, order, N, out)
    else:
        lib().orc_bspline_lut_f64(dim, order, N, out)
    return out


def blob_derivative_lut(radius=2.0, alpha=10.83, order=2, N=100, dtype=np.float32, normalized=False):
    """ProjectedBlob::get_derivative_lut() (normalized: get_normalized_gradient_lut())."""
    out = np.empty(N, dtype)
    f = lib().orc_blob_deriv_lut_f32 if np.dtype(dtype) == np.float32 else lib().orc_blob_deriv_lut_f64
    f(radius, alpha, order, N, int(normalized), out)
    return out


def bspline_derivative_lut(dim, order=2, N=100, dtype=np.float32, normalized=False):
    """ProjectedBSpline::get_derivative_lut() (normalized: get_normalized_gradient_lut())."""
    out = np.empty(N, dtype)
    f = lib().orc_bspline_deriv_lut_f32 if np.dtype(dtype) == np.float32 else lib().orc_bspline_deriv_lut_f64
    f(dim, order, N, int(normalized), out)
    return out


def lut_eval(lut, support, distance):
    lut = np.ascontiguousarray(lut)
    if lut.dtype == np.float64:
        return lib().orc_lut_eval_f64(lut, lut.shape[0], support, distance)
    return lib().orc_lut_eval_f32(lut, lut.shape[0], support, distance)


# --------------------------------------------------------------------------------------
# Solvers (elsa/solvers/CGLS.cpp, APGD.cpp, PowerIterations.cpp) over any pair of callables
# A(x) -> sinogram, At(y) -> volume, in float64 unless the inputs say otherwise. Reference
# implementations for the device solvers; written for small problems.


def _dot(a, b):
    return float(np.dot(a.ravel().astype(np.float64), b.ravel().astype(np.float64)))


def cgls(A, At, b, shape, iterations, eps=0.0, tol=1e-4, x0=None, dtype=np.float64):
    """CGLS.cpp:19-91. Returns (x, residual norms ||s|| at the start and after each iteration
    taken, iterations taken)."""
    damp = eps * eps
    b = np.asarray(b, dtype)
    if x0 is not None:
        x = np.array(x0, dtype).reshape(shape)
        s = b - A(x)                                   # :31-35
        r = At(s) - damp * x                           # :37-38
    else:
        x = np.zeros(shape, dtype)
        s = b.copy()
        r = At(s)                                      # :42
    c = r.copy()
    kold = _dot(r, r)                                  # :48
    res = [np.sqrt(_dot(s, s))]
    taken = 0
    for _ in range(iterations):
        if kold < tol:                                 # :54-58
            break
        q = A(c)                                       # :60
        delta = _dot(q, q)
        alpha = kold / delta if damp == 0 else kold / (delta + damp * _dot(c, c))
        x = x + dtype(alpha) * c                       # :71
        s = s - dtype(alpha) * q                       # :72
        r = At(s)
        if damp != 0.0:
            r = r - dtype(damp) * x
        k = _dot(r, r)
        beta = k / kold
        c = r + dtype(beta) * c                        # :82
        kold = k
        taken += 1
        res.append(np.sqrt(_dot(s, s)))
    return x, np.array(res), taken


def power_iterations(A, At, shape, niters=5, dtype=np.float64):
    """PowerIterations.cpp:10-27 on A^T A: u = 1; niters times u = A^T A u / ||.||;
    returns u . (A^T A u) / ||u||."""
    u = np.ones(shape, dtype)
    for _ in range(niters):
        u = At(A(u))
        u = u / np.sqrt(_dot(u, u))
    return _dot(u, At(A(u))) / np.sqrt(_dot(u, u))


def finite_differences(x, dims, adjoint=False):
    """elsa's FiniteDifferences(domain) with DiffType::FORWARD and every axis active.
    dims: voxels per axis, axis 0 fastest (elsa's order); x flat. Forward
    (FiniteDifferences.cpp:106-155): one block of nvox values per axis, axis 0 first, block d
    at voxel i = -x[i] + x[i + e_d], -x[i] on the last plane of axis d. Adjoint: per voxel, axis
    by axis in that order, + y_d[i - e_d] (not on the first plane) - y_d[i] -- the transpose of
    the forward. (elsa's applyAdjointHelper, FiniteDifferences.cpp:157-206, reads block d at
    offset d * _coordDelta[d], which is d * nvox only for the last axis: its 3D adjoint takes
    axis 1's block from offset n0 * n1. This restatement uses nvox, the transpose.)"""
    dims = tuple(int(n) for n in dims)
    D, nvox = len(dims), int(np.prod(dims))
    shp = dims[::-1]                                   # C order: axis 0 is the last numpy axis
    if not adjoint:
        v = np.asarray(x).reshape(shp)
        out = np.empty((D,) + shp, v.dtype)
        for d in range(D):
            ax = D - 1 - d
            y = -v                                     # :136
            lo = [slice(None)] * D
            hi = [slice(None)] * D
            lo[ax], hi[ax] = slice(0, -1), slice(1, None)
            y[tuple(lo)] += v[tuple(hi)]               # :137-138
            out[d] = y
        return out.reshape(-1)
    y = np.asarray(x).reshape((D,) + shp)
    acc = np.zeros(shp, y.dtype)
    for d in range(D):
        ax = D - 1 - d
        lo = [slice(None)] * D
        hi = [slice(None)] * D
        lo[ax], hi[ax] = slice(0, -1), slice(1, None)
        acc[tuple(hi)] += y[d][tuple(lo)]              # :185-186
        acc -= y[d]                                    # :188
    return acc.reshape(-1)


def admml2(A, At, b, shape, iterations, tau=0.9, sigma=None, ninner=5, x0=None, dtype=np.float64, reg=None):
    """ADMML2.cpp:46-97: x = CGNE(K = [A; lam R], w = [b; z - u], ninner, x)
    (RegularizedInversion.cpp:11-61, CGNE.cpp:22-69) with lam = 1/sqrt(tau) and R the identity
    (reg None) or reg = (R, Rt) callables on flat arrays; z = prox(R x + u) (soft thresholding of
    threshold tau * sigma, or the identity when sigma is None); u += R x; u -= z. Returns
    (x, ||b - A x|| at the start and after each iteration)."""
    t = dtype(tau)
    lam = dtype(1) / np.sqrt(t)
    b = np.asarray(b, dtype)
    x = np.zeros(shape, dtype) if x0 is None else np.array(x0, dtype).reshape(shape)
    if reg is None:
        R = Rt = lambda v: v
        ktk = lambda v: At(A(v)) + lam * (lam * v)
    else:
        R = lambda v: reg[0](v.reshape(-1))
        Rt = lambda y: reg[1](y).reshape(shape)
        ktk = lambda v: At(A(v)) + lam * Rt(lam * R(v))
    z = np.zeros_like(R(x))
    u = np.zeros_like(z)
    atb = At(b)
    res = [np.sqrt(_dot(b - A(x), b - A(x)))]
    for _ in range(iterations):
        rhs = atb + lam * Rt(z - u)
        r = rhs - ktk(x)
        c = r.copy()
        kold = _dot(r, r)
        for _ in range(ninner):
            ac = ktk(c)
            alpha = kold / _dot(c, ac)
            x = x + dtype(alpha) * c
            r = r - dtype(alpha) * ac
            k = _dot(r, r)
            c = r + dtype(k / kold) * c
            kold = k
        rx = R(x)
        v = rx + u
        if sigma is None:
            z = v
        else:
            tmp = np.abs(v) - t * dtype(sigma)
            tmp = dtype(0.5) * (tmp + np.abs(tmp))
            z = np.sign(v) * tmp
        u = (u + rx) - z
        res.append(np.sqrt(_dot(b - A(x), b - A(x))))
    return x, np.array(res)


def apgd(A, At, b, shape, iterations, mu=None, sigma=None, epsilon=None, x0=None, dtype=np.float64):
    """APGD.cpp:33-102 (FISTA) with prox = soft thresholding of threshold mu * sigma
    (ProximalL1.cpp:32-43), or the identity when sigma is None. mu None: 0.45 / L with L from
    5 power iterations (APGD.cpp:19-27). Returns (x, ||grad||^2 per iteration taken,
    iterations taken, mu)."""
    if mu is None:
        mu = float(dtype(0.45) / dtype(power_iterations(A, At, shape, 5, dtype)))
    if epsilon is None:
        epsilon = float(np.finfo(dtype).eps)
    x = np.zeros(shape, dtype) if x0 is None else np.array(x0, dtype).reshape(shape)
    xprev, y = x.copy(), x.copy()
    tprev = dtype(1)
    atb = At(np.asarray(b, dtype))
    grad = At(A(y)) - atb                              # :55-60
    gn = []
    taken = 0
    m = dtype(mu)
    for _ in range(iterations):
        z = y - m * grad                               # :69
        if sigma is None:
            x = z
        else:
            tmp = np.abs(z) - m * dtype(sigma)
            tmp = dtype(0.5) * (tmp + np.abs(tmp))
            x = np.sign(z) * tmp                       # :72 prox
        t = (dtype(1) + np.sqrt(dtype(1) + dtype(4) * tprev * tprev)) / dtype(2)
        y = x + ((tprev - dtype(1)) / t) * (x - xprev)
        xprev = x
        tprev = t
        g2 = _dot(grad, grad)
        gn.append(g2)
        taken += 1
        if g2 <= epsilon:                              # :83-87
            break
        grad = At(A(y)) - atb
    return x, np.array(gn), taken, mu
