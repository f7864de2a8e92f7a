"""Descriptors, geometry and trajectories with elsa's names and meaning.

Mirrors the parts of elsa's core and generators that the projectors consume:
  VolumeDescriptor / DataDescriptor   elsa/core/Descriptors/DataDescriptor.cpp:9-57
  Geometry                            elsa/core/Geometry.cpp:12-109, :225-270
  PlanarDetectorDescriptor            elsa/core/Descriptors/PlanarDetectorDescriptor.cpp:5-22
  CircleTrajectoryGenerator           elsa/generators/CircleTrajectoryGenerator.cpp:9-24
The matrices are computed by the C++ host code of libelsahip.so (ehp_geometry_pose,
ehp_circle_trajectory) in FP32, elsa's real_t.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from . import _lib


def _fp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _i64p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


class DataDescriptor:
    """Sizes, spacing and origin of a data block; axis 0 is the fastest (elsa convention)."""

    def __init__(self, numberOfCoefficientsPerDimension: Sequence[int], spacingPerDimension=None):
        n = np.asarray(numberOfCoefficientsPerDimension, np.int64).reshape(-1)
        if n.size == 0 or np.any(n <= 0):
            raise ValueError("DataDescriptor: non-positive number of coefficients not allowed")
        sp = np.ones(n.size, np.float32) if spacingPerDimension is None else \
            np.asarray(spacingPerDimension, np.float32).reshape(-1)
        if sp.size != n.size:
            raise ValueError("DataDescriptor: mismatch between numberOfCoefficientsPerDimension and "
                             "spacingPerDimension")
        if np.any(sp < 0):
            raise ValueError("DataDescriptor: non-positive spacing not allowed")
        self._n = n
        self._spacing = sp
        self._origin = (np.float32(0.5) * (n.astype(np.float32) * sp)).astype(np.float32)

    def getNumberOfDimensions(self) -> int:
        return int(self._n.size)

    def getNumberOfCoefficientsPerDimension(self) -> np.ndarray:
        return self._n.copy()

    def getNumberOfCoefficients(self) -> int:
        return int(np.prod(self._n))

    def getSpacingPerDimension(self) -> np.ndarray:
        return self._spacing.copy()

    def getLocationOfOrigin(self) -> np.ndarray:
        return self._origin.copy()

    @property
    def shape(self) -> tuple:
        """numpy/torch shape of a C-contiguous array of this descriptor (slowest axis first)."""
        return tuple(int(v) for v in self._n[::-1])

    def __eq__(self, other) -> bool:
        return (type(self) is type(other) and np.array_equal(self._n, other._n)
                and np.array_equal(self._spacing, other._spacing))


class VolumeDescriptor(DataDescriptor):
    pass


class Geometry:
    """One pose: projection matrix P, its inverse Pinv, camera centre C, extrinsics ext and
    the source-detector distance (elsa/core/Geometry.h). 2D takes one angle (radians);
    3D takes YZY Euler angles (gamma, beta, alpha)."""

    def __init__(self, sourceToCenterOfRotation: float, centerOfRotationToDetector: float,
                 volumeDescriptor: DataDescriptor, detectorSize: Sequence[int], gamma: float = 0.0,
                 beta: float = 0.0, alpha: float = 0.0, detectorSpacing=None, principalPointOffset=None,
                 centerOfRotationOffset=None):
        dim = volumeDescriptor.getNumberOfDimensions()
        if dim not in (2, 3):
            raise ValueError("Geometry: only 2D and 3D geometries are supported")
        det = np.asarray(detectorSize, np.int64).reshape(-1)[: dim - 1]
        dsp = np.ones(dim - 1, np.float32) if detectorSpacing is None else \
            np.asarray(detectorSpacing, np.float32).reshape(-1)[: dim - 1]
        dorg = (np.float32(0.5) * (det.astype(np.float32) * dsp)).astype(np.float32)
        vsp = volumeDescriptor.getSpacingPerDimension()
        vorg = volumeDescriptor.getLocationOfOrigin()
        ppo = np.zeros(dim - 1, np.float32) if principalPointOffset is None else \
            np.asarray(principalPointOffset, np.float32).reshape(-1)
        cor = np.zeros(dim, np.float32) if centerOfRotationOffset is None else \
            np.asarray(centerOfRotationOffset, np.float32).reshape(-1)
        nm = dim * (dim + 1)
        self.dim = dim
        self.P = np.zeros(nm, np.float32)
        self.Pinv = np.zeros(nm, np.float32)
        self.C = np.zeros(dim, np.float32)
        self.ext = np.zeros(nm, np.float32)
        sdd = np.zeros(1, np.float32)
        L = _lib.load()
        _lib.check(L.ehp_geometry_pose(dim, sourceToCenterOfRotation, centerOfRotationToDetector, _fp(vsp), _fp(vorg),
                                       _fp(dsp), _fp(dorg), gamma, beta, alpha, _fp(ppo), _fp(cor), _fp(self.P),
                                       _fp(self.Pinv), _fp(self.C), _fp(self.ext), _fp(sdd)), "Geometry")
        self.sdd = np.float32(sdd[0])

    @classmethod
    def _from_arrays(cls, dim, P, Pinv, C, ext, sdd) -> "Geometry":
        g = cls.__new__(cls)
        g.dim, g.P, g.Pinv, g.C, g.ext, g.sdd = dim, P, Pinv, C, ext, np.float32(sdd)
        return g

    def getProjectionMatrix(self) -> np.ndarray:
        return self.P.reshape(self.dim + 1, self.dim).T.copy()

    def getInverseProjectionMatrix(self) -> np.ndarray:
        return self.Pinv.reshape(self.dim, self.dim + 1).T.copy()

    def getCameraCenter(self) -> np.ndarray:
        return self.C.copy()

    def getExtrinsicMatrix(self) -> np.ndarray:
        return self.ext.reshape(self.dim + 1, self.dim).T.copy()

    def getSourceDetectorDistance(self) -> float:
        return float(self.sdd)


class DetectorDescriptor(DataDescriptor):
    """A detector of size (u[, v], poses) with one Geometry per pose."""

    def __init__(self, numberOfCoefficientsPerDimension, geometryList: Sequence[Geometry], spacingPerDimension=None):
        super().__init__(numberOfCoefficientsPerDimension, spacingPerDimension)
        self._geometry = list(geometryList)
        if self._n[-1] != len(self._geometry):
            raise ValueError("DetectorDescriptor: the last dimension must equal the number of poses")

    def getGeometry(self):
        return list(self._geometry)

    def getNumberOfGeometryPoses(self) -> int:
        return len(self._geometry)

    def pose_arrays(self):
        """Stacked (P, Pinv, C, ext, sdd) of all poses, as the C-ABI takes them."""
        g = self._geometry
        return (np.ascontiguousarray(np.stack([x.P for x in g]), np.float32),
                np.ascontiguousarray(np.stack([x.Pinv for x in g]), np.float32),
                np.ascontiguousarray(np.stack([x.C for x in g]), np.float32),
                np.ascontiguousarray(np.stack([x.ext for x in g]), np.float32),
                np.ascontiguousarray(np.array([x.sdd for x in g], np.float32)))


class PlanarDetectorDescriptor(DetectorDescriptor):
    pass


class CurvedDetectorDescriptor(DetectorDescriptor):
    """A detector curved around the source (elsa CurvedDetectorDescriptor,
    CurvedDetectorDescriptor.h/.cpp): its numberOfCoefficientsPerDimension[0] pixels span
    `angle` radians on a circle of radius n / angle at source-detector distance s2d. Each pixel
    centre maps to a planar detector coordinate (getPlanarCoords), through which the
    Joseph and Siddon projectors trace its ray."""

    def __init__(self, numberOfCoefficientsPerDimension, geometryList, angle: float, s2d: float,
                 spacingPerDimension=None):
        super().__init__(numberOfCoefficientsPerDimension, geometryList, spacingPerDimension)
        self._angle = float(angle)
        self._s2d = float(s2d)
        dim = self.getNumberOfDimensions()
        det = np.asarray(self._n[: dim - 1], np.int64)
        npix = int(np.prod(det))
        self._coords = np.zeros(npix * (dim - 1), np.float32)
        _lib.check(_lib.load().ehp_curved_detector_coords(dim, _i64p(det), self._angle, self._s2d, _fp(self._coords)),
                   "CurvedDetectorDescriptor")

    def getPlanarCoords(self) -> np.ndarray:
        """(pixels, dim - 1): the planar coordinate of every pixel centre, u fastest."""
        return self._coords.reshape(-1, self.getNumberOfDimensions() - 1).copy()

    def getRadius(self) -> float:
        return float(np.float32(self._n[0]) / np.float32(self._angle))

    def pixel_coords(self):
        return self._coords


class CurvedCircleTrajectoryGenerator:
    """elsa CurvedCircleTrajectoryGenerator (CurvedCircleTrajectoryGenerator.cpp:8-29): the
    circular trajectory of CircleTrajectoryGenerator with a curved detector spanning `angle`
    radians at source-detector distance sourceToCenter + centerToDetector."""

    @staticmethod
    def createTrajectory(numberOfPoses: int, volumeDescriptor: DataDescriptor, arcDegrees: float,
                         sourceToCenter: float, centerToDetector: float, angle: float,
                         detectorSize=None) -> "CurvedDetectorDescriptor":
        planar = CircleTrajectoryGenerator.createTrajectory(numberOfPoses, volumeDescriptor, arcDegrees, sourceToCenter,
                                                            centerToDetector, detectorSize=detectorSize)
        return CurvedDetectorDescriptor(list(planar.getNumberOfCoefficientsPerDimension()), planar.getGeometry(), angle,
                                        float(np.float32(sourceToCenter) + np.float32(centerToDetector)),
                                        list(planar.getSpacingPerDimension()))


class CircleTrajectoryGenerator:
    """Circular trajectory (elsa/generators/CircleTrajectoryGenerator.cpp:9-24): poses at
    LinSpaced(numberOfPoses, 0, arcDegrees) -- both ends included -- about the last volume
    axis in 2D and about y in 3D; detector floor(N*sqrt(2)) unless given."""

    @staticmethod
    def createTrajectory(numberOfPoses: int, volumeDescriptor: DataDescriptor, arcDegrees: float,
                         sourceToCenter: float, centerToDetector: float, detectorSize=None) -> PlanarDetectorDescriptor:
        dim = volumeDescriptor.getNumberOfDimensions()
        vol = volumeDescriptor.getNumberOfCoefficientsPerDimension()
        if not np.all(volumeDescriptor.getSpacingPerDimension() == 1):
            raise NotImplementedError("CircleTrajectoryGenerator: only unit volume spacing is supported here")
        nm = dim * (dim + 1)
        n = int(numberOfPoses)
        P = np.zeros((n, nm), np.float32)
        Pinv = np.zeros((n, nm), np.float32)
        C = np.zeros((n, dim), np.float32)
        ext = np.zeros((n, nm), np.float32)
        sdd = np.zeros(n, np.float32)
        det_out = np.zeros(2, np.int64)
        det_in = None if detectorSize is None else np.asarray(detectorSize, np.int64).reshape(-1)
        L = _lib.load()
        _lib.check(L.ehp_circle_trajectory(dim, n, _i64p(vol), arcDegrees, sourceToCenter, centerToDetector,
                                           None if det_in is None else _i64p(det_in), _i64p(det_out), _fp(P), _fp(Pinv),
                                           _fp(C), _fp(ext), _fp(sdd)), "CircleTrajectoryGenerator")
        geoms = [Geometry._from_arrays(dim, P[i], Pinv[i], C[i], ext[i], sdd[i]) for i in range(n)]
        det = [int(v) for v in det_out[: dim - 1]] + [n]
        return PlanarDetectorDescriptor(det, geoms, list(volumeDescriptor.getSpacingPerDimension()[: dim - 1]) + [1.0])
