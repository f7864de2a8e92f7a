"""Ordered subsets of a trajectory: elsa's SubsetSampler (elsa/projectors/SubsetSampler.h,
SubsetSampler.cpp) over the GPU projectors.

  SubsetSampler(volumeDescriptor, detectorDescriptor, nSubsets, samplingStrategy="round_robin")
      .getPartitionedData(sinogram)     -> per-subset sinograms (poses of the subset, in order)
      .getSubsetProjectors(cls, **kw)   -> one projector per subset (its poses only)
      .getProjector(cls, **kw)          -> the projector of the full trajectory
  SubsetSampler.splitRoundRobin(indices, nSubsets)                 (SubsetSampler.cpp:53-64)
  SubsetSampler.splitRotationalClustering(detectorDescriptor, n)   (SubsetSampler.cpp:66-97)

A subset's sinogram and projector are those of a detector descriptor with the subset's poses,
so solvers (e.g. an ordered-subsets loop) can apply them independently; a ProjectorBatch of
the subset projectors runs them in one launch per pass.
"""
from __future__ import annotations

import math

import numpy as np

from .descriptors import PlanarDetectorDescriptor


def _rotation(g, dim):
    """The rotation of a pose: the first dim columns of the extrinsic matrix [R | t] S with the
    volume spacing S divided out (each column of R has unit length)."""
    M = g.getExtrinsicMatrix()[:, :dim].astype(np.float64)
    return M / np.linalg.norm(M, axis=0, keepdims=True)


class SubsetSampler:
    ROUND_ROBIN = "round_robin"
    ROTATIONAL_CLUSTERING = "rotational_clustering"

    def __init__(self, volumeDescriptor, detectorDescriptor, nSubsets, samplingStrategy="round_robin"):
        n = int(nSubsets)
        if n <= 1:
            raise ValueError("SubsetSampler: nSubsets must be >= 2")
        poses = detectorDescriptor.getNumberOfGeometryPoses()
        if samplingStrategy == self.ROUND_ROBIN:
            self.indexMapping = self.splitRoundRobin(list(range(poses)), n)
        elif samplingStrategy == self.ROTATIONAL_CLUSTERING:
            self.indexMapping = self.splitRotationalClustering(detectorDescriptor, n)
        else:
            raise ValueError("SubsetSampler: unsupported sampling strategy")
        self._vd = volumeDescriptor
        self._dd = detectorDescriptor
        geoms = detectorDescriptor.getGeometry()
        coeffs = list(detectorDescriptor.getNumberOfCoefficientsPerDimension())
        spacing = list(detectorDescriptor.getSpacingPerDimension())
        self.detectorDescriptors = []
        for idx in self.indexMapping:
            c = coeffs[:-1] + [len(idx)]
            self.detectorDescriptors.append(PlanarDetectorDescriptor(c, [geoms[i] for i in idx], spacing))

    @staticmethod
    def splitRoundRobin(indices, nSubsets):
        out = [[] for _ in range(int(nSubsets))]
        for i, v in enumerate(indices):
            out[i % int(nSubsets)].append(v)
        return out

    @staticmethod
    def splitRotationalClustering(detectorDescriptor, nSubsets):
        """Poses sorted by their rotation relative to the first pose (the signed angle in 2D, the
        rotation angle in 3D), then split round robin."""
        geoms = detectorDescriptor.getGeometry()
        dim = detectorDescriptor.getNumberOfDimensions()
        r0 = _rotation(geoms[0], dim).astype(np.float32)

        def dist(g):
            prod = (r0 @ _rotation(g, dim).astype(np.float32).T).astype(np.float32)
            if dim == 2:
                return math.atan2(float(prod[1, 0]), float(prod[0, 0]))
            return math.acos(max(-1.0, min(1.0, (float(np.trace(prod)) - 1.0) / 2.0)))

        keys = [dist(g) for g in geoms]
        order = sorted(range(len(geoms)), key=lambda i: keys[i])
        return SubsetSampler.splitRoundRobin(order, nSubsets)

    def getPartitionedData(self, sinogram):
        """Per-subset sinograms (numpy or torch, poses along the first axis)."""
        return [sinogram[list(idx)] for idx in self.indexMapping]

    def getProjector(self, cls, **kw):
        return cls(self._vd, self._dd, **kw)

    def getSubsetProjectors(self, cls, **kw):
        return [cls(self._vd, dd, **kw) for dd in self.detectorDescriptors]
