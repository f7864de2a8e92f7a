"""Two interchangeable projector back ends for the known-answer tests.

OracleBackend runs the CPU restatement (oracle/), GpuBackend runs the HIP kernels through
elsa_amd and its C-ABI. The KATs in kats.py take either one, the way elsa's Siddon tests are
templated over SiddonsMethod and SiddonsMethodCUDA
(elsa/projectors_cuda/tests/test_SiddonsMethodCUDA.cpp:30-32).

A pose is a dict with keys s2c, c2d and optionally gamma, beta, alpha (radians), ppo
(principal point offset) and cor (centre-of-rotation offset), as elsa's Geometry takes them.
"""
from __future__ import annotations

import numpy as np

import oracle


class OracleBackend:
    name = "oracle"

    def __init__(self, dtype=np.float32, mode=oracle.MODE_ELSA):
        self.dtype = np.dtype(dtype)
        self.mode = mode

    def setup(self, vol, det, poses, spacing=None):
        dim = len(vol)
        ps = [oracle.geometry_pose(dim, p["s2c"], p["c2d"], vol, det, angle_rad=p.get("gamma", 0.0),
                                   beta=p.get("beta", 0.0), alpha=p.get("alpha", 0.0), vol_spacing=spacing,
                                   ppo=p.get("ppo"), cor=p.get("cor"), dtype=self.dtype) for p in poses]
        return oracle.geom_from_poses(dim, vol, det, ps, self.dtype)

    def joseph(self, g, x, adjoint=False):
        return oracle.joseph(g, np.asarray(x, self.dtype), adjoint, self.mode)

    def siddon(self, g, x, adjoint=False):
        return oracle.siddon(g, np.asarray(x, self.dtype), adjoint, mode=self.mode)

    def voxel(self, g, x, kind="blob", adjoint=False, **kw):
        """kind: blob, bspline, or pc_blob / pc_bspline for the phase-contrast projectors (the
        derivative table in 2D, the normalized-gradient table in 3D, differential weight)."""
        pc = kind.startswith("pc_")
        base = kind[3:] if pc else kind
        norm = pc and g.dim == 3
        if base == "blob":
            args = (kw.get("radius", 2.0), kw.get("alpha", 10.83), kw.get("order", 2), 100, self.dtype)
            lut = oracle.blob_derivative_lut(*args, normalized=norm) if pc else oracle.blob_lut(*args)
            sup = kw.get("radius", 2.0)
        else:
            order = kw.get("order", 2)
            lut = (oracle.bspline_derivative_lut(g.dim, order, 100, self.dtype, normalized=norm) if pc
                   else oracle.bspline_lut(g.dim, order, 100, self.dtype))
            sup = (order + 1) * 0.5
        return oracle.voxel(g, np.asarray(x, self.dtype), lut, sup, adjoint, differential=pc)


class GpuBackend:
    """HIP projectors of libelsahip.so via elsa_amd (numpy in, numpy out)."""

    name = "gpu"

    def __init__(self, dtype=np.float32, kernels="fast"):
        self.dtype = np.dtype(dtype)
        self.kernels = kernels
        self._ops = {}

    def setup(self, vol, det, poses, spacing=None):
        import elsa_amd as E

        vd = E.VolumeDescriptor(list(vol), spacing)
        geoms = [E.Geometry(p["s2c"], p["c2d"], vd, list(det), gamma=p.get("gamma", 0.0), beta=p.get("beta", 0.0),
                            alpha=p.get("alpha", 0.0), principalPointOffset=p.get("ppo"),
                            centerOfRotationOffset=p.get("cor")) for p in poses]
        dd = E.PlanarDetectorDescriptor(list(det) + [len(poses)], geoms)
        return vd, dd

    def _op(self, cls, g, **kw):
        import elsa_amd as E

        key = (cls, id(g[1]), tuple(sorted(kw.items())))
        if key not in self._ops:
            self._ops[key] = getattr(E, cls)(g[0], g[1], dtype=self.dtype, kernels=self.kernels, **kw)
        return self._ops[key]

    def joseph(self, g, x, adjoint=False):
        op = self._op("JosephsMethodCUDA", g)
        x = np.asarray(x, self.dtype)
        return op.applyAdjoint(x) if adjoint else op.apply(x)

    def siddon(self, g, x, adjoint=False):
        op = self._op("SiddonsMethodCUDA", g)
        x = np.asarray(x, self.dtype)
        return op.applyAdjoint(x) if adjoint else op.apply(x)

    VOXEL = {"blob": "BlobVoxelProjectorCUDA", "bspline": "BSplineVoxelProjectorCUDA",
             "pc_blob": "PhaseContrastBlobVoxelProjectorCUDA", "pc_bspline": "PhaseContrastBSplineVoxelProjectorCUDA"}

    def voxel(self, g, x, kind="blob", adjoint=False, **kw):
        op = self._op(self.VOXEL[kind], g, **kw)
        x = np.asarray(x, self.dtype)
        return op.applyAdjoint(x) if adjoint else op.apply(x)


def geom_to_descriptors(g):
    """elsa_amd descriptors carrying exactly the matrices of an oracle Geom (for parity runs
    that feed both sides the same FP32 geometry)."""
    import elsa_amd as E

    vd = E.VolumeDescriptor(list(g.vol))
    gs = [E.Geometry._from_arrays(g.dim, np.ascontiguousarray(g.P[i], np.float32),
                                  np.ascontiguousarray(g.Pinv[i], np.float32), np.ascontiguousarray(g.C[i], np.float32),
                                  np.ascontiguousarray(g.ext[i], np.float32), np.float32(g.sdd[i]))
          for i in range(g.nviews)]
    dd = E.PlanarDetectorDescriptor(list(g.det) + [g.nviews], gs)
    return vd, dd
