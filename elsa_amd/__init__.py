"""elsa_amd -- MI355X-native X-ray projectors with elsa's operator interface.

The projectors (Joseph, Siddon, blob / B-spline voxel) are HIP kernels for gfx950 in
libelsahip.so, reached through the C-ABI of include/elsa_hip.h. This package is the Python
face of that library, with pyelsa's class names.
"""
from .descriptors import (CircleTrajectoryGenerator, DataDescriptor, DetectorDescriptor, Geometry,
                          PlanarDetectorDescriptor, VolumeDescriptor)
from .projectors import (BlobVoxelProjectorCUDA, BlobVoxelProjectorCUDAd, BlobVoxelProjectorCUDAf,
                         BSplineVoxelProjectorCUDA, BSplineVoxelProjectorCUDAd, BSplineVoxelProjectorCUDAf,
                         JosephsMethodCUDA, JosephsMethodCUDAd, JosephsMethodCUDAf, LinearOperator,
                         SiddonsMethodCUDA, SiddonsMethodCUDAd, SiddonsMethodCUDAf, blob_lut, bspline_lut)
from .solvers import APGD, CGLS, ProximalIdentity, ProximalL1, powerIterations
from . import autograd
from .autotune import autotune

__all__ = [
    "CircleTrajectoryGenerator", "DataDescriptor", "DetectorDescriptor", "Geometry", "PlanarDetectorDescriptor",
    "VolumeDescriptor", "LinearOperator", "JosephsMethodCUDA", "JosephsMethodCUDAf", "JosephsMethodCUDAd",
    "SiddonsMethodCUDA", "SiddonsMethodCUDAf", "SiddonsMethodCUDAd", "BlobVoxelProjectorCUDA",
    "BlobVoxelProjectorCUDAf", "BlobVoxelProjectorCUDAd", "BSplineVoxelProjectorCUDA", "BSplineVoxelProjectorCUDAf",
    "BSplineVoxelProjectorCUDAd", "blob_lut", "bspline_lut", "CGLS", "APGD", "ProximalL1", "ProximalIdentity",
    "powerIterations", "autotune",
]
