"""elsa_amd -- MI355X-native X-ray projectors with elsa's operator interface.

The projectors (Joseph, Siddon, blob / B-spline voxel) are HIP kernels for gfx950 in
libelsahip.so, reached through the C-ABI of include/elsa_hip.h. This package is the Python
face of that library, with pyelsa's class names.
"""
from .descriptors import (CircleTrajectoryGenerator, CurvedCircleTrajectoryGenerator, CurvedDetectorDescriptor,
                          DataDescriptor, DetectorDescriptor, Geometry, PlanarDetectorDescriptor, VolumeDescriptor)
from .projectors import (BlobVoxelProjectorCUDA, BlobVoxelProjectorCUDAd, BlobVoxelProjectorCUDAf,
                         BSplineVoxelProjectorCUDA, BSplineVoxelProjectorCUDAd, BSplineVoxelProjectorCUDAf,
                         JosephsMethodCUDA, JosephsMethodCUDAd, JosephsMethodCUDAf, LinearOperator,
                         PhaseContrastBlobVoxelProjectorCUDA, PhaseContrastBlobVoxelProjectorCUDAd,
                         PhaseContrastBlobVoxelProjectorCUDAf, PhaseContrastBSplineVoxelProjectorCUDA,
                         PhaseContrastBSplineVoxelProjectorCUDAd, PhaseContrastBSplineVoxelProjectorCUDAf,
                         SiddonsMethodCUDA, SiddonsMethodCUDAd, SiddonsMethodCUDAf, blob_derivative_lut, blob_lut,
                         bspline_derivative_lut, bspline_lut)
from .solvers import ADMML2, APGD, CGLS, CGLSRefined, ProximalIdentity, ProximalL1, powerIterations
from . import autograd, io
from .autotune import autotune, autotune_batch
from .batch import ProjectorBatch
from .subsets import SubsetSampler

__all__ = [
    "CircleTrajectoryGenerator", "CurvedCircleTrajectoryGenerator", "CurvedDetectorDescriptor", "DataDescriptor", "DetectorDescriptor", "Geometry", "PlanarDetectorDescriptor",
    "VolumeDescriptor", "LinearOperator", "JosephsMethodCUDA", "JosephsMethodCUDAf", "JosephsMethodCUDAd",
    "SiddonsMethodCUDA", "SiddonsMethodCUDAf", "SiddonsMethodCUDAd", "BlobVoxelProjectorCUDA",
    "BlobVoxelProjectorCUDAf", "BlobVoxelProjectorCUDAd", "BSplineVoxelProjectorCUDA", "BSplineVoxelProjectorCUDAf",
    "BSplineVoxelProjectorCUDAd", "PhaseContrastBlobVoxelProjectorCUDA", "PhaseContrastBlobVoxelProjectorCUDAf",
    "PhaseContrastBlobVoxelProjectorCUDAd", "PhaseContrastBSplineVoxelProjectorCUDA",
    "PhaseContrastBSplineVoxelProjectorCUDAf", "PhaseContrastBSplineVoxelProjectorCUDAd", "blob_lut", "bspline_lut",
    "blob_derivative_lut", "bspline_derivative_lut", "CGLS", "CGLSRefined", "APGD", "ADMML2", "ProximalL1", "ProximalIdentity",
    "powerIterations", "autotune", "autotune_batch", "ProjectorBatch", "SubsetSampler",
]
