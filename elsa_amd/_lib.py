"""ctypes binding of libelsahip.so (include/elsa_hip.h).

The library is loaded from this package directory only. If it is missing the import of the
projector classes fails with an ImportError naming the build command: there is no fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libelsahip.so")

EHP_OK, EHP_ERR_ARGUMENT, EHP_ERR_HIP, EHP_ERR_ALLOC, EHP_ERR_UNSUPPORTED = range(5)
EHP_F32, EHP_F64 = 0, 1
EHP_SIMPLE_KERNELS = 1

# every symbol include/elsa_hip.h declares
EXPORTS = (
    "ehp_last_error", "ehp_version", "ehp_geometry_create", "ehp_geometry_destroy",
    "ehp_joseph_create", "ehp_siddon_create", "ehp_voxel_create", "ehp_destroy", "ehp_forward",
    "ehp_adjoint", "ehp_op_sizes", "ehp_sample_count", "ehp_tile", "ehp_joseph_ray_info", "ehp_geometry_pose",
    "ehp_circle_trajectory", "ehp_blob_lut", "ehp_bspline_lut", "ehp_cgls", "ehp_apgd", "ehp_power_iterations",
    "ehp_batch_create", "ehp_batch_destroy", "ehp_batch_layout", "ehp_batch_forward", "ehp_batch_adjoint",
    "ehp_batch_cgls", "ehp_batch_apgd", "ehp_batch_power_iterations", "ehp_blob_derivative_lut",
    "ehp_bspline_derivative_lut", "ehp_pose_batch", "ehp_cgls_refined", "ehp_admml2", "ehp_batch_admml2",
    "ehp_geometry_set_pixel_coords", "ehp_curved_detector_coords", "ehp_adjoint_noalloc",
)
EHP_SOLVE_GRAPH = 1
EHP_ADMM_FINITE_DIFFERENCES = 2
EHP_DIFFERENTIAL = 2
EHP_BF16 = 2

_lib = None


class EhpError(RuntimeError):
    pass


def _hip_runtimes():
    """Paths of the HIP runtime libraries mapped into this process."""
    paths = set()
    try:
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64" in line:
                    paths.add(os.path.realpath(line.split()[-1]))
    except OSError:
        pass
    return paths


def load():
    """Load libelsahip.so and declare the C signatures (no GPU call happens here).

    torch (ROCm builds) ships its own HIP runtime. libelsahip.so must bind to that same
    runtime, or torch's streams and ours would belong to different runtimes and each side
    could even fail to see the GPU: torch is therefore imported first, so that the library's
    libamdhip64 dependency resolves to the copy already loaded. Two different HIP runtimes
    in the process afterwards are an error."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m elsa_amd.build`")
    try:
        import torch  # noqa: F401  (loads torch's HIP runtime first)
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    rts = _hip_runtimes()
    if len(rts) > 1:
        raise ImportError("two different HIP runtimes are loaded (" + ", ".join(sorted(rts)) +
                          "); libelsahip.so must share torch's: import torch before loading it")
    vp, i64, c_int, c_f, c_d = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_float, ctypes.c_double
    fp = ctypes.POINTER(ctypes.c_float)
    i64p = ctypes.POINTER(ctypes.c_int64)
    L.ehp_last_error.restype = ctypes.c_char_p
    L.ehp_last_error.argtypes = []
    L.ehp_version.restype = ctypes.c_char_p
    L.ehp_version.argtypes = []
    L.ehp_geometry_create.argtypes = [c_int, i64, i64p, i64p, fp, fp, fp, fp, fp, ctypes.POINTER(vp)]
    L.ehp_geometry_destroy.argtypes = [vp]
    L.ehp_geometry_set_pixel_coords.argtypes = [vp, fp]
    L.ehp_curved_detector_coords.argtypes = [c_int, i64p, c_f, c_f, fp]
    L.ehp_geometry_destroy.restype = None
    L.ehp_joseph_create.argtypes = [vp, c_int, c_int, ctypes.POINTER(vp)]
    L.ehp_siddon_create.argtypes = [vp, c_int, c_int, ctypes.POINTER(vp)]
    L.ehp_voxel_create.argtypes = [vp, c_int, vp, c_int, c_d, c_int, ctypes.POINTER(vp)]
    L.ehp_destroy.argtypes = [vp]
    L.ehp_destroy.restype = None
    L.ehp_forward.argtypes = [vp, vp, vp, vp]
    L.ehp_adjoint.argtypes = [vp, vp, vp, vp]
    L.ehp_adjoint_noalloc.argtypes = [vp, vp, vp, vp]
    L.ehp_op_sizes.argtypes = [vp, i64p, i64p]
    L.ehp_sample_count.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
    L.ehp_joseph_ray_info.argtypes = [vp, vp, vp]
    ip = ctypes.POINTER(ctypes.c_int)
    L.ehp_tile.argtypes = [vp, c_int, c_int, c_int, ip, ip, ip]
    L.ehp_pose_batch.argtypes = [vp, c_int, ip]
    L.ehp_geometry_pose.argtypes = [c_int, c_f, c_f, fp, fp, fp, fp, c_f, c_f, c_f, fp, fp, fp, fp, fp, fp, fp]
    L.ehp_circle_trajectory.argtypes = [c_int, i64, i64p, c_f, c_f, c_f, i64p, i64p, fp, fp, fp, fp, fp]
    L.ehp_blob_lut.argtypes = [c_int, c_d, c_d, i64, c_int, vp, ctypes.POINTER(c_d)]
    L.ehp_bspline_lut.argtypes = [c_int, i64, i64, c_int, vp, ctypes.POINTER(c_d)]
    L.ehp_blob_derivative_lut.argtypes = [c_int, c_d, c_d, i64, c_int, c_int, vp, ctypes.POINTER(c_d)]
    L.ehp_bspline_derivative_lut.argtypes = [c_int, i64, i64, c_int, c_int, vp, ctypes.POINTER(c_d)]
    dp = ctypes.POINTER(c_d)
    L.ehp_cgls.argtypes = [vp, vp, vp, c_int, c_int, c_d, c_d, c_int, dp, ip, vp]
    L.ehp_apgd.argtypes = [vp, vp, vp, c_int, c_int, c_d, c_d, c_d, c_int, dp, ip, dp, vp]
    L.ehp_power_iterations.argtypes = [vp, c_int, dp, vp]
    L.ehp_cgls_refined.argtypes = [vp, vp, vp, vp, c_int, c_int, c_int, c_int, dp, vp]
    L.ehp_admml2.argtypes = [vp, vp, vp, c_int, c_int, c_d, c_d, c_int, c_int, dp, dp, vp]
    L.ehp_batch_admml2.argtypes = [vp, vp, vp, c_int, c_int, c_d, c_d, c_int, c_int, dp, dp, vp]
    L.ehp_batch_create.argtypes = [ctypes.POINTER(vp), c_int, ctypes.POINTER(vp)]
    L.ehp_batch_destroy.argtypes = [vp]
    L.ehp_batch_destroy.restype = None
    L.ehp_batch_layout.argtypes = [vp, i64p, i64p, ip]
    L.ehp_batch_forward.argtypes = [vp, vp, vp, vp]
    L.ehp_batch_adjoint.argtypes = [vp, vp, vp, vp]
    L.ehp_batch_cgls.argtypes = [vp, vp, vp, c_int, c_int, c_d, c_d, c_int, dp, ip, vp]
    L.ehp_batch_apgd.argtypes = [vp, vp, vp, c_int, c_int, dp, c_d, c_d, c_int, dp, ip, dp, vp]
    L.ehp_batch_power_iterations.argtypes = [vp, c_int, dp, vp]
    for name in EXPORTS:
        if name not in ("ehp_last_error", "ehp_version", "ehp_geometry_destroy", "ehp_destroy", "ehp_batch_destroy"):
            getattr(L, name).restype = c_int
    _lib = L
    return L


def check(rc: int, what: str = "") -> None:
    """Raise the Python counterpart of elsa's exception for a failing status code."""
    if rc == EHP_OK:
        return
    msg = load().ehp_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == EHP_ERR_ARGUMENT:
        raise ValueError(text)
    if rc == EHP_ERR_ALLOC:
        raise MemoryError(text)
    if rc == EHP_ERR_UNSUPPORTED:
        raise NotImplementedError(text)
    raise EhpError(text)
