"""Tile autotuner: times the operator's LDS-tiled kernel for a set of tile shapes on the
GPU and keeps the fastest (ehp_tile of the C-ABI is the knob).

The tiled kernels sum in 64-bit fixed point, where each contribution is rounded on its own
and integer sums do not depend on their order, so every tile shape gives bit-identical
results: the tuner changes speed only.
"""
from __future__ import annotations

import numpy as np

from .projectors import BlobVoxelProjectorCUDA, BSplineVoxelProjectorCUDA, JosephsMethodCUDA, SiddonsMethodCUDA

CANDIDATES = {
    ("joseph", 3): [(8, 15, 15), (4, 15, 15), (12, 15, 15), (16, 15, 15), (8, 23, 23), (8, 31, 15), (8, 11, 11)],
    ("joseph", 2): [(4, 15, 1), (8, 15, 1), (2, 15, 1), (8, 31, 1), (4, 7, 1), (16, 31, 1)],
    ("siddon", 3): [(8, 16, 16), (4, 16, 16), (16, 16, 16), (8, 32, 16), (8, 8, 8), (12, 16, 16)],
    ("siddon", 2): [(4, 16, 1), (8, 16, 1), (2, 16, 1), (8, 32, 1), (4, 8, 1)],
    ("voxel", 3): [(32, 32, 1), (64, 32, 1), (32, 16, 1), (128, 16, 1), (16, 16, 1)],
    ("voxel", 2): [(256, 1, 1), (128, 1, 1), (512, 1, 1), (64, 1, 1)],
}


def _kind(op):
    if isinstance(op, JosephsMethodCUDA):
        return "joseph"
    if isinstance(op, SiddonsMethodCUDA):
        return "siddon"
    if isinstance(op, (BlobVoxelProjectorCUDA, BSplineVoxelProjectorCUDA)):
        return "voxel"
    raise TypeError(f"no tiled kernel for {type(op).__name__}")


def autotune(op, candidates=None, reps=2, seed=0, x=None):
    """Time each candidate tile (LA, TB, TC) -- (width, height, -) for the voxel forward --
    reps times, set the fastest on op and return (best, {tile: ms}).

    x: the input of the tiled direction to time on (a volume for the voxel projectors, a
    sinogram otherwise), as a GPU tensor; default uniform random data. The voxel forward skips
    zero voxels, so a representative volume tunes it better than random data."""
    import torch

    kind = _kind(op)
    dim = op.getDomainDescriptor().getNumberOfDimensions()
    cands = list(candidates) if candidates is not None else CANDIDATES[(kind, dim)]
    g = torch.Generator(device="cuda").manual_seed(seed)
    tdt = torch.float32 if op._dtype == np.float32 else torch.float64
    if kind == "voxel":  # the forward is tiled
        inp = x if x is not None else torch.rand(op.getDomainDescriptor().shape, dtype=tdt, device="cuda",
                                                  generator=g)
        out = torch.empty(op.getRangeDescriptor().shape, dtype=tdt, device="cuda")
        run = lambda: op.apply(inp, out)
    else:  # the back projection is tiled
        inp = x if x is not None else torch.rand(op.getRangeDescriptor().shape, dtype=tdt, device="cuda",
                                                  generator=g)
        out = torch.empty(op.getDomainDescriptor().shape, dtype=tdt, device="cuda")
        run = lambda: op.applyAdjoint(inp, out)
    original = op.tile()
    times = {}
    for t in cands:
        try:
            op.tile(*t)
        except ValueError:  # does not fit the device's LDS
            continue
        run()  # warm-up
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            run()
        e1.record()
        torch.cuda.synchronize()
        times[tuple(op.tile())] = e0.elapsed_time(e1) / reps
    if not times:
        op.tile(*original)
        return tuple(original), times
    best = min(times, key=times.get)
    op.tile(*best)
    return best, times
