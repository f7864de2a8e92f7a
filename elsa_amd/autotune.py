"""Tile autotuner: times the operator's LDS-tiled kernel for a set of tile shapes (and, for
the tiled back projections, poses per step) on the GPU and keeps the fastest (ehp_tile and
ehp_pose_batch of the C-ABI are the knobs).

The tiled kernels sum in 64-bit fixed point, where each contribution is rounded on its own
and integer sums do not depend on their order, so every tile shape and pose batch gives
bit-identical results: the tuner changes speed only.
"""
from __future__ import annotations

import numpy as np

from .projectors import BlobVoxelProjectorCUDA, BSplineVoxelProjectorCUDA, JosephsMethodCUDA, SiddonsMethodCUDA

CANDIDATES = {
    ("joseph", 3): [(8, 17, 17), (8, 15, 15), (8, 13, 13), (6, 19, 19), (4, 23, 23), (8, 19, 19), (12, 15, 15)],
    ("joseph", 2): [(4, 15, 1), (8, 15, 1), (2, 15, 1), (8, 31, 1), (4, 7, 1), (16, 31, 1), (16, 15, 1)],
    ("siddon", 3): [(8, 16, 16), (4, 16, 16), (16, 16, 16), (8, 32, 16), (8, 8, 8), (12, 16, 16)],
    ("siddon", 2): [(4, 16, 1), (8, 16, 1), (2, 16, 1), (8, 32, 1), (4, 8, 1), (16, 16, 1)],
    ("voxel", 3): [(32, 32, 1), (64, 32, 1), (32, 16, 1), (128, 16, 1), (16, 16, 1)],
    ("voxel", 2): [(256, 1, 1), (128, 1, 1), (512, 1, 1), (64, 1, 1)],
}
POSE_BATCHES = (4, 8, 16, 32)


def _kind(op):
    if isinstance(op, JosephsMethodCUDA):
        return "joseph"
    if isinstance(op, SiddonsMethodCUDA):
        return "siddon"
    if isinstance(op, (BlobVoxelProjectorCUDA, BSplineVoxelProjectorCUDA)):
        return "voxel"
    raise TypeError(f"no tiled kernel for {type(op).__name__}")


def _timer(run, reps):
    import torch

    run()  # warm-up
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        run()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _sweep(ops, kind, cands, batches, run, reps, rebuild=None):
    """Time every tile (set on all ops), keep the best, then every pose batch; returns the best
    tile, the best batch and {key: ms}. rebuild(), when given, runs after each change of the
    settings and before its timing (for objects that capture them at creation)."""
    rebuild = rebuild or (lambda: None)
    times = {}
    original = [op.tile() for op in ops]
    for t in cands:
        try:
            for op in ops:
                op.tile(*t)
        except ValueError:  # does not fit the device's LDS
            continue
        rebuild()
        times[tuple(ops[0].tile())] = _timer(run, reps)
    if not times:
        for op, t in zip(ops, original):
            op.tile(*t)
        rebuild()
        return tuple(original[0]), None, times
    best = min(times, key=times.get)
    for op in ops:
        op.tile(*best)
    best_batch = None
    if kind != "voxel" and batches:
        bt = {}
        for b in batches:
            for op in ops:
                op.pose_batch(b)
            rebuild()
            bt[b] = _timer(run, reps)
            times[f"pose_batch={b}"] = bt[b]
        best_batch = min(bt, key=bt.get)
        for op in ops:
            op.pose_batch(best_batch)
    rebuild()
    return best, best_batch, times


def _inputs(op, kind, x, seed):
    import torch

    g = torch.Generator(device="cuda").manual_seed(seed)
    tdt = torch.float32 if op._dtype == np.float32 else torch.float64
    dom, rng = op.getDomainDescriptor().shape, op.getRangeDescriptor().shape
    shape_in, shape_out = (dom, rng) if kind == "voxel" else (rng, dom)
    inp = x if x is not None else torch.rand(shape_in, dtype=tdt, device="cuda", generator=g)
    return inp, torch.empty(shape_out, dtype=tdt, device="cuda")


def autotune(op, candidates=None, reps=2, seed=0, x=None, batches=POSE_BATCHES):
    """Time each candidate tile (LA, TB, TC) -- (width, height, -) for the voxel forward --
    reps times, then each pose batch of the tiled back projection; set the fastest on op and
    return (best_tile, {tile or "pose_batch=n": ms}).

    x: the input of the tiled direction to time on (a volume for the voxel projectors, a
    sinogram otherwise), as a GPU tensor; default uniform random data. The voxel forward skips
    zero voxels, so a representative volume tunes it better than random data."""
    kind = _kind(op)
    dim = op.getDomainDescriptor().getNumberOfDimensions()
    cands = list(candidates) if candidates is not None else CANDIDATES[(kind, dim)]
    inp, out = _inputs(op, kind, x, seed)
    run = (lambda: op.apply(inp, out)) if kind == "voxel" else (lambda: op.applyAdjoint(inp, out))
    best, _, times = _sweep([op], kind, cands, batches, run, reps)
    return best, times


def autotune_batch(ops, candidates=None, reps=2, seed=0, batches=POSE_BATCHES):
    """The same search for the operators of a batch as a whole: one tile and one pose batch for
    all of them, timed on the fused launches of a ProjectorBatch. Returns (ProjectorBatch built
    with the best settings, best_tile, best_pose_batch, {key: ms})."""
    import torch

    from .batch import ProjectorBatch

    kinds = {_kind(op) for op in ops}
    if len(kinds) != 1:
        raise ValueError("autotune_batch needs operators of one kind")
    kind = kinds.pop()
    dim = ops[0].getDomainDescriptor().getNumberOfDimensions()
    cands = list(candidates) if candidates is not None else CANDIDATES[(kind, dim)]
    state = {"B": ProjectorBatch(ops)}
    B0 = state["B"]
    tdt = torch.float32 if B0._dtype == np.float32 else torch.float64
    g = torch.Generator(device="cuda").manual_seed(seed)
    n_in, n_out = (B0._n_vol, B0._n_sino) if kind == "voxel" else (B0._n_sino, B0._n_vol)
    inp = torch.rand(n_in, dtype=tdt, device="cuda", generator=g)
    out = torch.empty(n_out, dtype=tdt, device="cuda")

    def rebuild():  # a batch takes the tiles and pose batches of its operators at creation
        state["B"] = ProjectorBatch(ops)

    def run():
        B = state["B"]
        (B.apply if kind == "voxel" else B.applyAdjoint)(inp, out)

    best, best_batch, times = _sweep(ops, kind, cands, batches, run, reps, rebuild)
    return state["B"], best, best_batch, times
