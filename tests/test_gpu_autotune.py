"""Tile autotuner and tile independence of the tiled kernels' results."""
import numpy as np
import pytest

import oracle
from backends import geom_to_descriptors

pytestmark = pytest.mark.gpu


def make(dim):
    if dim == 2:
        g = oracle.circle_geometry(2, (64, 48), 30, 360, 6400.0, 320.0, det=(90,))
    else:
        g = oracle.circle_geometry(3, (40, 32, 36), 24, 360, 4000.0, 200.0, det=(60, 48))
    return g, geom_to_descriptors(g)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("cls", ["JosephsMethodCUDA", "SiddonsMethodCUDA", "BlobVoxelProjectorCUDA"])
def test_tiles_give_bitwise_identical_results_and_autotune_keeps_them(cls, dim):
    """Integer sums do not depend on their order, so every tile shape gives the same bits;
    the autotuner picks one of its candidates and leaves the results unchanged."""
    import elsa_amd as E

    g, (vd, dd) = make(dim)
    op = getattr(E, cls)(vd, dd)
    rng = np.random.default_rng(1)
    x = rng.random(g.vol_shape, dtype=np.float32)
    y = rng.random(g.sino_shape, dtype=np.float32)
    ref_f, ref_b = op.apply(x), op.applyAdjoint(y)
    from elsa_amd.autotune import CANDIDATES

    cands = CANDIDATES[("voxel" if "Voxel" in cls else cls[:6].lower(), dim)]
    for t in cands[:3]:
        op.tile(*t)
        np.testing.assert_array_equal(op.apply(x), ref_f)
        np.testing.assert_array_equal(op.applyAdjoint(y), ref_b)
    if "Voxel" not in cls:
        for b in (1, 5, 32):
            assert op.pose_batch(b) == b
            np.testing.assert_array_equal(op.applyAdjoint(y), ref_b)
        with pytest.raises(ValueError):
            op.pose_batch(33)
        op.pose_batch(8)
    best, times = E.autotune(op, reps=1)
    assert best in times and len(times) >= 2
    assert tuple(op.tile()) == best
    np.testing.assert_array_equal(op.apply(x), ref_f)
    np.testing.assert_array_equal(op.applyAdjoint(y), ref_b)


def test_autotune_batch_keeps_results():
    """One tile and pose batch for all operators of a batch, tuned on the fused launch."""
    import torch

    import elsa_amd as E

    geoms = [make(2)[0], oracle.circle_geometry(2, (40, 40), 17, 360, 4000.0, 200.0)]
    ops = [E.JosephsMethodCUDA(*geom_to_descriptors(g)) for g in geoms]
    rng = np.random.default_rng(2)
    ys = [rng.random(g.sino_shape, dtype=np.float32) for g in geoms]
    refs = [op.applyAdjoint(y) for op, y in zip(ops, ys)]
    B, tile, pb, times = E.autotune_batch(ops, reps=1)
    assert B.fused and tile in times and f"pose_batch={pb}" in times
    assert all(tuple(op.tile()) == tile and op.pose_batch() == pb for op in ops)
    got = B.unpack(B.applyAdjoint(B.pack(ys, adjoint_space=True)))
    torch.cuda.synchronize()
    for g_, r in zip(got, refs):
        np.testing.assert_array_equal(g_.cpu().numpy(), r)
