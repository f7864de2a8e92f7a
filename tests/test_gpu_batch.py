"""Batches of independent problems (ProjectorBatch): every problem's forward, back
projection and solver run equal what its own operator gives, bit for bit, whether the batch
fuses the passes into single launches (FP32 Joseph, FP32 Siddon, FP32 voxel projectors) or
calls the operators in turn."""
import numpy as np
import pytest

import oracle
from backends import geom_to_descriptors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E():
    import elsa_amd

    return elsa_amd


def problems(dim, n, seed):
    """Uneven problems: sizes, view counts and start angles drawn at random; source and
    detector distances 100 N and 5 N, detector floor(N sqrt 2)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        if dim == 2:
            N = int(rng.choice([24, 32, 48, 64, 96]))
            vol = (N, int(rng.choice([N, N // 2 + 8])))
        else:
            N = int(rng.choice([12, 16, 24]))
            vol = (N, int(rng.choice([N, N + 4])), int(rng.choice([N, N - 4])))
        views = int(rng.integers(8, 40))
        start = float(rng.uniform(0, 360))
        angles = (start + np.linspace(0, 360, views)).astype(np.float32)
        out.append(oracle.circle_geometry(dim, vol, views, 360, 100.0 * N, 5.0 * N, angles_deg=angles))
    return out


def make_ops(E, geoms, cls="JosephsMethodCUDA", dtype=np.float32):
    ops = []
    for g in geoms:
        vd, dd = geom_to_descriptors(g)
        ops.append(getattr(E, cls)(vd, dd, dtype=dtype))
    return ops


def data(geoms, seed, dtype=np.float32):
    rng = np.random.default_rng(seed)
    xs = [rng.random(g.vol_shape).astype(dtype) for g in geoms]
    ys = [rng.random(g.sino_shape).astype(dtype) for g in geoms]
    return xs, ys


def check_passes(E, ops, xs, ys, fused, adj_rtol=0.0):
    """adj_rtol: 0 (bit for bit) unless the operators' back projections sum with atomics in no
    fixed order (FP64 and the simple kernels), then a relative L2 bound."""
    import torch

    B = E.ProjectorBatch(ops)
    assert B.fused == fused
    assert B.vol_offsets[-1] == sum(x.size for x in xs) and B.sino_offsets[-1] == sum(y.size for y in ys)
    fx = B.apply(B.pack(xs))
    by = B.applyAdjoint(B.pack(ys, adjoint_space=True))
    torch.cuda.synchronize()
    for j, op in enumerate(ops):
        np.testing.assert_array_equal(B.unpack(fx, adjoint_space=True)[j].cpu().numpy(), op.apply(xs[j]))
        got, ref = B.unpack(by)[j].cpu().numpy(), op.applyAdjoint(ys[j])
        if adj_rtol == 0.0:
            np.testing.assert_array_equal(got, ref)
        else:
            assert np.linalg.norm(got - ref) <= adj_rtol * np.linalg.norm(ref)
    return B


@pytest.mark.parametrize("dim,n", [(2, 12), (3, 5)])
def test_fused_joseph_batch_equals_each_operator(E, dim, n):
    geoms = problems(dim, n, seed=dim)
    xs, ys = data(geoms, 7)
    check_passes(E, make_ops(E, geoms), xs, ys, fused=True)


@pytest.mark.parametrize("dim,n", [(2, 12), (3, 5)])
def test_fused_siddon_batch_equals_each_operator(E, dim, n):
    geoms = problems(dim, n, seed=10 + dim)
    xs, ys = data(geoms, 8)
    check_passes(E, make_ops(E, geoms, "SiddonsMethodCUDA"), xs, ys, fused=True)


@pytest.mark.parametrize("dim,n", [(2, 9), (3, 4)])
def test_fused_voxel_batch_equals_each_operator(E, dim, n):
    """Blob, B-spline and phase-contrast blob operators in one batch: the tiled forward and the
    gather run as one launch each over all problems (the forward's fixed-point scale per
    problem, from its own volume), bit for bit each operator's own result."""
    geoms = problems(dim, n, seed=20 + dim)
    kinds = ["BlobVoxelProjectorCUDA", "BSplineVoxelProjectorCUDA", "PhaseContrastBlobVoxelProjectorCUDA"]
    ops = [getattr(E, kinds[j % 3])(*geom_to_descriptors(g)) for j, g in enumerate(geoms)]
    xs, ys = data(geoms, 9)
    xs[1][:] = 0.0
    check_passes(E, ops, xs, ys, fused=True)


def test_fused_siddon_batch_zero_nonfinite_and_cgls(E):
    """Per-problem fixed-point scales in a fused Siddon batch (a zero and a NaN sinogram stay in
    their own problems), and a batched CGLS equal to each problem's own, bit for bit."""
    import torch

    geoms = problems(2, 5, seed=13)
    ops = make_ops(E, geoms, "SiddonsMethodCUDA")
    xs, ys = data(geoms, 4)
    ys[1][:] = 0.0
    ys[2][0, 0] = np.nan
    B = E.ProjectorBatch(ops)
    assert B.fused
    by = B.unpack(B.applyAdjoint(B.pack(ys, adjoint_space=True)))
    assert np.all(by[1].cpu().numpy() == 0.0)
    assert np.isnan(by[2].cpu().numpy()).all()
    for j in (0, 3, 4):
        np.testing.assert_array_equal(by[j].cpu().numpy(), ops[j].applyAdjoint(ys[j]))
    bs = [op.apply(x) for op, x in zip(ops, phantoms(geoms, 5))]
    cg = E.CGLS(B, B.pack(bs, adjoint_space=True), graph=True)
    x = B.unpack(cg.solve(4))
    torch.cuda.synchronize()
    for j, op in enumerate(ops):
        np.testing.assert_array_equal(x[j].cpu().numpy(), E.CGLS(op, bs[j], graph=True).solve(4))


def test_fused_batch_with_zero_and_nonfinite_problems(E):
    """A zero sinogram and a NaN in another problem's sinogram stay in their own problems: the
    fixed-point scale is chosen per problem."""
    geoms = problems(2, 4, seed=11)
    xs, ys = data(geoms, 3)
    ys[1][:] = 0.0
    ys[2][0, 0] = np.nan
    ops = make_ops(E, geoms)
    B = E.ProjectorBatch(ops)
    assert B.fused
    by = B.unpack(B.applyAdjoint(B.pack(ys, adjoint_space=True)))
    assert np.all(by[1].cpu().numpy() == 0.0)
    assert np.isnan(by[2].cpu().numpy()).all()
    for j in (0, 3):
        np.testing.assert_array_equal(by[j].cpu().numpy(), ops[j].applyAdjoint(ys[j]))


@pytest.mark.parametrize("kind", ["mixed_kinds", "mixed_dims", "fp64", "simple"])
def test_unfused_batches_equal_each_operator(E, kind):
    if kind == "fp64":
        geoms = problems(2, 4, seed=5)
        ops = make_ops(E, geoms, dtype=np.float64)
        xs, ys = data(geoms, 1, np.float64)
    elif kind == "mixed_dims":
        geoms = problems(2, 2, seed=6) + problems(3, 2, seed=6)
        ops = make_ops(E, geoms)
        xs, ys = data(geoms, 1)
    elif kind == "simple":
        geoms = problems(2, 3, seed=8)
        ops = [E.JosephsMethodCUDA(*geom_to_descriptors(g), kernels="simple") for g in geoms]
        xs, ys = data(geoms, 1)
    else:
        geoms = problems(2, 4, seed=9)
        ops = make_ops(E, geoms[:2]) + make_ops(E, geoms[2:3], "SiddonsMethodCUDA") + \
            make_ops(E, geoms[3:], "BlobVoxelProjectorCUDA")
        xs, ys = data(geoms, 1)
    rtol = {"fp64": 1e-13, "simple": 1e-6}.get(kind, 0.0)
    check_passes(E, ops, xs, ys, fused=False, adj_rtol=rtol)


@pytest.mark.parametrize("cls", ["JosephsMethodCUDA", "SiddonsMethodCUDA", "BlobVoxelProjectorCUDA",
                                 "PhaseContrastBSplineVoxelProjectorCUDA"])
def test_clone_equals_the_original(E, cls):
    """clone() (elsa's LinearOperator::clone) makes a new operator with its own device state and
    the same results bit for bit; its tiled kernels take the original's tile."""
    g = problems(3, 1, seed=41)[0]
    op = getattr(E, cls)(*geom_to_descriptors(g))
    c = op.clone()
    assert c.handle.value != op.handle.value and type(c) is type(op)
    xs, ys = data([g], 2)
    np.testing.assert_array_equal(c.apply(xs[0]), op.apply(xs[0]))
    np.testing.assert_array_equal(c.applyAdjoint(ys[0]), op.applyAdjoint(ys[0]))
    assert c.tile() == op.tile()


def test_batch_rejects_bad_operator_lists(E):
    geoms = problems(2, 2, seed=1)
    a = make_ops(E, geoms[:1])[0]
    b = make_ops(E, geoms[1:], dtype=np.float64)[0]
    with pytest.raises(ValueError):
        E.ProjectorBatch([])
    with pytest.raises(ValueError):
        E.ProjectorBatch([a, b])
    with pytest.raises(ValueError):
        E.ProjectorBatch([a, a])


def phantoms(geoms, seed):
    rng = np.random.default_rng(seed)
    out = []
    for g in geoms:
        grids = np.meshgrid(*[np.linspace(-1, 1, s) for s in g.vol_shape], indexing="ij")
        r2 = sum(q * q for q in grids)
        out.append(((r2 < rng.uniform(0.2, 0.6)) * 1.0 + 0.1 * rng.random(g.vol_shape)).astype(np.float32))
    return out


@pytest.mark.parametrize("graph", [False, True])
def test_batch_cgls_equals_each_problem_solved_alone(E, graph):
    """Each problem keeps its own step lengths and its own stopping test: problem 1 has b = 0
    and stops before its first iteration while the others go on."""
    import torch

    geoms = problems(2, 6, seed=21)
    ops = make_ops(E, geoms)
    bs = [op.apply(x) for op, x in zip(ops, phantoms(geoms, 2))]
    bs[1][:] = 0.0
    B = E.ProjectorBatch(ops)
    assert B.fused
    it = 6
    cg = E.CGLS(B, B.pack(bs, adjoint_space=True), graph=graph)
    x = cg.solve(it)
    torch.cuda.synchronize()
    xs = B.unpack(x)
    assert cg.iterations_done[1] == 0 and np.all(xs[1].cpu().numpy() == 0.0)
    assert all(cg.iterations_done[j] == it for j in range(len(ops)) if j != 1)
    for j, op in enumerate(ops):
        single = E.CGLS(op, bs[j], graph=graph)
        xj = single.solve(it)
        np.testing.assert_array_equal(xs[j].cpu().numpy(), xj)
        np.testing.assert_array_equal(cg.residuals[j], single.residuals)
        assert cg.iterations_done[j] == single.iterations_done


def test_batch_apgd_and_power_iterations_equal_each_problem(E):
    import torch

    geoms = problems(2, 5, seed=31)
    ops = make_ops(E, geoms)
    bs = [op.apply(x) for op, x in zip(ops, phantoms(geoms, 4))]
    B = E.ProjectorBatch(ops)
    L = E.powerIterations(B)
    for j, op in enumerate(ops):
        assert L[j] == E.powerIterations(op)
    it = 5
    ap = E.APGD(B, B.pack(bs, adjoint_space=True), prox=E.ProximalL1(1e-3), graph=True)
    x = ap.solve(it)
    torch.cuda.synchronize()
    xs = B.unpack(x)
    for j, op in enumerate(ops):
        single = E.APGD(op, bs[j], prox=E.ProximalL1(1e-3), graph=True)
        xj = single.solve(it)
        assert ap.mu_used[j] == single.mu_used
        np.testing.assert_array_equal(xs[j].cpu().numpy(), xj)
        np.testing.assert_array_equal(ap.residuals[j], single.residuals)
