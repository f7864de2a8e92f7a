"""GPU parity of the device-resident solvers (CGLS, APGD, power iterations) against the
oracle's restatements of elsa's CGLS.cpp, APGD.cpp and PowerIterations.cpp.

The FP64 operator computes with the same FP32 geometry and weights as the oracle's
acc64 mode (FP32 geometry, FP64 sums), so FP64 solves are compared tightly; FP32 solves
are compared with a tolerance and, between graph and eager runs, bit for bit."""
import numpy as np
import pytest

import oracle
from backends import geom_to_descriptors

pytestmark = pytest.mark.gpu

CASES = {
    "fan2d_64": dict(dim=2, vol=(64, 64), views=60, det=(92,), s2c=6400.0, c2d=320.0),
    "cone3d_24": dict(dim=3, vol=(24, 24, 24), views=30, det=(34, 34), s2c=2400.0, c2d=120.0),
}


def make_case(name):
    c = CASES[name]
    return oracle.circle_geometry(c["dim"], c["vol"], c["views"], 360, c["s2c"], c["c2d"], det=c["det"])


def phantom(shape, seed):
    rng = np.random.default_rng(seed)
    grids = np.meshgrid(*[np.linspace(-1, 1, s) for s in shape], indexing="ij")
    r2 = sum(g * g for g in grids)
    return ((r2 < 0.5) * 1.0 + 0.2 * rng.random(shape)).astype(np.float32)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def oracle_ops(g):
    """FP32 geometry and weights, FP64 data and sums, the GPU's position rule."""
    A = lambda x: oracle.joseph(g, x, mode=oracle.MODE_DIRECT, acc64=True)
    At = lambda y: oracle.joseph(g, y, adjoint=True, mode=oracle.MODE_DIRECT, acc64=True)
    return A, At


@pytest.fixture(scope="module")
def E():
    import elsa_amd

    return elsa_amd


def gpu_op(E, g, dtype):
    vd, dd = geom_to_descriptors(g)
    return E.JosephsMethodCUDA(vd, dd, dtype=dtype)


@pytest.mark.parametrize("name", list(CASES))
def test_cgls_fp64_matches_oracle(E, name):
    g = make_case(name)
    A, At = oracle_ops(g)
    b = A(phantom(g.vol_shape, 1).astype(np.float64))
    x_ref, res_ref, n_ref = oracle.cgls(A, At, b, g.vol_shape, 8, tol=0.0)
    op = gpu_op(E, g, np.float64)
    solver = E.CGLS(op, b, tol=0.0)
    x = solver.solve(8)
    assert solver.iterations_done == n_ref == 8
    assert rel(x, x_ref) <= 1e-10, rel(x, x_ref)
    np.testing.assert_allclose(solver.residuals, res_ref, rtol=1e-10)


@pytest.mark.parametrize("name", list(CASES))
def test_cgls_fp32_close_to_exact_and_monotone(E, name):
    """FP32 vectors with FP64 scalars: within 1e-4 of the FP64 iteration on the same FP32
    weights after 4 iterations, and ||b - Ax|| never grows over 8. (Later iterates of the
    2D case are too sensitive for a tight comparison: on it, the oracle's own FP64 CGLS
    driven by the FP32 oracle operator, 1.7e-7 away from the exact one, drifts from the exact
    iteration by 1e-6, 6e-5 and 4e-3 after 4, 6 and 8 iterations.)"""
    g = make_case(name)
    A, At = oracle_ops(g)
    b32 = A(phantom(g.vol_shape, 2).astype(np.float64)).astype(np.float32)
    x_ref, res_ref, _ = oracle.cgls(A, At, b32.astype(np.float64), g.vol_shape, 4, tol=0.0)
    op = gpu_op(E, g, np.float32)
    solver = E.CGLS(op, b32, tol=0.0)
    x = solver.solve(4)
    assert rel(x, x_ref) <= 1e-4, rel(x, x_ref)
    np.testing.assert_allclose(solver.residuals, res_ref, rtol=1e-4)
    solver.solve(8)
    assert np.all(np.diff(solver.residuals) <= 1e-6 * solver.residuals[0])


def test_cgls_graph_equals_eager_bitwise(E):
    """One iteration captured into a hipGraph and replayed gives the same bits as the eager
    run (every kernel on the path is deterministic), also after eager work in between."""
    import torch

    g = make_case("cone3d_24")
    op = gpu_op(E, g, np.float32)
    b = torch.from_numpy(op.apply(phantom(g.vol_shape, 3))).cuda()
    eager = E.CGLS(op, b, tol=0.0)
    x_e = eager.solve(6).cpu().numpy()
    graph = E.CGLS(op, b, tol=0.0, graph=True)
    x_g1 = graph.solve(6).cpu().numpy()
    res_g1 = graph.residuals.copy()
    _ = op.applyAdjoint(op.apply(x_e))  # eager work between the replays
    torch.cuda.synchronize()
    x_g2 = graph.solve(6).cpu().numpy()
    np.testing.assert_array_equal(x_g1, x_e)
    np.testing.assert_array_equal(x_g2, x_e)
    np.testing.assert_array_equal(res_g1, eager.residuals)


def test_cgls_tolerance_stops_where_elsa_stops(E):
    g = make_case("fan2d_64")
    A, At = oracle_ops(g)
    b = A(phantom(g.vol_shape, 4).astype(np.float64))
    # ||A^T s||^2 before each of the first iterations of the exact iteration
    ks = []
    x, s = np.zeros(g.vol_shape), b.copy()
    r = At(s)
    c = r.copy()
    kold = float(np.sum(r * r))
    for _ in range(4):
        ks.append(kold)
        q = A(c)
        alpha = kold / float(np.sum(q * q))
        x, s = x + alpha * c, s - alpha * q
        r = At(s)
        k = float(np.sum(r * r))
        c = r + (k / kold) * c
        kold = k
    tol = 0.5 * (ks[3] + kold)  # a threshold between two of the values the iteration meets
    x_ref, _, n_ref = oracle.cgls(A, At, b, g.vol_shape, 10, tol=tol)
    assert 1 <= n_ref < 10
    op = gpu_op(E, g, np.float64)
    for graph in (False, True):
        solver = E.CGLS(op, b, tol=tol, graph=graph)
        xg = solver.solve(10)
        assert solver.iterations_done == n_ref
        assert rel(xg, x_ref) <= 1e-10
        assert np.all(solver.residuals[n_ref + 1:] == -1.0)


def test_cgls_start_value_and_damping(E):
    g = make_case("fan2d_64")
    A, At = oracle_ops(g)
    b = A(phantom(g.vol_shape, 5).astype(np.float64))
    x0 = phantom(g.vol_shape, 6).astype(np.float64) * 0.5
    x_ref, res_ref, _ = oracle.cgls(A, At, b, g.vol_shape, 6, eps=3.0, tol=0.0, x0=x0)
    op = gpu_op(E, g, np.float64)
    solver = E.CGLS(op, b, eps=3.0, tol=0.0)
    x = solver.solve(6, x0=x0)
    assert rel(x, x_ref) <= 1e-10, rel(x, x_ref)
    np.testing.assert_allclose(solver.residuals, res_ref, rtol=1e-10)


@pytest.mark.parametrize("name", list(CASES))
def test_power_iterations_match_oracle(E, name):
    g = make_case(name)
    A, At = oracle_ops(g)
    L_ref = oracle.power_iterations(A, At, g.vol_shape, 5)
    L = E.powerIterations(gpu_op(E, g, np.float64), 5)
    assert abs(L - L_ref) <= 1e-11 * L_ref, (L, L_ref)
    L32 = E.powerIterations(gpu_op(E, g, np.float32), 5)
    assert abs(L32 - L_ref) <= 1e-5 * L_ref, (L32, L_ref)


@pytest.mark.parametrize("sigma", [None, 0.05])
def test_apgd_fp64_matches_oracle(E, sigma):
    g = make_case("fan2d_64")
    A, At = oracle_ops(g)
    b = A(phantom(g.vol_shape, 7).astype(np.float64))
    x_ref, gn_ref, n_ref, mu_ref = oracle.apgd(A, At, b, g.vol_shape, 12, sigma=sigma)
    op = gpu_op(E, g, np.float64)
    prox = E.ProximalIdentity() if sigma is None else E.ProximalL1(sigma)
    solver = E.APGD(op, b, prox)
    x = solver.solve(12)
    assert abs(solver.mu_used - mu_ref) <= 1e-11 * mu_ref
    assert solver.iterations_done == n_ref == 12
    assert rel(x, x_ref) <= 1e-10, rel(x, x_ref)
    np.testing.assert_allclose(solver.residuals[:n_ref], gn_ref, rtol=1e-10)


def test_apgd_fp32_graph_equals_eager_and_close_to_exact(E):
    import torch

    g = make_case("cone3d_24")
    A, At = oracle_ops(g)
    op = gpu_op(E, g, np.float32)
    b = op.apply(phantom(g.vol_shape, 8))
    x_ref, _, _, mu_ref = oracle.apgd(A, At, b.astype(np.float64), g.vol_shape, 10, sigma=0.01)
    bt = torch.from_numpy(b).cuda()
    xe = E.APGD(op, bt, E.ProximalL1(0.01), mu=mu_ref).solve(10).cpu().numpy()
    xg = E.APGD(op, bt, E.ProximalL1(0.01), mu=mu_ref, graph=True).solve(10).cpu().numpy()
    np.testing.assert_array_equal(xg, xe)
    assert rel(xe, x_ref) <= 1e-4, rel(xe, x_ref)


def test_solver_argument_errors(E):
    g = make_case("fan2d_64")
    op = gpu_op(E, g, np.float32)
    with pytest.raises(ValueError):
        E.CGLS(op, np.zeros(7, np.float32)).solve(3)
    with pytest.raises(ValueError):
        E.powerIterations(op, 0)


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("sigma", [None, 0.05])
def test_admml2_fp64_matches_oracle(E, name, sigma):
    """ADMML2 (identity regulariser, CGNE x-updates) against the oracle's restatement of
    ADMML2.cpp on the same FP32 weights in FP64, with and without the L1 prox; the graph replay
    equals the recorded eager run bit for bit (the FP64 back projection sums in fixed point, in
    no order-dependent way)."""
    g = make_case(name)
    A, At = oracle_ops(g)
    b = A(phantom(g.vol_shape, 7).astype(np.float64))
    x_ref, res_ref = oracle.admml2(A, At, b, g.vol_shape, 4, tau=0.9, sigma=sigma, ninner=3)
    op = gpu_op(E, g, np.float64)
    prox = E.ProximalL1(sigma) if sigma is not None else None
    s = E.ADMML2(op, b, prox=prox, ninneriters=3, record=True)
    x = s.solve(4)
    assert s.tau_used == 0.9
    assert rel(x, x_ref) <= 1e-10, rel(x, x_ref)
    np.testing.assert_allclose(s.residuals, res_ref, rtol=1e-10)
    x_g = E.ADMML2(op, b, prox=prox, ninneriters=3, graph=True).solve(4)
    np.testing.assert_array_equal(np.asarray(x_g), np.asarray(x))


def test_batch_admml2_equals_each_problem(E):
    import torch

    geoms = [make_case("fan2d_64"), oracle.circle_geometry(2, (48, 40), 33, 360, 4800.0, 240.0)]
    ops = [gpu_op(E, g, np.float32) for g in geoms]
    bs = [op.apply(phantom(g.vol_shape, 9)) for op, g in zip(ops, geoms)]
    B = E.ProjectorBatch(ops)
    s = E.ADMML2(B, B.pack(bs, adjoint_space=True), prox=E.ProximalL1(0.01), ninneriters=4, graph=True)
    xs = B.unpack(s.solve(3))
    torch.cuda.synchronize()
    for j, op in enumerate(ops):
        xj = E.ADMML2(op, bs[j], prox=E.ProximalL1(0.01), ninneriters=4, graph=True).solve(3)
        np.testing.assert_array_equal(xs[j].cpu().numpy(), xj)


def fd_ops(g):
    dims = tuple(g.vol)
    R = lambda v: oracle.finite_differences(v, dims)
    Rt = lambda y: oracle.finite_differences(y, dims, adjoint=True)
    return R, Rt


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("sigma", [None, 0.02])
def test_admml2_finite_differences_fp64_matches_oracle(E, name, sigma):
    """ADMML2 with forward finite differences as regularisation operator (elsa's
    FiniteDifferences(domain), TV with the L1 prox) against the oracle in FP64: elsa's default
    tau = 0.9 / powerIterations(R^T R) from 5 device power iterations, the iterates and the
    residual history; the graph replay equals the eager run bit for bit (the FP64 back
    projection's fixed-point sums do not depend on the order of the additions)."""
    g = make_case(name)
    A, At = oracle_ops(g)
    R, Rt = fd_ops(g)
    b = A(phantom(g.vol_shape, 11).astype(np.float64))
    L = oracle.power_iterations(lambda v: R(v.reshape(-1)), lambda y: Rt(y).reshape(g.vol_shape), g.vol_shape)
    tau = 0.9 * (1.0 / L)
    op = gpu_op(E, g, np.float64)
    prox = E.ProximalL1(sigma) if sigma is not None else None
    s = E.ADMML2(op, b, prox=prox, ninneriters=3, record=True, regularizer="finite_differences")
    x = s.solve(4)
    assert abs(s.tau_used - tau) <= 1e-13 * tau, (s.tau_used, tau)
    x_ref, res_ref = oracle.admml2(A, At, b, g.vol_shape, 4, tau=s.tau_used, sigma=sigma, ninner=3, reg=(R, Rt))
    assert rel(x, x_ref) <= 1e-10, rel(x, x_ref)
    np.testing.assert_allclose(s.residuals, res_ref, rtol=1e-10)
    x_g = E.ADMML2(op, b, prox=prox, ninneriters=3, graph=True, regularizer="finite_differences").solve(4)
    np.testing.assert_array_equal(np.asarray(x_g), np.asarray(x))


def test_admml2_finite_differences_fp32(E):
    """FP32 TV-ADMML2: graph replay equals the eager run bit for bit, and the result is close
    to the FP64 oracle's."""
    g = make_case("fan2d_64")
    A, At = oracle_ops(g)
    R, Rt = fd_ops(g)
    b = A(phantom(g.vol_shape, 12).astype(np.float64)).astype(np.float32)
    op = gpu_op(E, g, np.float32)
    kw = dict(prox=E.ProximalL1(0.02), ninneriters=4, tau=0.1, regularizer="finite_differences")
    x_e = E.ADMML2(op, b, **kw).solve(5)
    x_g = E.ADMML2(op, b, graph=True, **kw).solve(5)
    np.testing.assert_array_equal(x_g, x_e)
    x_ref, _ = oracle.admml2(A, At, b.astype(np.float64), g.vol_shape, 5, tau=0.1, sigma=0.02, ninner=4,
                             reg=(R, Rt))
    assert rel(x_e, x_ref) <= 1e-4, rel(x_e, x_ref)


def test_batch_admml2_finite_differences_equals_each_problem(E):
    """A batch of volumes of different shapes (2D and 3D) goes through one finite-difference
    launch per application; each problem's result equals its own solve. Without tau a batch
    is refused (elsa's default comes from one operator's power iterations)."""
    import torch

    geoms = [make_case("fan2d_64"), oracle.circle_geometry(2, (48, 40), 33, 360, 4800.0, 240.0),
             make_case("cone3d_24")]
    ops = [gpu_op(E, g, np.float32) for g in geoms]
    bs = [op.apply(phantom(g.vol_shape, 13)) for op, g in zip(ops, geoms)]
    B = E.ProjectorBatch(ops)
    kw = dict(prox=E.ProximalL1(0.01), ninneriters=3, tau=0.2, regularizer="finite_differences")
    xs = B.unpack(E.ADMML2(B, B.pack(bs, adjoint_space=True), graph=True, **kw).solve(3))
    torch.cuda.synchronize()
    for j, op in enumerate(ops):
        xj = E.ADMML2(op, bs[j], graph=True, **kw).solve(3)
        np.testing.assert_array_equal(xs[j].cpu().numpy(), xj)
    with pytest.raises(ValueError):
        E.ADMML2(B, B.pack(bs, adjoint_space=True), regularizer="finite_differences").solve(1)
