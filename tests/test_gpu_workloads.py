"""GPU parity at the full sizes of the benchmarked workloads (the smaller cases of
test_gpu_projectors.py exercise the same kernels on shapes the oracle finishes quickly).

  cone256   3D cone beam, 256^3 FP32 modified Shepp-Logan, 360 poses over 360 deg, 384x384
            detector, source-centre 25600, centre-detector 1280: Joseph and Siddon
  grad128   4 x 128^3 random-ellipsoid volumes, 180 poses, 192x192 detector, 12800 / 640:
            the autograd wrapper
  batch64   64 independent 2D fan-beam problems, N in {128,...,512}, 90-360 poses from a
            random start, detector floor(N sqrt 2), 100 N / 5 N: fused passes and solvers
  mixed512  512^3, 720 poses, 768x768 detector, 51200 / 2560: Joseph and the blob voxel
            projector in FP32 and bfloat16 storage

Bars (the oracle runs on a subset of the poses, spread over the orbit, where the full
trajectory would take it minutes; whole-trajectory properties are checked on the GPU alone):
  * Siddon forward: bit for bit against the oracle with the GPU's position rule; Joseph forward
    (lerp-form interpolation): the FP32-rounding bars of fp32_bounds;
  * fixed-point back projections: rel. L2 <= 1e-6 against the oracle, within 1e-6 of the
    one-thread-per-ray atomic kernels, and bitwise equal run to run;
  * dot-product test |<Ax,y> - <x,A^T y>| / (|Ax| |y|) <= 1e-6 over all poses;
  * bfloat16 operators: equal to the FP32 operator on the same bfloat16 inputs, rounded once;
  * batches: every problem bit for bit equal to its own operator, passes and solvers.
"""
import numpy as np
import pytest

import oracle
from backends import geom_to_descriptors
from fp32_bounds import check_forward

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def subset(g, idx):
    idx = np.asarray(idx)
    return oracle.Geom(g.dim, g.vol, g.det, g.P[idx], g.Pinv[idx], g.C[idx], g.ext[idx], g.sdd[idx])


def dot_test(op, shape_x, shape_y, seed):
    import torch

    gen = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.randn(shape_x, device="cuda", generator=gen)
    y = torch.randn(shape_y, device="cuda", generator=gen)
    ax = op.apply(x).double()
    aty = op.applyAdjoint(y).double()
    lhs = float((ax * y.double()).sum())
    rhs = float((x.double() * aty).sum())
    return abs(lhs - rhs) / (float(torch.linalg.norm(ax)) * float(torch.linalg.norm(y.double())))


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import elsa_amd  # noqa: F401


# --------------------------------------------------------------------------------------
# 3D cone beam: 256^3, 360 poses, 384x384 detector

CONE256 = dict(vol=(256, 256, 256), views=360, det=(384, 384), s2c=25600.0, c2d=1280.0)
CONE256_POSES = np.arange(0, 360, 15)  # 24 poses, every 15 degrees


@pytest.fixture(scope="module")
def cone256():
    import torch

    import elsa_amd as E
    from elsa_amd.phantoms import modified_shepp_logan

    c = CONE256
    g = oracle.circle_geometry(3, c["vol"], c["views"], 360, c["s2c"], c["c2d"], det=c["det"])
    vd, dd = geom_to_descriptors(g)
    x = modified_shepp_logan(c["vol"][::-1])
    state = {"g": g, "vd": vd, "dd": dd, "x": x, "xd": torch.from_numpy(x).cuda(), "E": E,
             "sub": subset(g, CONE256_POSES)}
    yield state
    state.clear()
    torch.cuda.empty_cache()


@pytest.mark.parametrize("cls", ["JosephsMethodCUDA", "SiddonsMethodCUDA"])
def test_cone256_forward_parity_on_24_poses(cone256, cls):
    """Forward of the phantom over all 360 poses; the rows of 24 poses spread over the orbit
    equal the oracle's, bit for bit for Siddon and to FP32 rounding for Joseph (its plane-major
    forward interpolates in lerp form; the bars of fp32_bounds)."""
    E, g, sub = cone256["E"], cone256["g"], cone256["sub"]
    op = getattr(E, cls)(cone256["vd"], cone256["dd"])
    y = op.apply(cone256["xd"]).cpu().numpy()
    assert y.shape == g.sino_shape
    fn = oracle.joseph if cls == "JosephsMethodCUDA" else oracle.siddon
    if cls == "JosephsMethodCUDA":
        check_forward(y[CONE256_POSES], sub, cone256["x"])
    else:
        np.testing.assert_array_equal(y[CONE256_POSES], fn(sub, cone256["x"], mode=oracle.MODE_DIRECT))
    # and elsa's own position rule (repeated addition) within its drift along 256-voxel rays
    assert rel_l2(y[CONE256_POSES], fn(sub, cone256["x"])) <= 1e-4


@pytest.mark.parametrize("cls", ["JosephsMethodCUDA", "SiddonsMethodCUDA"])
def test_cone256_adjoint_parity_on_24_poses(cone256, cls):
    """Back projection of the phantom's sinogram over the 24 poses (an operator of those poses
    on the full volume and detector): rel. L2 <= 1e-6 against the oracle."""
    E, sub = cone256["E"], cone256["sub"]
    vd, dd = geom_to_descriptors(sub)
    op = getattr(E, cls)(vd, dd)
    y = op.apply(cone256["xd"])
    rng = np.random.default_rng(24)
    y = (y.cpu().numpy() * (0.5 + rng.random(sub.sino_shape, dtype=np.float32))).astype(np.float32)
    back = op.applyAdjoint(y)
    fn = oracle.joseph if cls == "JosephsMethodCUDA" else oracle.siddon
    ref = fn(sub, y, adjoint=True, mode=oracle.MODE_DIRECT)
    assert rel_l2(back, ref) <= 1e-6, rel_l2(back, ref)


@pytest.mark.parametrize("cls", ["JosephsMethodCUDA", "SiddonsMethodCUDA"])
def test_cone256_full_orbit_adjointness_and_reproducibility(cone256, cls):
    """All 360 poses: dot-product test, the tiled back projection within 1e-6 of the
    one-thread-per-ray atomic kernel, and a repeated back projection bitwise equal."""
    import torch

    E, g = cone256["E"], cone256["g"]
    op = getattr(E, cls)(cone256["vd"], cone256["dd"])
    assert dot_test(op, g.vol_shape, g.sino_shape, 11) <= 1e-6
    y = op.apply(cone256["xd"])
    tiled = op.applyAdjoint(y)
    again = op.applyAdjoint(y)
    assert torch.equal(tiled, again)
    simple = getattr(E, cls)(cone256["vd"], cone256["dd"], kernels="simple")
    ref = simple.applyAdjoint(y)
    err = float(torch.linalg.norm((tiled - ref).double()) / torch.linalg.norm(ref.double()))
    assert err <= 1e-6, err
    # the plane-major forward (lerp form, FP64 sums) and the one-thread-per-ray forward (elsa's
    # weight products): the same samples and taps, FP32 rounding apart; Siddon's are the same sums
    ys = simple.apply(cone256["xd"])
    if cls == "SiddonsMethodCUDA":
        assert torch.equal(y, ys)
    else:
        assert float(torch.linalg.norm((y - ys).double()) / torch.linalg.norm(ys.double())) <= 1e-6


def test_cone256_cgls_20_iterations_decrease_the_residual(cone256):
    """20 CGLS iterations (elsa's CGLS.cpp) on b = A x_phantom: ||b - A x_k|| decreases at
    every iteration, and the graph-replayed solve gives the eager one bit for bit."""
    import torch

    E = cone256["E"]
    op = E.JosephsMethodCUDA(cone256["vd"], cone256["dd"])
    b = op.apply(cone256["xd"])
    cg = E.CGLS(op, b, tol=0.0)
    x = cg.solve(20)
    r = np.asarray(cg.residuals[: cg.iterations_done + 1])
    assert cg.iterations_done == 20
    assert np.all(np.diff(r) < 0), r
    assert r[-1] / r[0] < 0.05
    cgg = E.CGLS(op, b, tol=0.0, graph=True)
    assert torch.equal(cgg.solve(20), x)


# --------------------------------------------------------------------------------------
# differentiable projector: 4 x 128^3, 180 poses, 192x192 detector

GRAD128 = dict(vol=(128, 128, 128), views=180, det=(192, 192), s2c=12800.0, c2d=640.0, batch=4)


@pytest.fixture(scope="module")
def grad128():
    import torch

    import elsa_amd as E
    from elsa_amd.phantoms import random_ellipsoids

    c = GRAD128
    g = oracle.circle_geometry(3, c["vol"], c["views"], 360, c["s2c"], c["c2d"], det=c["det"])
    vd, dd = geom_to_descriptors(g)
    truth = torch.stack([torch.from_numpy(random_ellipsoids(c["vol"], 20, seed=1 + i))
                         for i in range(c["batch"])]).cuda()
    state = {"g": g, "vd": vd, "dd": dd, "truth": truth, "E": E, "op": E.JosephsMethodCUDA(vd, dd)}
    yield state
    state.clear()
    torch.cuda.empty_cache()


def tv(v):
    dz = v[:, 1:, :-1, :-1] - v[:, :-1, :-1, :-1]
    dy = v[:, :-1, 1:, :-1] - v[:, :-1, :-1, :-1]
    dx = v[:, :-1, :-1, 1:] - v[:, :-1, :-1, :-1]
    return (dx * dx + dy * dy + dz * dz + 1e-8).sqrt().sum()


def test_grad128_backward_is_the_adjoint_bitwise(grad128):
    """For the batch of 4 volumes, the gradient of <A x_i, w_i> is A^T w_i, bit for bit the
    operator's applyAdjoint; the dot-product test holds on every volume."""
    import torch

    from elsa_amd.autograd import project

    op, g = grad128["op"], grad128["g"]
    x = grad128["truth"].clone().requires_grad_(True)
    gen = torch.Generator(device="cuda").manual_seed(3)
    w = torch.randn((GRAD128["batch"],) + g.sino_shape, device="cuda", generator=gen)
    y = project(op, x)
    for i in range(GRAD128["batch"]):
        assert torch.equal(y[i], op.apply(grad128["truth"][i]))
    (y * w).sum().backward()
    for i in range(GRAD128["batch"]):
        assert torch.equal(x.grad[i], op.applyAdjoint(w[i]))
    assert dot_test(op, g.vol_shape, g.sino_shape, 5) <= 1e-6


def test_grad128_fp64_central_differences(grad128):
    """The loss 0.5 ||A x - b||^2 + 0.01 TV(x) in FP64 (FP64 operator): its autograd
    gradient (backward through the hand-written adjoint) against central differences with
    h = 1e-3 along 8 random directions, relative error <= 1e-5 (the quadratic part's central
    difference is exact up to rounding; TV's third derivative bounds the rest)."""
    import torch

    from elsa_amd.autograd import project

    E, vd, dd = grad128["E"], grad128["vd"], grad128["dd"]
    op64 = E.JosephsMethodCUDA(vd, dd, dtype=np.float64)
    truth = grad128["truth"][:1].double()
    with torch.no_grad():
        b = project(op64, truth)

    def loss(v):
        r = project(op64, v) - b
        return 0.5 * (r * r).sum() + 0.01 * tv(v)

    x = (truth * 0.5).requires_grad_(True)
    grad, = torch.autograd.grad(loss(x), x)
    gen = torch.Generator(device="cuda").manual_seed(9)
    h = 1e-3
    worst = 0.0
    with torch.no_grad():
        for _ in range(8):
            d = torch.randn(x.shape, dtype=torch.float64, device="cuda", generator=gen)
            fd = float(loss(x + h * d) - loss(x - h * d)) / (2 * h)
            an = float((grad * d).sum())
            worst = max(worst, abs(fd - an) / abs(an))
    assert worst <= 1e-5, worst


def test_grad128_oracle_parity_on_12_poses(grad128):
    """The 128^3 operator's rows of 12 poses equal the oracle's forward (fp32_bounds)."""
    g = grad128["g"]
    idx = np.arange(0, 180, 15)
    y = grad128["op"].apply(grad128["truth"][1]).cpu().numpy()
    check_forward(y[idx], subset(g, idx), grad128["truth"][1].cpu().numpy())


# --------------------------------------------------------------------------------------
# batches: 64 uneven 2D fan-beam problems

BATCH64 = dict(n=64, sizes=(128, 192, 256, 384, 512), views=(90, 360), seed=2)


@pytest.fixture(scope="module")
def batch64():
    """The operators and phantoms of the batched workload, drawn as the benchmark draws them:
    size, pose count and start angle per problem from one seeded generator, random-ellipse
    phantoms, sinograms with 1% Gaussian noise."""
    import torch

    import elsa_amd as E
    from elsa_amd.phantoms import random_ellipsoids

    c = BATCH64
    rng = np.random.default_rng(c["seed"])
    geoms, ops, xs, bs = [], [], [], []
    for j in range(c["n"]):
        N = int(rng.choice(c["sizes"]))
        V = int(rng.integers(c["views"][0], c["views"][1] + 1))
        start = float(rng.uniform(0, 360))
        det = int(np.float32(N) * np.float32(np.sqrt(np.float32(2))))
        deg = np.float32(start) + np.linspace(0, 360, V, dtype=np.float32)
        g = oracle.circle_geometry(2, (N, N), V, 360, 100.0 * N, 5.0 * N, det=(det,), angles_deg=deg)
        vd, dd = geom_to_descriptors(g)
        op = E.JosephsMethodCUDA(vd, dd)
        x = torch.from_numpy(random_ellipsoids((N, N), n=10, seed=1000 + j)).cuda()
        y = op.apply(x)
        gen = torch.Generator(device="cuda").manual_seed(2000 + j)
        y = (y + 0.01 * float(y.max()) * torch.randn(y.shape, device="cuda", generator=gen)).contiguous()
        geoms.append(g)
        ops.append(op)
        xs.append(x)
        bs.append(y)
    B = E.ProjectorBatch(ops)
    state = {"E": E, "geoms": geoms, "ops": ops, "xs": xs, "bs": bs, "B": B}
    yield state
    state.clear()
    torch.cuda.empty_cache()


def test_batch64_fused_passes_equal_each_operator(batch64):
    import torch

    B, ops, xs, bs = batch64["B"], batch64["ops"], batch64["xs"], batch64["bs"]
    assert B.fused and len(B) == 64
    sino = B.apply(B.pack(xs))
    back = B.applyAdjoint(B.pack(bs, adjoint_space=True))
    for j, op in enumerate(ops):
        s0, s1 = int(B.sino_offsets[j]), int(B.sino_offsets[j + 1])
        v0, v1 = int(B.vol_offsets[j]), int(B.vol_offsets[j + 1])
        assert torch.equal(sino[s0:s1], op.apply(xs[j]).reshape(-1))
        assert torch.equal(back[v0:v1], op.applyAdjoint(bs[j]).reshape(-1))


def test_batch64_solvers_equal_each_operator(batch64):
    """50 CGLS and 50 FISTA (L1, lambda 1e-3) iterations in hipGraphs on the whole batch give
    every problem's own solve bit for bit."""
    import torch

    E, B, ops, bs = batch64["E"], batch64["B"], batch64["ops"], batch64["bs"]
    b = B.pack(bs, adjoint_space=True)
    xc = B.unpack(E.CGLS(B, b, tol=0.0, graph=True).solve(50))
    xa = B.unpack(E.APGD(B, b, prox=E.ProximalL1(1e-3), graph=True).solve(50))
    for j, op in enumerate(ops):
        assert torch.equal(xc[j], E.CGLS(op, bs[j], tol=0.0, graph=True).solve(50)), j
        assert torch.equal(xa[j], E.APGD(op, bs[j], prox=E.ProximalL1(1e-3), graph=True).solve(50)), j


def test_batch64_oracle_parity_on_3_problems(batch64):
    """Three problems of different sizes against the oracle: forward to FP32 rounding
    (fp32_bounds), back projection rel. L2 <= 1e-6."""
    geoms, ops, xs, bs = batch64["geoms"], batch64["ops"], batch64["xs"], batch64["bs"]
    sizes = [g.vol[0] for g in geoms]
    picks = [int(np.argmin(sizes)), int(np.argmax(sizes)), next(j for j, n in enumerate(sizes) if n == 256)]
    for j in picks:
        g = geoms[j]
        x = xs[j].cpu().numpy()
        check_forward(ops[j].apply(xs[j]).cpu().numpy(), g, x, what=f"problem {j}")
        y = bs[j].cpu().numpy()
        back = ops[j].applyAdjoint(bs[j]).cpu().numpy()
        ref = oracle.joseph(g, y, adjoint=True, mode=oracle.MODE_DIRECT)
        assert rel_l2(back, ref) <= 1e-6, (j, rel_l2(back, ref))


# --------------------------------------------------------------------------------------
# mixed precision at scale: 512^3, 720 poses, 768x768 detector

MIXED512 = dict(vol=(512, 512, 512), views=720, det=(768, 768), s2c=51200.0, c2d=2560.0)
MIXED512_POSES = np.array([0, 90, 270, 450])  # 0, 45, 135 and 225 degrees


@pytest.fixture(scope="module")
def mixed512():
    import torch

    import elsa_amd as E
    from elsa_amd.phantoms import modified_shepp_logan

    c = MIXED512
    g = oracle.circle_geometry(3, c["vol"], c["views"], 360, c["s2c"], c["c2d"], det=c["det"])
    vd, dd = geom_to_descriptors(g)
    x = torch.from_numpy(modified_shepp_logan(c["vol"][::-1])).cuda()
    state = {"E": E, "g": g, "vd": vd, "dd": dd, "x16": x.to(torch.bfloat16), "sub": subset(g, MIXED512_POSES)}
    yield state
    state.clear()
    torch.cuda.empty_cache()


def test_mixed512_joseph_bf16_equals_fp32_rounded_once(mixed512):
    import torch

    E, vd, dd = mixed512["E"], mixed512["vd"], mixed512["dd"]
    A32 = E.JosephsMethodCUDA(vd, dd)
    A16 = E.JosephsMethodCUDA(vd, dd, dtype="bfloat16")
    x16 = mixed512["x16"]
    f32 = A32.apply(x16.float())
    f16 = A16.apply(x16)
    assert torch.equal(f16, f32.to(torch.bfloat16))
    y16 = f16
    b16 = A16.applyAdjoint(y16)
    assert torch.equal(b16, A32.applyAdjoint(y16.float()).to(torch.bfloat16))
    # FP32 rows of 4 poses (0, 45, 135, 225 degrees) against the oracle (fp32_bounds)
    sub = mixed512["sub"]
    check_forward(f32[torch.from_numpy(MIXED512_POSES).cuda()].cpu().numpy(), sub, x16.float().cpu().numpy())


def test_mixed512_joseph_adjoint_parity_on_4_poses(mixed512):
    import torch

    E, sub = mixed512["E"], mixed512["sub"]
    vd, dd = geom_to_descriptors(sub)
    op = E.JosephsMethodCUDA(vd, dd)
    y = op.apply(mixed512["x16"].float())
    back = op.applyAdjoint(y).cpu().numpy()
    ref = oracle.joseph(sub, y.cpu().numpy(), adjoint=True, mode=oracle.MODE_DIRECT)
    assert rel_l2(back, ref) <= 1e-6, rel_l2(back, ref)
    del op
    torch.cuda.empty_cache()


def test_mixed512_blob_bf16_equals_fp32_rounded_once(mixed512):
    """Blob voxel projector (radius 2, alpha 10.83, order 2) over all 720 poses: bfloat16
    storage equals FP32 on the same bfloat16 inputs rounded once, both directions."""
    import torch

    E, vd, dd = mixed512["E"], mixed512["vd"], mixed512["dd"]
    B32 = E.BlobVoxelProjectorCUDA(vd, dd)
    B16 = E.BlobVoxelProjectorCUDA(vd, dd, dtype="bfloat16")
    x16 = mixed512["x16"]
    f16 = B16.apply(x16)
    assert torch.equal(f16, B32.apply(x16.float()).to(torch.bfloat16))
    b16 = B16.applyAdjoint(f16)
    assert torch.equal(b16, B32.applyAdjoint(f16.float()).to(torch.bfloat16))


def test_mixed512_blob_parity_on_2_poses(mixed512):
    """The blob operator of 2 poses (0 and 45 degrees) on the full volume and detector
    against the oracle: back projection of the phantom's sinogram over the whole volume, and
    the forward of a sparse volume (0.1% of the voxels set; the oracle splats every voxel
    once per pose, which a dense 512^3 volume would keep it at for minutes); rel. L2 <= 1e-6."""
    import torch

    E, g = mixed512["E"], mixed512["g"]
    sub = subset(g, [0, 90])
    vd, dd = geom_to_descriptors(sub)
    op = E.BlobVoxelProjectorCUDA(vd, dd)
    lut = oracle.blob_lut(2.0, 10.83, 2, 100, np.float32)
    y = op.apply(mixed512["x16"].float())
    back = op.applyAdjoint(y).cpu().numpy()
    ref = oracle.voxel(sub, y.cpu().numpy(), lut, 2.0, adjoint=True)
    assert rel_l2(back, ref) <= 1e-6, rel_l2(back, ref)
    rng = np.random.default_rng(512)
    xs = np.zeros(sub.vol_shape, np.float32)
    flat = xs.reshape(-1)
    hit = rng.choice(flat.size, flat.size // 1000, replace=False)
    flat[hit] = rng.random(hit.size, dtype=np.float32)
    fwd = op.apply(torch.from_numpy(xs).cuda()).cpu().numpy()
    ref = oracle.voxel(sub, xs, lut, 2.0)
    assert rel_l2(fwd, ref) <= 1e-6, rel_l2(fwd, ref)
