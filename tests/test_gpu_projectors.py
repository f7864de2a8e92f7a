"""GPU parity of the HIP projectors (through elsa_amd and the C-ABI) against the oracle.

Bars:
  * integer work (per-ray driving axis and sample counts, total sample counts): bit-exact;
  * Siddon forward and voxel back projection (same arithmetic, same summation order as the
    CPU restatement): bit-exact in FP32;
  * Joseph against the oracle evaluating sample positions the same way (pos0 + s*step):
    the one-thread-per-ray forward bit for bit; the plane-major forward (lerp-form
    interpolation) rel. L2 <= 1e-6 and at least as close to the exact FP64 sums of the FP32
    terms as elsa's FP32 form (fp32_bounds); back projections rel. L2 <= 1e-6;
  * Joseph against elsa's own rule (repeated addition of the step): rel. L2 <= 1e-4 (the
    accumulated step drifts by up to ~1e-4 voxel after 256 steps in FP32; measured at most
    5.4e-5 on these cases), and the GPU's error against the FP64 oracle is no larger than
    elsa-FP32's;
  * atomically accumulated back projections (Joseph, Siddon) and the voxel forward:
    rel. L2 <= 1e-6 against the oracle;
  * matched adjoints: |<Ax,y> - <x,A^T y>| / (|Ax| |y|) <= 1e-6 in FP32, 1e-12 in FP64.
"""
import numpy as np
import pytest

import kats
import oracle
from backends import GpuBackend, geom_to_descriptors
from fp32_bounds import check_forward

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import elsa_amd  # noqa: F401  (loads libelsahip.so; fails loudly when missing)


# --------------------------------------------------------------------------------------
# known-answer tests of the reference, through the C-ABI


@pytest.mark.parametrize("kat", kats.JOSEPH_KATS, ids=lambda k: k.__name__)
@pytest.mark.parametrize("dt", [np.float32, np.float64], ids=["f32", "f64"])
def test_joseph_kats_gpu(kat, dt):
    kat(GpuBackend(dt))


@pytest.mark.parametrize("kat", kats.SIDDON_KATS, ids=lambda k: k.__name__)
@pytest.mark.parametrize("dt", [np.float32, np.float64], ids=["f32", "f64"])
def test_siddon_kats_gpu(kat, dt):
    kat(GpuBackend(dt))


@pytest.mark.parametrize("kat", kats.VOXEL_KATS, ids=lambda k: k.__name__)
@pytest.mark.parametrize("dt,kernels", [(np.float32, "fast"), (np.float32, "simple"), (np.float64, "fast")],
                         ids=["f32", "f32-simple", "f64"])
def test_voxel_kats_gpu(kat, dt, kernels):
    kat(GpuBackend(dt, kernels))


@pytest.mark.parametrize("kat", kats.PHASE_CONTRAST_KATS, ids=lambda k: k.__name__)
@pytest.mark.parametrize("dt", [np.float32, np.float64], ids=["f32", "f64"])
def test_phase_contrast_kats_gpu(kat, dt):
    kat(GpuBackend(dt))


# --------------------------------------------------------------------------------------
# parity on seeded inputs

CASES = {
    # 2D fan beam of the 2D CPU reference workload: 256^2, 180 poses over 360 deg, det 384
    "fan2d_256": dict(dim=2, vol=(256, 256), views=180, det=(384,), s2c=25600.0, c2d=1280.0),
    # cone beam at a reduced size with the same magnification as the 3D drop-in workload
    "cone3d_64": dict(dim=3, vol=(64, 64, 64), views=90, det=(96, 96), s2c=6400.0, c2d=320.0),
    # strongly divergent cone with a non-cubic volume and ragged detector
    "cone3d_wide": dict(dim=3, vol=(40, 24, 33), views=13, det=(61, 29), s2c=60.0, c2d=40.0),
    # a small fan with a detector wider than the fan (many rays miss)
    "fan2d_wide": dict(dim=2, vol=(31, 17), views=7, det=(80,), s2c=40.0, c2d=30.0),
}


def make_case(name, dtype=np.float32):
    c = CASES[name]
    return oracle.circle_geometry(c["dim"], c["vol"], c["views"], 360, c["s2c"], c["c2d"], det=c["det"], dtype=dtype)


def phantom(shape, seed, dtype=np.float32):
    rng = np.random.default_rng(seed)
    x = rng.random(shape)
    # plus a smooth part so that interpolation weights matter
    grids = np.meshgrid(*[np.linspace(-1, 1, s) for s in shape], indexing="ij")
    r2 = sum(g * g for g in grids)
    return (x * 0.3 + (r2 < 0.6)).astype(dtype)


@pytest.fixture(scope="module")
def ops():
    import elsa_amd as E

    cache = {}

    def get(cls, name, dtype=np.float32, **kw):
        key = (cls, name, np.dtype(dtype).name, tuple(sorted(kw.items())))
        if key not in cache:
            g32 = make_case(name, np.float32)
            vd, dd = geom_to_descriptors(g32)
            cache[key] = getattr(E, cls)(vd, dd, dtype=dtype, **kw)
        return cache[key]

    return get


@pytest.mark.parametrize("name", list(CASES))
def test_joseph_ray_state_bit_exact(name, ops):
    """Driving axis and sample count of every ray equal the oracle's (integer parity)."""
    g = make_case(name)
    op = ops("JosephsMethodCUDA", name)
    axis, nsteps = op.ray_info()
    oaxis, onsteps, _, _, _ = oracle.joseph_rays(g)
    np.testing.assert_array_equal(nsteps, onsteps)
    np.testing.assert_array_equal(axis, oaxis)
    assert op.sample_count() == int(onsteps.sum())


@pytest.mark.parametrize("kernels", ["fast", "simple"])
@pytest.mark.parametrize("name", list(CASES))
def test_joseph_forward_parity(name, kernels, ops):
    g = make_case(name)
    x = phantom(g.vol_shape, 1)
    kw = {} if kernels == "fast" else {"kernels": "simple"}
    gpu = ops("JosephsMethodCUDA", name, **kw).apply(x)
    direct = oracle.joseph(g, x, mode=oracle.MODE_DIRECT)
    elsa = oracle.joseph(g, x, mode=oracle.MODE_ELSA)
    ref64 = oracle.joseph(make_case(name, np.float64), x.astype(np.float64), mode=oracle.MODE_DIRECT)
    if kernels == "simple":
        # same positions, taps and summation order as the oracle's direct rule: bit for bit
        np.testing.assert_array_equal(gpu, direct)
    else:
        # same positions, floors and taps, interpolation in lerp form: FP32 rounding of the sums
        check_forward(gpu, g, x)
    assert rel_l2(gpu, elsa) <= 1e-4, rel_l2(gpu, elsa)
    # error budget against FP64: the GPU is at least as close as elsa's FP32 arithmetic
    assert rel_l2(gpu, ref64) <= max(1.05 * rel_l2(elsa, ref64), 1e-6), (rel_l2(gpu, ref64), rel_l2(elsa, ref64))


@pytest.mark.parametrize("name", list(CASES))
def test_joseph_adjoint_parity(name, ops):
    g = make_case(name)
    y = phantom(g.sino_shape, 2)
    gpu = ops("JosephsMethodCUDA", name).applyAdjoint(y)
    direct = oracle.joseph(g, y, adjoint=True, mode=oracle.MODE_DIRECT)
    elsa = oracle.joseph(g, y, adjoint=True, mode=oracle.MODE_ELSA)
    ref64 = oracle.joseph(make_case(name, np.float64), y.astype(np.float64), adjoint=True, mode=oracle.MODE_DIRECT)
    assert rel_l2(gpu, direct) <= 1e-6, rel_l2(gpu, direct)
    assert rel_l2(gpu, elsa) <= 1e-4, rel_l2(gpu, elsa)
    assert rel_l2(gpu, ref64) <= max(1.05 * rel_l2(elsa, ref64), 1e-6), (rel_l2(gpu, ref64), rel_l2(elsa, ref64))


@pytest.mark.parametrize("kernels", ["fast", "simple"])
@pytest.mark.parametrize("name", list(CASES))
def test_siddon_parity(name, kernels, ops):
    g = make_case(name)
    x = phantom(g.vol_shape, 3)
    y = phantom(g.sino_shape, 4)
    kw = {} if kernels == "fast" else {"kernels": "simple"}
    op = ops("SiddonsMethodCUDA", name, **kw)
    fwd = op.apply(x)
    # same traversal (crossing parameters T0 + k tDelta), same order: bit for bit
    np.testing.assert_array_equal(fwd, oracle.siddon(g, x, mode=oracle.MODE_DIRECT))
    # elsa's repeated addition of tDelta drifts from the direct rule along long rays
    assert rel_l2(fwd, oracle.siddon(g, x)) <= 5e-5
    adj = op.applyAdjoint(y)
    direct = oracle.siddon(g, y, adjoint=True, mode=oracle.MODE_DIRECT)
    assert rel_l2(adj, direct) <= 1e-6, rel_l2(adj, direct)
    assert rel_l2(adj, oracle.siddon(g, y, adjoint=True)) <= 5e-5


@pytest.mark.parametrize("name", ["fan2d_256", "cone3d_wide", "fan2d_wide"])
@pytest.mark.parametrize("kind", ["blob", "bspline"])
def test_voxel_fp64_tiled_forward(name, kind, ops):
    """VoxelProjectorCUDA<double>'s forward over detector tiles (the FP64 kernel's weights, each
    tap's x * w rounded once to the fixed-point unit): the oracle's FP32-geometry / FP64-data
    forward and the one-thread-per-voxel FP64 kernel (FP64 atomics) to FP64 rounding, bitwise
    reproducible; the back projection (a gather) equals the oracle."""
    g = make_case(name)
    x = phantom(g.vol_shape, 15, np.float64)
    y = phantom(g.sino_shape, 16, np.float64)
    if kind == "blob":
        op = ops("BlobVoxelProjectorCUDA", name, dtype=np.float64)
        simple = ops("BlobVoxelProjectorCUDA", name, dtype=np.float64, kernels="simple")
        lut, sup = oracle.blob_lut(2.0, 10.83, 2, 100, np.float64), 2.0
    else:
        op = ops("BSplineVoxelProjectorCUDA", name, dtype=np.float64)
        simple = ops("BSplineVoxelProjectorCUDA", name, dtype=np.float64, kernels="simple")
        lut, sup = oracle.bspline_lut(g.dim, 2, 100, np.float64), 1.5
    assert op.tile()[0] > 0  # the tiled forward runs
    fwd = op.apply(x)
    ref = oracle.voxel(g, x, lut, sup, acc64=True)
    assert rel_l2(fwd, ref) <= 1e-13, rel_l2(fwd, ref)
    assert rel_l2(fwd, simple.apply(x)) <= 1e-13
    np.testing.assert_array_equal(op.apply(x), fwd)
    adj = op.applyAdjoint(y)
    assert rel_l2(adj, oracle.voxel(g, y, lut, sup, adjoint=True, acc64=True)) <= 1e-13


@pytest.mark.parametrize("name", ["fan2d_256", "cone3d_wide", "fan2d_wide"])
@pytest.mark.parametrize("kind", ["blob", "bspline"])
@pytest.mark.parametrize("kernels", ["fast", "simple"])
def test_voxel_parity(name, kind, kernels, ops):
    g = make_case(name)
    x = phantom(g.vol_shape, 5)
    y = phantom(g.sino_shape, 6)
    kw = {} if kernels == "fast" else {"kernels": "simple"}
    if kind == "blob":
        op = ops("BlobVoxelProjectorCUDA", name, **kw)
        lut, sup = oracle.blob_lut(2.0, 10.83, 2, 100, np.float32), 2.0
    else:
        op = ops("BSplineVoxelProjectorCUDA", name, **kw)
        lut, sup = oracle.bspline_lut(g.dim, 2, 100, np.float32), 1.5
    adj_ref = oracle.voxel(g, y, lut, sup, adjoint=True)
    if kernels == "simple":
        # elsa's tap arithmetic in elsa's order: the oracle bit for bit
        np.testing.assert_array_equal(op.applyAdjoint(y), adj_ref)
    else:
        # the tiled forward's tap arithmetic (a reciprocal of the scaling per voxel and pose):
        # within a few ulp per weight
        assert rel_l2(op.applyAdjoint(y), adj_ref) <= 1e-6, rel_l2(op.applyAdjoint(y), adj_ref)
    fwd = op.apply(x)
    ref = oracle.voxel(g, x, lut, sup)
    assert rel_l2(fwd, ref) <= 1e-6, rel_l2(fwd, ref)
    if kernels == "fast":
        # fixed-point sums: bitwise reproducible, and no further from the exact sum of the FP32
        # terms than the oracle's own FP32 sum (up to one rounding of each output)
        np.testing.assert_array_equal(op.apply(x), fwd)


@pytest.mark.parametrize("name", ["fan2d_256", "cone3d_wide"])
def test_voxel_table_with_nonzero_tail(name):
    """A table whose last entry is not 0 (elsa's blob and B-spline tables vanish at the support,
    and the fast kernels then drop their range test): the tiled forward and the gather weigh taps
    past the table's end with 0, as the oracle does."""
    import elsa_amd as E
    from elsa_amd.projectors import _VoxelProjector

    class RaisedBlob(E.BlobVoxelProjectorCUDA):
        def __init__(self, vd, dd):
            self.radius, self.alpha, self.order = 2.0, 10.83, 2
            lut, self.support = E.blob_lut(2.0, 10.83, 2, 100, np.float32)
            self.lut = (lut + np.float32(0.05)).astype(np.float32)
            self._set_kernels("fast")
            _VoxelProjector.__init__(self, vd, dd, np.float32)

    g = make_case(name)
    vd, dd = geom_to_descriptors(g)
    op = RaisedBlob(vd, dd)
    assert op.lut[-1] != 0.0
    x = phantom(g.vol_shape, 9)
    y = phantom(g.sino_shape, 10)
    assert rel_l2(op.apply(x), oracle.voxel(g, x, op.lut, 2.0)) <= 1e-6
    assert rel_l2(op.applyAdjoint(y), oracle.voxel(g, y, op.lut, 2.0, adjoint=True)) <= 1e-6


@pytest.mark.parametrize("name", ["fan2d_256", "cone3d_wide", "fan2d_wide"])
@pytest.mark.parametrize("kind", ["blob", "bspline"])
@pytest.mark.parametrize("kernels", ["fast", "simple"])
def test_phase_contrast_parity(name, kind, kernels, ops):
    """Phase-contrast projectors (derivative table in 2D, normalized gradient in 3D, weight
    times sgn(d_u) or d_u): the same bars as the classic voxel projectors."""
    g = make_case(name)
    x = phantom(g.vol_shape, 7)
    y = phantom(g.sino_shape, 8)
    kw = {} if kernels == "fast" else {"kernels": "simple"}
    norm = g.dim == 3
    if kind == "blob":
        op = ops("PhaseContrastBlobVoxelProjectorCUDA", name, **kw)
        lut, sup = oracle.blob_derivative_lut(2.0, 10.83, 2, 100, np.float32, normalized=norm), 2.0
    else:
        op = ops("PhaseContrastBSplineVoxelProjectorCUDA", name, **kw)
        lut, sup = oracle.bspline_derivative_lut(g.dim, 2, 100, np.float32, normalized=norm), 1.5
    np.testing.assert_array_equal(op.lut, lut)
    adj = op.applyAdjoint(y)
    adj_ref = oracle.voxel(g, y, lut, sup, adjoint=True, differential=True)
    # the differential weights change sign across the footprint, so the sums cancel: rounding
    # is measured against the sum of the terms' magnitudes (|weight| <= |table| in 2D and
    # <= |table| * support in 3D)
    scale = sup if g.dim == 3 else 1.0
    if kernels == "simple":
        np.testing.assert_array_equal(adj, adj_ref)
    else:
        adj_mag = oracle.voxel(g, np.abs(y), np.abs(lut), sup, adjoint=True) * scale
        err = np.linalg.norm(np.asarray(adj, np.float64) - adj_ref) / np.linalg.norm(adj_mag)
        assert err <= 1e-6, err
    fwd = op.apply(x)
    ref = oracle.voxel(g, x, lut, sup, differential=True)
    mag = oracle.voxel(g, np.abs(x), np.abs(lut), sup) * scale
    err = np.linalg.norm(np.asarray(fwd, np.float64) - ref) / np.linalg.norm(mag)
    assert err <= 1e-6, err
    if kernels == "fast":
        np.testing.assert_array_equal(op.apply(x), fwd)


@pytest.mark.parametrize("cls", ["JosephsMethodCUDA", "SiddonsMethodCUDA", "BlobVoxelProjectorCUDA"])
@pytest.mark.parametrize("name", ["fan2d_256", "cone3d_64", "cone3d_wide"])
@pytest.mark.parametrize("dt", [np.float32, np.float64], ids=["f32", "f64"])
def test_dot_product_adjointness(cls, name, dt, ops):
    g = make_case(name)
    op = ops(cls, name, dtype=dt)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(g.vol_shape).astype(dt)
    y = rng.standard_normal(g.sino_shape).astype(dt)
    ax = op.apply(x).astype(np.float64)
    aty = op.applyAdjoint(y).astype(np.float64)
    lhs = np.vdot(ax, y.astype(np.float64))
    rhs = np.vdot(x.astype(np.float64), aty)
    tol = 1e-6 if dt == np.float32 else 1e-12
    assert abs(lhs - rhs) <= tol * np.linalg.norm(ax) * np.linalg.norm(y), (lhs, rhs)


@pytest.mark.parametrize("name", ["cone3d_64", "fan2d_wide"])
def test_fp64_data_projectors(name, ops):
    """data_t = double keeps elsa's FP32 geometry (real_t is float) and sums in double: equal
    to the oracle with FP32 geometry and FP64 data (only the order of the adjoint's atomic
    additions differs), and within FP32 rounding of the all-FP32 oracle."""
    g = make_case(name)
    x = phantom(g.vol_shape, 8, np.float64)
    gpu = ops("JosephsMethodCUDA", name, dtype=np.float64).apply(x)
    ref = oracle.joseph(g, x, mode=oracle.MODE_DIRECT, acc64=True)
    assert rel_l2(gpu, ref) <= 1e-13, rel_l2(gpu, ref)
    assert rel_l2(gpu, oracle.joseph(g, x.astype(np.float32), mode=oracle.MODE_DIRECT)) <= 1e-6
    y = phantom(g.sino_shape, 9, np.float64)
    gpu_b = ops("SiddonsMethodCUDA", name, dtype=np.float64).applyAdjoint(y)
    ref_b = oracle.siddon(g, y, adjoint=True, mode=oracle.MODE_DIRECT, acc64=True)
    assert rel_l2(gpu_b, ref_b) <= 1e-13, rel_l2(gpu_b, ref_b)
    assert rel_l2(gpu_b, oracle.siddon(g, y.astype(np.float32), adjoint=True, mode=oracle.MODE_DIRECT)) <= 5e-6


@pytest.mark.parametrize("name", list(CASES))
def test_joseph_fp64_plane_forward_bitwise(name, ops):
    """JosephsMethodCUDA<double>'s forward over FP64 plane-major views (elsa's FP32 weights, tap
    order and FP64 sums): the oracle's FP32-geometry / FP64-data forward bit for bit, and the
    one-thread-per-ray FP64 kernel's bit for bit."""
    g = make_case(name)
    x = phantom(g.vol_shape, 23, np.float64)
    fast = ops("JosephsMethodCUDA", name, dtype=np.float64).apply(x)
    np.testing.assert_array_equal(fast, oracle.joseph(g, x, mode=oracle.MODE_DIRECT, acc64=True))
    np.testing.assert_array_equal(fast, ops("JosephsMethodCUDA", name, dtype=np.float64, kernels="simple").apply(x))


@pytest.mark.parametrize("name", list(CASES))
def test_siddon_fp64_plane_forward_bitwise(name, ops):
    """SiddonsMethodCUDA<double>'s forward over FP64 plane-major views (FP32 traversal and
    lengths, FP64 sums in traversal order): the oracle's FP32-geometry / FP64-data forward bit for
    bit, and the one-thread-per-ray FP64 kernel's bit for bit."""
    g = make_case(name)
    x = phantom(g.vol_shape, 24, np.float64)
    fast = ops("SiddonsMethodCUDA", name, dtype=np.float64).apply(x)
    np.testing.assert_array_equal(fast, oracle.siddon(g, x, mode=oracle.MODE_DIRECT, acc64=True))
    np.testing.assert_array_equal(fast, ops("SiddonsMethodCUDA", name, dtype=np.float64, kernels="simple").apply(x))


@pytest.mark.parametrize("name", list(CASES))
def test_joseph_fp64_tiled_back_projection(name, ops):
    """JosephsMethodCUDA<double>'s back projection (LDS tiles, 64-bit fixed-point sums of the
    FP64 contributions): within a few units of FP64 rounding of the oracle's sequential FP64 sums
    of the same FP32-geometry terms, equal to the one-thread-per-ray atomic kernel to the same
    bar, and bitwise reproducible run to run (the atomic kernel is not)."""
    g = make_case(name)
    y = phantom(g.sino_shape, 21, np.float64)
    op = ops("JosephsMethodCUDA", name, dtype=np.float64)
    tiled = op.applyAdjoint(y)
    ref = oracle.joseph(g, y, adjoint=True, mode=oracle.MODE_DIRECT, acc64=True)
    assert rel_l2(tiled, ref) <= 1e-13, rel_l2(tiled, ref)
    simple = ops("JosephsMethodCUDA", name, dtype=np.float64, kernels="simple").applyAdjoint(y)
    assert rel_l2(tiled, simple) <= 1e-13, rel_l2(tiled, simple)
    for _ in range(2):
        np.testing.assert_array_equal(op.applyAdjoint(y), tiled)


@pytest.mark.parametrize("name", list(CASES))
def test_siddon_fp64_tiled_back_projection(name, ops):
    """SiddonsMethodCUDA<double>'s back projection (LDS tiles, each crossing's FP64 y * length
    rounded once to the fixed-point unit): within a few units of FP64 rounding of the oracle's
    FP64 sums and of the one-thread-per-ray atomic kernel, bitwise reproducible run to run."""
    g = make_case(name)
    y = phantom(g.sino_shape, 22, np.float64)
    op = ops("SiddonsMethodCUDA", name, dtype=np.float64)
    assert op.tile()[0] > 0  # the tiled kernel runs
    tiled = op.applyAdjoint(y)
    ref = oracle.siddon(g, y, adjoint=True, mode=oracle.MODE_DIRECT, acc64=True)
    assert rel_l2(tiled, ref) <= 1e-13, rel_l2(tiled, ref)
    simple = ops("SiddonsMethodCUDA", name, dtype=np.float64, kernels="simple").applyAdjoint(y)
    assert rel_l2(tiled, simple) <= 1e-13, rel_l2(tiled, simple)
    for _ in range(2):
        np.testing.assert_array_equal(op.applyAdjoint(y), tiled)


def test_outputs_overwritten_not_accumulated(ops):
    """elsa's projectors overwrite the output (TraverseJosephsCUDA.cu:69, SiddonsMethodCUDA.cpp:75)."""
    import torch

    name = "cone3d_wide"
    g = make_case(name)
    x = torch.from_numpy(phantom(g.vol_shape, 9)).cuda()
    for cls in ("JosephsMethodCUDA", "SiddonsMethodCUDA", "BlobVoxelProjectorCUDA"):
        op = ops(cls, name)
        out = torch.full(g.sino_shape, 7.0, device="cuda")
        op.apply(x, out)
        ref = op.apply(x)
        assert torch.equal(out, ref) or float((out - ref).abs().max()) <= 1e-5 * float(ref.abs().max())
        vol = torch.full(g.vol_shape, -3.0, device="cuda")
        op.applyAdjoint(ref, vol)
        ref_v = op.applyAdjoint(ref)
        assert float((vol - ref_v).abs().max()) <= 1e-5 * float(ref_v.abs().max())


def test_views_of_a_larger_buffer(ops):
    """Input and output may be views at an offset inside a bigger allocation (the begin
    pointer is used, never the parent storage)."""
    import torch

    name = "fan2d_wide"
    g = make_case(name)
    op = ops("JosephsMethodCUDA", name)
    x = phantom(g.vol_shape, 10)
    n_sino = int(np.prod(g.sino_shape))
    big = torch.zeros(3 * n_sino, device="cuda")
    view = big[n_sino:2 * n_sino].view(g.sino_shape)
    op.apply(torch.from_numpy(x).cuda(), view)
    torch.cuda.synchronize()
    assert float(big[:n_sino].abs().sum()) == 0.0 and float(big[2 * n_sino:].abs().sum()) == 0.0
    np.testing.assert_allclose(view.cpu().numpy(), op.apply(x), rtol=0, atol=0)


@pytest.mark.parametrize("name", list(CASES))
def test_siddon_tiled_equals_simple_kernels(name, ops):
    """The LDS-tiled Siddon back projection and the atomic one sum the same terms; the tiled
    one is bitwise reproducible run to run."""
    g = make_case(name)
    y = phantom(g.sino_shape, 13)
    op = ops("SiddonsMethodCUDA", name)
    tiled = op.applyAdjoint(y)
    simple = ops("SiddonsMethodCUDA", name, kernels="simple").applyAdjoint(y)
    assert rel_l2(tiled, simple) <= 1e-6
    np.testing.assert_array_equal(op.applyAdjoint(y), tiled)


@pytest.mark.parametrize("name", list(CASES))
def test_joseph_tiled_equals_simple_kernels(name, ops):
    """The LDS-tiled back projection and the one-thread-per-ray atomic one compute the same
    sums of the same bit-identical terms; the plane-major forward (lerp-form interpolation) and
    the one-thread-per-ray forward (elsa's weight products) agree to FP32 rounding."""
    g = make_case(name)
    y = phantom(g.sino_shape, 11)
    tiled = ops("JosephsMethodCUDA", name).applyAdjoint(y)
    simple = ops("JosephsMethodCUDA", name, kernels="simple").applyAdjoint(y)
    assert rel_l2(tiled, simple) <= 1e-6
    x = phantom(g.vol_shape, 12)
    fast_fwd = ops("JosephsMethodCUDA", name).apply(x)
    assert rel_l2(fast_fwd, ops("JosephsMethodCUDA", name, kernels="simple").apply(x)) <= 1e-6


def test_size_mismatch_raises(ops):
    op = ops("JosephsMethodCUDA", "fan2d_wide")
    with pytest.raises(ValueError):
        op.apply(np.zeros((5, 5), np.float32))
