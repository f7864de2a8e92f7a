"""Curved detectors (elsa CurvedDetectorDescriptor): the planar coordinates of the pixel
centres against elsa's known answers (elsa/core/tests/test_CurvedDetectorDescriptor.cpp), and
on the GPU the Joseph and Siddon projectors tracing rays through them, against the oracle."""
import ctypes

import numpy as np
import pytest

import kats
import oracle


def product_coords(dim, det, angle, s2d):
    from elsa_amd import _lib

    L = _lib.load()
    out = np.zeros(int(np.prod(det)) * (dim - 1), np.float32)
    d = np.asarray(det, np.int64)
    rc = L.ehp_curved_detector_coords(dim, d.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), angle, s2d,
                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    assert rc == 0
    return out.reshape(-1, dim - 1)


@pytest.mark.parametrize("source", ["product", "oracle"])
def test_planar_coords_known_answers_2d(source):
    """test_CurvedDetectorDescriptor.cpp:254-320 (angle 2, s2d 14)."""
    get = (lambda det: product_coords(2, det, 2.0, 14.0)) if source == "product" else \
        (lambda det: oracle.curved_planar_coords(det, 2.0, 14.0))
    c5 = get((5,))[:, 0]
    assert c5[2] == 2.5
    for got, exp in zip(c5[[0, 1, 3, 4]], (0.60392, 1.51253, 3.48747, 4.39608)):
        assert kats.approx(got, exp, 1e-5), (got, exp)
    assert kats.approx(2.5 - c5[0], abs(2.5 - c5[4]))
    c6 = get((6,))[:, 0]
    for got, exp in zip(c6, (0.61183, 1.52298, 2.50083, 3.49917, 4.47702, 5.38817)):
        assert kats.approx(got, exp, 1e-5), (got, exp)


@pytest.mark.parametrize("source", ["product", "oracle"])
def test_planar_coords_known_answers_3d(source):
    """test_CurvedDetectorDescriptor.cpp:322-370 (5 x 5 pixels)."""
    c = product_coords(3, (5, 5), 2.0, 14.0) if source == "product" else oracle.curved_planar_coords((5, 5), 2.0, 14.0)
    for k, row in enumerate((2, 7, 12, 17, 22)):
        assert kats.approx(c[row, 0], 2.5) and kats.approx(c[row, 1], 0.5 + k)
    assert kats.approx(2.5 - c[20, 0], abs(2.5 - c[24, 0]))
    assert kats.approx(2.5 - c[9, 1], abs(2.5 - c[19, 1]))
    assert c[1, 0] - c[0, 0] < c[2, 0] - c[1, 0] and c[0, 1] < c[1, 1] and c[20, 1] > c[21, 1]


def test_product_and_oracle_coords_agree():
    for dim, det in ((2, (37,)), (3, (21, 9))):
        np.testing.assert_allclose(product_coords(dim, det, 0.7, 300.0),
                                   oracle.curved_planar_coords(det, 0.7, 300.0), rtol=0, atol=2e-6)


def _curved_case(dim):
    import elsa_amd as E

    if dim == 2:
        vd = E.VolumeDescriptor([40, 36])
        dd = E.CurvedCircleTrajectoryGenerator.createTrajectory(23, vd, 360, 800.0, 200.0, 0.9, detectorSize=[64])
    else:
        vd = E.VolumeDescriptor([24, 20, 22])
        dd = E.CurvedCircleTrajectoryGenerator.createTrajectory(13, vd, 360, 480.0, 120.0, 0.8,
                                                                detectorSize=[40, 34])
    P, Pinv, C, ext, sdd = dd.pose_arrays()
    g = oracle.Geom(dim, tuple(vd.getNumberOfCoefficientsPerDimension()),
                    tuple(dd.getNumberOfCoefficientsPerDimension()[:-1]), P, Pinv, C, ext, sdd)
    return vd, dd, g


@pytest.mark.gpu
@pytest.mark.parametrize("kernels", ["fast", "simple"])
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("cls", ["JosephsMethodCUDA", "SiddonsMethodCUDA"])
def test_curved_projectors_match_oracle(dim, cls, kernels):
    """Rays through the planar coordinates: forward bit for bit for Siddon (same rays, same
    traversal) and within 1e-6 for Joseph (lerp-form interpolation), back projection within
    1e-6 (fast: the LDS-tiled kernel, its candidate pixels found through the planar centres;
    simple: one thread per ray); and different from the planar detector of the same
    trajectory."""
    import elsa_amd as E

    vd, dd, g = _curved_case(dim)
    op = getattr(E, cls)(vd, dd, kernels=kernels)
    if kernels == "fast":
        assert op.tile()[0] > 0  # the tiled back projection serves the curved detector
    pix = dd.getPlanarCoords()
    rng = np.random.default_rng(dim)
    x = rng.random(g.vol_shape, dtype=np.float32)
    y = rng.random(g.sino_shape, dtype=np.float32)
    fn = oracle.joseph if cls == "JosephsMethodCUDA" else oracle.siddon
    fwd = op.apply(x)
    ref_fwd = fn(g, x, mode=oracle.MODE_DIRECT, pix=pix)
    if cls == "JosephsMethodCUDA":
        # the plane-major forward interpolates in lerp form: FP32 rounding of elsa's sums
        assert np.linalg.norm(fwd - ref_fwd) <= 1e-6 * np.linalg.norm(ref_fwd)
    else:
        np.testing.assert_array_equal(fwd, ref_fwd)
    adj = op.applyAdjoint(y)
    ref = fn(g, y, adjoint=True, mode=oracle.MODE_DIRECT, pix=pix)
    assert np.linalg.norm(adj - ref) <= 1e-6 * np.linalg.norm(ref)
    planar = fn(g, x, mode=oracle.MODE_DIRECT)
    assert np.linalg.norm(planar - fwd) > 1e-3 * np.linalg.norm(fwd)


@pytest.mark.gpu
@pytest.mark.parametrize("dim", [2, 3])
def test_curved_tiled_equals_simple_adjoint(dim):
    """The tiled back projection and the one-thread-per-ray kernel agree to FP32 rounding of
    the sums on a curved detector (both sum the same terms; the tiled kernel exactly, in
    fixed point), and the tiled result is deterministic."""
    import elsa_amd as E

    vd, dd, g = _curved_case(dim)
    fast = E.JosephsMethodCUDA(vd, dd)
    simple = E.JosephsMethodCUDA(vd, dd, kernels="simple")
    y = np.random.default_rng(7).standard_normal(g.sino_shape).astype(np.float32)
    a = fast.applyAdjoint(y)
    b = simple.applyAdjoint(y)
    assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(b)
    np.testing.assert_array_equal(a, fast.applyAdjoint(y))


@pytest.mark.gpu
def test_curved_bf16_is_fp32_rounded_once():
    """bfloat16 storage on a curved detector: the FP32 operator's results on the same
    (bfloat16-valued) data, rounded once to bfloat16."""
    import torch

    import elsa_amd as E

    vd, dd, g = _curved_case(3)
    A32 = E.JosephsMethodCUDA(vd, dd)
    A16 = E.JosephsMethodCUDA(vd, dd, dtype="bfloat16")
    gen = torch.Generator(device="cuda").manual_seed(5)
    x = (torch.rand(g.vol_shape, device="cuda", generator=gen) * 3 - 1).to(torch.bfloat16)
    y = torch.rand(g.sino_shape, device="cuda", generator=gen).to(torch.bfloat16)
    assert torch.equal(A16.apply(x), A32.apply(x.float()).to(torch.bfloat16))
    assert torch.equal(A16.applyAdjoint(y), A32.applyAdjoint(y.float()).to(torch.bfloat16))


@pytest.mark.gpu
def test_curved_detector_rejected_where_unsupported():
    import elsa_amd as E

    vd, dd, g = _curved_case(2)
    with pytest.raises(NotImplementedError):
        E.BlobVoxelProjectorCUDA(vd, dd)
