"""CPU checks of libelsahip.so's host side: it loads, exports what include/elsa_hip.h
declares, rejects bad arguments like elsa's constructors, and its geometry and basis-function
tables equal the oracle's (and through it the reference's golden vectors). No GPU call."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "elsa_hip.h")


@pytest.fixture(scope="module")
def lib():
    from elsa_amd import _lib

    return _lib.load()


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ehp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(lib):
    from elsa_amd import _lib

    names = declared_symbols()
    assert len(names) >= 17
    assert sorted(_lib.EXPORTS) == names
    for n in names:
        assert hasattr(lib, n), n
    assert b"gfx950" in lib.ehp_version()


def test_library_has_gfx950_code_object():
    """The device code of libelsahip.so is built for gfx950 only."""
    from elsa_amd import _lib

    blob = open(_lib.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_geometry_create_rejects_bad_arguments(lib):
    from elsa_amd import _lib

    g = ctypes.c_void_p()
    fp = ctypes.POINTER(ctypes.c_float)
    z = np.zeros(64, np.float32)
    vol = np.array([4, 4, 4], np.int64)
    det = np.array([3, 3], np.int64)
    i64p = ctypes.POINTER(ctypes.c_int64)
    args = [z.ctypes.data_as(fp)] * 5
    assert lib.ehp_geometry_create(4, 1, vol.ctypes.data_as(i64p), det.ctypes.data_as(i64p), *args,
                                   ctypes.byref(g)) == _lib.EHP_ERR_ARGUMENT
    assert b"2D and 3D" in lib.ehp_last_error()
    assert lib.ehp_geometry_create(3, 0, vol.ctypes.data_as(i64p), det.ctypes.data_as(i64p), *args,
                                   ctypes.byref(g)) == _lib.EHP_ERR_ARGUMENT
    assert b"empty" in lib.ehp_last_error()
    bad = np.array([4, 0, 4], np.int64)
    assert lib.ehp_geometry_create(3, 1, bad.ctypes.data_as(i64p), det.ctypes.data_as(i64p), *args,
                                   ctypes.byref(g)) == _lib.EHP_ERR_ARGUMENT
    assert lib.ehp_geometry_create(3, 1, vol.ctypes.data_as(i64p), det.ctypes.data_as(i64p), *args,
                                   ctypes.byref(g)) == _lib.EHP_OK
    # creating an operator with a bad dtype or flags fails before any device work
    op = ctypes.c_void_p()
    assert lib.ehp_joseph_create(g, 7, 0, ctypes.byref(op)) == _lib.EHP_ERR_ARGUMENT
    assert lib.ehp_joseph_create(g, 0, 3, ctypes.byref(op)) == _lib.EHP_ERR_ARGUMENT
    assert lib.ehp_voxel_create(g, 0, None, 100, 2.0, 0, ctypes.byref(op)) == _lib.EHP_ERR_ARGUMENT
    assert lib.ehp_siddon_create(g, 0, 2, ctypes.byref(op)) == _lib.EHP_ERR_ARGUMENT
    assert lib.ehp_joseph_create(g, 0, 2, ctypes.byref(op)) == _lib.EHP_ERR_ARGUMENT  # differential: voxel only
    lut = np.ones(100, np.float32)
    assert lib.ehp_voxel_create(g, 0, lut.ctypes.data_as(ctypes.c_void_p), 100, 2.0, 4,
                                ctypes.byref(op)) == _lib.EHP_ERR_ARGUMENT
    lib.ehp_geometry_destroy(g)
    assert lib.ehp_forward(None, None, None, None) == _lib.EHP_ERR_ARGUMENT


def test_python_descriptors_raise_like_elsa():
    import elsa_amd as E

    with pytest.raises(ValueError):
        E.VolumeDescriptor([4, 0])
    vd = E.VolumeDescriptor([4, 4])
    with pytest.raises(ValueError):
        E.PlanarDetectorDescriptor([3, 2], [E.Geometry(10.0, 5.0, vd, [3])])


@pytest.mark.parametrize("dim,vol,views,det", [(2, (256, 256), 180, (384,)), (3, (256, 256, 256), 360, (384, 384)),
                                               (3, (64, 48, 80), 17, (70, 50)), (2, (37, 53), 9, None)])
def test_circle_trajectory_equals_oracle(dim, vol, views, det):
    """ehp_circle_trajectory (product C++) and the oracle's restatement agree bit for bit."""
    import elsa_amd as E

    vd = E.VolumeDescriptor(list(vol))
    dd = E.CircleTrajectoryGenerator.createTrajectory(views, vd, 360, 100.0 * vol[0], 5.0 * vol[0],
                                                      detectorSize=None if det is None else list(det))
    g = oracle.circle_geometry(dim, vol, views, 360, 100.0 * vol[0], 5.0 * vol[0], det=det)
    for a, b in zip(dd.pose_arrays(), (g.P, g.Pinv, g.C, g.ext, g.sdd)):
        np.testing.assert_array_equal(a, b)
    assert tuple(dd.getNumberOfCoefficientsPerDimension()[:-1]) == g.det


def test_geometry_pose_with_offsets_equals_oracle():
    import elsa_amd as E

    rng = np.random.default_rng(0)
    for dim in (2, 3):
        for _ in range(20):
            vol = tuple(int(v) for v in rng.integers(3, 40, dim))
            det = tuple(int(v) for v in rng.integers(1, 50, dim - 1))
            gamma, beta, alpha = (float(np.float32(a)) for a in rng.uniform(-4, 4, 3))
            ppo = rng.uniform(-2, 2, dim - 1).astype(np.float32)
            cor = rng.uniform(-3, 3, dim).astype(np.float32)
            s2c, c2d = float(rng.uniform(10, 500)), float(rng.uniform(1, 100))
            vd = E.VolumeDescriptor(list(vol))
            if dim == 2:
                beta = alpha = 0.0
            pg = E.Geometry(s2c, c2d, vd, list(det), gamma, beta, alpha, principalPointOffset=ppo,
                            centerOfRotationOffset=cor)
            og = oracle.geometry_pose(dim, s2c, c2d, vol, det, gamma, beta, alpha, ppo=ppo, cor=cor)
            for a, b in zip((pg.P, pg.Pinv, pg.C, pg.ext, pg.sdd), og):
                np.testing.assert_array_equal(a, b)


def test_basis_tables_equal_oracle_and_golden():
    import elsa_amd as E

    for dt in (np.float32, np.float64):
        for order in (0, 1, 2):
            a, sa = E.blob_lut(2.0, 10.83, order, 100, dt)
            np.testing.assert_array_equal(a, oracle.blob_lut(2.0, 10.83, order, 100, dt))
            assert sa == 2.0
        for dim in (2, 3):
            for order in (0, 1, 2, 3):
                b, sb = E.bspline_lut(dim, order, 100, dt)
                np.testing.assert_array_equal(b, oracle.bspline_lut(dim, order, 100, dt))
                assert sb == (order + 1) * 0.5
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "luts.json")))
    lut, _ = E.blob_lut(2.0, 10.83, 2, 101, np.float64)
    np.testing.assert_allclose(lut[:99], gold["projected_blob"]["expected"][:99], rtol=1e-6, atol=1e-7)
    lut, _ = E.bspline_lut(2, 3, 101, np.float64)
    np.testing.assert_allclose(lut, gold["projected_bspline"]["expected"], rtol=1e-6, atol=1e-7)


def test_projector_without_gpu_fails_loudly():
    """On a machine without a GPU, constructing a projector raises; nothing falls back to
    the CPU."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import elsa_amd as E

    vd = E.VolumeDescriptor([8, 8])
    dd = E.CircleTrajectoryGenerator.createTrajectory(4, vd, 360, 100.0, 10.0)
    with pytest.raises(Exception):
        E.JosephsMethodCUDA(vd, dd)


def test_library_shares_torchs_hip_runtime(lib):
    """libelsahip.so binds to the HIP runtime torch already loaded (one libamdhip64 in the
    process), so torch's streams and device memory are the library's too."""
    from elsa_amd import _lib

    rts = _lib._hip_runtimes()
    assert len(rts) == 1, rts


def test_derivative_tables_equal_oracle_and_golden():
    """The phase-contrast tables of the product (ehp_blob_derivative_lut,
    ehp_bspline_derivative_lut) equal the oracle's bit for bit, and through it the reference's
    golden vectors (test_BlobDerivative.cpp, test_BSplines.cpp)."""
    import elsa_amd as E

    for dt in (np.float32, np.float64):
        for norm in (False, True):
            for order in (0, 1, 2):
                for a, alpha in ((2.0, 10.83), (1.0, 2.5)):
                    t, sup = E.blob_derivative_lut(a, alpha, order, 100, dt, normalized=norm)
                    np.testing.assert_array_equal(t, oracle.blob_derivative_lut(a, alpha, order, 100, dt,
                                                                                normalized=norm))
                    assert sup == a
            for dim in (2, 3):
                for order in (0, 1, 2, 3):
                    t, sup = E.bspline_derivative_lut(dim, order, 100, dt, normalized=norm)
                    np.testing.assert_array_equal(t, oracle.bspline_derivative_lut(dim, order, 100, dt,
                                                                                   normalized=norm))
                    assert sup == (order + 1) * 0.5
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "derivatives.json")))
    t, _ = E.bspline_derivative_lut(2, 3, 101, np.float64)
    np.testing.assert_allclose(t, gold["bspline_derivative_lut"]["expected"], rtol=1e-6, atol=1e-9)
    tab = [x for x in gold["blob_derivative"]["tables"] if x["alpha"] == 10.83 and x["m"] == 2][0]
    t, _ = E.blob_derivative_lut(1.0, 10.83, 2, 101, np.float64)
    np.testing.assert_allclose(t, tab["expected"], rtol=1e-5, atol=1e-7)
