"""Pins the CPU oracle before it is trusted as the checker.

Each check is chosen so that a plausible slip in the restatement (a dropped term, a swapped
index, a wrong sign) fails it: golden vectors of the reference's own tests, its known-answer
tests, and what the mathematics fixes (adjointness, linearity, closed forms, brute force).
"""
import json
import math
import os

import numpy as np
import pytest

import kats
import oracle
from backends import OracleBackend

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# --------------------------------------------------------------------------------------
# golden vectors


def test_projected_blob_lut_golden():
    """elsa/projectors/tests/test_Luts.cpp:125-247 (ProjectedBlob<data_t, 101>, rel 1e-6)."""
    gold = json.load(open(os.path.join(GOLDEN, "luts.json")))["projected_blob"]
    for dt in (np.float32, np.float64):
        lut = oracle.blob_lut(2.0, 10.83, 2, 101, dt)
        for i in range(99):
            got = oracle.lut_eval(lut, 2.0, i / 50.0) if dt == np.float64 else oracle.lut_eval(lut, 2.0, np.float32(i / 50.0))
            assert kats.approx(got, gold["expected"][i], 1e-6), (dt, i, got, gold["expected"][i])


def test_projected_bspline_lut_golden():
    """elsa/projectors/tests/test_Luts.cpp:249-306 (ProjectedBSpline order 3, dim 2)."""
    gold = json.load(open(os.path.join(GOLDEN, "luts.json")))["projected_bspline"]
    for dt in (np.float32, np.float64):
        lut = oracle.bspline_lut(2, 3, 101, dt)
        for i in range(101):
            got = oracle.lut_eval(lut, 2.0, i / 50.0) if dt == np.float64 else oracle.lut_eval(lut, 2.0, np.float32(i / 50.0))
            assert kats.approx(got, gold["expected"][i], 1e-6), (dt, i, got)


def test_projected_blob_weights_golden():
    """elsa/projectors/tests/test_BlobVoxelProjector.cpp:2289-2336 ("Test weights"): the
    closed-form projected blob (a = 2, alpha = 10.83, m = 2) at distance i / 25 (doctest
    Approx), in both of the oracle's precisions."""
    gold = json.load(open(os.path.join(GOLDEN, "blob_weights.json")))
    assert len(gold["expected"]) == 51
    for i, e in enumerate(gold["expected"]):
        got = oracle.blob_projected(i / 25.0, gold["radius"], gold["alpha"], gold["order"])
        assert kats.approx(got, e), (i, got, e)


def test_bessel_golden():
    """elsa/core/tests/test_Bessel.cpp:29-89 (std::cyl_bessel_i tables, doctest Approx)."""
    gold = json.load(open(os.path.join(GOLDEN, "bessel.json")))
    for m in range(5):
        exp = gold[f"I{m}"]
        x = np.linspace(0.0, 4.0, len(exp))
        for xi, e in zip(x, exp):
            assert kats.approx(oracle.bessi(m, float(xi)), e), (m, xi)


def test_bessel_vs_scipy_and_recurrence():
    """I_0, I_1 are the A&S 9.8.1-9.8.4 fits (relative error below 1e-6 of scipy); higher
    orders come from upward recurrence (orders 2-4) or Miller's algorithm, which elsa accepts
    to doctest Approx, i.e. 1.2e-5 on scale 1 (cancellation at small x)."""
    sp = pytest.importorskip("scipy.special")
    for m in range(0, 7):
        for x in (0.3, 1.0, 3.7, 3.8, 7.5, 10.83):
            got, ref = oracle.bessi(m, x), sp.iv(m, x)
            if m < 2:
                assert got == pytest.approx(ref, rel=1e-6), (m, x)
            else:
                assert kats.approx(got, ref), (m, x, got, ref)


def test_blob_profile_vs_scipy():
    """Blob profile of elsa/projectors/Blobs.h:17-30 against the definition with scipy's I_m
    (the formula elsa's test_Blobs.cpp:112-122 generated its tables from)."""
    sp = pytest.importorskip("scipy.special")
    for a in (1.0, 2.0, 3.0):
        for alpha in (2.5, 5.0, 7.5, 10.83):
            for m in (0, 1, 2, 3):
                for r in np.linspace(0, a * 1.05, 23):
                    arg = math.sqrt(max(0.0, 1 - (r / a) ** 2))
                    ref = 0.0 if r > a else sp.iv(m, alpha * arg) / sp.iv(m, alpha) * arg ** m
                    assert oracle.blob_evaluate(float(r), a, alpha, m) == pytest.approx(ref, rel=1e-5, abs=1e-7)


def test_projected_blob_is_line_integral_of_profile():
    """Closed form of Blobs.h:67-98 against numerical integration of the profile along the
    line (elsa test_Blobs.cpp:48-120 does the same with Simpson's rule)."""
    from scipy.integrate import quad

    for a in (1.0, 2.0, 3.0):
        for alpha in (2.5, 10.83):
            for m in (0, 1, 2):
                for s in np.linspace(0, a * 0.98, 9):
                    L = math.sqrt(a * a - s * s)
                    num = 2 * quad(lambda t: oracle.blob_evaluate(math.sqrt(t * t + s * s), a, alpha, m), 0, L)[0]
                    assert oracle.blob_projected(float(s), a, alpha, m) == pytest.approx(num, rel=1e-4, abs=1e-9)


def test_bspline_partition_of_unity():
    """Cardinal B-splines sum to one over integer shifts; a wrong binomial or sign breaks it.
    (Off the knots: the order-0 box is closed at +-1/2 in elsa's definition.)"""
    for order in range(0, 5):
        lut = oracle.bspline_lut(2, order, 2001, np.float64)
        sup = (order + 1) * 0.5
        for x in np.linspace(0, 1, 8):
            tot = sum(oracle.lut_eval(lut, sup, abs(x - k)) for k in range(-4, 5))
            assert tot == pytest.approx(1.0, abs=2e-3), (order, x)


# --------------------------------------------------------------------------------------
# geometry


def _closed_form_pose(dim, s2c, c2d, vol, det, gamma):
    """P, Pinv, C from the closed forms of the circle geometry, in double."""
    if dim == 2:
        c, s = math.cos(gamma), math.sin(gamma)
        R = np.array([[c, -s], [s, c]])
    else:
        c, s = math.cos(gamma), math.sin(gamma)
        R = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    org = 0.5 * np.asarray(vol, float)
    t = R @ (-org)
    t[-1] += s2c
    K = np.eye(dim)
    for i in range(dim - 1):
        K[i, i] = s2c + c2d
        K[i, -1] = 0.5 * det[i]
    P = K @ np.hstack([R, t[:, None]])
    C = -R.T @ t
    return P, C


@pytest.mark.parametrize("dim", [2, 3])
def test_geometry_closed_forms(dim):
    vol = (16,) * dim
    det = (23,) * (dim - 1)
    for gamma in (0.0, 0.3, 1.7, -2.9):
        P, Pinv, C, ext, sdd = oracle.geometry_pose(dim, 250.0, 40.0, vol, det, gamma, dtype=np.float64)
        Pm = P.reshape(dim + 1, dim).T
        Pi = Pinv.reshape(dim, dim + 1).T
        Pc, Cc = _closed_form_pose(dim, 250.0, 40.0, vol, det, gamma)
        np.testing.assert_allclose(Pm, Pc, rtol=1e-12, atol=1e-9)
        np.testing.assert_allclose(C, Cc, rtol=1e-12, atol=1e-9)
        # Pinv inverts the 3x3 block: P[:, :d] @ Pinv[:d, :] = I
        np.testing.assert_allclose(Pm[:, :dim] @ Pi[:dim, :], np.eye(dim), atol=1e-12)
        # the camera centre projects to the homogeneous zero
        np.testing.assert_allclose(Pm @ np.append(C, 1.0), 0, atol=1e-9)
        assert sdd == 290.0
        # FP32 geometry agrees with FP64 to float rounding
        P32, Pinv32, C32, _, _ = oracle.geometry_pose(dim, 250.0, 40.0, vol, det, gamma, dtype=np.float32)
        np.testing.assert_allclose(P32, P, rtol=1e-5, atol=1e-4)
        np.testing.assert_allclose(C32, C, rtol=1e-6, atol=1e-4)


def test_geometry_kat_3d_rotations():
    """elsa/core/tests/test_Geometry.cpp: with gamma = pi/2 the source sits on +x at s2c
    (rotation about y), the principal ray runs along -x."""
    P, Pinv, C, ext, sdd = oracle.geometry_pose(3, 10.0, 5.0, (4, 4, 4), (3, 3), float(np.pi / 2), dtype=np.float64)
    # volume centre is (2, 2, 2); the camera centre lies s2c away from it
    assert np.linalg.norm(C - 2.0) == pytest.approx(10.0)
    o, d = C, (Pinv.reshape(3, 4).T @ np.array([1.5, 1.5, 1.0]))[:3]
    d = d / np.linalg.norm(d)
    # the central pixel's ray passes through the volume centre
    t = np.dot(np.array([2.0, 2.0, 2.0]) - o, d)
    assert np.linalg.norm(o + t * d - 2.0) < 1e-9


def test_circle_trajectory_linspaced_inclusive():
    """LinSpaced(n, 0, arc) includes both ends (CircleTrajectoryGenerator.cpp:16-19):
    for arc 360 the first and last poses coincide."""
    g = oracle.circle_geometry(2, (8, 8), 5, 360, 100.0, 20.0)
    np.testing.assert_allclose(g.C[0], g.C[-1], atol=1e-4)
    th = oracle.linspaced_f32(5, 360)
    assert list(th) == [0.0, 90.0, 180.0, 270.0, 360.0]
    # default detector floor(N*sqrt(2))
    assert oracle.circle_geometry(3, (256, 256, 256), 2, 180, 100.0, 20.0).det == (362, 362)


def test_intersection_kats():
    """elsa/projectors/tests/test_Intersection.cpp:21-150 (corners and edges of boxes)."""
    r2 = math.sqrt(0.5)
    assert oracle.intersect([-3, -3], [r2, r2], (1, 0)) is not None
    assert oracle.intersect([1, 2], [r2, -r2], (1, 0)) is None
    assert oracle.intersect([3, -2], [-r2, r2], (1, 0)) is not None
    assert oracle.intersect([3, 1], [-r2, -r2], (1, 0)) is None
    assert oracle.intersect([132, 30], [-1, 0], (30, 31)) is not None
    assert oracle.intersect([132, 30], [-1, 0], (40, 30)) is not None   # along the upper edge
    assert oracle.intersect([132, 30], [-1, 0], (40, 29)) is None
    assert oracle.intersect([30, -35], [0, 1], (31, 30)) is not None
    assert oracle.intersect([30, -35], [0, 1], (30, 40)) is not None
    assert oracle.intersect([30, -35], [0, 1], (29, 40)) is None
    assert oracle.intersect([1.5, -10], [0, 1], (0, 1)) is None
    assert oracle.intersect([1.5, -10], [0, 1], (1, 1)) is None
    assert oracle.intersect([1.5, -10], [0, 1], (2, 1)) is not None
    # entry/exit parameters of an axis ray
    assert oracle.intersect([-1.0, 0.5], [1.0, 0.0], (2, 2)) == pytest.approx((1.0, 3.0))
    assert oracle.intersect([1.0, 1.0], [1.0, 0.0], (2, 2)) == pytest.approx((-1.0, 1.0))


# --------------------------------------------------------------------------------------
# Joseph / Siddon / voxel on the oracle


@pytest.mark.parametrize("kat", kats.JOSEPH_KATS, ids=lambda k: k.__name__)
@pytest.mark.parametrize("mode", [oracle.MODE_ELSA, oracle.MODE_DIRECT], ids=["elsa", "direct"])
@pytest.mark.parametrize("dt", [np.float32, np.float64], ids=["f32", "f64"])
def test_joseph_kats_oracle(kat, mode, dt):
    kat(OracleBackend(dt, mode))


@pytest.mark.parametrize("kat", kats.JOSEPH_KATS_FP64_ONLY, ids=lambda k: k.__name__)
def test_joseph_tie_kats_oracle(kat):
    kat(OracleBackend(np.float64))
    with pytest.raises(AssertionError):
        kat(OracleBackend(np.float32))  # the exit lands exactly on the sample plane in FP32


@pytest.mark.parametrize("kat", kats.SIDDON_KATS, ids=lambda k: k.__name__)
@pytest.mark.parametrize("mode", [oracle.MODE_ELSA, oracle.MODE_DIRECT], ids=["elsa", "direct"])
@pytest.mark.parametrize("dt", [np.float32, np.float64], ids=["f32", "f64"])
def test_siddon_kats_oracle(kat, mode, dt):
    kat(OracleBackend(dt, mode))


@pytest.mark.parametrize("dim", [2, 3])
def test_siddon_direct_rule_close_to_repeated_addition(dim):
    """The two rules for the crossing parameters differ only by rounding: FP32 results agree to
    1e-5, and both are within FP32 rounding of the FP64 traversal."""
    n = 32 if dim == 3 else 64
    det = (45, 40) if dim == 3 else (90,)
    g = _rand_geom(dim, n, 9, det, np.float32, seed=11)
    g64 = _rand_geom(dim, n, 9, det, np.float64, seed=11)
    x = np.random.default_rng(12).random(g.vol_shape).astype(np.float32)
    a = oracle.siddon(g, x, mode=oracle.MODE_ELSA)
    b = oracle.siddon(g, x, mode=oracle.MODE_DIRECT)
    ref = oracle.siddon(g64, x.astype(np.float64), mode=oracle.MODE_DIRECT)
    rel = lambda u, v: np.linalg.norm(u - v) / np.linalg.norm(v)
    assert rel(a, b) <= 1e-5
    assert rel(b, ref) <= 1e-4 and rel(a, ref) <= 1e-4


@pytest.mark.parametrize("kat", kats.VOXEL_KATS, ids=lambda k: k.__name__)
@pytest.mark.parametrize("dt", [np.float32, np.float64], ids=["f32", "f64"])
def test_voxel_kats_oracle(kat, dt):
    kat(OracleBackend(dt))


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("adjoint", [False, True], ids=["fwd", "adj"])
def test_voxel_fp32_geometry_fp64_data(dim, adjoint):
    """The FP32-geometry / FP64-data voxel projector (elsa's VoxelProjectorCUDA<double>) visits
    the FP32 projector's footprints with the same weights: on FP32-valued data and table it
    equals the all-FP32 result to FP32 rounding of the sums, it is linear to FP64 rounding, and
    the all-FP64 geometry differs from it by the FP32 geometry's rounding only."""
    det = (23,) if dim == 2 else (19, 15)
    g = _rand_geom(dim, 12, 7, det, np.float32, seed=5, s2c=40.0)
    g64 = _rand_geom(dim, 12, 7, det, np.float64, seed=5, s2c=40.0)
    lut = oracle.blob_lut(2.0, 10.83, 2, 100, np.float32)
    rng = np.random.default_rng(6)
    x = rng.standard_normal(g.sino_shape if adjoint else g.vol_shape).astype(np.float32)
    f32 = oracle.voxel(g, x, lut, 2.0, adjoint=adjoint)
    f64 = oracle.voxel(g, x.astype(np.float64), lut.astype(np.float64), 2.0, adjoint=adjoint, acc64=True)
    rel = lambda u, v: np.linalg.norm(u - v) / np.linalg.norm(v)
    assert rel(f32, f64) <= 1e-6
    x2 = rng.standard_normal(x.shape)
    f2 = oracle.voxel(g, x2, lut.astype(np.float64), 2.0, adjoint=adjoint, acc64=True)
    f12 = oracle.voxel(g, x.astype(np.float64) + 3.0 * x2, lut.astype(np.float64), 2.0, adjoint=adjoint, acc64=True)
    assert rel(f12, f64 + 3.0 * f2) <= 1e-14
    full64 = oracle.voxel(g64, x.astype(np.float64), lut.astype(np.float64), 2.0, adjoint=adjoint)
    assert 0 < rel(f64, full64) <= 1e-4


def _rand_geom(dim, n, nviews, det, dtype, seed=0, s2c=None):
    rng = np.random.default_rng(seed)
    angles = np.sort(rng.uniform(0, 360, nviews)).astype(np.float32)
    return oracle.circle_geometry(dim, (n,) * dim, nviews, 360, s2c or 20.0 * n, 2.0 * n, det=det, dtype=dtype,
                                  angles_deg=angles)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kind", ["joseph", "siddon", "blob"])
def test_adjointness_fp64(dim, kind):
    """<Ax, y> = <x, A^T y> to FP64 rounding: a transposed tap or a dropped weight breaks it."""
    n = 12 if dim == 3 else 24
    det = (17, 15) if dim == 3 else (31,)
    g = _rand_geom(dim, n, 7, det, np.float64, seed=3)
    rng = np.random.default_rng(4)
    x = rng.standard_normal(g.vol_shape)
    y = rng.standard_normal(g.sino_shape)
    if kind == "joseph":
        A = lambda v: oracle.joseph(g, v)
        At = lambda v: oracle.joseph(g, v, adjoint=True)
    elif kind == "siddon":
        A = lambda v: oracle.siddon(g, v, mode=oracle.MODE_DIRECT)
        At = lambda v: oracle.siddon(g, v, adjoint=True, mode=oracle.MODE_DIRECT)
    else:
        lut = oracle.blob_lut(2.0, 10.83, 2, 100, np.float64)
        A = lambda v: oracle.voxel(g, v, lut, 2.0)
        At = lambda v: oracle.voxel(g, v, lut, 2.0, adjoint=True)
    lhs = np.vdot(A(x), y)
    rhs = np.vdot(x, At(y))
    assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(A(x)) * np.linalg.norm(y) * 10


def test_joseph_linearity_and_modes():
    g = _rand_geom(3, 10, 5, (15, 13), np.float32, seed=5)
    rng = np.random.default_rng(6)
    a = rng.standard_normal(g.vol_shape).astype(np.float32)
    b = rng.standard_normal(g.vol_shape).astype(np.float32)
    lhs = oracle.joseph(g, 2 * a + 3 * b)
    rhs = 2 * oracle.joseph(g, a) + 3 * oracle.joseph(g, b)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-4, atol=1e-4)
    # the two position rules agree to rounding of the accumulated step
    e = oracle.joseph(g, a, mode=oracle.MODE_ELSA)
    d = oracle.joseph(g, a, mode=oracle.MODE_DIRECT)
    assert np.linalg.norm(e - d) <= 1e-5 * np.linalg.norm(e)


def test_joseph_constant_volume_chord():
    """A constant volume projects to value x chord length up to the half-voxel boundary
    sampling: rays parallel to an axis through the box centre see exactly N samples."""
    n = 16
    g = oracle.geom_from_poses(2, (n, n), (1,), [oracle.geometry_pose(2, 1e5, 10.0, (n, n), (1,), 0.0,
                                                                      dtype=np.float64)], np.float64)
    y = oracle.joseph(g, np.full(g.vol_shape, 2.0))
    assert y.ravel()[0] == pytest.approx(2.0 * n, rel=1e-9)


def test_siddon_lengths_sum_to_chord():
    """Siddon: the intersection lengths of a ray sum to its chord through the box."""
    n = 9
    g = _rand_geom(3, n, 6, (13, 11), np.float64, seed=8)
    ones = oracle.siddon(g, np.ones(g.vol_shape), mode=oracle.MODE_DIRECT)
    axis, nsteps, pos0, step, ln = oracle.joseph_rays(g)
    for r in np.random.default_rng(9).choice(ones.size, 40, replace=False):
        view, rem = divmod(int(r), 13 * 11)
        v, u = divmod(rem, 13)
        o, d = oracle.make_ray(g.astype(np.float32), view, u, v)
        hit = oracle.intersect(o, d, (n, n, n))
        chord = 0.0 if hit is None else max(0.0, hit[1] - max(hit[0], 0.0))
        assert ones.ravel()[r] == pytest.approx(chord, abs=2e-4), r


def test_joseph_fp64_brute_force_tiny():
    """FP64 brute force: the forward value of each ray equals len * sum over the planes of the
    bilinear interpolation, recomputed here independently from the ray parameters."""
    n = 6
    g = _rand_geom(3, n, 3, (7, 5), np.float64, seed=10, s2c=30.0)
    rng = np.random.default_rng(11)
    x = rng.standard_normal(g.vol_shape)
    y = oracle.joseph(g, x, mode=oracle.MODE_DIRECT).ravel()
    axis, nsteps, pos0, step, ln = oracle.joseph_rays(g)

    def val(c):
        if all(0 <= c[i] < n for i in range(3)):
            return x[c[2], c[1], c[0]]
        return 0.0

    for r in range(y.size):
        acc = 0.0
        a = axis[r]
        bs = [i for i in range(3) if i != a]
        for s in range(nsteps[r]):
            p = pos0[r] + s * step[r]
            f = [int(math.floor(p[i] - 0.5)) for i in range(3)]
            t = [p[i] - 0.5 - f[i] for i in range(3)]
            for d1 in (0, 1):
                for d2 in (0, 1):
                    c = list(f)
                    c[bs[0]] += d1
                    c[bs[1]] += d2
                    w = (t[bs[0]] if d1 else 1 - t[bs[0]]) * (t[bs[1]] if d2 else 1 - t[bs[1]])
                    acc += w * val(c)
        assert y[r] == pytest.approx(acc * ln[r], rel=1e-9, abs=1e-9), r


@pytest.mark.parametrize("dim", [2, 3])
def test_joseph_exact_weights_brute_force_fp32_points(dim):
    """MODE_EXACT_WEIGHTS with FP64 data (the reference of the lerp-form forward): for an FP32
    geometry, the value of each ray equals len * sum over its samples of the bilinear (2D:
    linear) interpolation at the FP32 sample points fma(s, step, pos0), with the weights formed
    exactly, recomputed here in FP64 from the oracle's own FP32 ray states; and it differs from
    elsa's FP32 weight products (MODE_DIRECT, acc64) only by their rounding."""
    n = 7
    det = (9, 6) if dim == 3 else (10,)
    g = _rand_geom(dim, n, 4, det, np.float32, seed=20, s2c=40.0)
    rng = np.random.default_rng(21)
    x = rng.standard_normal(g.vol_shape)
    y = oracle.joseph(g, x, mode=oracle.MODE_EXACT_WEIGHTS, acc64=True).ravel()
    elsa_w = oracle.joseph(g, x, mode=oracle.MODE_DIRECT, acc64=True).ravel()
    assert np.linalg.norm(y - elsa_w) <= 1e-6 * np.linalg.norm(y)
    axis, nsteps, pos0, step, ln = oracle.joseph_rays(g)

    def val(c):
        if all(0 <= c[i] < n for i in range(dim)):
            return x[tuple(c[::-1])]
        return 0.0

    for r in range(y.size):
        acc = 0.0
        a = axis[r]
        bs = [i for i in range(dim) if i != a]
        for s in range(nsteps[r]):
            # fma(s, step, pos0) in FP32: the product and sum are exact in FP64, rounded once
            p = [np.float32(float(pos0[r, i]) + (s * float(step[r, i]) if i != a else 0.0)) if i != a
                 else np.float32(np.float32(pos0[r, i]) + np.float32(np.float32(s) * step[r, i])) for i in range(dim)]
            f = [int(math.floor(np.float32(p[i] - np.float32(0.5)))) for i in range(dim)]
            t = [float(p[i]) - 0.5 - f[i] for i in range(dim)]
            for d1 in (0, 1):
                for d2 in ((0, 1) if dim == 3 else (0,)):
                    c = list(f)
                    c[bs[0]] += d1
                    w = t[bs[0]] if d1 else 1 - t[bs[0]]
                    if dim == 3:
                        c[bs[1]] += d2
                        w *= t[bs[1]] if d2 else 1 - t[bs[1]]
                    acc += w * val(c)
        assert y[r] == pytest.approx(acc * float(ln[r]), rel=1e-12, abs=1e-12), r


# --------------------------------------------------------------------------------------
# solver restatements, pinned by closed forms on small dense problems


def _dense(seed, m=40, n=15):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((m, n))
    b = rng.standard_normal(m)
    return M, b, (lambda x: M @ x), (lambda y: M.T @ y)


def test_cgls_reaches_least_squares_and_ridge_solutions():
    M, b, A, At = _dense(1)
    x, res, taken = oracle.cgls(A, At, b, (15,), 40, tol=1e-24)
    ref = np.linalg.lstsq(M, b, rcond=None)[0]
    assert np.linalg.norm(x - ref) <= 1e-9 * np.linalg.norm(ref)
    assert np.all(np.diff(res) <= 1e-12)  # ||b - Ax|| never grows
    eps = 0.7
    xd, _, _ = oracle.cgls(A, At, b, (15,), 60, eps=eps, tol=1e-24)
    ridge = np.linalg.solve(M.T @ M + eps * eps * np.eye(15), M.T @ b)
    assert np.linalg.norm(xd - ridge) <= 1e-9 * np.linalg.norm(ridge)
    # a start value: the same minimiser
    x0 = np.random.default_rng(2).standard_normal(15)
    x2, _, _ = oracle.cgls(A, At, b, (15,), 40, tol=1e-24, x0=x0)
    assert np.linalg.norm(x2 - ref) <= 1e-9 * np.linalg.norm(ref)


def test_cgls_stops_on_tolerance():
    M, b, A, At = _dense(3)
    kold = float(np.sum((M.T @ b) ** 2))
    x, res, taken = oracle.cgls(A, At, b, (15,), 10, tol=2 * kold)
    assert taken == 0 and not np.any(x) and len(res) == 1


def test_power_iterations_and_apgd_closed_forms():
    d = np.array([5.0, 3.0, 1.0, 0.5])
    A = lambda x: np.sqrt(d) * x  # A^T A = diag(d)
    L = oracle.power_iterations(A, A, (4,), niters=60)
    assert abs(L - 5.0) <= 1e-9
    # A = I: the L1 problem's minimiser is soft thresholding of b
    b = np.array([3.0, -0.2, 0.5, -4.0, 0.0, 1.5])
    ident = lambda x: x
    x, gn, taken, mu = oracle.apgd(ident, ident, b, (6,), 200, mu=0.45, sigma=1.0, epsilon=0.0)
    assert np.allclose(x, np.sign(b) * np.maximum(np.abs(b) - 1.0, 0.0), atol=1e-12)
    # identity prox: least squares
    M, bb, A2, At2 = _dense(4)
    x2, gn2, taken2, mu2 = oracle.apgd(A2, At2, bb, (15,), 3000, epsilon=1e-26)
    ref = np.linalg.lstsq(M, bb, rcond=None)[0]
    assert np.linalg.norm(x2 - ref) <= 1e-6 * np.linalg.norm(ref)
    assert abs(mu2 - 0.45 / oracle.power_iterations(A2, At2, (15,), 5)) <= 1e-15


def test_admml2_closed_forms():
    """ADMML2 with the identity regulariser. elsa's x-update (RegularizedInversion.cpp) solves
    min 0.5 ||Ax - b||^2 + 0.5 ||lam x - (z - u)||^2 with lam = 1/sqrt(tau), while z and u
    track x itself (ADMML2.cpp:77-85). Its fixed points are therefore:
      * no prox (u -> 0, z = x): (A^T A + lam (lam - 1) I) x = A^T b, a ridge solution that is
        least squares exactly at tau = 1;
      * L1 prox, tau = 1: the optimality conditions of 0.5 ||Ax - b||^2 + sigma ||x||_1
        (A^T (b - A x) = sigma sign(x) where x != 0, |.| <= sigma where x = 0).
    Zero inner CGNE steps keep x at its start."""
    M, b, A, At = _dense(5)
    for tau in (0.9, 1.0, 0.5):
        lam = 1.0 / np.sqrt(tau)
        ref = np.linalg.solve(M.T @ M + lam * (lam - 1.0) * np.eye(15), M.T @ b)
        x, res = oracle.admml2(A, At, b, (15,), 400, tau=tau, ninner=5)
        assert np.linalg.norm(x - ref) <= 1e-9 * np.linalg.norm(ref), tau
    assert np.allclose(np.linalg.lstsq(M, b, rcond=None)[0], oracle.admml2(A, At, b, (15,), 400, tau=1.0)[0])
    sigma = 2.0
    x1, _ = oracle.admml2(A, At, b, (15,), 3000, tau=1.0, sigma=sigma, ninner=10)
    g = M.T @ (b - M @ x1)
    nz = np.abs(x1) > 1e-8
    assert nz.any() and (~nz).any()
    np.testing.assert_allclose(g[nz], sigma * np.sign(x1[nz]), atol=1e-6)
    assert np.all(np.abs(g[~nz]) <= sigma + 1e-6)
    x0 = np.ones(15)
    x2, _ = oracle.admml2(A, At, b, (15,), 3, ninner=0, x0=x0)
    np.testing.assert_array_equal(x2, x0)


@pytest.mark.parametrize("case", range(len(kats.FD_KATS)))
def test_finite_differences_kats(case):
    """elsa's FiniteDifferences test vectors (forward differences, every axis active), and the
    adjoint vector of its 1D test."""
    dims, data, fwd, adj = kats.FD_KATS[case]
    y = oracle.finite_differences(np.array(data, np.float64), dims)
    np.testing.assert_array_equal(y, np.array(fwd, np.float64))
    if adj is not None:
        np.testing.assert_array_equal(oracle.finite_differences(y, dims, adjoint=True), np.array(adj, np.float64))


@pytest.mark.parametrize("dims", [(7,), (4, 5), (3, 4, 5), (1, 6), (2, 1, 3)])
def test_finite_differences_adjoint_is_transpose(dims):
    """The adjoint is the transpose of the forward matrix, built column by column from unit
    vectors (in 3D this is where elsa's adjoint, which reads axis 1's block at offset n0 * n1,
    is not); constants lie in the null space; the forward is linear."""
    n = int(np.prod(dims))
    F = np.stack([oracle.finite_differences(np.eye(n)[i], dims) for i in range(n)], axis=1)
    G = np.stack([oracle.finite_differences(np.eye(len(dims) * n)[i], dims, adjoint=True)
                  for i in range(len(dims) * n)], axis=1)
    np.testing.assert_array_equal(G, F.T)
    assert set(np.unique(F)) <= {-1.0, 0.0, 1.0}
    rng = np.random.default_rng(5)
    x = rng.standard_normal(n)
    np.testing.assert_array_equal(F @ x, oracle.finite_differences(x, dims))
    # the last plane of each axis keeps -x alone: F @ 1 is -1 there and 0 elsewhere
    ones = oracle.finite_differences(np.ones(n), dims).reshape((len(dims),) + tuple(dims[::-1]))
    for d in range(len(dims)):
        last = np.moveaxis(ones[d], len(dims) - 1 - d, 0)
        assert np.all(last[-1] == -1.0) and np.all(last[:-1] == 0.0)


def test_admml2_finite_differences_closed_form():
    """ADMML2 with R = finite differences and no prox: u -> 0 and z = R x, so the fixed point
    solves (A^T A + lam (lam - 1) R^T R) x = A^T b with lam = 1/sqrt(tau) -- least squares at
    tau = 1. A transposed or unscaled regularisation block fails it."""
    M, b, A, At = _dense(6)
    dims = (5, 3)
    R = lambda v: oracle.finite_differences(v, dims)
    Rt = lambda y: oracle.finite_differences(y, dims, adjoint=True)
    F = np.stack([R(np.eye(15)[i]) for i in range(15)], axis=1)
    for tau in (0.5, 0.8):
        lam = 1.0 / np.sqrt(tau)
        ref = np.linalg.solve(M.T @ M + lam * (lam - 1.0) * F.T @ F, M.T @ b)
        x, _ = oracle.admml2(A, At, b, (15,), 600, tau=tau, ninner=5, reg=(R, Rt))
        assert np.linalg.norm(x - ref) <= 1e-9 * np.linalg.norm(ref), tau
    x1, _ = oracle.admml2(A, At, b, (15,), 300, tau=1.0, ninner=5, reg=(R, Rt))
    np.testing.assert_allclose(x1, np.linalg.lstsq(M, b, rcond=None)[0], rtol=1e-8, atol=1e-10)
    # with the identity as R the regularised form is the identity form, step for step
    I = lambda v: v
    xa, ra = oracle.admml2(A, At, b, (15,), 5, tau=0.7, sigma=0.1, ninner=3, reg=(I, I))
    xb, rb = oracle.admml2(A, At, b, (15,), 5, tau=0.7, sigma=0.1, ninner=3)
    np.testing.assert_array_equal(xa, xb)
    np.testing.assert_array_equal(ra, rb)


# --------------------------------------------------------------------------------------
# phase contrast: derivative tables and the differential weight


def _golden(name):
    import json
    import os

    return json.load(open(os.path.join(os.path.dirname(__file__), "golden", name)))


@pytest.mark.parametrize("dt", [np.float32, np.float64], ids=["f32", "f64"])
def test_blob_derivative_lut_golden(dt):
    """ProjectedBlob derivative table against test_BlobDerivative.cpp's vectors (a = 1 at
    s = i/100 and a = 2 at s = i/50, where s/a is the same), and the normalized gradient
    against the same numbers divided by s (test_BlobGradientHelper.cpp), doctest Approx."""
    gold = _golden("derivatives.json")["blob_derivative"]
    for t in gold["tables"]:
        exp = t["expected"]
        for a in (1.0, 2.0):
            d = oracle.blob_derivative_lut(a, t["alpha"], t["m"], 101, dt)
            g = oracle.blob_derivative_lut(a, t["alpha"], t["m"], 101, dt, normalized=True)
            for i in range(101):
                assert kats.approx(d[i], exp[i]), (a, t["alpha"], t["m"], i, d[i], exp[i])
                if i >= 1:
                    s = i * a / 100.0
                    assert kats.approx(g[i], exp[i] / s), (a, t["alpha"], t["m"], i)


@pytest.mark.parametrize("dt", [np.float32, np.float64], ids=["f32", "f64"])
def test_bspline_derivative_lut_golden(dt):
    """ProjectedBSpline(2, 3) derivative table against test_BSplines.cpp:144-164."""
    gold = _golden("derivatives.json")["bspline_derivative_lut"]
    lut = oracle.bspline_derivative_lut(2, 3, 101, dt)
    for i, e in enumerate(gold["expected"]):
        assert kats.approx(lut[i], e), (i, lut[i], e)


def test_derivative_luts_are_derivatives():
    """The derivative tables are the derivatives of the projected basis functions (central
    differences of the projected tables), the 3D normalized gradient times s is the 3D
    derivative, and the 3D B-spline derivative is the 2D one times b(0): a dropped factor or a
    wrong sign fails one of these."""
    n, h = 2001, None
    for a, alpha, m in ((2.0, 10.83, 2), (1.0, 2.5, 1), (1.5, 7.5, 2)):
        f = oracle.blob_lut(a, alpha, m, n, np.float64)
        d = oracle.blob_derivative_lut(a, alpha, m, n, np.float64)
        h = a / (n - 1)
        cd = (f[2:] - f[:-2]) / (2 * h)
        assert np.max(np.abs(cd[5:-40] - d[1:-1][5:-40])) <= 1e-4 * np.max(np.abs(d))
        g = oracle.blob_derivative_lut(a, alpha, m, n, np.float64, normalized=True)
        s = np.arange(n) * h
        np.testing.assert_allclose(g[1:] * s[1:], d[1:], rtol=1e-12, atol=1e-14)
    for dim in (2, 3):
        for order in (1, 2, 3):
            sup = (order + 1) * 0.5
            f = oracle.bspline_lut(dim, order, n, np.float64)
            d = oracle.bspline_derivative_lut(dim, order, n, np.float64)
            h = sup / (n - 1)
            cd = (f[2:] - f[:-2]) / (2 * h)
            # away from the knots (multiples of 1/2), where lower orders have kinks
            x = np.arange(1, n - 1) * h
            ok = np.abs(x * 2 - np.round(x * 2)) > 4 * h
            assert np.max(np.abs(cd - d[1:-1])[ok]) <= 1e-5, (dim, order)
            g = oracle.bspline_derivative_lut(dim, order, n, np.float64, normalized=True)
            np.testing.assert_allclose(g[1:] * (np.arange(1, n) * h), d[1:], rtol=1e-12, atol=1e-14)
        b0 = oracle.bspline_lut(2, 2, 3, np.float64)[0]  # b(0) of order 2 (2D table is b itself)
        np.testing.assert_allclose(oracle.bspline_derivative_lut(3, 2, 101, np.float64),
                                   oracle.bspline_derivative_lut(2, 2, 101, np.float64) * b0, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("kat", kats.PHASE_CONTRAST_KATS, ids=lambda k: k.__name__)
@pytest.mark.parametrize("dt", [np.float32, np.float64], ids=["f32", "f64"])
def test_phase_contrast_kats_oracle(kat, dt):
    kat(OracleBackend(dt))


@pytest.mark.parametrize("dim", [2, 3])
def test_phase_contrast_adjoint_and_antisymmetry(dim):
    """The differential projector's adjoint is matched, it is linear in the table, and it
    differs from the classic projector with the same table."""
    n = 9 if dim == 3 else 15
    det = (11, 9) if dim == 3 else (17,)
    g = _rand_geom(dim, n, 4, det, np.float64, seed=8)
    lut = oracle.blob_derivative_lut(2.0, 10.83, 2, 100, np.float64, normalized=dim == 3)
    rng = np.random.default_rng(9)
    x = rng.standard_normal(g.vol_shape)
    y = rng.standard_normal(g.sino_shape)
    A = oracle.voxel(g, x, lut, 2.0, differential=True)
    At = oracle.voxel(g, y, lut, 2.0, adjoint=True, differential=True)
    assert abs(np.vdot(A, y) - np.vdot(x, At)) <= 1e-12 * np.linalg.norm(A) * np.linalg.norm(y)
    np.testing.assert_allclose(oracle.voxel(g, x, -lut, 2.0, differential=True), -A, rtol=0, atol=0)
    assert np.linalg.norm(oracle.voxel(g, x, lut, 2.0) - A) > 1e-3 * np.linalg.norm(A)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kind", ["blob", "bspline"])
def test_phase_contrast_is_u_derivative_of_classic(dim, kind):
    """For one voxel, the phase-contrast projection is the derivative of the classic
    projection with respect to the detector u coordinate in units of the scaled distance:
    shifting the detector by +-delta pixels (principal point offset) moves d_u by
    +-delta / scaling, so (classic(+delta) - classic(-delta)) * scaling / (2 delta) must equal
    the differential projection (2D: g'(|d|) sgn(d_u), 3D: g'(|d|) / |d| * d_u). Fine tables
    (2001 entries) keep the interpolation error small."""
    n = 7
    vol = (n,) * dim
    det = (13, 11) if dim == 3 else (13,)
    N = 2001
    if kind == "blob":
        sup = 2.0
        classic = oracle.blob_lut(2.0, 10.83, 2, N, np.float64)
        diff = oracle.blob_derivative_lut(2.0, 10.83, 2, N, np.float64, normalized=dim == 3)
    else:
        sup = 1.5
        classic = oracle.bspline_lut(dim, 2, N, np.float64)
        diff = oracle.bspline_derivative_lut(dim, 2, N, np.float64, normalized=dim == 3)
    delta = 1e-3
    for gamma in (0.3, 1.9):
        for vox in ((3, 3, 3)[:dim], (2, 4, 5)[:dim]):
            x = np.zeros(vol[::-1]); x[vox[::-1]] = 1.0

            def geom(off):
                ppo = np.array([off] + [0.0] * (dim - 2))
                p = oracle.geometry_pose(dim, 60.0, 6.0, vol, det, angle_rad=gamma, ppo=ppo, dtype=np.float64)
                return oracle.geom_from_poses(dim, vol, det, [p], np.float64)

            g0 = geom(0.0)
            hv = np.array([v + 0.5 for v in vox] + [1.0])
            E = g0.ext[0].reshape(dim + 1, dim).T
            scaling = g0.sdd[0] / np.linalg.norm(E @ hv)
            fd = (oracle.voxel(geom(delta), x, classic, sup) - oracle.voxel(geom(-delta), x, classic, sup))
            fd *= scaling / (2 * delta)
            pc = oracle.voxel(g0, x, diff, sup, differential=True)
            assert np.max(np.abs(fd - pc)) <= 2e-3 * np.max(np.abs(pc)), (kind, dim, gamma, vox)
