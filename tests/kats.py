"""Known-answer tests transcribed from elsa's projector tests, run against any back end.

Sources (paths relative to the elsa tree):
  Joseph  elsa/projectors/tests/test_JosephsMethod.cpp
  Siddon  elsa/projectors/tests/test_SiddonsMethod.cpp and
          elsa/projectors_cuda/tests/test_SiddonsMethodCUDA.cpp
  voxel   elsa/projectors_cuda/tests/test_VoxelProjectorCUDA.cpp
  phase contrast  elsa/projectors_cuda/tests/test_PhaseContrastVoxelProjectorCUDA.cpp
Volumes are indexed vol[z, y, x] here (elsa's volume(x, y, z) with x fastest).
Tolerances are those of the reference tests: doctest Approx (relative 100*FLT_EPSILON on
scale 1), isApprox (||a-b|| <= prec*min(||a||,||b||), prec 1e-5 or the test's epsilon 0.01)
and isCwiseApprox (elementwise relative 0.01).
"""
from __future__ import annotations

import math

import numpy as np

PI = np.float32(np.pi)
EPS = 0.01  # elsa test_routines/testHelpers.h:20


def approx(a, b, eps=100 * np.finfo(np.float32).eps):
    """doctest::Approx(b).epsilon(eps) == a."""
    a, b = float(a), float(b)
    return abs(a - b) < eps * (1.0 + max(abs(a), abs(b)))


def is_approx(x, y, prec=1e-5):
    """elsa isApprox (testHelpers.cpp:20-34)."""
    x = np.asarray(x, np.float64).ravel()
    y = np.asarray(y, np.float64).ravel()
    return np.linalg.norm(x - y) <= prec * min(np.linalg.norm(x), np.linalg.norm(y))


def is_zero(x, eps=EPS):
    """isApprox against an all-zero container: holds only for an exactly zero result."""
    return bool(np.all(np.asarray(x) == 0))


def cwise_approx(x, y, eps=EPS):
    x = np.asarray(x, np.float64).ravel()
    y = np.asarray(y, np.float64).ravel()
    return all(approx(a, b, eps) for a, b in zip(x, y))


def pose(s2c, c2d, gamma=0.0, beta=0.0, alpha=0.0, ppo=None, cor=None):
    return {"s2c": float(s2c), "c2d": float(c2d), "gamma": float(gamma), "beta": float(beta), "alpha": float(alpha),
            "ppo": ppo, "cor": cor}


def deg(d):
    """Degree{d}.to_radian() in float: d * pi_t / 180."""
    return float(np.float32(np.float32(np.float32(d) * PI) / np.float32(180)))


# --------------------------------------------------------------------------------------
# Joseph, test_JosephsMethod.cpp


def joseph_single_ray_2d(B):
    """test_JosephsMethod.cpp:17-202: 5x5 ones, one ray, s2c 100, c2d 5."""
    x = np.ones((5, 5))
    col = np.zeros((5, 5)); col[:, 2] = 5
    row = np.zeros((5, 5)); row[2, :] = 5
    diag = np.diag(np.full(5, 10.0))
    for angle, expect, cmp, tol in [(0.0, 5, col, None), (deg(180), 5, col, 1e-5), (deg(90), 5, row, 1e-5),
                                    (deg(270), 5, row, 1e-5), (deg(45), None, diag, EPS), (deg(225), None, diag, EPS)]:
        g = B.setup((5, 5), (1,), [pose(100, 5, angle)])
        y = B.joseph(g, x)
        if expect is not None:
            assert approx(y.ravel()[0], expect), (angle, y)
        back = B.joseph(g, y, adjoint=True)
        assert (cwise_approx(back, cmp) if tol is None else is_approx(back, cmp, tol)), (angle, back)


def joseph_four_rays_2d(B):
    """test_JosephsMethod.cpp:205-265: poses 0, 90, 180, 270 degrees."""
    g = B.setup((5, 5), (1,), [pose(100, 5, deg(a)) for a in (0, 90, 180, 270)])
    y = B.joseph(g, np.ones((5, 5)))
    back = B.joseph(g, y, adjoint=True)
    cmp = np.array([0, 0, 10, 0, 0, 0, 0, 10, 0, 0, 10, 10, 20, 10, 10, 0, 0, 10, 0, 0, 0, 0, 10, 0, 0],
                   np.float64).reshape(5, 5)
    assert is_approx(back, cmp), back


def joseph_rays_missing(B):
    """test_JosephsMethod.cpp:438-640: rays beside the volume give zero both ways."""
    n = 5
    cases2d = [(0.0, (n, 0)), (0.0, (-n, 0)), (float(PI / 2), (0, n)), (float(PI / 2), (0, -n))]
    for ang, cor in cases2d:
        g = B.setup((n, n), (1,), [pose(20 * n, n, ang, cor=cor)])
        assert is_zero(B.joseph(g, np.ones((n, n)))), (ang, cor)
        assert is_zero(B.joseph(g, np.ones((1, 1)), adjoint=True)), (ang, cor)
    h = float(PI / 2)
    beta = [0, 0, 0, 0, 0, 0, h, h, h]
    gamma = [0, 0, 0, h, h, h, h, h, h]
    ox = [n, 0, n, 0, 0, 0, n, 0, n]
    oy = [0, n, n, n, 0, n, 0, 0, 0]
    oz = [0, 0, 0, 0, n, n, 0, n, n]
    for i in range(9):
        g = B.setup((n, n, n), (1, 1), [pose(20 * n, n, gamma[i], beta[i], 0.0, ppo=(0, 0),
                                              cor=(-ox[i], -oy[i], -oz[i]))])
        assert is_zero(B.joseph(g, np.ones((n, n, n)))), i
        assert is_zero(B.joseph(g, np.ones((1, 1, 1)), adjoint=True)), i


def joseph_axis_aligned_2d(B):
    """test_JosephsMethod.cpp:642-828: axis-aligned rays through and beside pixel centres."""
    n = 5
    angles = [0.0, float(PI / 2), float(PI), float(3 * PI / 2)]
    bp = [np.zeros((n, n)), np.zeros((n, n))]
    bp[0][:, 2] = 1
    bp[1][2, :] = 1
    for i, a in enumerate(angles):
        g = B.setup((n, n), (1,), [pose(20 * n, n, a)])
        for j in range(n):
            v = np.zeros((n, n))
            if i % 2 == 0:
                v[j, n // 2] = 1
            else:
                v[n // 2, j] = 1
            y = B.joseph(g, v)
            assert approx(y.ravel()[0], 1), (i, j, y)
        assert is_approx(B.joseph(g, y, adjoint=True), bp[i % 2]), i
    offx = [0.25, 0.0, 0.25, 0.0]
    offy = [0.0, 0.25, 0.0, 0.25]
    bp[0] = np.array([0, 0.25, 0.75, 0, 0] * 5, np.float64).reshape(n, n)
    bp[1] = np.array([0] * 5 + [0.25] * 5 + [0.75] * 5 + [0] * 10, np.float64).reshape(n, n)
    for i, a in enumerate(angles):
        g = B.setup((n, n), (1,), [pose(n * 2000, n, a, ppo=0.0, cor=(-offx[i], -offy[i]))])
        for j in range(n):
            v = np.zeros((n, n))
            if i % 2 == 0:
                v[j, n // 2] = 1
            else:
                v[n // 2, j] = 1
            assert approx(B.joseph(g, v).ravel()[0], 0.75), (i, j)
        back = B.joseph(g, np.ones((1, 1)), adjoint=True)
        assert is_approx(back, bp[i % 2]), (i, back)
    # rays along the right / left volume boundary
    right = np.zeros((n, n)); right[:, n - 1] = 0.5
    left = np.zeros((n, n)); left[:, 0] = 0.5
    for cor, col, expect in [((n * 0.5, 0), n - 1, right), ((-n / 2.0, 0), 0, left)]:
        g = B.setup((n, n), (1,), [pose(n * 2000, n, 0.0, ppo=0.0, cor=cor)])
        for j in range(n):
            v = np.zeros((n, n)); v[j, col] = 1
            assert approx(B.joseph(g, v).ravel()[0], 0.5), (cor, j)
        assert is_approx(B.joseph(g, np.ones((1, 1)), adjoint=True), expect), cor


def joseph_non_square_2d(B):
    """test_JosephsMethod.cpp:830-883: 5x2 volume, ray through the short dimension."""
    g = B.setup((5, 2), (1,), [pose(100, 5, 0.0, ppo=0.0, cor=(1.6, 0))])
    v = np.zeros((2, 5)); v[0, 4] = 1; v[1, 4] = 1
    assert approx(B.joseph(g, v).ravel()[0], 1.2)
    exp = np.zeros((2, 5)); exp[:, 4] = 0.6; exp[:, 3] = 0.4
    assert is_approx(B.joseph(g, np.ones((1, 1)), adjoint=True), exp)


def joseph_single_ray_3d(B):
    """test_JosephsMethod.cpp:885-1090: axis-aligned rays in a 3^3 volume, centred and offset."""
    n = 3
    h = float(PI / 2)
    beta = [0.0, 0.0, 0.0, 0.0, h, float(3 * PI / 2)]
    gamma = [0.0, float(PI), h, float(3 * PI / 2), h, float(3 * PI / 2)]

    def placed(i, j):
        v = np.zeros((n, n, n))
        if i // 2 == 0:
            v[j, 1, 1] = 1      # volume(1, 1, j)
        elif i // 2 == 1:
            v[1, 1, j] = 1      # volume(j, 1, 1)
        else:
            v[1, j, 1] = 1      # volume(1, j, 1)
        return v

    bp = [np.zeros(27) for _ in range(3)]
    bp[2][[10, 13, 16]] = 1
    bp[1][[12, 13, 14]] = 1
    bp[0][[4, 13, 22]] = 1
    for i in range(6):
        g = B.setup((n, n, n), (1, 1), [pose(20 * n, n, gamma[i], beta[i])])
        for j in range(n):
            y = B.joseph(g, placed(i, j))
            assert approx(y.ravel()[0], 1.0), (i, j, y)
        assert is_approx(B.joseph(g, y, adjoint=True), bp[i // 2].reshape(n, n, n)), i
    ox = [0.25, 0.25, 0.0, 0.0, 0.0, 0.0]
    oy = [0.0, 0.0, 0.25, 0.25, 0.0, 0.0]
    oz = [0.0, 0.0, 0.0, 0.0, 0.25, 0.25]
    bp[2] = np.array([0, .25, 0, 0, .25, 0, 0, .25, 0, 0, .75, 0, 0, .75, 0, 0, .75, 0] + [0] * 9, np.float64)
    bp[1] = np.array([0] * 9 + [.25, .25, .25, .75, .75, .75, 0, 0, 0] + [0] * 9, np.float64)
    bp[0] = np.array([0, 0, 0, .25, .75, 0, 0, 0, 0] * 3, np.float64)
    for i in range(6):
        g = B.setup((n, n, n), (1, 1), [pose(n * 2000, n, gamma[i], beta[i], ppo=(0, 0), cor=(-ox[i], -oy[i], -oz[i]))])
        for j in range(n):
            assert approx(B.joseph(g, placed(i, j)).ravel()[0], 0.75), (i, j)
        assert is_approx(B.joseph(g, np.ones((1, 1, 1)), adjoint=True), bp[i // 2].reshape(n, n, n)), i


def joseph_angle_30_2d(B):
    """test_JosephsMethod.cpp:1296-1391: -30 degree ray through the centre and through a corner."""
    n = 4
    g = B.setup((n, n), (1,), [pose(20 * n, n, float(-PI / 6))])
    w = float(np.float32(2 / np.sqrt(np.float32(3))))
    v = np.ones((n, n))
    for (xx, yy) in [(3, 0), (1, 1), (1, 2), (1, 3), (2, 0), (2, 1), (2, 2), (0, 3)]:
        v[yy, xx] = 0
    assert is_zero(B.joseph(g, v))
    for (xx, yy) in [(3, 0), (1, 1), (1, 2), (1, 3)]:
        v[yy, xx] = 1
    assert approx(B.joseph(g, v).ravel()[0], 2 * w)
    s3 = math.sqrt(3.0)
    exp = np.array([0, 0, (3 - s3) / 2, (s3 - 1) / 2, 0, (s3 - 1) / (2 * s3), (s3 + 1) / (2 * s3), 0, 0,
                    (s3 + 1) / (2 * s3), (s3 - 1) / (2 * s3), 0, (s3 - 1) / 2, (3 - s3) / 2, 0, 0]) * w
    assert is_approx(B.joseph(g, np.ones((1, 1)), adjoint=True), exp.reshape(n, n))


def joseph_corner_ray_2d(B):
    """test_JosephsMethod.cpp:1353-1391: -30 degree ray that only clips pixel (0, 0). The ray
    leaves the volume exactly on the plane y = 0.5 of the one sample, so whether that sample
    exists (numSteps = ceil(|exit_y - 0.5|)) hangs on the last bit of the camera centre; elsa
    computes that with a float QR solve, this build with a double solve rounded to float, and
    in FP32 the exit lands exactly on 0.5 (no sample). FP64 takes the reference's branch."""
    n = 4
    g = B.setup((n, n), (1,), [pose(20 * n, n, float(-PI / 6), ppo=0.0, cor=(-2 - math.sqrt(3.0) / 2, 0))])
    v = np.ones((n, n)); v[0, 0] = 0
    assert is_zero(B.joseph(g, v))
    v[0, 0] = 1
    assert approx(B.joseph(g, v).ravel()[0], 1 / math.sqrt(3.0), EPS)
    exp = np.zeros(16); exp[0] = 1 / math.sqrt(3.0)
    assert is_approx(B.joseph(g, np.ones((1, 1)), adjoint=True), exp.reshape(n, n))


JOSEPH_KATS = [joseph_single_ray_2d, joseph_four_rays_2d, joseph_rays_missing, joseph_axis_aligned_2d,
               joseph_non_square_2d, joseph_single_ray_3d, joseph_angle_30_2d]
# cases whose expected value depends on an exact floating-point tie (see the docstring)
JOSEPH_KATS_FP64_ONLY = [joseph_corner_ray_2d]


# --------------------------------------------------------------------------------------
# Siddon, elsa/projectors_cuda/tests/test_SiddonsMethodCUDA.cpp (shared CPU/GPU KATs)


def siddon_axis_aligned_2d(B):
    """test_SiddonsMethodCUDA.cpp:284-447."""
    n = 5
    bp = [np.zeros((n, n)), np.zeros((n, n))]
    bp[0][:, 2] = 1
    bp[1][2, :] = 1
    for i, a in enumerate((0, 90, 180, 270)):
        g = B.setup((n, n), (1,), [pose(20 * n, n, deg(a))])
        for j in range(n):
            v = np.zeros((n, n))
            if i % 2 == 0:
                v[j, n // 2] = 1
            else:
                v[n // 2, j] = 1
            y = B.siddon(g, v)
            assert approx(y.ravel()[0], 1), (a, j, y)
        assert is_approx(B.siddon(g, y, adjoint=True), bp[i % 2]), a
    # along the left voxel boundary of the centre column: the higher index wins
    g = B.setup((n, n), (1,), [pose(n * 2000, n, 0.0, ppo=0.0, cor=(-0.5, 0))])
    for j in range(n):
        v = np.zeros((n, n)); v[j, n // 2] = 1
        assert approx(B.siddon(g, v).ravel()[0], 1.0), j
    assert is_approx(B.siddon(g, np.ones((1, 1)), adjoint=True), bp[0])
    # along the right volume boundary: ignored
    g = B.setup((n, n), (1,), [pose(n * 2000, n, 0.0, ppo=0.0, cor=(n * 0.5, 0))])
    assert approx(B.siddon(g, np.ones((n, n))).ravel()[0], 0)
    assert is_zero(B.siddon(g, np.ones((1, 1)), adjoint=True))
    # along the left volume boundary: the first column
    g = B.setup((n, n), (1,), [pose(n * 2000, n, 0.0, ppo=0.0, cor=(-n * 0.5, 0))])
    left = np.zeros((n, n)); left[:, 0] = 1
    for j in range(n):
        v = np.zeros((n, n)); v[j, 0] = 1
        assert approx(B.siddon(g, v).ravel()[0], 1), j
    assert is_approx(B.siddon(g, np.ones((1, 1)), adjoint=True), left)


def siddon_angle_30_2d(B):
    """test_SiddonsMethodCUDA.cpp:478-597: -30 degrees through the centre and entering through
    the right border (exact intersection lengths)."""
    n = 4
    s3 = math.sqrt(3.0)
    g = B.setup((n, n), (1,), [pose(20 * n, n, deg(-30))])
    v = np.zeros((n, n))
    v[0, 3] = 1; v[0, 2] = 2; v[1, 2] = 3
    assert approx(B.siddon(g, v).ravel()[0], 2 * s3 + 2)
    v = np.zeros((n, n))
    v[2, 1] = 3; v[3, 1] = 2; v[3, 0] = 1
    assert approx(B.siddon(g, v).ravel()[0], 2 * s3 + 2)
    exp = np.array([0, 0, 2 - 2 / s3, 4 / s3 - 2, 0, 0, 2 / s3, 0, 0, 2 / s3, 0, 0, 4 / s3 - 2, 2 - 2 / s3, 0, 0])
    assert is_approx(B.siddon(g, np.ones((1, 1)), adjoint=True), exp.reshape(n, n), EPS)
    g = B.setup((n, n), (1,), [pose(20 * n, n, deg(-30), ppo=0.0, cor=(float(np.sqrt(np.float32(3))), 0))])
    v = np.ones((n, n))
    for (xx, yy) in [(3, 1), (3, 2), (3, 3), (2, 3)]:
        v[yy, xx] = 0
    assert is_zero(B.siddon(g, v))
    for (xx, yy), val in zip([(3, 1), (3, 2), (3, 3), (2, 3)], (4, 3, 2, 1)):
        v[yy, xx] = val
    assert approx(B.siddon(g, v).ravel()[0], 14 - 4 * s3)
    exp = np.array([0, 0, 0, 0, 0, 0, 0, 4 - 2 * s3, 0, 0, 0, 2 / s3, 0, 0, 2 - 2 / s3, 4 / s3 - 2])
    assert is_approx(B.siddon(g, np.ones((1, 1)), adjoint=True), exp.reshape(n, n), EPS)


def siddon_multiple_rays_2d(B):
    """test_SiddonsMethodCUDA.cpp:928-1000: four axis-aligned poses."""
    n = 5
    g = B.setup((n, n), (1,), [pose(20 * n, n, deg(a)) for a in (0, 90, 180, 270)])
    v = np.zeros((n, n))
    v[:, n // 2] = 1
    v[n // 2, :] = 1
    y = B.siddon(g, v)
    assert all(approx(s, 5.0) for s in y.ravel()), y
    cmp = np.array([0, 0, 10, 0, 0, 0, 0, 10, 0, 0, 10, 10, 20, 10, 10, 0, 0, 10, 0, 0, 0, 0, 10, 0, 0],
                   np.float64).reshape(n, n)
    assert is_approx(B.siddon(g, y, adjoint=True), cmp)


def siddon_axis_aligned_3d(B):
    """test_SiddonsMethodCUDA.cpp:1005-1120: empty inputs and axis-aligned rays in 3^3."""
    n = 3
    g = B.setup((n, n, n), (1, 1), [pose(20 * n, n, 0.0)])
    assert is_zero(B.siddon(g, np.zeros((n, n, n))))
    assert is_zero(B.siddon(g, np.zeros((1, 1, 1)), adjoint=True))
    h = float(PI / 2)
    beta = [0.0, 0.0, 0.0, 0.0, h, float(3 * PI / 2)]
    gamma = [0.0, float(PI), h, float(3 * PI / 2), h, float(3 * PI / 2)]
    bp = [np.zeros(27) for _ in range(3)]
    bp[0][[4, 13, 22]] = 1
    bp[1][[12, 13, 14]] = 1
    bp[2][[10, 13, 16]] = 1
    for i in range(6):
        g = B.setup((n, n, n), (1, 1), [pose(20 * n, n, gamma[i], beta[i])])
        for j in range(n):
            v = np.zeros((n, n, n))
            if i // 2 == 0:
                v[j, 1, 1] = 1
            elif i // 2 == 1:
                v[1, 1, j] = 1
            else:
                v[1, j, 1] = 1
            y = B.siddon(g, v)
            assert approx(y.ravel()[0], 1.0), (i, j, y)
        assert is_approx(B.siddon(g, y, adjoint=True), bp[i // 2].reshape(n, n, n)), i


SIDDON_KATS = [siddon_axis_aligned_2d, siddon_angle_30_2d, siddon_multiple_rays_2d, siddon_axis_aligned_3d]


# --------------------------------------------------------------------------------------
# Voxel projectors, elsa/projectors_cuda/tests/test_VoxelProjectorCUDA.cpp


def _blob0(B):
    import oracle

    return float(oracle.blob_lut(2.0, 10.83, 2, 100, B.dtype)[0])


def voxel_centre_voxel_3d(B):
    """test_VoxelProjectorCUDA.cpp:96-139: the centre voxel of 5^3 projects to blob(0) on a
    one-pixel detector for poses Gamma = 0, 4, 8, ... radians (epsilon 0.005)."""
    w0 = _blob0(B)
    x = np.zeros((5, 5, 5)); x[2, 2, 2] = 1
    for i in range(0, 360, 4):
        g = B.setup((5, 5, 5), (1, 1), [pose(100, 5, float(i))])
        y = B.voxel(g, x, "blob")
        assert approx(w0, y.ravel()[0], 0.005), (i, y)


def voxel_centre_voxel_2d(B):
    """test_VoxelProjectorCUDA.cpp:141-190 (unit spacing): centre voxel -> blob(0)."""
    w0 = _blob0(B)
    x = np.zeros((5, 5)); x[2, 2] = 1
    for i in range(0, 360, 4):
        g = B.setup((5, 5), (1,), [pose(100, 5, deg(i))])
        y = B.voxel(g, x, "blob")
        assert approx(w0, y.ravel()[0]), (i, y)


def voxel_backward_2d(B):
    """test_VoxelProjectorCUDA.cpp:945-1010: back projection of a single detector value: the
    central column gets blob(0), the neighbouring columns blob(d/s) of their projected
    distance (projectAndScaleVoxelOnDetector, DetectorDescriptor.cpp:130-151)."""
    import oracle

    g = B.setup((5, 5), (1,), [pose(100, 5, 0.0)])
    x = B.voxel(g, np.ones((1, 1)), "blob", adjoint=True)
    lut = oracle.blob_lut(2.0, 10.83, 2, 100, np.float64)
    w0 = _blob0(B)
    for s in range(5):
        assert approx(x[s, 2], w0), (s, x[s, 2])
    og = oracle.geom_from_poses(2, (5, 5), (1,), [oracle.geometry_pose(2, 100.0, 5.0, (5, 5), (1,), 0.0,
                                                                       dtype=np.float64)], np.float64)
    P = og.P[0].reshape(3, 2).T
    E = og.ext[0].reshape(3, 2).T
    for col in (1, 3):
        for s in range(5):
            hv = np.array([col + 0.5, s + 0.5, 1.0])
            pc = P @ hv
            det = pc[0] / pc[1] - 0.5
            scaling = og.sdd[0] / np.linalg.norm(E @ hv)
            w = oracle.lut_eval(lut, 2.0, abs(det / scaling))
            assert approx(x[s, col], w, 1e-4), (col, s, x[s, col], w)


# ---- the remaining 2D known answers of test_VoxelProjectorCUDA.cpp, for both basis functions
# (blob :191-1278, B-spline :1280-2462; the two halves are the same tests with the B-spline's
# table) and the CPU tests test_BlobVoxelProjector.cpp / test_BSplineVoxelProjector.cpp.


def _basis(kind, dtype):
    """op.blob(s) / op.bspline(s) of elsa (Blobs.h:266-272, BSplines.h:154-160): the 100-entry
    table at |s|, in the back end's dtype (2D tables)."""
    import oracle

    if kind == "blob":
        lut, sup = oracle.blob_lut(2.0, 10.83, 2, 100, dtype), 2.0
    else:
        lut, sup = oracle.bspline_lut(2, 2, 100, dtype), 1.5
    return lambda s: float(oracle.lut_eval(lut, sup, abs(float(s))))


def _projected_weight(kind, angle_deg, coord, spacing=None, dtype=np.float64):
    """The weight elsa's tests expect for voxel (x, y) on the one-pixel detector of the
    5^2 / s2c 100 / c2d 5 geometry: basis((u - 0.5) / scaling) with u and scaling from
    projectAndScaleVoxelOnDetector (DetectorDescriptor.cpp:130-151), evaluated in FP64."""
    import oracle

    og = oracle.geom_from_poses(2, (5, 5), (1,), [oracle.geometry_pose(2, 100.0, 5.0, (5, 5), (1,),
                                                                       float(np.radians(angle_deg)),
                                                                       vol_spacing=spacing, dtype=np.float64)],
                                np.float64)
    P = og.P[0].reshape(3, 2).T
    E = og.ext[0].reshape(3, 2).T
    hv = np.array([coord[0] + 0.5, coord[1] + 0.5, 1.0])
    pc = P @ hv
    u = pc[0] / pc[1] - 0.5
    return _basis(kind, dtype)(u / (og.sdd[0] / np.linalg.norm(E @ hv)))


def _vol2d(coords):
    """A 5^2 volume with the voxels elsa addresses as x(i, j) set to 1 (vol[j, i] here)."""
    x = np.zeros((5, 5))
    for i, j in coords:
        x[j, i] = 1
    return x


def _centre_voxel_spacing(B, kind):
    """:141-190 (blob) and :1325-1374 (B-spline): with volume spacing 0.5, 1 and 1.5 the centre
    voxel of 5^2 projects to basis(0) on a one-pixel detector, poses 0, 4, ..., 356 deg."""
    w0 = _basis(kind, B.dtype)(0.0)
    for k in (1, 2, 3):
        sp = (k / 2.0, k / 2.0)
        for i in range(0, 360, 4):
            g = B.setup((5, 5), (1,), [pose(100, 5, deg(i))], spacing=sp)
            y = B.voxel(g, _vol2d([(2, 2)]), kind)
            assert approx(w0, y.ravel()[0]), (sp, i, y)


def _two_pixels(B, kind):
    """:191-239 / :1375-1423: centre voxel, two-pixel detector: both pixels get
    basis(0.4761878) (epsilon 0.01), poses 0, 4, ..., 356 deg."""
    w = _basis(kind, B.dtype)(0.4761878)
    for i in range(0, 360, 4):
        g = B.setup((5, 5), (2,), [pose(100, 5, deg(i))])
        y = B.voxel(g, _vol2d([(2, 2)]), kind).ravel()
        assert approx(w, y[0], 0.01) and approx(w, y[1], 0.01), (i, y, w)


def _three_pixels(B, kind):
    """:240-296 / :1424-1480: three pixels, poses 0, 5, ..., 355 deg: the centre pixel gets
    basis(0), the outer ones basis(0.95233786) (epsilon 0.01)."""
    f = _basis(kind, B.dtype)
    w0, w1 = f(0.0), f(0.95233786)
    for i in range(0, 360, 5):
        g = B.setup((5, 5), (3,), [pose(100, 5, deg(i))])
        y = B.voxel(g, _vol2d([(2, 2)]), kind).ravel()
        assert approx(w0, y[1]), (i, y)
        assert approx(w1, y[0], 0.01) and approx(w1, y[2], 0.01), (i, y, w1)


# :297-944 / :1481-2128, one-pixel detector at the given pose: for each voxel set (elsa's
# x(i, j) coordinates) set to 1, the detector value equals the sum of the projected weights of
# its voxels; the sets on the ray are elsa's "5 * weight of distance 0" cases.
SINGLE_PIXEL_SETS = {
    0: [[(2, s) for s in range(5)], [(3, s) for s in range(5)], [(1, s) for s in range(5)]],
    45: [[(s, s) for s in range(5)], [(s, s + 1) for s in range(4)], [(s + 1, s) for s in range(4)],
         [(s, s + 2) for s in range(3)], [(s + 2, s) for s in range(3)]],
    90: [[(s, 2) for s in range(5)], [(s, 3) for s in range(5)], [(s, 1) for s in range(5)]],
    135: [[(s, 4 - s) for s in range(5)], [(s, 5 - s) for s in range(1, 5)], [(s, 3 - s) for s in range(4)],
          [(s, 6 - s) for s in range(2, 5)], [(s, 2 - s) for s in range(3)]],
    180: [[(2, s) for s in range(5)], [(3, s) for s in range(5)], [(1, s) for s in range(5)]],
    225: [[(s, s) for s in range(5)]],
    270: [[(s, 2) for s in range(5)]],
    315: [[(s, 4 - s) for s in range(5)]],
}


def _single_pixel(B, kind):
    f = _basis(kind, B.dtype)
    for a, sets in SINGLE_PIXEL_SETS.items():
        g = B.setup((5, 5), (1,), [pose(100, 5, deg(a))])
        for k, vs in enumerate(sets):
            y = float(B.voxel(g, _vol2d(vs), kind).ravel()[0])
            if k == 0:  # the voxels on the ray
                assert approx(y, 5 * f(0.0)), (a, y)
            exp = sum(_projected_weight(kind, a, v, dtype=B.dtype) for v in vs)
            assert approx(y, exp), (a, k, y, exp)


# :945-1278 / :2129-2462, back projection of a one-pixel detector holding 1: the voxels on the
# ray get basis(0), the listed voxels beside it their projected weight. Every voxel of the
# volume is checked against its projected weight as well (elsa lists 13-17 of them per pose).
BACKWARD_SETS = {
    0: [[(2, s) for s in range(5)], [(3, s) for s in range(5)], [(1, s) for s in range(5)]],
    180: [[(2, s) for s in range(5)], [(3, s) for s in range(5)], [(1, s) for s in range(5)]],
    90: [[(s, 2) for s in range(5)], [(s, 3) for s in range(5)], [(s, 1) for s in range(5)]],
    270: [[(s, 2) for s in range(5)], [(s, 3) for s in range(5)], [(s, 1) for s in range(5)]],
    45: [[(s, s) for s in range(5)], [(s, s + 1) for s in range(4)], [(s + 1, s) for s in range(4)],
         [(s, 2 - s) for s in range(3)], [(2 - s, s) for s in range(3)]],
}


def _backward(B, kind):
    f = _basis(kind, B.dtype)
    for a, sets in BACKWARD_SETS.items():
        g = B.setup((5, 5), (1,), [pose(100, 5, deg(a))])
        x = B.voxel(g, np.ones((1, 1)), kind, adjoint=True)
        for i, j in sets[0]:
            assert approx(x[j, i], f(0.0)), (a, i, j, x[j, i])
        for vs in sets[1:]:
            for i, j in vs:
                assert approx(x[j, i], _projected_weight(kind, a, (i, j), dtype=B.dtype)), (a, i, j, x[j, i])
        for j in range(5):
            for i in range(5):
                assert approx(x[j, i], _projected_weight(kind, a, (i, j), dtype=B.dtype)), (a, i, j, x[j, i])


def voxel_blob_centre_voxel_spacing_2d(B):
    _centre_voxel_spacing(B, "blob")


def voxel_bspline_centre_voxel_spacing_2d(B):
    _centre_voxel_spacing(B, "bspline")


def voxel_blob_two_pixels_2d(B):
    _two_pixels(B, "blob")


def voxel_bspline_two_pixels_2d(B):
    _two_pixels(B, "bspline")


def voxel_blob_three_pixels_2d(B):
    _three_pixels(B, "blob")


def voxel_bspline_three_pixels_2d(B):
    _three_pixels(B, "bspline")


def voxel_blob_single_pixel_2d(B):
    _single_pixel(B, "blob")


def voxel_bspline_single_pixel_2d(B):
    _single_pixel(B, "bspline")


def voxel_blob_backward_2d(B):
    _backward(B, "blob")


def voxel_bspline_backward_2d(B):
    _backward(B, "bspline")


def voxel_bspline_centre_voxel_3d(B):
    """test_VoxelProjectorCUDA.cpp:1280-1323: the centre voxel of 5^3 projects to bspline(0)
    (the 3D table) on a one-pixel detector, Gamma = 0, 4, 8, ... radians (epsilon 0.005)."""
    import oracle

    w0 = float(oracle.bspline_lut(3, 2, 100, B.dtype)[0])
    x = np.zeros((5, 5, 5)); x[2, 2, 2] = 1
    for i in range(0, 360, 4):
        g = B.setup((5, 5, 5), (1, 1), [pose(100, 5, float(i))])
        y = B.voxel(g, x, "bspline")
        assert approx(w0, y.ravel()[0], 0.005), (i, y)


def voxel_bspline_45_degree_cpu(B):
    """test_BSplineVoxelProjector.cpp:1860-1938: at 45 deg the voxels on the ray give
    5 bspline(0); the four voxels just above or just below it give 4 bspline(sqrt(1/2)) to
    epsilon 0.001 (their projected distance is sqrt(1/2) up to the fan's divergence)."""
    f = _basis("bspline", B.dtype)
    g = B.setup((5, 5), (1,), [pose(100, 5, deg(45))])
    y = float(B.voxel(g, _vol2d([(s, s) for s in range(5)]), "bspline").ravel()[0])
    assert approx(y, 5 * f(0.0)), y
    for vs in ([(s, s + 1) for s in range(4)], [(s + 1, s) for s in range(4)]):
        y = float(B.voxel(g, _vol2d(vs), "bspline").ravel()[0])
        assert approx(y, 4 * f(math.sqrt(0.5)), 0.001), (vs, y)


def voxel_off_centre_3d(B):
    """A closed form for the 3D footprint, independent of the projection matrices: one voxel
    off centre along v (and u) in a non-cubic volume, a distant source, a non-square detector,
    poses about y. With the pinhole model of Geometry.cpp:53-109 written out (camera
    coordinates x_c = R(c - origin) + s2c e_z, pixel coordinate sdd x_c / z_c + det / 2),
    every pixel (i, j) inside the footprint must hold blob(|(p - 0.5 - (i, j))| / s) with
    s = sdd / |x_c|, and every other pixel 0; the back projection of each such pixel alone
    gives the voxel that weight."""
    import oracle

    vol, det, s2c, c2d = (7, 9, 5), (13, 21), 400.0, 40.0
    sdd = s2c + c2d
    lut = oracle.blob_lut(2.0, 10.83, 2, 100, np.float64)
    f = lambda d: float(oracle.lut_eval(lut, 2.0, d))
    vox = (4, 7, 1)  # elsa's x(4, 7, 1): 1.5 voxels right of the axis, 3 above it
    x = np.zeros(vol[::-1]); x[vox[2], vox[1], vox[0]] = 1
    for gamma in (0.0, 0.5, 2.0, 4.0):
        cg, sg = math.cos(gamma), math.sin(gamma)
        R = np.array([[cg, 0, sg], [0, 1, 0], [-sg, 0, cg]])
        xc = R @ (np.array(vox) + 0.5 - 0.5 * np.array(vol)) + np.array([0, 0, s2c])
        p = sdd * xc[:2] / xc[2] + 0.5 * np.array(det) - 0.5
        s = sdd / np.linalg.norm(xc)
        exp = np.zeros((det[1], det[0]))
        for j in range(det[1]):
            for i in range(det[0]):
                d = math.hypot(p[0] - i, p[1] - j) / s
                if abs(p[0] - i) <= 2.0 * s and abs(p[1] - j) <= 2.0 * s:
                    exp[j, i] = f(d)
        assert np.count_nonzero(exp) >= 9, (gamma, p, s)
        g = B.setup(vol, det, [pose(s2c, c2d, gamma)])
        y = B.voxel(g, x, "blob").reshape(det[1], det[0])
        for j in range(det[1]):
            for i in range(det[0]):
                assert approx(y[j, i], exp[j, i], 1e-4), (gamma, i, j, y[j, i], exp[j, i])
        for j, i in zip(*np.nonzero(exp)):
            e = np.zeros((1, det[1], det[0])); e[0, j, i] = 1
            b = B.voxel(g, e, "blob", adjoint=True)
            assert approx(b[vox[2], vox[1], vox[0]], exp[j, i], 1e-4), (gamma, i, j)


VOXEL_KATS = [voxel_centre_voxel_3d, voxel_centre_voxel_2d, voxel_backward_2d, voxel_blob_centre_voxel_spacing_2d,
              voxel_bspline_centre_voxel_spacing_2d, voxel_blob_two_pixels_2d, voxel_bspline_two_pixels_2d,
              voxel_blob_three_pixels_2d, voxel_bspline_three_pixels_2d, voxel_blob_single_pixel_2d,
              voxel_bspline_single_pixel_2d, voxel_blob_backward_2d, voxel_bspline_backward_2d,
              voxel_bspline_centre_voxel_3d, voxel_bspline_45_degree_cpu, voxel_off_centre_3d]


# --------------------------------------------------------------------------------------
# Phase-contrast voxel projectors, test_PhaseContrastVoxelProjectorCUDA.cpp


def _derivative(B, kind):
    """blob.derivative(s) / bspline.derivative(s) of elsa (Blobs.h:280-287, BSplines.h:168-175):
    the 2D derivative table at |s| times sgn(s), in the back end's dtype."""
    import oracle

    if kind == "pc_blob":
        lut, sup = oracle.blob_derivative_lut(2.0, 10.83, 2, 100, B.dtype), 2.0
    else:
        lut, sup = oracle.bspline_derivative_lut(2, 2, 100, B.dtype), 1.5
    return lambda s: float(oracle.lut_eval(lut, sup, abs(s))) * float(np.sign(s))


def _pc_two_pixels(B, kind):
    """:40-88 (blob) and :2174-2222 (B-spline): the centre voxel of 5^2 on a two-pixel detector
    gives derivative(0.4761878) and its negative (epsilon 0.01), poses 0, 4, ..., 356 deg."""
    dv = _derivative(B, kind)
    w = dv(0.4761878)
    x = np.zeros((5, 5)); x[2, 2] = 1
    for i in range(0, 360, 4):
        g = B.setup((5, 5), (2,), [pose(100, 5, deg(i))])
        y = B.voxel(g, x, kind).ravel()
        assert approx(w, y[0], 0.01) and approx(-w, y[1], 0.01), (i, y, w)


def _pc_three_pixels(B, kind):
    """:90-146 (blob) and :797-853 (B-spline): three pixels, poses 0, 5, ..., 355 deg: the
    centre pixel gets derivative(0), the outer ones +-derivative(0.95233786) (epsilon 0.01)."""
    dv = _derivative(B, kind)
    w0, w1 = dv(0.0), dv(0.95233786)
    x = np.zeros((5, 5)); x[2, 2] = 1
    for i in range(0, 360, 5):
        g = B.setup((5, 5), (3,), [pose(100, 5, deg(i))])
        y = B.voxel(g, x, kind).ravel()
        assert approx(w0, y[1]), (i, y)
        assert approx(w1, y[0], 0.01) and approx(-w1, y[2], 0.01), (i, y, w1)


def _pc_single_pixel(B, kind):
    """:148-246 (blob) and :855-953 (B-spline), pose 0: the column on the ray sums to
    5 derivative(0); the columns beside it to the sum of derivative(u / scaling) at their
    projected, scaled positions (projectAndScaleVoxelOnDetector, DetectorDescriptor.cpp:
    130-151, evaluated here in FP64)."""
    import oracle

    dv = _derivative(B, kind)
    g = B.setup((5, 5), (1,), [pose(100, 5, 0.0)])
    og = oracle.geom_from_poses(2, (5, 5), (1,), [oracle.geometry_pose(2, 100.0, 5.0, (5, 5), (1,), 0.0,
                                                                       dtype=np.float64)], np.float64)
    P = og.P[0].reshape(3, 2).T
    E = og.ext[0].reshape(3, 2).T
    for col in (2, 3, 1):
        x = np.zeros((5, 5)); x[:, col] = 1
        y = B.voxel(g, x, kind).ravel()
        total = 0.0
        for s in range(5):
            hv = np.array([col + 0.5, s + 0.5, 1.0])
            pc = P @ hv
            u = pc[0] / pc[1] - 0.5
            total += dv(u / (og.sdd[0] / np.linalg.norm(E @ hv)))
        assert approx(y[0], total), (col, y, total)


def pc_blob_two_pixels_2d(B):
    _pc_two_pixels(B, "pc_blob")


def pc_bspline_two_pixels_2d(B):
    _pc_two_pixels(B, "pc_bspline")


def pc_blob_three_pixels_2d(B):
    _pc_three_pixels(B, "pc_blob")


def pc_bspline_three_pixels_2d(B):
    _pc_three_pixels(B, "pc_bspline")


def pc_blob_single_pixel_2d(B):
    _pc_single_pixel(B, "pc_blob")


def pc_bspline_single_pixel_2d(B):
    _pc_single_pixel(B, "pc_bspline")


PHASE_CONTRAST_KATS = [pc_blob_two_pixels_2d, pc_bspline_two_pixels_2d, pc_blob_three_pixels_2d,
                       pc_bspline_three_pixels_2d, pc_blob_single_pixel_2d, pc_bspline_single_pixel_2d]


# ---- FiniteDifferences, forward differences on every axis (elsa/operators/tests/
# test_FiniteDifferences.cpp): (voxels per axis, axis 0 first; data; expected apply; expected
# applyAdjoint of that apply's output, or None where elsa's test has none)
FD_KATS = [
    ((5,), [30, 3, 2, -1, 7], [-27, -1, -3, 8, -7], [27, -26, 2, -11, 15]),                 # :59-89
    ((4, 4), [16, 5, 9, 4, 2, 11, 7, 14, 3, 10, 6, 15, 13, 8, 12, 1],                          # :139-161
     [-11, 4, -5, -4, 9, -4, 7, -14, 7, -4, 9, -15, -5, 4, -11, -1,
      -14, 6, -2, 10, 1, -1, -1, 1, 10, -2, 6, -14, -13, -8, -12, -1], None),
    ((2, 2, 2), [1, 2, 3, 4, 5, 6, 7, 8],                                                       # :295-318
     [1, -2, 1, -4, 1, -6, 1, -8, 2, 2, -3, -4, 2, 2, -7, -8, 4, 4, 4, 4, -5, -6, -7, -8], None),
    ((2, 3, 2), [6, 3, 5, 8, 7, 5, 1, 0, 3, 10, -4, 1],                                         # :321-343
     [-3, -3, 3, -8, -2, -5, -1, -0, 7, -10, 5, -1, -1, 5, 2, -3, -7, -5, 2, 10,
      -7, -9, 4, -1, -5, -3, -2, 2, -11, -4, -1, -0, -3, -10, 4, -1], None),
]
