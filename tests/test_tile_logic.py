"""CPU run of the LDS-tiled Joseph kernels' index logic (the same functions the HIP kernels
call, built with g++): every LDS slot and volume index in range, every voxel owned by exactly
one tile per pass, and the assembled back projection equal to the oracle's; the plane-major
forward's views and loops against the oracle (fp32_bounds) and each other."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle
from fp32_bounds import assert_forward_close, oracle_forward_refs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# ELSA_SANITIZE=1 builds the checks with AddressSanitizer and UndefinedBehaviorSanitizer under
# their own names (tests/test_sanitizers.py runs this file that way, with the runtimes preloaded)
SAN = os.environ.get("ELSA_SANITIZE") == "1"
SFX = "_san" if SAN else ""
CXXFLAGS = ["g++", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math"] + (
    ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
    if SAN else ["-O2"])
SRC = os.path.join(ROOT, "tests", "cxx", "tile_logic_check.cpp")
OUT = os.path.join(ROOT, "tests", "cxx", "_build", f"libtile_logic{SFX}.so")


def _build(src, out):
    """Compile one check into out, through a file of this process renamed into place, so that
    test processes building the same check at once (pytest -n) never load a partial library."""
    os.makedirs(os.path.dirname(out), exist_ok=True)
    tmp = f"{out}.{os.getpid()}.tmp"
    subprocess.run([*CXXFLAGS, "-I", os.path.join(ROOT, "elsa_amd", "csrc"), src, "-o", tmp], check=True)
    os.replace(tmp, out)


@pytest.fixture(scope="module")
def lib():
    _build(SRC, OUT)
    L = ctypes.CDLL(OUT)
    f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    i64 = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
    f64 = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    L.tile_adjoint.argtypes = [ctypes.c_int, i64, i64, ctypes.c_int, f32, f32, f32, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, f32, f32, f64]
    L.tile_adjoint64.argtypes = [ctypes.c_int, i64, i64, ctypes.c_int, f32, f32, f32, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, f64, f64, f64]
    L.tile_adjoint_px.argtypes = [ctypes.c_int, i64, i64, ctypes.c_int, f32, f32, f32, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, f32, ctypes.c_int, f32, f32, f64]
    L.plane_forward.argtypes = [ctypes.c_int, i64, i64, ctypes.c_int, f32, f32, f32,
                                np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS"), f32, f32, ctypes.c_int]
    L.plane_forward_mixed_waves.restype = ctypes.c_int64
    L.plane_forward64.argtypes = [ctypes.c_int, i64, i64, ctypes.c_int, f32, f32, f32,
                                  np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS"), f64, f64, ctypes.c_int]
    return L


CASES = [
    # dim, vol, views, det, s2c, c2d, tile (LA, TB, TC)
    (3, (40, 24, 33), 13, (61, 29), 60.0, 40.0, (32, 15, 15)),
    (3, (32, 32, 32), 24, (48, 48), 3200.0, 160.0, (32, 15, 15)),
    (3, (17, 9, 23), 7, (30, 20), 25.0, 25.0, (8, 5, 3)),
    (3, (64, 64, 64), 9, (96, 96), 6400.0, 320.0, (16, 7, 9)),
    (3, (64, 64, 64), 9, (96, 96), 6400.0, 320.0, (8, 17, 17)),
    (2, (31, 17), 11, (80,), 40.0, 30.0, (8, 5, 1)),
    (2, (64, 64), 30, (90,), 6400.0, 320.0, (64, 125, 1)),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}d-{'x'.join(map(str, c[1]))}-t{c[6][0]}")
def test_tiled_adjoint_logic_matches_oracle(lib, case):
    dim, vol, views, det, s2c, c2d, tile = case
    rng = np.random.default_rng(1)
    angles = np.sort(rng.uniform(0, 360, views)).astype(np.float32)
    g = oracle.circle_geometry(dim, vol, views, 360, s2c, c2d, det=det, angles_deg=angles)
    y = rng.standard_normal(g.sino_shape).astype(np.float32)
    out = np.zeros(g.vol_shape, np.float32)
    info = np.zeros(6, np.float64)
    rc = lib.tile_adjoint(dim, np.array(list(vol) + [1] * (3 - dim), np.int64),
                          np.array(list(det) + [1] * (2 - len(det)), np.int64), views,
                          np.ascontiguousarray(g.P), np.ascontiguousarray(g.Pinv), np.ascontiguousarray(g.C),
                          tile[0], tile[1], tile[2], y, out, info)
    assert rc == 0, rc
    ref = oracle.joseph(g, y, adjoint=True, mode=oracle.MODE_DIRECT)
    err = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    assert err <= 1e-6, err
    # each term is elsa's FP32 product y * len times the exact weights of the FP32 fractions,
    # rounded once to a multiple of q: against the exact-weight sums in FP64 (y * len exact), a
    # voxel is off by at most 2^-24 of the same sums over |y| per term (the FP32 product) and per
    # pass (the FP32 volume), plus half a unit q per term
    y64 = y.astype(np.float64)
    exact = oracle.joseph(g, y64, adjoint=True, mode=oracle.MODE_EXACT_WEIGHTS, acc64=True)
    mag = oracle.joseph(g, np.abs(y64), adjoint=True, mode=oracle.MODE_EXACT_WEIGHTS, acc64=True)
    bound = (1 + dim) * 2.0 ** -24 * mag * (1 + 1e-9) + dim * 0.5 * info[1] * info[0]
    assert np.all(np.abs(out - exact) <= bound), np.max(np.abs(out - exact) / bound)
    # and closer to it than the oracle's FP32 restatement (elsa's rounded weights and sums)
    assert np.linalg.norm(out - exact) < np.linalg.norm(ref - exact)
    assert info[3:].sum() >= 1
    # slot sums stay well inside int64 (the bound is a worst case)
    assert 0 < info[2] < 0.5, info


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}d-{'x'.join(map(str, c[1]))}-t{c[6][0]}")
def test_tiled_adjoint_fp64_logic_matches_oracle(lib, case):
    """The FP64 tiled back projection's logic: each contribution (y * len) * weight in FP64, the
    weight elsa's FP32 product, rounded to a multiple of q = 2^(eg - 62) and summed in 64-bit
    integers. Against the oracle's FP64 sums of the same FP32-geometry terms (acc64), the
    difference is at most half a unit q per contribution plus one FP64 rounding of the voxel."""
    dim, vol, views, det, s2c, c2d, tile = case
    rng = np.random.default_rng(2)
    angles = np.sort(rng.uniform(0, 360, views)).astype(np.float32)
    g = oracle.circle_geometry(dim, vol, views, 360, s2c, c2d, det=det, angles_deg=angles)
    y = rng.standard_normal(g.sino_shape)
    out = np.zeros(g.vol_shape, np.float64)
    info = np.zeros(6, np.float64)
    rc = lib.tile_adjoint64(dim, np.array(list(vol) + [1] * (3 - dim), np.int64),
                            np.array(list(det) + [1] * (2 - len(det)), np.int64), views,
                            np.ascontiguousarray(g.P), np.ascontiguousarray(g.Pinv), np.ascontiguousarray(g.C),
                            tile[0], tile[1], tile[2], np.ascontiguousarray(y), out, info)
    assert rc == 0, rc
    exact = oracle.joseph(g, y, adjoint=True, mode=oracle.MODE_DIRECT, acc64=True)
    q = info[1]
    # contributions per voxel and pass: at most the slot bound info[0] (poses times the rays
    # whose footprint can cover one slot, times |len| >= 1); one FP64 rounding per pass
    bound = dim * (0.5 * q * info[0] + 2.0 ** -52 * np.abs(exact))
    assert np.all(np.abs(out - exact) <= bound), np.max(np.abs(out - exact) / bound)
    err = np.linalg.norm(out - exact) / np.linalg.norm(exact)
    assert err <= 1e-14, err
    assert 0 < info[2] < 0.5, info


CURVED_CASES = [
    # dim, vol, views, det, s2c, c2d, fan angle, tile (LA, TB, TC), trace the rays outside
    (3, (24, 20, 22), 13, (40, 34), 48.0, 36.0, 1.2, (8, 7, 5), True),
    (3, (32, 32, 32), 9, (48, 40), 60.0, 60.0, 0.9, (16, 7, 9), True),
    (3, (40, 24, 33), 13, (61, 29), 60.0, 40.0, 1.5, (32, 15, 15), False),
    # one tile over the whole volume, magnified onto a wide fan: rows bend by pixels across it;
    # at 45 degrees an edge of the volume projects to the middle column, where the lowest
    # and highest rows of the projection then lie
    (3, (16, 16, 16), 5, (64, 128), 20.0, 40.0, 2.0, (16, 16, 16), True),
    (2, (31, 17), 11, (80,), 40.0, 30.0, 1.0, (8, 5, 1), True),
]


@pytest.mark.parametrize("case", CURVED_CASES, ids=lambda c: f"{c[0]}d-{'x'.join(map(str, c[1]))}-a{c[6]}")
def test_tiled_adjoint_curved_detector(lib, case):
    """Curved detectors (rays through the planar centres of CurvedDetectorDescriptor): the tiles'
    candidate rectangles, found by binary search over the planar centres, hold every ray that
    reaches the tile (checked by tracing every other ray, where enabled), and the assembled back
    projection matches the oracle's."""
    dim, vol, views, det, s2c, c2d, angle, tile, outside = case
    rng = np.random.default_rng(5)
    angles = np.sort(rng.uniform(0, 360, views)).astype(np.float32)
    if tile[0] == vol[0]:
        angles = np.array([45.0, 135.0, 225.0, 315.0, 10.0][:views], np.float32)
    g = oracle.circle_geometry(dim, vol, views, 360, s2c, c2d, det=det, angles_deg=angles)
    pix = np.ascontiguousarray(oracle.curved_planar_coords(det, angle, s2c + c2d).reshape(-1))
    y = rng.standard_normal(g.sino_shape).astype(np.float32)
    out = np.zeros(g.vol_shape, np.float32)
    info = np.zeros(8, np.float64)
    rc = lib.tile_adjoint_px(dim, np.array(list(vol) + [1] * (3 - dim), np.int64),
                             np.array(list(det) + [1] * (2 - len(det)), np.int64), views,
                             np.ascontiguousarray(g.P), np.ascontiguousarray(g.Pinv), np.ascontiguousarray(g.C),
                             tile[0], tile[1], tile[2], pix, int(outside), y, out, info)
    assert rc == 0, rc
    ref = oracle.joseph(g, y, adjoint=True, mode=oracle.MODE_DIRECT, pix=pix.reshape(-1, dim - 1))
    err = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    assert err <= 1e-6, err
    assert info[3:6].sum() >= 1 and 0 < info[2] < 0.5, info
    if dim == 3:
        # the row search's column is the middle one (least spread of planar v), and the
        # rows of a curved detector bend monotonically on either side of it: no extra margin
        assert info[6] in ((det[0] - 1) // 2, det[0] // 2) and info[7] == 0.0, info


def test_curved_candidates_reject_unstructured_coords(lib):
    """Planar centres without the structure of a curved detector (planar u varying down a
    column) are refused by the tiled path's check (the operator then keeps the one-thread-per-
    ray kernels)."""
    dim, vol, det = 3, (12, 12, 12), (16, 12)
    g = oracle.circle_geometry(dim, vol, 3, 360, 40.0, 30.0, det=det)
    pix = oracle.curved_planar_coords(det, 0.8, 70.0).copy()
    pix[5 * det[0] + 3, 0] += 0.01
    y = np.ones(g.sino_shape, np.float32)
    out = np.zeros(g.vol_shape, np.float32)
    rc = lib.tile_adjoint_px(dim, np.array(vol, np.int64), np.array(det, np.int64), 3, np.ascontiguousarray(g.P),
                             np.ascontiguousarray(g.Pinv), np.ascontiguousarray(g.C), 4, 5, 5,
                             np.ascontiguousarray(pix.reshape(-1)), 0, y, out, np.zeros(8, np.float64))
    assert rc == 11


def test_rect_split_exact_at_large_rectangles(lib):
    """The tiled kernels split a candidate index into (row, column) through a float reciprocal
    and one correction step; exact for rectangles far beyond 2^22 entries (here up to 2^31 - 1),
    where the reciprocal alone goes wrong."""
    L = lib
    for nu in (1, 3, 7, 640, 1023, 2048, 4095, 8191, 65535, (1 << 22) + 1):
        for loc in np.unique(np.concatenate([np.arange(0, 5000), rng_locs(nu)])):
            row, col = ctypes.c_int(), ctypes.c_int()
            L.rect_split_host(int(loc), nu, ctypes.byref(row), ctypes.byref(col))
            assert (row.value, col.value) == divmod(int(loc), nu), (loc, nu, row.value, col.value)


def rng_locs(nu):
    rng = np.random.default_rng(nu)
    hi = min(2 ** 31 - 1, nu * ((1 << 21) - 1))
    loc = rng.integers(0, hi, 3000, dtype=np.int64)
    near = (loc // nu) * nu
    return np.concatenate([loc, near, np.maximum(near - 1, 0), np.minimum(near + 1, hi - 1), [hi - 1]])


FWD_CASES = [
    # dim, vol, views, det, s2c, c2d, fan axes to try per driving axis
    (3, (40, 24, 33), 13, (61, 29), 60.0, 40.0, [(2, 2, 1), (1, 0, 0), (2, 0, 1)]),
    (3, (17, 9, 23), 7, (30, 20), 25.0, 25.0, [(2, 2, 1), (1, 0, 0)]),
    (3, (64, 64, 64), 9, (96, 96), 6400.0, 320.0, [(2, 2, 1)]),
    (2, (31, 17), 11, (80,), 40.0, 30.0, [(1, 0)]),
    (2, (64, 64), 30, (90,), 6400.0, 320.0, [(1, 0)]),
]


# how the forward's rays are grouped: one plane loop per ray, or wavefronts of wu x (64 / wu)
# pixels whose rays of one driving axis and direction share one plane loop (the GPU's grouping)
WAVES = [0, 2 | (64 << 8), 2 | (16 << 8), 2 | (8 << 8)]
WAVE_IDS = ["per_ray", "wave64x1", "wave16x4", "wave8x8"]


@pytest.mark.parametrize("waves", WAVES, ids=WAVE_IDS)
@pytest.mark.parametrize("case", FWD_CASES, ids=lambda c: f"{c[0]}d-{'x'.join(map(str, c[1]))}")
def test_plane_forward_logic_matches_oracle(lib, case, waves):
    """Every zero-padded plane-major view a driving axis can get gives the oracle's forward to
    FP32 rounding (same positions, floors and taps, lerp-form interpolation; the bars of
    fp32_bounds: at least as close to the exact sums as elsa's FP32 form), with every read inside
    its view -- both with one plane loop per ray and with the GPU's grouping, where the rays of a
    wavefront of one driving axis and direction share one plane loop. Every grouping and view
    layout gives the same result bit for bit."""
    dim, vol, views, det, s2c, c2d, fans = case
    rng = np.random.default_rng(2)
    angles = np.sort(rng.uniform(0, 360, views)).astype(np.float32)
    g = oracle.circle_geometry(dim, vol, views, 360, s2c, c2d, det=det, angles_deg=angles)
    x = rng.standard_normal(g.vol_shape).astype(np.float32)
    refs = oracle_forward_refs(g, x)
    first = None
    for fan in fans:
        out = np.full(g.sino_shape, np.nan, np.float32)
        rc = lib.plane_forward(dim, np.array(list(vol) + [1] * (3 - dim), np.int64),
                               np.array(list(det) + [1] * (2 - len(det)), np.int64), views,
                               np.ascontiguousarray(g.P), np.ascontiguousarray(g.Pinv), np.ascontiguousarray(g.C),
                               np.array(list(fan) + [0] * (3 - len(fan)), np.int32), x, out, waves)
        assert rc == 0, (fan, rc)
        assert_forward_close(out, *refs, what=f"fan axes {fan}")
        if first is None:
            first = out
        np.testing.assert_array_equal(out, first, err_msg=f"fan axes {fan}")


@pytest.mark.parametrize("waves", WAVES, ids=WAVE_IDS)
@pytest.mark.parametrize("case", [c for c in FWD_CASES if c[0] == 3], ids=lambda c: "x".join(map(str, c[1])))
@pytest.mark.parametrize("b16", [False, True], ids=["f32", "bf16"])
def test_plane_forward_row_pair_views_match_oracle(lib, case, waves, b16):
    """Row-pair views (the 3D forward's layout: every voxel stored twice, each sample's 2 x 2
    footprint one read, 16 bytes in FP32 and 8 in bfloat16) give the oracle's forward of the
    (rounded) volume to FP32 rounding (fp32_bounds), and the same sums bit for bit as views with
    one element per voxel; every read inside its view and aligned, every view element written
    at most once."""
    dim, vol, views, det, s2c, c2d, fans = case
    rng = np.random.default_rng(4)
    angles = np.sort(rng.uniform(0, 360, views)).astype(np.float32)
    g = oracle.circle_geometry(dim, vol, views, 360, s2c, c2d, det=det, angles_deg=angles)
    x = rng.standard_normal(g.vol_shape).astype(np.float32)
    if b16:
        x = _to_bf16(x)
    refs = oracle_forward_refs(g, x)
    for fan in fans:
        args = (dim, np.array(list(vol), np.int64), np.array(list(det), np.int64), views,
                np.ascontiguousarray(g.P), np.ascontiguousarray(g.Pinv), np.ascontiguousarray(g.C),
                np.array(list(fan), np.int32), x)
        out = np.full(g.sino_shape, np.nan, np.float32)
        rc = lib.plane_forward(*args, out, waves | 4 | int(b16))
        assert rc == 0, (fan, rc)
        assert_forward_close(out, *refs, what=f"fan axes {fan}")
        single = np.full(g.sino_shape, np.nan, np.float32)
        assert lib.plane_forward(*args, single, waves | int(b16)) == 0
        np.testing.assert_array_equal(out, single, err_msg=f"fan axes {fan}")


@pytest.mark.parametrize("wu", [0, 64, 8], ids=["per_ray", "wave64x1", "wave8x8"])
@pytest.mark.parametrize("case", FWD_CASES, ids=lambda c: f"{c[0]}d-{'x'.join(map(str, c[1]))}")
def test_plane_forward_fp64_matches_oracle_bitwise(lib, case, wu):
    """FP64 operators (elsa's JosephsMethod with data_t = double, FP32 geometry): the forward
    through FP64 plane-major views, with elsa's FP32 weights, tap order and FP64 sums, equals the
    oracle's FP32-geometry / FP64-data forward bit for bit, for every view layout and grouping,
    with every read inside its view."""
    dim, vol, views, det, s2c, c2d, fans = case
    rng = np.random.default_rng(8)
    angles = np.sort(rng.uniform(0, 360, views)).astype(np.float32)
    g = oracle.circle_geometry(dim, vol, views, 360, s2c, c2d, det=det, angles_deg=angles)
    x = rng.standard_normal(g.vol_shape)
    ref = oracle.joseph(g, x, mode=oracle.MODE_DIRECT, acc64=True)
    for fan in fans:
        out = np.full(g.sino_shape, np.nan, np.float64)
        rc = lib.plane_forward64(dim, np.array(list(vol) + [1] * (3 - dim), np.int64),
                                 np.array(list(det) + [1] * (2 - len(det)), np.int64), views,
                                 np.ascontiguousarray(g.P), np.ascontiguousarray(g.Pinv), np.ascontiguousarray(g.C),
                                 np.array(list(fan) + [0] * (3 - len(fan)), np.int32), x, out, wu)
        assert rc == 0, (fan, rc)
        np.testing.assert_array_equal(out, ref, err_msg=f"fan axes {fan}")


def test_plane_forward_wavefront_groups_cover_mixed_waves(lib):
    """A strongly divergent cone (source 60, detector 40 from the centre) puts rays of several
    driving axes and directions into one wavefront; each group's shared plane loop still gives
    the oracle's forward (fp32_bounds) and the per-ray loop's sums bit for bit."""
    dim, vol, views, det = 3, (40, 24, 33), 13, (61, 29)
    rng = np.random.default_rng(7)
    angles = np.sort(rng.uniform(0, 360, views)).astype(np.float32)
    g = oracle.circle_geometry(dim, vol, views, 360, 20.0, 40.0, det=det, angles_deg=angles)
    x = rng.standard_normal(g.vol_shape).astype(np.float32)
    args = (dim, np.array(vol, np.int64), np.array(det, np.int64), views, np.ascontiguousarray(g.P),
            np.ascontiguousarray(g.Pinv), np.ascontiguousarray(g.C), np.array([2, 2, 1], np.int32), x)
    out = np.full(g.sino_shape, np.nan, np.float32)
    rc = lib.plane_forward(*args, out, 2 | (16 << 8))
    assert rc == 0
    assert lib.plane_forward_mixed_waves() > 0
    check = oracle_forward_refs(g, x)
    assert_forward_close(out, *check)
    per_ray = np.full(g.sino_shape, np.nan, np.float32)
    assert lib.plane_forward(*args, per_ray, 0) == 0
    np.testing.assert_array_equal(out, per_ray)


def _to_bf16(x):
    """Round float32 to bfloat16 (nearest even) and back to float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("waves", WAVES, ids=WAVE_IDS)
@pytest.mark.parametrize("case", FWD_CASES, ids=lambda c: f"{c[0]}d-{'x'.join(map(str, c[1]))}")
def test_plane_forward_bf16_views_match_oracle(lib, case, waves):
    """bfloat16 views (half-width offsets, two-element loads) give the oracle's forward of the
    bfloat16-rounded volume (fp32_bounds) and the FP32 views' sums of it bit for bit, with every
    read inside its view."""
    dim, vol, views, det, s2c, c2d, fans = case
    rng = np.random.default_rng(5)
    angles = np.sort(rng.uniform(0, 360, views)).astype(np.float32)
    g = oracle.circle_geometry(dim, vol, views, 360, s2c, c2d, det=det, angles_deg=angles)
    x = _to_bf16(rng.standard_normal(g.vol_shape).astype(np.float32))
    refs = oracle_forward_refs(g, x)
    for fan in fans:
        args = (dim, np.array(list(vol) + [1] * (3 - dim), np.int64), np.array(list(det) + [1] * (2 - len(det)), np.int64),
                views, np.ascontiguousarray(g.P), np.ascontiguousarray(g.Pinv), np.ascontiguousarray(g.C),
                np.array(list(fan) + [0] * (3 - len(fan)), np.int32), x)
        out = np.full(g.sino_shape, np.nan, np.float32)
        rc = lib.plane_forward(*args, out, waves | 1)
        assert rc == 0, (fan, rc)
        assert_forward_close(out, *refs, what=f"fan axes {fan}")
        f32 = np.full(g.sino_shape, np.nan, np.float32)
        assert lib.plane_forward(*args, f32, waves) == 0
        np.testing.assert_array_equal(out, f32, err_msg=f"fan axes {fan}")


def test_tiled_adjoint_logic_high_dynamic_range(lib):
    """One ray a million times larger than the rest. Each term is rounded in its own ray's
    scale, so the documented bound holds voxel by voxel: |error| <= 2^-31 * sum over the rays
    that reach the voxel of |y * len| (times 1.01), plus one FP32 rounding of the result; and
    overall the sums are closer to the exact sums than sequential FP32 sums are."""
    dim, vol, views, det, s2c, c2d, tile = CASES[1]
    rng = np.random.default_rng(5)
    angles = np.sort(rng.uniform(0, 360, views)).astype(np.float32)
    g = oracle.circle_geometry(dim, vol, views, 360, s2c, c2d, det=det, angles_deg=angles)
    y = rng.standard_normal(g.sino_shape).astype(np.float32)
    y.flat[len(y.flat) // 2 + 17] = 3.0e6
    out = np.zeros(g.vol_shape, np.float32)
    info = np.zeros(6, np.float64)
    rc = lib.tile_adjoint(dim, np.array(list(vol), np.int64), np.array(list(det), np.int64), views,
                          np.ascontiguousarray(g.P), np.ascontiguousarray(g.Pinv), np.ascontiguousarray(g.C),
                          tile[0], tile[1], tile[2], y, out, info)
    assert rc == 0, rc
    exact = oracle.joseph(g, y, adjoint=True, mode=oracle.MODE_DIRECT, acc64=True)
    ref32 = oracle.joseph(g, y, adjoint=True, mode=oracle.MODE_DIRECT)
    # a ray reaches a voxel with at most one sample (len <= sqrt 3), and at most info[0] rays
    # (the slot bound) reach one voxel: the large ray plus that many ordinary ones
    big, rest = np.sort(np.abs(y).ravel())[-1:-3:-1]
    bound = 1.01 * 2.0 ** -31 * np.sqrt(3.0) * (float(big) + info[0] * float(rest))
    err = np.abs(out - exact)
    assert np.all(err <= bound + np.spacing(np.abs(exact).astype(np.float32))), float(err.max())
    assert np.linalg.norm(out - exact) <= np.linalg.norm(ref32 - exact)


# --------------------------------------------------------------------------------------
# Siddon: the traversal and the tiled back projection of ehp_siddon.hpp

SID_SRC = os.path.join(ROOT, "tests", "cxx", "siddon_tile_check.cpp")
SID_OUT = os.path.join(ROOT, "tests", "cxx", "_build", f"libsiddon_tile{SFX}.so")


@pytest.fixture(scope="module")
def sid():
    _build(SID_SRC, SID_OUT)
    L = ctypes.CDLL(SID_OUT)
    f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    i64 = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
    f64 = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    L.siddon_forward.argtypes = [ctypes.c_int, i64, i64, ctypes.c_int, f32, f32, f32, f32, f32]
    L.siddon_tile_cover.argtypes = [ctypes.c_int, i64, i64, ctypes.c_int, f32, f32, f32, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, f64]
    L.siddon_tile_cover_px.argtypes = [ctypes.c_int, i64, i64, ctypes.c_int, f32, f32, f32, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, f32, f64]
    L.siddon_tile_adjoint.argtypes = [ctypes.c_int, i64, i64, ctypes.c_int, f32, f32, f32, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, f32, f32]
    L.siddon_tile_adjoint64.argtypes = [ctypes.c_int, i64, i64, ctypes.c_int, f32, f32, f32, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, f64, f64]
    L.siddon_plane_forward.argtypes = [ctypes.c_int, i64, i64, ctypes.c_int, f32, f32, f32,
                                       np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS"), f32, f32, ctypes.c_int]
    L.siddon_plane_forward64.argtypes = [ctypes.c_int, i64, i64, ctypes.c_int, f32, f32, f32,
                                         np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS"), f64, f64,
                                         ctypes.c_int]
    return L


SID_CASES = [
    # dim, vol, views, det, s2c, c2d, tile (LA, TB, TC)
    (3, (40, 24, 33), 13, (61, 29), 60.0, 40.0, (8, 16, 16)),
    (3, (17, 9, 23), 7, (30, 20), 25.0, 25.0, (4, 5, 3)),
    (3, (32, 32, 32), 12, (48, 48), 3200.0, 160.0, (8, 16, 16)),
    (2, (31, 17), 11, (80,), 40.0, 30.0, (4, 7, 1)),
    (2, (64, 64), 30, (90,), 6400.0, 320.0, (4, 16, 1)),
]


def _sid_args(case, seed):
    dim, vol, views, det, s2c, c2d, tile = case
    rng = np.random.default_rng(seed)
    angles = np.sort(rng.uniform(0, 360, views)).astype(np.float32)
    g = oracle.circle_geometry(dim, vol, views, 360, s2c, c2d, det=det, angles_deg=angles)
    head = (dim, np.array(list(vol) + [1] * (3 - dim), np.int64), np.array(list(det) + [1] * (2 - len(det)), np.int64),
            views, np.ascontiguousarray(g.P), np.ascontiguousarray(g.Pinv), np.ascontiguousarray(g.C))
    return g, head, rng


@pytest.mark.parametrize("case", SID_CASES, ids=lambda c: f"{c[0]}d-{'x'.join(map(str, c[1]))}")
def test_siddon_traversal_matches_oracle_bitwise(sid, case):
    g, head, rng = _sid_args(case, 3)
    x = rng.standard_normal(g.vol_shape).astype(np.float32)
    out = np.full(g.sino_shape, np.nan, np.float32)
    assert sid.siddon_forward(*head, x, out) == 0
    np.testing.assert_array_equal(out, oracle.siddon(g, x, mode=oracle.MODE_DIRECT))


@pytest.mark.parametrize("wu", [64, 8])
@pytest.mark.parametrize("case", SID_CASES, ids=lambda c: f"{c[0]}d-{'x'.join(map(str, c[1]))}")
def test_siddon_plane_forward_matches_oracle_bitwise(sid, case, wu):
    """The GPU forward's scheme: zero-padded plane-major views, the rays of a wavefront of one
    driving axis and direction walking its planes together, each ray taking its own steps on
    the current plane. Same crossings and order as the whole traversal, so the oracle's sums
    bit for bit, with every read inside its view and on the current plane; for every layout a
    driving axis can get."""
    g, head, rng = _sid_args(case, 6)
    x = rng.standard_normal(g.vol_shape).astype(np.float32)
    ref = oracle.siddon(g, x, mode=oracle.MODE_DIRECT)
    fans = [(2, 2, 1), (1, 0, 0), (2, 0, 1)] if case[0] == 3 else [(1, 0)]
    for fan in fans:
        out = np.full(g.sino_shape, np.nan, np.float32)
        rc = sid.siddon_plane_forward(*head, np.array(list(fan) + [0] * (3 - len(fan)), np.int32), x, out, wu)
        assert rc == 0, (fan, rc)
        np.testing.assert_array_equal(out, ref, err_msg=f"fan axes {fan}")


@pytest.mark.parametrize("wu", [64, 8])
@pytest.mark.parametrize("case", SID_CASES, ids=lambda c: f"{c[0]}d-{'x'.join(map(str, c[1]))}")
def test_siddon_plane_forward_fp64_matches_oracle_bitwise(sid, case, wu):
    """The same plane walk over FP64 views (SiddonsMethodCUDA<double>: FP32 traversal and
    lengths, FP64 data and sums): the oracle's FP32-geometry / FP64-data forward bit for bit, for
    every view layout, and within FP32 rounding of the all-FP32 oracle."""
    g, head, rng = _sid_args(case, 16)
    x = rng.standard_normal(g.vol_shape)
    ref = oracle.siddon(g, x, mode=oracle.MODE_DIRECT, acc64=True)
    assert ref.dtype == np.float64
    ref32 = oracle.siddon(g, x.astype(np.float32), mode=oracle.MODE_DIRECT)
    fans = [(2, 2, 1), (1, 0, 0), (2, 0, 1)] if case[0] == 3 else [(1, 0)]
    for fan in fans:
        out = np.full(g.sino_shape, np.nan, np.float64)
        rc = sid.siddon_plane_forward64(*head, np.array(list(fan) + [0] * (3 - len(fan)), np.int32), x, out, wu)
        assert rc == 0, (fan, rc)
        np.testing.assert_array_equal(out, ref, err_msg=f"fan axes {fan}")
    scale = oracle.siddon(g, np.abs(x), mode=oracle.MODE_DIRECT, acc64=True)
    assert np.all(np.abs(out - ref32) <= 1e-5 * scale + 1e-30)


@pytest.mark.parametrize("case", SID_CASES, ids=lambda c: f"{c[0]}d-{'x'.join(map(str, c[1]))}")
def test_siddon_tiles_restart_rays_exactly(sid, case):
    """For each ray, the crossings of the tiles of its pass are the crossings of the whole
    traversal -- the same voxels with bit-identical lengths, none twice -- and the ray lies in
    the candidate rectangle of every tile it crosses."""
    g, head, rng = _sid_args(case, 4)
    tile = case[6]
    stats = np.zeros(2)
    rc = sid.siddon_tile_cover(*head, tile[0], tile[1], tile[2], 3, stats)
    assert rc == 0, rc
    assert stats[0] > 0 and stats[1] > stats[0]


@pytest.mark.parametrize("case", CURVED_CASES, ids=lambda c: f"{c[0]}d-{'x'.join(map(str, c[1]))}-a{c[6]}")
def test_siddon_tiles_curved_detector(sid, case):
    """Curved detectors: every ray through a planar centre lies in the candidate rectangle of
    every tile it crosses, and the tiles' crossings are those of the whole traversal."""
    dim, vol, views, det, s2c, c2d, angle, tile, _ = case
    g, head, rng = _sid_args((dim, vol, views, det, s2c, c2d, tile), 4)
    pix = np.ascontiguousarray(oracle.curved_planar_coords(det, angle, s2c + c2d).reshape(-1))
    stats = np.zeros(2)
    rc = sid.siddon_tile_cover_px(*head, max(tile[0] // 2, 1), tile[1], tile[2], 1, pix, stats)
    assert rc == 0, rc
    assert stats[0] > 0 and stats[1] > stats[0]


@pytest.mark.parametrize("case", SID_CASES, ids=lambda c: f"{c[0]}d-{'x'.join(map(str, c[1]))}")
def test_siddon_tiled_adjoint_logic_matches_oracle(sid, case):
    g, head, rng = _sid_args(case, 5)
    tile = case[6]
    y = rng.standard_normal(g.sino_shape).astype(np.float32)
    out = np.zeros(g.vol_shape, np.float32)
    rc = sid.siddon_tile_adjoint(*head, tile[0], tile[1], tile[2], y, out)
    assert rc == 0, rc
    ref = oracle.siddon(g, y, adjoint=True, mode=oracle.MODE_DIRECT)
    exact = oracle.siddon(g, y, adjoint=True, mode=oracle.MODE_DIRECT, acc64=True)
    assert np.linalg.norm(out - ref) <= 1e-6 * np.linalg.norm(ref)
    assert np.linalg.norm(out - exact) <= np.linalg.norm(ref - exact)


@pytest.mark.parametrize("case", SID_CASES, ids=lambda c: f"{c[0]}d-{'x'.join(map(str, c[1]))}")
def test_siddon_tiled_adjoint_fp64_logic_matches_oracle(sid, case):
    """SiddonsMethodCUDA<double>'s tiled back projection: each crossing's y * length (FP64 y,
    FP32 length) rounded once to the fixed-point unit q and summed in 64-bit integers. Against
    the oracle's sequential FP64 sums of the same terms: within 1e-14 in relative L2 and 1e-12
    of the largest voxel everywhere (q is about 2^-50 of the largest term here)."""
    g, head, rng = _sid_args(case, 9)
    tile = case[6]
    y = rng.standard_normal(g.sino_shape)
    out = np.zeros(g.vol_shape, np.float64)
    assert sid.siddon_tile_adjoint64(*head, tile[0], tile[1], tile[2], np.ascontiguousarray(y), out) == 0
    exact = oracle.siddon(g, y, adjoint=True, mode=oracle.MODE_DIRECT, acc64=True)
    err = np.linalg.norm(out - exact) / np.linalg.norm(exact)
    assert err <= 1e-14, err
    # the same crossings as the FP32 logic: a wrong voxel or length moves a voxel by O(y)
    assert np.max(np.abs(out - exact)) <= 1e-12 * np.max(np.abs(exact))


# --------------------------------------------------------------------------------------
# voxel projectors: the detector-tiled forward of ehp_voxel.hpp

VOX_SRC = os.path.join(ROOT, "tests", "cxx", "voxel_tile_check.cpp")
VOX_OUT = os.path.join(ROOT, "tests", "cxx", "_build", f"libvoxel_tile{SFX}.so")


@pytest.fixture(scope="module")
def vox():
    _build(VOX_SRC, VOX_OUT)
    L = ctypes.CDLL(VOX_OUT)
    f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    i64 = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
    f64 = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    L.voxel_tile_forward.argtypes = [ctypes.c_int, i64, i64, ctypes.c_int, f32, f32, f32, f32, f32, f32, ctypes.c_int,
                                     ctypes.c_float, ctypes.c_int, ctypes.c_int, f32, f32, f64]
    L.voxel_tile_forward64.argtypes = [ctypes.c_int, i64, i64, ctypes.c_int, f32, f32, f32, f32, f32, f64,
                                       ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int, f64, f64, f64]
    return L


VOX_CASES = [
    # dim, vol, views, det, s2c, c2d, tile (W, H), basis
    (3, (20, 14, 17), 9, (40, 26), 60.0, 40.0, (16, 8), "blob"),
    (3, (16, 16, 16), 7, (24, 24), 1600.0, 80.0, (8, 8), "bspline"),
    (3, (12, 9, 11), 5, (21, 13), 25.0, 25.0, (7, 5), "blob"),
    (2, (31, 17), 11, (80,), 40.0, 30.0, (16, 1), "blob"),
    (2, (48, 48), 20, (70,), 4800.0, 240.0, (32, 1), "bspline"),
    # a table whose last entry is not 0 (the kernels' range-tested taps)
    (3, (20, 14, 17), 9, (40, 26), 60.0, 40.0, (16, 8), "raised"),
    (2, (31, 17), 11, (80,), 40.0, 30.0, (16, 1), "raised"),
]


@pytest.mark.parametrize("case", VOX_CASES, ids=lambda c: f"{c[0]}d-{'x'.join(map(str, c[1]))}-{c[7]}")
def test_voxel_tiled_forward_logic_matches_oracle(vox, case):
    """Every footprint tap is visited by exactly one (pose, tile) -- the frustum of a tile
    catches every voxel that reaches it, none twice -- and the fixed-point sums give the
    oracle's forward projection."""
    dim, vol, views, det, s2c, c2d, tile, basis = case
    rng = np.random.default_rng(7)
    angles = np.sort(rng.uniform(0, 360, views)).astype(np.float32)
    g = oracle.circle_geometry(dim, vol, views, 360, s2c, c2d, det=det, angles_deg=angles)
    if basis == "blob":
        lut, support = oracle.blob_lut(2.0, 10.83, 2, 100, np.float32), 2.0
    elif basis == "raised":
        lut, support = (oracle.blob_lut(2.0, 10.83, 2, 100, np.float32) + np.float32(0.05)).astype(np.float32), 2.0
        assert lut[-1] != 0.0
    else:
        lut, support = oracle.bspline_lut(dim, 2, 100, np.float32), 1.5
    x = rng.standard_normal(g.vol_shape).astype(np.float32)
    out = np.full(g.sino_shape, np.nan, np.float32)
    stats = np.zeros(2)
    rc = vox.voxel_tile_forward(dim, np.array(list(vol) + [1] * (3 - dim), np.int64),
                                np.array(list(det) + [1] * (2 - len(det)), np.int64), views,
                                np.ascontiguousarray(g.P), np.ascontiguousarray(g.ext), np.ascontiguousarray(g.C),
                                np.ascontiguousarray(g.sdd), np.ascontiguousarray(g.Pinv), np.ascontiguousarray(lut),
                                lut.shape[0], support, tile[0], tile[1], x, out, stats)
    assert rc == 0, rc
    assert stats[0] > 0
    ref = oracle.voxel(g, x, lut, support)
    assert np.linalg.norm(out - ref) <= 1e-6 * np.linalg.norm(ref), np.linalg.norm(out - ref) / np.linalg.norm(ref)


@pytest.mark.parametrize("case", VOX_CASES, ids=lambda c: f"{c[0]}d-{'x'.join(map(str, c[1]))}-{c[7]}")
def test_voxel_tiled_forward_fp64_logic_matches_oracle(vox, case):
    """The FP64 tiled forward (FP64 table, the one-thread-per-voxel FP64 kernel's weights, each
    tap's x * w rounded once to the fixed-point unit): the oracle's FP64 forward to a few units
    of FP64 rounding, and every tap visited by exactly one (pose, tile)."""
    dim, vol, views, det, s2c, c2d, tile, basis = case
    rng = np.random.default_rng(8)
    angles = np.sort(rng.uniform(0, 360, views)).astype(np.float32)
    g = oracle.circle_geometry(dim, vol, views, 360, s2c, c2d, det=det, angles_deg=angles)
    if basis == "blob":
        lut, support = oracle.blob_lut(2.0, 10.83, 2, 100, np.float64), 2.0
    elif basis == "raised":
        lut, support = oracle.blob_lut(2.0, 10.83, 2, 100, np.float64) + 0.05, 2.0
    else:
        lut, support = oracle.bspline_lut(dim, 2, 100, np.float64), 1.5
    x = rng.standard_normal(g.vol_shape)
    out = np.full(g.sino_shape, np.nan, np.float64)
    stats = np.zeros(2)
    rc = vox.voxel_tile_forward64(dim, np.array(list(vol) + [1] * (3 - dim), np.int64),
                                  np.array(list(det) + [1] * (2 - len(det)), np.int64), views,
                                  np.ascontiguousarray(g.P), np.ascontiguousarray(g.ext), np.ascontiguousarray(g.C),
                                  np.ascontiguousarray(g.sdd), np.ascontiguousarray(g.Pinv), np.ascontiguousarray(lut),
                                  lut.shape[0], support, tile[0], tile[1], x, out, stats)
    assert rc == 0, rc
    assert stats[0] > 0
    # the oracle with FP32 geometry and the FP64 table and data
    ref = oracle.voxel(g, x, lut, support, acc64=True)
    err = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    assert err <= 1e-14, err
