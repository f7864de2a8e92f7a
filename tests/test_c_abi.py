"""A plain C99 program includes include/elsa_hip.h, links libelsahip.so and calls the host-only
entry points (tests/cxx/c_abi_consumer.c): the header compiles as C with -pedantic -Werror, the
library links and runs without a GPU, and the trajectories, poses and tables it prints equal the
oracle's bit for bit (the same bars as the ctypes checks of test_host.py)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cxx", "c_abi_consumer.c")
OUT = os.path.join(ROOT, "tests", "cxx", "_build", "c_abi_consumer")
LIBDIR = os.path.join(ROOT, "elsa_amd")


@pytest.fixture(scope="module")
def printed():
    if shutil.which("gcc") is None:
        pytest.skip("gcc is not installed")
    if not os.path.exists(os.path.join(LIBDIR, "libelsahip.so")):
        pytest.fail("libelsahip.so is not built (python -m elsa_amd.build)")
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    # a program of this process (test processes of pytest -n may build it at the same time)
    exe = f"{OUT}.{os.getpid()}"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-I", os.path.join(ROOT, "include"),
                    SRC, "-o", exe, "-L", LIBDIR, "-lelsahip", f"-Wl,-rpath,{LIBDIR}"], check=True)
    try:
        res = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    finally:
        os.remove(exe)
    assert res.returncode == 0, res.stderr
    lines = {}
    for line in res.stdout.splitlines():
        name, _, rest = line.partition(" ")
        lines[name] = rest
    assert "done" in lines
    return lines


def floats(s):
    return np.array([np.float32(float(v)) for v in s.split()], np.float32)


def test_c_consumer_trajectories_equal_oracle(printed):
    assert "gfx950" in printed["version"]
    g2 = oracle.circle_geometry(2, (256, 256), 180, 360, 25600.0, 1280.0, det=(384,))
    assert int(printed["det2"]) == 384
    np.testing.assert_array_equal(floats(printed["P2_0"]), g2.P[0])
    np.testing.assert_array_equal(floats(printed["P2_45"]), g2.P[45])
    np.testing.assert_array_equal(floats(printed["C2_179"]), g2.C[179])
    np.testing.assert_array_equal(floats(printed["S2_0"]), g2.sdd[:1])
    g3 = oracle.circle_geometry(3, (64, 48, 40), 4, 360, 6400.0, 320.0, det=(96, 80))
    np.testing.assert_array_equal(floats(printed["P3_1"]), g3.P[1])
    np.testing.assert_array_equal(floats(printed["Pinv3_2"]), g3.Pinv[2])


def test_c_consumer_pose_and_tables_equal_oracle(printed):
    P, Pinv, C, ext, sdd = oracle.geometry_pose(3, 6400.0, 320.0, (64, 48, 40), (96, 80), 0.5, 0.25, -0.125)
    np.testing.assert_array_equal(floats(printed["pose_P"]), P)
    np.testing.assert_array_equal(floats(printed["pose_C"]), C)
    np.testing.assert_array_equal(floats(printed["pose_sdd"]), np.array([sdd], np.float32))
    np.testing.assert_array_equal(floats(printed["blob_lut"]), oracle.blob_lut(2.0, 10.83, 2, 100, np.float32))
    assert float(printed["blob_support"]) == 2.0
    np.testing.assert_array_equal(floats(printed["bspline_lut"]), oracle.bspline_lut(3, 2, 100, np.float32))
    assert float(printed["bspline_support"]) == 1.5


def test_c_consumer_error_path(printed):
    rc, handle, has_msg = printed["bad_dim"].split()
    assert int(rc) == 1 and handle == "null" and has_msg == "1"  # EHP_ERR_ARGUMENT, message set
    assert int(printed["no_poses"]) == 1
