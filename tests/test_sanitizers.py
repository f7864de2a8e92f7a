"""The oracle and the CPU runs of the tiled kernels' index logic under AddressSanitizer and
UndefinedBehaviorSanitizer (host code only; the reference's own CI runs its tests with these,
cmake/Sanitizer.cmake and the sanitizer jobs of .gitlab-ci.yml).

Builds oracle/_san/liboracle_san.so (`make -C oracle san`) and the tests/cxx checks with
-fsanitize=address,undefined -fno-sanitize-recover=all, then runs test_oracle.py and
test_tile_logic.py against those builds in a child interpreter with the sanitizer runtimes
preloaded: any out-of-bounds access, use after free or undefined behaviour aborts the child.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _runtime(name):
    path = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.mark.timeout(900)
def test_oracle_and_tile_logic_under_asan_ubsan():
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("g++ has no sanitizer runtimes here")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], check=True)
    env = dict(os.environ)
    env.update({
        # the sanitizer runtimes go first; whatever the environment already preloads stays
        "LD_PRELOAD": ":".join(p for p in (asan, ubsan, os.environ.get("LD_PRELOAD", "")) if p),
        # Python's own allocations are not instrumented and live until exit
        "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1:halt_on_error=1",
        "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1",
        "ELSA_SANITIZE": "1",
        "ELSA_ORACLE_LIB": os.path.join(ROOT, "oracle", "_san", "liboracle_san.so"),
        "OMP_NUM_THREADS": "4",
    })
    res = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                          os.path.join(ROOT, "tests", "test_oracle.py"),
                          os.path.join(ROOT, "tests", "test_tile_logic.py")],
                         cwd=ROOT, env=env, capture_output=True, text=True, timeout=850)
    out = res.stdout + res.stderr
    assert res.returncode == 0, out[-6000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-6000:]
