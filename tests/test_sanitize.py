"""AddressSanitizer / UndefinedBehaviorSanitizer run of the FP64 oracle.

Builds oracle/scmc_oracle.c together with tests/sanitize_driver.c under
``-fsanitize=address,undefined -fno-sanitize-recover=all`` and runs every
oracle entry point (covariance, field, replay, wave, tabulated phase, Fourier
field, voxel media) on tiny problems.  Any out-of-bounds access, use after
free, leak, signed overflow or invalid shift ends the driver with a non-zero
status and a sanitizer report.  CPU only.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    out = str(tmp_path_factory.mktemp("asan") / "sanitize_driver")
    cmd = ["gcc", "-O1", "-g", "-std=gnu11", "-fopenmp", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-I", ORACLE, os.path.join(HERE, "sanitize_driver.c"), os.path.join(ORACLE, "scmc_oracle.c"),
           "-o", out, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and "sanitizer" in r.stderr.lower() and "cannot find" in r.stderr.lower():
        pytest.skip("sanitizer runtime libraries not installed: " + r.stderr[-400:])
    assert r.returncode == 0, r.stderr[-4000:]
    return out


def test_oracle_under_asan_ubsan(driver):
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["OMP_NUM_THREADS"] = "2"
    r = subprocess.run([driver], capture_output=True, text=True, env=env, timeout=600)
    report = (r.stdout + r.stderr)[-6000:]
    assert r.returncode == 0, report
    assert "runtime error" not in r.stderr, report
    assert "ERROR: AddressSanitizer" not in r.stderr, report
    assert "all entry points ran" in r.stdout
