"""The solver's host code (SELL layouts, sweep schedules, overlapped stage-2
order, chunk layouts, FP32/FP16 conversion, BILU(0) triangles, GMRES control
loop) run without a GPU: linked against a host-memory stand-in of the device
layer and checked by AddressSanitizer/UBSan over every option combination."""
import os
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hoststub")


@pytest.fixture(scope="module")
def driver():
    r = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stdout + r.stderr)
    return os.path.join(HERE, "_build", "drive")


@pytest.mark.parametrize("dims", [(8, 8, 6, 3), (12, 10, 6, 2), (6, 6, 4, 4), (3, 2, 1, 3)])
def test_solver_host_code_under_asan(driver, dims):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([driver] + [str(v) for v in dims], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "runtime error" not in r.stderr, r.stderr[-4000:]
