"""The oracle's C restatement under AddressSanitizer + UndefinedBehaviorSanitizer.

oracle/san_check.c runs every routine of oracle/refcpu.c once (pattern,
assembly with self weight, all four boundary-condition kinds including a
function-valued segment, PCG, CG, Gauss-Seidel, strain, stress, a damage
step, SIMP energies, sensitivities, filter and one OC step) on structured
cantilevers; any out-of-bounds access, leak or undefined behaviour aborts
it.  The host-side import of the reference's FEM objects runs under the
same sanitizers in tests/test_dropin.py.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


@pytest.fixture(scope="module")
def san_check():
    if shutil.which("gcc") is None or shutil.which("make") is None:
        pytest.skip("gcc and make are needed for the sanitizer build")
    r = subprocess.run(["make", "-s", "-C", ORACLE, "san"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return os.path.join(ORACLE, "_san", "san_check")


@pytest.mark.parametrize("nx,ny", [(12, 6), (5, 9), (1, 1)])
def test_oracle_clean_under_asan_ubsan(san_check, nx, ny):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([san_check, str(nx), str(ny)], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith(f"san_check {nx}x{ny}:"), r.stdout
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, r.stderr
