"""The smoothed-aggregation hierarchy on the CPU (tests/amg/amg_host_check).

The symbolic setup of nbots_amd/csrc/amg_setup.cpp runs under
AddressSanitizer + UBSan, followed by an FP64 mirror of the device numeric
phase (smoothed prolongator, Galerkin products through the W = A P pattern,
the coarsest inverse) and of the V-cycle, as the preconditioner of a
flexible CG.  Systems come from the oracle (pattern, assembly, boundary
conditions); the bars:
  * the flexible CG reaches a true relative residual of 1e-8 within the
    stated iteration counts (the method's, not the mesh size's: a model of
    the same method in scipy took 24-35 iterations on these problems),
    against 65-91 outer iterations of degree-40 Chebyshev ones;
  * the hierarchy coarsens to at most 200 unknowns;
  * without the prolongator's smoothing (plain aggregation) the same solve
    takes clearly more iterations, so the check sees the smoothing;
  * the coarsest level's dense inverse (Cholesky with a pivot guard) inverts
    an SPD matrix of order 200 and, on a singular Laplacian, drops exactly the
    unknown its null space leaves and inverts the rest.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
AMG = os.path.join(ROOT, "tests", "amg")


@pytest.fixture(scope="module")
def host_check():
    if shutil.which("g++") is None or shutil.which("make") is None:
        pytest.skip("g++ and make are needed for the host check")
    r = subprocess.run(["make", "-s", "-C", AMG], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return os.path.join(AMG, "_bin", "amg_host_check")


def _system(kind, nx, ny, simp_rho=None):
    from nbots_amd import synth
    if kind == "cantilever":
        p = synth.cantilever_traction(nx, ny, seed=20261202)
    else:
        p = synth.mbb(nx, ny, rho=0.5)
        if simp_rho is not None:
            p.elem_E = synth.simp_modulus(simp_rho, 3.0, 1.0, 1e-9)
    rp, col = oracle.pattern(p.N_nod, p.conn)
    val, F, st = oracle.assemble(p, rp, col)
    assert st == 0
    oracle.set_bconditions(p, rp, col, val, F)
    return p, rp, col, val, F


def _write(path, p, rp, col, val, F):
    n = F.size
    with open(path, "wb") as f:
        f.write(np.array([n, col.size], np.uint64).tobytes())
        f.write(np.asarray(rp, np.uint32).tobytes())
        f.write(np.asarray(col, np.uint32).tobytes())
        f.write(np.asarray(val, np.float64).tobytes())
        f.write(np.asarray(p.nod, np.float64).reshape(-1).tobytes())
        f.write(np.asarray(F, np.float64).tobytes())


def _run(exe, path, max_iter=500, omega_scale=1.0):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, path, str(max_iter), str(omega_scale)], capture_output=True,
                       text=True, timeout=600, env=env)
    kv = dict(tok.split("=", 1) for tok in r.stdout.split() if "=" in tok)
    return r, kv


@pytest.mark.parametrize("kind,nx,ny,bound", [("cantilever", 64, 32, 37),
                                              ("cantilever", 128, 64, 43),
                                              ("mbb", 96, 48, 25)])
def test_amg_flexible_cg_converges(host_check, tmp_path, kind, nx, ny, bound):
    p, rp, col, val, F = _system(kind, nx, ny)
    path = str(tmp_path / "sys.bin")
    _write(path, p, rp, col, val, F)
    r, kv = _run(host_check, path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert int(kv["iters"]) <= bound, r.stdout
    assert float(kv["rel_res"]) <= 1.0000001e-8
    sizes = [int(s) for s in kv["sizes"].split(",")]
    assert sizes[0] == F.size and sizes[-1] <= 200 and len(sizes) >= 3, r.stdout


def test_amg_on_a_simp_design(host_check, tmp_path):
    """The design after 12 iterations of the oracle's SIMP loop on the 96x48
    MBB beam: void regions (E = 1e-9) beside solid members, the contrast that
    makes these systems hard for Jacobi-type methods."""
    from nbots_amd import synth
    from test_gpu_simp import oracle_simp
    hist = oracle_simp(synth.mbb(96, 48, rho=0.5), 12)
    p, rp, col, val, F = _system("mbb", 96, 48, simp_rho=hist[-1][3])
    path = str(tmp_path / "sys.bin")
    _write(path, p, rp, col, val, F)
    r, kv = _run(host_check, path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert float(kv["rel_res"]) <= 1.0000001e-8
    assert int(kv["iters"]) <= 39, r.stdout


def test_prolongator_smoothing_matters(host_check, tmp_path):
    p, rp, col, val, F = _system("cantilever", 128, 64)
    path = str(tmp_path / "sys.bin")
    _write(path, p, rp, col, val, F)
    r1, kv1 = _run(host_check, path)
    r0, kv0 = _run(host_check, path, 500, 0.0)
    assert r1.returncode == 0 and r0.returncode == 0, r1.stdout + r0.stdout
    assert int(kv0["iters"]) >= 1.5 * int(kv1["iters"]), (kv0, kv1)


def test_dense_inverse(host_check):
    r = subprocess.run([host_check, "--dense"], capture_output=True, text=True, timeout=120)
    kv = dict(tok.split("=", 1) for tok in r.stdout.split() if "=" in tok)
    assert r.returncode == 0, r.stdout + r.stderr
    assert int(kv["spd_dropped"]) == 0 and float(kv["spd_err"]) < 1e-12, r.stdout
    assert int(kv["lap_dropped"]) == 1 and float(kv["lap_err"]) < 1e-9, r.stdout
