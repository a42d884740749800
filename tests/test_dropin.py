"""The drop-in link: C programs built against the reference's own headers
and objects, linked -lnbots_amd ahead of the compiled reference library
(tests/dropin/).

* CPU: the import path (nbots_amd/csrc/import.cpp) reads the reference's
  own mesh and bcond objects through their accessors, under AddressSanitizer
  and UBSan.  The mesh must equal the fixture the reference wrote, and the
  boundary-condition plan, applied to the oracle's K and F with the
  reference's elimination (sparse.c:416-430), must give the oracle's K and
  F after its set_bconditions bit for bit.
* GPU: the reference's FEM unit-test flow (static_elasticity2D.c:246-303),
  unchanged, answers the test's known values through this library; a
  matrix passes between this library and the reference's LU and in-place
  transpose.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, ROOT, load_golden, problem_from_golden

BIN = os.path.join(ROOT, "tests", "dropin", "_bin")
FILES = {"beam": "beam_cantilever", "plate": "plate_with_hole"}


def _binary(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.skip(f"{path} is not built (make -C tests/dropin, where the reference exists)")
    return path


def _read_plan(path):
    kinds = [np.float64, np.uint32, np.uint32, np.uint32, np.float64, np.uint32, np.uint32,
             np.float64, np.float64]
    out, raw, pos = [], open(path, "rb").read(), 0
    for k in kinds:
        n = int(np.frombuffer(raw, np.uint64, 1, pos)[0])
        pos += 8
        out.append(np.frombuffer(raw, k, n, pos).copy())
        pos += n * np.dtype(k).itemsize
    assert pos == len(raw)
    return out


def _apply_plan(rp, col, val, F, n_dof, n_ptr, n_add, d_dof, d_rank, vfirst, vlast):
    """The plan as the reference would apply it: Neumann additions per dof in
    order, then each dof's first elimination in rank order (sparse.c:416-430);
    a repeated elimination of a dof only leaves its last value in F."""
    val, F = val.copy(), F.copy()
    for k, dof in enumerate(n_dof):
        for a in n_add[n_ptr[k]:n_ptr[k + 1]]:
            F[dof] += a
    for k in np.argsort(d_rank, kind="stable"):
        c, v = int(d_dof[k]), vfirst[k]
        for s in range(rp[c], rp[c + 1]):
            j = int(col[s])
            if j == c:
                val[s] = 1.0
                F[c] = v
            else:
                val[s] = 0.0
                t = rp[j] + int(np.searchsorted(col[rp[j]:rp[j + 1]], c))
                var = val[t]
                val[t] = 0.0
                F[j] -= var * v
    for c, v in zip(d_dof, vlast):
        F[c] = v
    return val, F


@pytest.mark.parametrize("case", ["beam", "plate"])
def test_import_under_sanitizers(case, tmp_path):
    exe = _binary("import_check")
    out = tmp_path / "plan.bin"
    r = subprocess.run([exe, case, os.path.join(GOLDEN, FILES[case] + ".txt"), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    nod, conn, *plan = _read_plan(out)
    g = load_golden(FILES[case])
    assert np.array_equal(nod, g["nod"]) and np.array_equal(conn, g["conn"])
    p = problem_from_golden(g)
    rp, col = oracle.pattern(p.N_nod, p.conn)
    val, F, st = oracle.assemble(p, rp, col)
    assert st == 0
    q, fptr, fval = oracle.fn_neumann(p)
    val_o, F_o = val.copy(), F.copy()
    oracle.set_bconditions(q, rp, col, val_o, F_o, ns_fptr=fptr, ns_fval=fval)
    val_p, F_p = _apply_plan(rp, col, val, F, *plan)
    assert np.array_equal(val_p, val_o) and np.array_equal(F_p, F_o)


def test_import_rejects_missing_segment():
    exe = _binary("import_check")
    r = subprocess.run([exe, "badid", os.path.join(GOLDEN, "beam_cantilever.txt")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "segment 4000 does not exist" in r.stdout


def _run_dropin(*args):
    exe = _binary("fem_dropin")
    print("[dropin]", *args, flush=True)
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=300)
    kv = dict(ln.split("=", 1) for ln in r.stdout.splitlines() if "=" in ln)
    assert r.returncode == 0 and kv.get("status") == "ok", r.stdout + r.stderr
    return kv


@pytest.mark.gpu
def test_reference_beam_flow_through_dropin():
    """The reference's beam test, its own objects and call sequence: max |u|
    = 1.00701e-1 +- 1e-6 (static_elasticity2D.c:118), and within 1e-9 of the
    value the reference's CPU driver computed on the same mesh."""
    kv = _run_dropin("beam", os.path.join(GOLDEN, "beam_cantilever.txt"))
    g = load_golden("beam_cantilever")
    assert int(kv["nodes"]) == g["nod"].size // 2 and int(kv["fem_status"]) == 0
    m = float(kv["max_disp"])
    assert abs(m - 1.00701e-1) < 1e-6
    assert abs(m - float(g["max_disp"])) <= 1e-9 * float(g["max_disp"])


@pytest.mark.gpu
def test_reference_plate_flow_through_dropin():
    """The reference's plate-with-a-hole test with its function-valued
    Kirsch tractions, evaluated by the reference's own bcond iterator: mean
    von Mises error against Kirsch below the test's 9.7e-3."""
    kv = _run_dropin("plate", os.path.join(GOLDEN, "plate_with_hole.txt"))
    assert int(kv["fem_status"]) == 0
    assert float(kv["kirsch_err"]) < 9.7e-3


@pytest.mark.gpu
def test_sparse_handles_cross_libraries():
    """nb_sparse_create here, the reference's LU on the same handle (its
    factors freed by this library's nb_sparse_destroy), the GPU PCG within
    1e-9 of the LU solution, and GPU products that follow the reference's
    in-place transpose exactly."""
    kv = _run_dropin("sparse")
    assert int(kv["lu_status"]) == 0 and int(kv["cg_status"]) == 0
    assert float(kv["lu_vs_cg"]) <= 1e-9
    assert float(kv["spmv_vs_rows"]) == 0.0
    assert float(kv["asym"]) > 0.0
