"""The smoothed-aggregation multigrid preconditioner (NBX_MIXED_AMG_FCG) on the GPU.

Bars:
  * the solve reaches the true relative residual 1e-8, checked with the
    oracle's reference-order SpMV, and the displacement agrees with a 1e-13
    oracle solve within rel-L2 1e-6 (the bar of the other mixed solvers);
  * the iteration count is the method's: within the host mirror's count
    (tests/test_amg_host.py runs the same hierarchy in FP64 on the CPU) plus a
    margin for the FP32 V-cycle;
  * new values (a reassembled K) refresh the hierarchy's numeric phase, new
    Dirichlet conditions rebuild it, and the solves still converge;
  * the SIMP loop with AMG solves follows the FP64 loop's compliance history
    within 1e-6;
  * a system without node coordinates (the sparse drop-in) refuses the variant;
  * the polling cadence taken from the previous solve changes no result.
"""
import dataclasses

import numpy as np
import pytest

from conftest import load_golden, problem_from_golden

pytestmark = pytest.mark.gpu

NBX_ERR_ARG = 4


def _amg_opts(**kw):
    from nbots_amd.fem import MIXED_AMG_FCG, solve_opts
    return solve_opts(MIXED_AMG_FCG, **kw)


@pytest.mark.parametrize("case,bound", [("traction_128x64", 45), ("c1_cantilever_64x32", 40)])
def test_amg_solve_against_oracle(case, bound):
    import oracle
    from nbots_amd.fem import Fem, SOLVE_MIXED
    p = problem_from_golden(load_golden(case))
    f = Fem.from_problem(p)
    print(f"[amg] {case}", flush=True)
    info = f.solve(SOLVE_MIXED, tol=1e-8, opts=_amg_opts(true_residual=True))
    assert info["status"] == 0 and info["layout"] & 16, info
    assert info["iters"] <= bound, info["iters"]
    u = f.displacement()
    rp, col, val, F = f.system()
    f.destroy()
    r = F - oracle.spmv(rp, col, val, u)
    assert np.linalg.norm(r) <= 1.0000001e-8 * np.linalg.norm(F)
    x_ref = oracle.pcg(rp, col, val, F, tol=1e-13 * np.linalg.norm(F), max_iter=200000)[0]
    assert np.linalg.norm(u - x_ref) <= 1e-6 * np.linalg.norm(x_ref)


@pytest.mark.parametrize("case", ["beam_cantilever", "plate_with_hole", "jittered_64x32",
                                  "cant_8x4_jit"])
def test_amg_other_meshes(case):
    """Unstructured meshes of the reference's tessellator (the beam; the
    plate with a hole, whose fp16 copy keeps one width per slice, layout bit
    3 clear), a jittered grid, and a system of 90 unknowns whose hierarchy is
    its dense coarsest level alone: each solve matches an FP64 solve."""
    from nbots_amd import synth
    from nbots_amd.fem import Fem, SOLVE_FP64, SOLVE_MIXED, solve_opts
    if case == "jittered_64x32":
        p = synth.cantilever_point(64, 32, jitter=0.3)
    else:
        p = problem_from_golden(load_golden(case))
    f, g = Fem.from_problem(p), Fem.from_problem(p)
    print(f"[amg] {case}", flush=True)
    ia = f.solve(SOLVE_MIXED, tol=1e-9, max_iter=500, opts=_amg_opts(true_residual=True))
    i64 = g.solve(SOLVE_FP64, tol=1e-11, opts=solve_opts(true_residual=True))
    assert ia["status"] == 0 and i64["status"] == 0, (ia, i64)
    assert ia["layout"] & 16 and ia["iters"] <= 60, ia
    if case == "cant_8x4_jit":
        assert not ia["layout"] & 2 and ia["iters"] <= 3, ia   # the dense solve alone
    else:
        assert ia["layout"] & 6 == 6 and bool(ia["layout"] & 8) == (case != "plate_with_hole"), ia
    ua, u64 = f.displacement(), g.displacement()
    f.destroy()
    g.destroy()
    assert np.linalg.norm(ua - u64) <= 1e-6 * np.linalg.norm(u64)


def test_amg_follows_new_values_and_conditions():
    """Reassembly with other moduli refreshes the numeric phase; new
    boundary conditions (a different Dirichlet plan) rebuild the hierarchy.
    Each solve matches an FP64 solve of the same system."""
    from nbots_amd import synth
    from nbots_amd.fem import Fem, SOLVE_FP64, SOLVE_MIXED, solve_opts
    p = synth.cantilever_traction(96, 48, seed=11)
    f = Fem.from_problem(p)
    g = Fem.from_problem(p)
    rng = np.random.default_rng(4)
    for step in range(3):
        if step == 1:
            E = np.exp(rng.standard_normal(p.N_elem))
            for h in (f, g):
                h.assemble(p.E, p.nu, p.thickness, E)
                h.set_bconditions(p)
        if step == 2:
            # other supports and load on the same mesh: the left edge
            # clamped, the right edge held in y and pulled in x
            q = dataclasses.replace(p, ds=synth._bc([3, 1], [1, 1, 0, 1], [0.0] * 4),
                                    ns=synth._bc([1], [1, 0], [48.0, 0.0]))
            for h in (f, g):
                h.assemble(q.E, q.nu, q.thickness, q.elem_E)
                h.set_bconditions(q)
        print(f"[amg] step {step}", flush=True)
        ia = f.solve(SOLVE_MIXED, tol=1e-9, opts=_amg_opts(true_residual=True))
        i64 = g.solve(SOLVE_FP64, tol=1e-10, opts=solve_opts(true_residual=True))
        assert ia["status"] == 0 and i64["status"] == 0, (ia, i64)
        ua, u64 = f.displacement(), g.displacement()
        assert np.linalg.norm(ua - u64) <= 1e-6 * np.linalg.norm(u64), step
    f.destroy()
    g.destroy()


def test_amg_simp_history_matches_fp64():
    from nbots_amd import synth
    from nbots_amd.fem import Fem, Simp, SOLVE_FP64, SOLVE_MIXED, solve_opts
    p = synth.mbb(96, 48, rho=0.5)
    hist = {}
    for name, mode, o in (("amg", SOLVE_MIXED, _amg_opts(true_residual=True)),
                          ("fp64", SOLVE_FP64, solve_opts(true_residual=True, simp=True))):
        f = Fem.from_problem(p)
        s = Simp(f, rmin=1.5, volfrac=0.5, penal=3.0, E0=1.0, Emin=1e-9, move=0.2)
        h = []
        for it in range(10):
            C, change, info = s.iterate(mode, tol=1e-8, opts=o)
            assert info["status"] == 0, (name, it, info)
            h.append(C)
        print(f"[amg] SIMP {name}: {h[0]:.6e} -> {h[-1]:.6e}", flush=True)
        hist[name] = np.array(h)
        s.destroy()
        f.destroy()
    rel = np.abs(hist["amg"] - hist["fp64"]) / np.abs(hist["fp64"])
    assert rel.max() <= 1e-6, rel.max()


def test_simp_default_solver_is_amg():
    """nbx_simp_iterate without options (and nbx_simp_opts_default) take the
    multigrid-preconditioned mixed solve."""
    from nbots_amd import synth
    from nbots_amd.fem import Fem, Simp, SOLVE_MIXED
    from nbots_amd.lib import SolveOpts, load
    o = SolveOpts()
    load().nbx_simp_opts_default(o)
    assert o.mixed_variant == 3
    f = Fem.from_problem(synth.mbb(48, 24, rho=0.5))
    s = Simp(f, rmin=1.5, volfrac=0.5, penal=3.0, E0=1.0, Emin=1e-9, move=0.2)
    for _ in range(2):
        C, change, info = s.iterate(SOLVE_MIXED, tol=1e-8)
        assert info["status"] == 0 and info["layout"] & 16, info
    s.destroy()
    f.destroy()


def test_polling_hint_leaves_results_unchanged():
    """A solve launches most of the previous solve's iteration count before it
    first reads the device state (host cadence only): the same system solved
    again from zero gives the same iterate bit for bit, as does a fresh system
    without a hint; a solve that starts at the solution stops at once (the
    hinted launches return immediately); a smaller max_iter than the hint still
    stops at max_iter."""
    from nbots_amd.fem import Fem, SOLVE_MIXED
    p = problem_from_golden(load_golden("traction_128x64"))
    f, g = Fem.from_problem(p), Fem.from_problem(p)
    o = _amg_opts(true_residual=True)
    i1 = f.solve(SOLVE_MIXED, tol=1e-8, opts=o)
    u1 = f.displacement()
    i2 = f.solve(SOLVE_MIXED, tol=1e-8, opts=o)
    u2 = f.displacement()
    i3 = g.solve(SOLVE_MIXED, tol=1e-8, opts=o)
    u3 = g.displacement()
    print(f"[amg] hint: {i1['iters']} {i2['iters']} {i3['iters']}", flush=True)
    assert i1["status"] == i2["status"] == i3["status"] == 0
    assert i1["iters"] == i2["iters"] == i3["iters"] >= 16, (i1, i2, i3)
    assert np.array_equal(u1, u2) and np.array_equal(u1, u3)
    i4 = f.solve(SOLVE_MIXED, tol=1e-8, u0=u1, opts=o)
    assert i4["status"] == 0 and i4["iters"] <= 2, i4
    assert np.linalg.norm(f.displacement() - u1) <= 1e-8 * np.linalg.norm(u1)
    f.solve(SOLVE_MIXED, tol=1e-8, opts=o)
    i5 = f.solve(SOLVE_MIXED, tol=1e-8, max_iter=10, opts=o)
    assert i5["status"] == 1 and i5["iters"] == 10, i5
    f.destroy()
    g.destroy()


def test_amg_refused_without_coordinates():
    import ctypes as C
    import oracle
    import nbots_amd as nb
    from nbots_amd import synth
    from nbots_amd.lib import SolveInfo, last_error
    p = synth.cantilever_traction(16, 8)
    rp, col = oracle.pattern(p.N_nod, p.conn)
    val, F, _ = oracle.assemble(p, rp, col)
    oracle.set_bconditions(p, rp, col, val, F)
    A = nb.Sparse.from_csr(rp, col, val, vars_per_node=2)
    x = np.zeros(A.N)
    info = SolveInfo()
    rc = A.L.nbx_sparse_solve_ex(A.h, F, x, 2, 1e-8, 1000, C.byref(_amg_opts()), C.byref(info))
    msg = last_error()
    A.destroy()
    assert rc == NBX_ERR_ARG and "AMG" in msg, (rc, msg)
