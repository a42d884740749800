"""Failure handling forced from the host: non-finite inputs.

The relative-tolerance solves (nbx_*_solve, FP64 and MIXED) return
NBX_ERR_BREAKDOWN for a right-hand side with a NaN or Inf entry, and for a
matrix value or initial guess that makes the residual non-finite, instead of
reporting convergence (their stop tests compare with `>`, which a NaN
fails).  The drop-in nb_sparse_solve_CG_precond_Jacobi keeps the
reference's semantics (cg_precond_jacobi.c:44, :86-89): a NaN residual
leaves the loop at once and the call returns 0 with tolerance_reached NaN,
which the oracle's restatement does too (test_oracle_golden pins that
against the compiled reference).  Every kernel runs to its normal end on
these inputs.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NBX_ERR_BREAKDOWN = 6


def _system(nx=24, ny=12):
    import oracle
    from nbots_amd import synth
    p = synth.cantilever_traction(nx, ny, seed=3)
    rp, col = oracle.pattern(p.N_nod, p.conn)
    val, F, _ = oracle.assemble(p, rp, col)
    oracle.set_bconditions(p, rp, col, val, F)
    return rp, col, val, F


def _solve_rc(rp, col, val, b, mode, opts):
    import ctypes as C
    import nbots_amd as nb
    from nbots_amd.lib import SolveInfo, last_error
    A = nb.Sparse.from_csr(rp, col, val, vars_per_node=2)
    x = np.zeros(A.N)
    info = SolveInfo()
    rc = A.L.nbx_sparse_solve_ex(A.h, np.ascontiguousarray(b, np.float64), x, mode, 1e-8,
                                 10 * A.N, C.byref(opts), C.byref(info))
    msg = last_error()
    A.destroy()
    return rc, msg


CASES = [("fp64", 1, {}), ("cheb_fcg", 2, {}), ("fp32_cg_ir", 2, dict(variant=0, inner_rtol=1e-4))]


@pytest.mark.parametrize("bad", [np.nan, np.inf])
@pytest.mark.parametrize("name,mode,kw", CASES)
def test_non_finite_rhs_is_a_breakdown(name, mode, kw, bad):
    from nbots_amd.fem import solve_opts
    rp, col, val, F = _system()
    b = F.copy()
    b[17] = bad
    print(f"[fail] {name} b[17] = {bad}", flush=True)
    rc, msg = _solve_rc(rp, col, val, b, mode, solve_opts(true_residual=True, **kw))
    assert rc == NBX_ERR_BREAKDOWN and "non-finite" in msg, (rc, msg)


@pytest.mark.parametrize("name,mode,kw", CASES)
def test_non_finite_matrix_value_is_a_breakdown(name, mode, kw):
    from nbots_amd.fem import solve_opts
    rp, col, val, F = _system()
    v = val.copy()
    v[rp[40] + 1] = np.inf
    print(f"[fail] {name} K value = inf", flush=True)
    rc, msg = _solve_rc(rp, col, v, F, mode, solve_opts(true_residual=True, **kw))
    assert rc == NBX_ERR_BREAKDOWN and "non-finite" in msg, (rc, msg)


@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_drop_in_keeps_reference_semantics(bad):
    """nb_sparse_solve_CG_precond_Jacobi with a NaN or Inf in b: the status,
    iteration count and non-finite tolerance_reached of the oracle's
    restatement (NaN: no iteration; Inf: two, the lagged test then sees a
    NaN), with status 0 as the reference returns."""
    import oracle
    import nbots_amd as nb
    rp, col, val, F = _system()
    b = F.copy()
    b[5] = bad
    A = nb.Sparse.from_csr(rp, col, val, vars_per_node=2)
    x, st, it, tr = A.solve_cg_jacobi(b, max_iter=1000, tol=1e-8)
    A.destroy()
    xo, ito, tro, sto = oracle.pcg(rp, col, val, b, max_iter=1000, tol=1e-8)
    assert (st, it) == (sto, ito) and st == 0
    assert not np.isfinite(tr) and not np.isfinite(tro)


def test_solver_usable_after_a_breakdown():
    """A breakdown leaves the system usable: the next solve of a finite
    right-hand side converges as on a fresh system."""
    import nbots_amd as nb
    from nbots_amd.fem import solve_opts
    rp, col, val, F = _system()
    A = nb.Sparse.from_csr(rp, col, val, vars_per_node=2)
    b = F.copy()
    b[3] = np.nan
    with pytest.raises(RuntimeError, match="non-finite"):
        A.solve(b, mode=2, tol=1e-8, opts=solve_opts(true_residual=True))
    x, info = A.solve(F, mode=2, tol=1e-8, opts=solve_opts(true_residual=True))
    A.destroy()
    B = nb.Sparse.from_csr(rp, col, val, vars_per_node=2)
    x2, info2 = B.solve(F, mode=2, tol=1e-8, opts=solve_opts(true_residual=True))
    B.destroy()
    assert info["status"] == 0 and info["iters"] == info2["iters"]
    assert np.array_equal(x, x2)
