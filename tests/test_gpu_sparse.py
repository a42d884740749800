"""GPU parity of the drop-in nb_sparse_t path (SpMV and Jacobi-PCG) against
the golden vectors written by the compiled reference and against the oracle.

Bars: SpMV bit-exact (the device sums every scalar row in the reference's
column order with no FMA); PCG with the reference's semantics within
relative 1e-10 on x and +-2 iterations on the small fixtures (the device
dot products use a different, fixed summation order)."""
import math

import numpy as np
import pytest

from conftest import SMALL_CASES, load_golden, load_json, problem_from_golden

pytestmark = pytest.mark.gpu


def _sparse(g):
    import nbots_amd as nb
    return nb.Sparse.from_csr(g["rowptr"], g["col"], g["val"], vars_per_node=2)


@pytest.mark.parametrize("case", SMALL_CASES)
def test_spmv_bit_exact(case):
    g = load_golden(case)
    A = _sparse(g)
    assert A.is_blocked() == 1
    x = np.sin(np.arange(A.N, dtype=np.float64))
    y = A.matvec(x)
    assert np.array_equal(y, g["spmv_sin"]), np.abs(y - g["spmv_sin"]).max()
    A.destroy()


@pytest.mark.parametrize("case", SMALL_CASES)
def test_symmetric_layout_bit_exact(case):
    """The assembled K is symmetric bit for bit, so a solve builds the
    symmetric layout (upper blocks, lower ones read transposed) and reports
    layout 1; nb_sparse_multiply_vector then runs on it and still equals the
    compiled reference's SpMV bit for bit, and the PCG result matches the
    reference's as on the full layout."""
    g = load_golden(case)
    A = _sparse(g)
    x, info = A.solve(g["F"], mode=0, tol=float(g["pcg_tol"]))
    assert info["layout"] == 1
    assert abs(info["iters"] - int(g["niter"])) <= 2
    assert np.linalg.norm(x - g["x"]) <= 1e-10 * np.linalg.norm(g["x"])
    y = A.matvec(np.sin(np.arange(A.N, dtype=np.float64)))
    assert np.array_equal(y, g["spmv_sin"]), np.abs(y - g["spmv_sin"]).max()
    A.destroy()


def test_symmetric_layout_spmv_medium():
    """256x128 cantilever with lognormal moduli (33k nodes, 129 workgroups):
    the drop-in SpMV on the full layout (before any solve) and on the
    symmetric layout (after a solve found K symmetric) agree bit for bit, and
    both equal the oracle's row sums (the reference's order)."""
    import oracle
    import nbots_amd as nb
    from nbots_amd import synth
    p = synth.cantilever_traction(256, 128)
    rp, col = oracle.pattern(p.N_nod, p.conn)
    val, F, _ = oracle.assemble(p, rp, col)
    oracle.set_bconditions(p, rp, col, val, F)
    A = nb.Sparse.from_csr(rp, col, val, vars_per_node=2)
    x = np.random.default_rng(3).standard_normal(A.N)
    y_full = A.matvec(x)
    _, info = A.solve(F, mode=1, tol=1e-6)
    assert info["layout"] == 1
    y_sym = A.matvec(x)
    assert np.array_equal(y_full, y_sym)
    assert np.array_equal(y_sym, oracle.spmv(rp, col, val, x))
    A.destroy()


@pytest.mark.parametrize("n_nod", [1, 2, 65])
def test_symmetric_layout_tiny(n_nod):
    """Chains of 1, 2 and 65 nodes (one and two slices; a single node has no
    lower block at all): the symmetric layout is built, solves match a dense
    solve and the SpMV stays exact."""
    import oracle
    import nbots_amd as nb
    adj = [[w for w in (v - 1, v + 1) if 0 <= w < n_nod] for v in range(n_nod)]
    A = nb.Sparse.from_graph([len(a) for a in adj], adj, 2)
    rp, col, _ = A.export()
    rng = np.random.default_rng(n_nod)
    dense = np.zeros((A.N, A.N))
    for i in range(A.N):
        for k in range(rp[i], rp[i + 1]):
            j = col[k]
            if j <= i:
                dense[i, j] = dense[j, i] = rng.uniform(-1, 1) if j < i else 0.0
    dense += np.diag(np.abs(dense).sum(1) + 1.0)
    val = np.array([dense[i, col[k]] for i in range(A.N) for k in range(rp[i], rp[i + 1])])
    A.import_values(val)
    b = rng.standard_normal(A.N)
    x, info = A.solve(b, mode=1, tol=1e-12)
    assert info["layout"] == 1
    assert np.linalg.norm(x - np.linalg.solve(dense, b)) <= 1e-10 * np.linalg.norm(x)
    xs = rng.standard_normal(A.N)
    assert np.array_equal(A.matvec(xs), oracle.spmv(rp, col, val, xs))
    A.destroy()


@pytest.mark.parametrize("mode", [1, 2])
def test_sparse_true_residual_stop(mode):
    """nbx_sparse_solve_ex with true_residual = 1 (FP64 PCG and the mixed
    Chebyshev flexible CG): the returned x meets the tolerance on the true
    residual, checked with the oracle's reference-order SpMV."""
    import oracle
    import nbots_amd as nb
    from nbots_amd import synth
    from nbots_amd.fem import solve_opts
    p = synth.cantilever_traction(128, 64)
    rp, col = oracle.pattern(p.N_nod, p.conn)
    val, F, _ = oracle.assemble(p, rp, col)
    oracle.set_bconditions(p, rp, col, val, F)
    A = nb.Sparse.from_csr(rp, col, val, vars_per_node=2)
    x, info = A.solve(F, mode=mode, tol=1e-10, opts=solve_opts(true_residual=True))
    assert info["status"] == 0 and info["layout"] == (15 if mode == 2 else 1)
    r = F - oracle.spmv(rp, col, val, x)
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(F)
    A.destroy()


@pytest.mark.parametrize("where", ["off_diagonal_block", "diagonal_block"])
def test_symmetric_layout_refused_for_asymmetric_values(where):
    """One entry changed on one side, in an off-diagonal block or in the
    x-y coupling of a diagonal block (which the symmetric layout stores
    once): the symmetry check fails, the solve runs on the full layout
    (layout 0), SpMV stays exact, and after the values are restored the next
    solve uses the symmetric copy again."""
    import oracle
    g = load_golden("cant_8x4_jit")
    A = _sparse(g)
    rp, col, val = g["rowptr"], g["col"], g["val"].copy()
    r = 6
    if where == "off_diagonal_block":
        j = [k for k in range(rp[r], rp[r + 1]) if col[k] // 2 != r // 2 and val[k] != 0.0][0]
        val[j] *= 1.0 + 2.0 ** -40
    else:
        j = [k for k in range(rp[r], rp[r + 1]) if col[k] == r + 1][0]
        val[j] = np.nextafter(val[j], np.inf)
    A.import_values(val)
    _, info = A.solve(g["F"], mode=0, tol=1e-8)
    assert info["layout"] == 0
    xs = np.cos(np.arange(A.N, dtype=np.float64))
    assert np.array_equal(A.matvec(xs), oracle.spmv(rp, col, val, xs))
    A.import_values(g["val"])
    _, info = A.solve(g["F"], mode=0, tol=1e-8)
    assert info["layout"] == 1
    A.destroy()


@pytest.mark.parametrize("vpn", [1, 3])
def test_spmv_scalar_csr_path(vpn):
    """vars_per_node != 2 goes through the scalar CSR kernels."""
    import oracle
    import nbots_amd as nb
    from nbots_amd import synth
    p = synth.cantilever_point(6, 3, jitter=0.1)
    rp, col = oracle.pattern(p.N_nod, p.conn, vpn=vpn)
    rng = np.random.default_rng(7)
    val = rng.standard_normal(col.size)
    A = nb.Sparse.from_csr(rp, col, val, vars_per_node=vpn)
    assert A.is_blocked() == 0
    x = rng.standard_normal(A.N)
    assert np.array_equal(A.matvec(x), oracle.spmv(rp, col, val, x))
    A.destroy()


@pytest.mark.parametrize("case", SMALL_CASES)
def test_pcg_reference_semantics(case):
    g = load_golden(case)
    A = _sparse(g)
    tol = float(g["pcg_tol"])
    x, st, it, tr = A.solve_cg_jacobi(g["F"], tol=tol)
    assert st == int(g["pcg_status"])
    assert abs(it - int(g["niter"])) <= 2, (it, int(g["niter"]))
    ref = g["x"]
    assert np.linalg.norm(x - ref) <= 1e-10 * np.linalg.norm(ref)
    # cut at 7 iterations: the lagged stop test and tol_reached semantics
    x7, st7, it7, tr7 = A.solve_cg_jacobi(g["F"], tol=tol, max_iter=7)
    assert (st7, it7) == (int(g["status7"]), int(g["niter7"]))
    assert math.isclose(tr7, float(g["tol_reached7"]), rel_tol=1e-9)
    assert np.linalg.norm(x7 - g["x7"]) <= 1e-11 * np.linalg.norm(g["x7"])
    A.destroy()


@pytest.mark.parametrize("max_iter", [0, 1, 2, 3, 4, 5, 6, 7, 8, 9])
def test_pcg_warm_start_cuts(max_iter):
    """From a nonzero x0, cut after 0..9 iterations: the device's x, niter,
    status and tol_reached follow the oracle's restatement of
    cg_precond_jacobi.c; with 0 iterations x comes back bit for bit.  The
    device applies x += alpha p four iterations at a time and the pending
    one to four terms after the stop test, so these cuts cover every
    remainder, before and after a group has been applied."""
    import oracle
    g = load_golden("cant_8x4_jit")
    A = _sparse(g)
    rng = np.random.default_rng(7)
    x0 = rng.standard_normal(g["F"].size)
    x, st, it, tr = A.solve_cg_jacobi(g["F"], x0=x0, max_iter=max_iter, tol=1e-10)
    rp, col, val = g["rowptr"], g["col"], g["val"]
    xr, itr, trr, str_ = oracle.pcg(rp, col, val, g["F"], x0=x0, max_iter=max_iter, tol=1e-10)
    assert (st, it) == (str_, itr)
    assert math.isclose(tr, trr, rel_tol=1e-12)
    if max_iter == 0:
        assert np.array_equal(x, x0)
    assert np.linalg.norm(x - xr) <= 1e-12 * np.linalg.norm(xr)
    # started at the converged solution: no iteration, x unchanged
    x2, st2, it2, _ = A.solve_cg_jacobi(g["F"], x0=xr, max_iter=max_iter, tol=1e30)
    assert (st2, it2) == (0, 0) and np.array_equal(x2, xr)
    A.destroy()


def test_pcg_c1_cantilever():
    """C1 (64x32 jittered cantilever, 4,290 DOF): the reference driver's solve."""
    g = load_golden("c1_cantilever_64x32")
    import oracle
    p = problem_from_golden(g)
    rp, col = oracle.pattern(p.N_nod, p.conn)
    val, F, _ = oracle.assemble(p, rp, col)
    oracle.set_bconditions(p, rp, col, val, F)
    import nbots_amd as nb
    A = nb.Sparse.from_csr(rp, col, val, vars_per_node=2)
    x, st, it, tr = A.solve_cg_jacobi(F, tol=1e-8)
    assert st == 0
    assert abs(it - int(g["niter"])) <= 3
    assert np.linalg.norm(x - g["x"]) <= 1e-9 * np.linalg.norm(g["x"])
    A.destroy()


def test_pcg_edge_cases():
    import nbots_amd as nb
    cases = load_json("pcg_edge_cases.json")
    # 1x1: exact convergence, then the lagged extra iteration divides 0/0
    A = nb.Sparse.from_csr(np.array([0, 1], np.uint32), np.array([0], np.uint32),
                           np.array([4.0]), vars_per_node=1)
    x, st, it, tr = A.solve_cg_jacobi(np.array([2.0]), max_iter=10, tol=1e-8)
    c = cases["one_by_one"]
    assert math.isnan(x[0]) and it == c["niter"] and tr == c["tol_reached"] and st == c["status"]
    A.destroy()
    rp = np.array([0, 2, 4], np.uint32)
    col = np.array([0, 1, 0, 1], np.uint32)
    A = nb.Sparse.from_csr(rp, col, np.array([4.0, 1.0, 1.0, 3.0]), vars_per_node=1)
    x, st, it, tr = A.solve_cg_jacobi(np.array([1.0, 2.0]), max_iter=1, tol=1e-12)
    c = cases["two_by_two_max1"]
    assert (st, it) == (c["status"], c["niter"])
    assert np.allclose(x, c["x"], rtol=1e-15, atol=0) and math.isclose(tr, c["tol_reached"])
    x, st, it, tr = A.solve_cg_jacobi(np.zeros(2), max_iter=5, tol=1e-8)
    c = cases["zero_rhs"]
    assert (st, it, tr) == (c["status"], c["niter"], c["tol_reached"]) and not x.any()
    A.destroy()


def test_solve_modes_small():
    """FP64 relative-tolerance and mixed FP32/FP64 modes reach 1e-8 relative
    residual on the traction fixture; displacement within 1e-6 of the
    reference solution (which itself stops at 1e-8)."""
    import nbots_amd as nb
    g = load_golden("traction_16x8")
    A = _sparse(g)
    F = g["F"]
    rp, col, val = g["rowptr"], g["col"], g["val"]
    import oracle
    for mode in (nb.SOLVE_FP64, nb.SOLVE_MIXED):
        x, info = A.solve(F, mode=mode, tol=1e-8)
        assert info["status"] == 0
        r = F - oracle.spmv(rp, col, val, x)
        assert np.linalg.norm(r) <= 1.0000001e-8 * np.linalg.norm(F), info
        assert np.linalg.norm(x - g["x"]) <= 1e-6 * np.linalg.norm(g["x"])
    A.destroy()


CG_CASES = {"traction_16x8": lambda s: s.cantilever_traction(16, 8),
            "c1_64x32": lambda s: s.cantilever_point(64, 32, jitter=0.1)}


@pytest.mark.parametrize("key", sorted(CG_CASES))
def test_plain_cg_reference_semantics(key):
    """nb_sparse_solve_conjugate_gradient (conjugate_gradient.c:13-77) against
    the compiled reference's run on the same K and F: status equal, niter
    within 2 and x within 1e-10 at convergence; cut at 7 iterations the
    status and niter are equal and tol_reached within 1e-9."""
    import nbots_amd as nb
    import oracle
    from nbots_amd import synth
    g = load_golden("cg_cases")
    p = CG_CASES[key](synth)
    rp, col = oracle.pattern(p.N_nod, p.conn)
    val, F, _ = oracle.assemble(p, rp, col)
    oracle.set_bconditions(p, rp, col, val, F)
    A = nb.Sparse.from_csr(rp, col, val, vars_per_node=2)
    tol = float(g[key + "_tol"])
    x, st, it, tr = A.solve_cg(F, tol=tol, max_iter=10 * p.N)
    assert st == int(g[key + "_status"])
    assert abs(it - int(g[key + "_niter"])) <= 2, (it, int(g[key + "_niter"]))
    ref = g[key + "_x"]
    assert np.linalg.norm(x - ref) <= 1e-10 * np.linalg.norm(ref)
    x7, st7, it7, tr7 = A.solve_cg(F, tol=tol, max_iter=7)
    assert (st7, it7) == (int(g[key + "_status7"]), int(g[key + "_niter7"]))
    assert math.isclose(tr7, float(g[key + "_tol_reached7"]), rel_tol=1e-9)
    assert np.linalg.norm(x7 - g[key + "_x7"]) <= 1e-11 * np.linalg.norm(g[key + "_x7"])
    A.destroy()


def test_plain_cg_edge_cases():
    """1x1 exact convergence then the lagged 0/0 iteration (x = NaN, niter
    2, as the compiled reference); empty right-hand side."""
    import nbots_amd as nb
    g = load_golden("cg_cases")
    A = nb.Sparse.from_csr(np.array([0, 1], np.uint32), np.array([0], np.uint32),
                           np.array([4.0]), vars_per_node=1)
    x, st, it, tr = A.solve_cg(np.array([2.0]), max_iter=10, tol=1e-8)
    assert math.isnan(x[0])
    assert (it, st, tr) == (int(g["one_niter"]), int(g["one_status"]), float(g["one_tol_reached"]))
    x, st, it, tr = A.solve_cg(np.zeros(1), max_iter=10, tol=1e-8)
    assert (st, it, tr) == (0, 0, 0.0) and x[0] == 0.0
    A.destroy()


def test_mixed_solve_wide_node_numbering():
    """Nodes numbered at random (512x128 cantilever, 66,177 nodes): block
    columns lie up to 66k rows from the diagonal, beyond the 16-bit deltas,
    so the fp16 preconditioner reads its 32-bit columns (layout 3, not 15).
    The mixed solve still meets the tolerance on the oracle's residual and
    agrees with the solve in the original numbering."""
    import scipy.sparse as sp
    import oracle
    import nbots_amd as nb
    from nbots_amd import synth
    from nbots_amd.fem import solve_opts
    p = synth.cantilever_traction(512, 128)
    rp, col = oracle.pattern(p.N_nod, p.conn)
    val, F, _ = oracle.assemble(p, rp, col)
    oracle.set_bconditions(p, rp, col, val, F)
    N = p.N
    perm = np.random.default_rng(7).permutation(p.N_nod)        # old node -> new node
    dof = (2 * perm[:, None] + np.arange(2)).ravel()            # old dof -> new dof
    K = sp.csr_matrix((val, col.astype(np.int64), rp.astype(np.int64)), shape=(N, N)).tocoo()
    Kp = sp.csr_matrix((K.data, (dof[K.row], dof[K.col])), shape=(N, N))
    Kp.sort_indices()
    rpp, colp, valp = Kp.indptr.astype(np.uint32), Kp.indices.astype(np.uint32), Kp.data.copy()
    assert valp.size == val.size
    rows = np.repeat(np.arange(N), np.diff(Kp.indptr))
    assert np.abs(colp.astype(np.int64) // 2 - rows // 2).max() >= 32768
    Fp = np.empty(N)
    Fp[dof] = F
    opts = solve_opts(true_residual=True)
    A = nb.Sparse.from_csr(rpp, colp, valp, vars_per_node=2)
    x, info = A.solve(Fp, mode=nb.SOLVE_MIXED, tol=1e-8, opts=opts)
    assert info["status"] == 0 and info["layout"] == 3, info
    r = Fp - oracle.spmv(rpp, colp, valp, x)
    assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(Fp)
    A.destroy()
    A0 = nb.Sparse.from_csr(rp, col, val, vars_per_node=2)
    x0, info0 = A0.solve(F, mode=nb.SOLVE_MIXED, tol=1e-8, opts=opts)
    assert info0["status"] == 0 and info0["layout"] == 15, info0
    assert np.linalg.norm(x[dof] - x0) <= 1e-6 * np.linalg.norm(x0)
    A0.destroy()


def test_gauss_seidel_drop_in_matches_reference():
    """nb_sparse_solve_Gauss_Seidel on the GPU against the compiled
    reference's (gauss_seidel.c:13-63, fixture gauss_seidel_16x8.npz) on K and
    F of the 16x8 cantilever: after 200 sweeps and after 20000, x bit for bit
    (the level-scheduled substitution sums each row in the reference's
    order), the same sweep count and status, the error within 1e-12 (its
    sum runs in another order).  A system that converges stops at the same
    sweep as the oracle's."""
    import nbots_amd as nb
    import oracle
    from nbots_amd import synth
    g = load_golden("gauss_seidel_16x8")
    p = synth.cantilever_traction(16, 8)
    rp, col = oracle.pattern(p.N_nod, p.conn)
    val, F, _ = oracle.assemble(p, rp, col)
    oracle.set_bconditions(p, rp, col, val, F)
    A = nb.Sparse.from_csr(rp, col, val, vars_per_node=2)
    for key in ("cut", "tol"):
        x, st, it, err = A.solve_gauss_seidel(F, max_iter=int(g[key + "_max_iter"]),
                                              tol=float(g[key + "_tol"]))
        assert np.array_equal(x, g[key + "_x"]), key
        assert (st, it) == (int(g[key + "_status"]), int(g[key + "_niter"]))
        assert abs(err - float(g[key + "_error"])) <= 1e-12 * float(g[key + "_error"])
    A.destroy()
    # diagonally dominant: converges; x, sweeps and status as the oracle's
    rp2 = np.array([0, 2, 5, 7], np.uint32)
    col2 = np.array([0, 1, 0, 1, 2, 1, 2], np.uint32)
    val2 = np.array([4.0, 1.0, 1.0, 5.0, 2.0, 2.0, 6.0])
    b2 = np.array([1.0, 2.0, 3.0])
    B = nb.Sparse.from_csr(rp2, col2, val2, vars_per_node=1)
    x, st, it, err = B.solve_gauss_seidel(b2, tol=1e-12)
    xo, ito, erro, sto = oracle.gauss_seidel(rp2, col2, val2, b2, tol=1e-12)
    assert st == sto == 0 and it == ito and np.array_equal(x, xo)
    B.destroy()


@pytest.mark.parametrize("case,uniform", [("traction_128x64", True), ("beam_cantilever", True),
                                          ("plate_with_hole", False)])
def test_fp16_off_diagonal_placements(case, uniform):
    """The fp16 preconditioner's two placements of the off-diagonal blocks:
    every slice as wide as the widest row (layout bit 3; the structured
    cantilever, whose interior nodes all have six neighbours, and the beam,
    whose widest rows pad 1.6 %) or one width per slice (the plate with a
    hole, whose nine-neighbour rows would pad 13 %).  Both solve
    K x = K x_true to the true-residual tolerance."""
    import oracle
    import nbots_amd as nb
    from nbots_amd.fem import solve_opts
    p = problem_from_golden(load_golden(case))
    rp, col = oracle.pattern(p.N_nod, p.conn)
    val, F, _ = oracle.assemble(p, rp, col)
    oracle.set_bconditions(p, rp, col, val, F)
    x_true = np.random.default_rng(5).standard_normal(len(F))
    b = oracle.spmv(rp, col, val, x_true)
    A = nb.Sparse.from_csr(rp, col, val, vars_per_node=2)
    x, info = A.solve(b, mode=2, tol=1e-10, opts=solve_opts(true_residual=True))
    assert info["status"] == 0, info
    assert info["layout"] & 7 == 7 and bool(info["layout"] & 8) == uniform, info["layout"]
    r = b - oracle.spmv(rp, col, val, x)
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(b)
    A.destroy()
