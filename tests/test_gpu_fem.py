"""GPU parity of the device FEM system and the drop-in driver.

Bars (see the module docstrings of the kernels for why they hold):
  * pattern, K and F before and after boundary conditions: bit-exact with
    the golden vectors the compiled reference wrote;
  * strain/stress of a given displacement: bit-exact with the oracle;
  * displacement of a full solve: relative L2 <= 1e-9 against the
    reference's own solve (same stopping rule, different dot-product order);
  * the reference's beam-cantilever unit test: max |u| = 1.00701e-1 +- 1e-6
    (utest/.../static_elasticity2D.c:118).
"""
import hashlib
import math

import numpy as np
import pytest

from conftest import SMALL_CASES, kirsch_vm_error, load_golden, problem_from_golden

pytestmark = pytest.mark.gpu


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("case", SMALL_CASES)
def test_assembly_and_bc_bit_exact(case):
    from nbots_amd.fem import Fem
    g = load_golden(case)
    p = problem_from_golden(g)
    f = Fem(p.nod, p.conn)
    assert f.assemble(p.E, p.nu, p.thickness, p.elem_E) == 0
    rp, col, val, F = f.system()
    assert np.array_equal(rp, g["rowptr"]) and np.array_equal(col, g["col"])
    assert np.array_equal(val, g["val_nobc"]), np.abs(val - g["val_nobc"]).max()
    assert np.array_equal(F, g["F_nobc"])
    f.set_bconditions(p)
    rp, col, val, F = f.system()
    assert np.array_equal(val, g["val"]), np.abs(val - g["val"]).max()
    assert np.array_equal(F, g["F"]), np.abs(F - g["F"]).max()
    f.destroy()


@pytest.mark.parametrize("case", ["c1_cantilever_64x32", "traction_128x64"])
def test_assembly_bit_exact_digests(case):
    from nbots_amd.fem import Fem
    g = load_golden(case)
    p = problem_from_golden(g)
    f = Fem.from_problem(p, bc=False)
    rp, col, val, F = f.system()
    assert sha(rp) == str(g["rowptr_sha"]) and sha(col) == str(g["col_sha"])
    assert sha(val) == str(g["val_nobc_sha"])
    f.set_bconditions(p)
    rp, col, val, F = f.system()
    assert sha(val) == str(g["val_sha"]) and sha(F) == str(g["F_sha"])
    f.destroy()


def test_beam_cantilever_known_answer():
    """The reference's own FEM test: 1000-vertex Delaunay mesh from the
    reference tessellator, prescribed u_y = -0.1 at a vertex, replayed
    through the drop-in nb_fem_compute_2D_Solid_Mechanics."""
    from nbots_amd.fem import compute_2D_solid_mechanics
    g = load_golden("beam_cantilever")
    p = problem_from_golden(g)
    st, disp, strain, stress = compute_2D_solid_mechanics(p)
    assert st == 0
    max_disp = np.sqrt((disp.reshape(-1, 2) ** 2).sum(1).max())
    assert abs(max_disp - 1.00701e-1) < 1e-6
    assert np.linalg.norm(disp - g["disp"]) <= 1e-9 * np.linalg.norm(g["disp"])
    assert np.linalg.norm(strain - g["strain"]) <= 1e-8 * np.linalg.norm(g["strain"])
    assert np.linalg.norm(stress - g["stress"]) <= 1e-8 * np.linalg.norm(g["stress"])


def test_plate_with_hole_known_answer():
    """The reference's plate-with-a-hole test (static_elasticity2D.c:120-136)
    through the drop-in: Kirsch tractions pushed as function-valued Neumann
    segments; fields match the compiled reference's, and the mean von Mises
    error against Kirsch stays under the test's 9.7e-3."""
    from nbots_amd.fem import compute_2D_solid_mechanics
    g = load_golden("plate_with_hole")
    p = problem_from_golden(g)
    st, disp, strain, stress = compute_2D_solid_mechanics(p)
    assert st == 0
    assert np.linalg.norm(disp - g["disp"]) <= 1e-9 * np.linalg.norm(g["disp"])
    assert np.linalg.norm(strain - g["strain"]) <= 1e-8 * np.linalg.norm(g["strain"])
    assert np.linalg.norm(stress - g["stress"]) <= 1e-8 * np.linalg.norm(g["stress"])
    assert kirsch_vm_error(p, stress) < 9.7e-3


@pytest.mark.parametrize("case", ["cant_2x1", "cant_4x2_jit", "cant_8x4_jit"])
def test_driver_matches_reference(case):
    from nbots_amd.fem import compute_2D_solid_mechanics
    g = load_golden(case)
    p = problem_from_golden(g)
    st, disp, strain, stress = compute_2D_solid_mechanics(p)
    assert st == 0
    assert np.linalg.norm(disp - g["disp"]) <= 1e-9 * np.linalg.norm(g["disp"])
    assert np.linalg.norm(strain - g["strain"]) <= 1e-8 * np.linalg.norm(g["strain"])
    assert np.linalg.norm(stress - g["stress"]) <= 1e-8 * np.linalg.norm(g["stress"])


def test_driver_distorted_element_status():
    """A clockwise triangle makes the reference's assembly fail: status 1
    (pipeline.c:158-159, static_elasticity2D.c:62-65)."""
    from nbots_amd import synth
    from nbots_amd.fem import compute_2D_solid_mechanics
    p = synth.cantilever_point(4, 2, jitter=0.0)
    c = p.conn.reshape(-1, 3).copy()
    c[3, 1], c[3, 2] = c[3, 2], c[3, 1]
    p.conn = c.reshape(-1)
    st, _, _, _ = compute_2D_solid_mechanics(p)
    assert st == 1


def test_strain_stress_bit_exact():
    """Strain and stress of the GPU's displacement equal the oracle's on it."""
    import oracle
    from nbots_amd import synth
    from nbots_amd.fem import Fem, SOLVE_FP64
    p = synth.cantilever_traction(24, 12)
    f = Fem.from_problem(p)
    f.solve(SOLVE_FP64, tol=1e-10)
    u = f.displacement()
    eps, sig = f.strain_stress()
    assert np.array_equal(eps, oracle.strain(p, u))
    assert np.array_equal(sig, oracle.stress(p, eps))
    f.destroy()


@pytest.mark.parametrize("mode,variant,bits", [(1, None, None), (2, 0, None), (2, 1, None),
                                               (2, 2, 16), (2, 2, 32), (2, 2, 64)])
def test_solve_modes_traction(mode, variant, bits):
    """FP64 PCG and the mixed-precision variants (FP32 CG sweeps in FP64
    refinement, FP32 CG with FP64 reliable updates, FP64 flexible CG with a
    Chebyshev-Jacobi preconditioner on fp16 or FP32 matrix values, and the
    same preconditioner all in FP64 for comparison) reach
    relative residual 1e-8 (checked with the bit-exact SpMV of the oracle
    on the device's K) and agree with a 1e-13 oracle solve within rel-L2
    1e-6 (the error bound of a 1e-8 residual times the condition number is
    far larger; 1e-6 is what all modes achieve here)."""
    import oracle
    g = load_golden("traction_128x64")
    p = problem_from_golden(g)
    from nbots_amd.fem import Fem, solve_opts
    f = Fem.from_problem(p)
    info = f.solve(mode, tol=1e-8,
                   opts=None if variant is None else solve_opts(variant, precond_bits=bits))
    assert info["status"] == 0, info
    assert info["layout"] & 1       # the assembled K is symmetric bit for bit
    # the fp16 preconditioner reads the copy without diagonal blocks, with
    # its columns as packed 16-bit deltas
    fp16 = mode == 2 and variant == 2 and bits == 16
    assert bool(info["layout"] & 2) == fp16 and bool(info["layout"] & 4) == fp16
    u = f.displacement()
    rp, col, val, F = f.system()
    r = F - oracle.spmv(rp, col, val, u)
    assert np.linalg.norm(r) <= 1.0000001e-8 * np.linalg.norm(F), (np.linalg.norm(r) / np.linalg.norm(F), info)
    x_ref, _, _, _ = oracle.pcg(rp, col, val, F, tol=1e-13 * np.linalg.norm(F), max_iter=100000)
    err = np.linalg.norm(u - x_ref) / np.linalg.norm(x_ref)
    print("mode", mode, "rel err vs 1e-13 oracle", err, info)
    assert err <= 1e-6
    f.destroy()


def test_fp64_preconditioner_matches_fp16_iterations():
    """The Chebyshev-FCG with the preconditioner all in FP64 (FP64 matrix and
    vectors) and with fp16 values, FP32 vectors and FP64 row sums both reach
    the true residual 1e-8; the low precision weakens the preconditioner a
    little (on this 128x64 plate 77 outer iterations against 65), never by
    more than a third, and never helps it.  The FP64 run is the reference
    point of the bench's precision factor."""
    from nbots_amd.fem import Fem, SOLVE_MIXED, solve_opts
    g = load_golden("traction_128x64")
    p = problem_from_golden(g)
    f = Fem.from_problem(p)
    its = {}
    for bits in (64, 16):
        info = f.solve(SOLVE_MIXED, tol=1e-8, opts=solve_opts(precond_bits=bits, true_residual=True))
        assert info["status"] == 0 and info["rel_res"] <= 1e-8, info
        assert f.true_residual(True) <= 1.0000001e-8
        its[bits] = info["iters"]
    print("outer iterations, FP64 / fp16 preconditioner:", its)
    assert its[64] - 2 <= its[16] <= 1.33 * its[64]
    f.destroy()


@pytest.mark.parametrize("tol", [1e-6, 1e-12])
def test_true_residual_compensated_is_exact(tol):
    """nbx_fem_true_residual: the compensated (Dot2) evaluation of
    ||F - K u|| / ||F|| equals the exact rational value for the device's
    displacement within 1e-12 relative; the plain FP64 evaluation is within
    its own rounding, about 16 eps * || |K| |u| + |F| || / ||F||."""
    from fractions import Fraction
    from nbots_amd.fem import Fem, SOLVE_FP64
    g = load_golden("traction_16x8")
    p = problem_from_golden(g)
    f = Fem.from_problem(p)
    f.solve(SOLVE_FP64, tol=tol)
    u = f.displacement()
    rp, col, val, F = f.system()
    r2, f2, absku = Fraction(0), Fraction(0), 0.0
    for i in range(F.size):
        ks = range(rp[i], rp[i + 1])
        s = Fraction(F[i]) - sum(Fraction(val[k]) * Fraction(u[col[k]]) for k in ks)
        r2 += s * s
        f2 += Fraction(F[i]) ** 2
        absku += (sum(abs(val[k] * u[col[k]]) for k in ks) + abs(F[i])) ** 2
    exact = math.sqrt(r2 / f2)
    assert abs(f.true_residual(True) - exact) <= 1e-12 * exact
    floor = np.finfo(float).eps * math.sqrt(absku / float(f2)) * 16
    assert abs(f.true_residual(False) - exact) <= floor
    f.destroy()


@pytest.mark.parametrize("mode", [1, 2])
def test_true_residual_stop(mode):
    """true_residual = 1: FP64 PCG and the Chebyshev flexible CG return only
    when ||F - K u|| / ||F||, computed afresh, meets the tolerance; with an
    unattainable tolerance they give up with NBX_NOT_CONVERGED once a
    restart stops gaining, after at most 8 passes."""
    from nbots_amd import synth
    from nbots_amd.fem import Fem, solve_opts
    p = synth.cantilever_traction(256, 128)
    f = Fem.from_problem(p)
    o = solve_opts(true_residual=True)
    info = f.solve(mode, tol=1e-8, opts=o, u0=np.zeros(p.N))
    assert info["status"] == 0
    assert f.true_residual(True) <= 1e-8
    assert abs(info["rel_res"] - f.true_residual(False)) <= 1e-3 * info["rel_res"]
    info = f.solve(mode, tol=1e-17, opts=o)
    assert info["status"] == 1
    assert f.true_residual(True) <= 1e-8
    f.destroy()


def test_compliance_sensitivities_small():
    """SIMP on the MBB beam: element energies bit-exact with the oracle on
    the same displacement; C = F.u (energy identity); dC/drho against
    central finite differences (step 1e-6, FP64 solves to 1e-13)."""
    import oracle
    from nbots_amd import synth
    from nbots_amd.fem import Fem, SOLVE_FP64
    p = synth.mbb(16, 8)
    f = Fem.from_problem(p)
    f.solve(SOLVE_FP64, tol=1e-13)
    u = f.displacement()
    C, dc, ce = f.compliance(p.rho)
    assert np.array_equal(ce, oracle.elem_energy(p, u))
    Cref, dcref = oracle.simp_sens(p.rho, ce)
    assert abs(C - Cref) <= 1e-12 * abs(Cref)
    assert np.allclose(dc, dcref, rtol=1e-13, atol=0)
    _, _, _, F = f.system()
    assert abs(C - F @ u) <= 1e-10 * abs(C)
    h = 1e-6
    rng = np.random.default_rng(3)
    for e in rng.choice(p.N_elem, 6, replace=False):
        Cs = []
        for s in (+1, -1):
            rho = p.rho.copy()
            rho[e] += s * h
            f.assemble(p.E, p.nu, p.thickness, synth.simp_modulus(rho))
            f.set_bconditions(p)
            f.solve(SOLVE_FP64, tol=1e-14)
            Cs.append(f.displacement() @ F)
        fd = (Cs[0] - Cs[1]) / (2 * h)
        assert abs(fd - dc[e]) <= 1e-5 * abs(dc[e]), (e, fd, dc[e])
    f.destroy()


def test_adjoint_general_rhs():
    """K lambda = r for an arbitrary right-hand side (zero on Dirichlet
    dofs); for r = F the adjoint of compliance is u itself."""
    import oracle
    from nbots_amd import synth
    from nbots_amd.fem import Fem, SOLVE_FP64
    p = synth.mbb(32, 16)
    f = Fem.from_problem(p)
    f.solve(SOLVE_FP64, tol=1e-12)
    u = f.displacement()
    rp, col, val, F = f.system()
    lam, info = f.adjoint(F, rtol=1e-12)
    assert np.linalg.norm(lam - u) <= 1e-9 * np.linalg.norm(u)
    rng = np.random.default_rng(5)
    r = rng.standard_normal(f.N)
    fixed = np.zeros(f.N, bool)
    fixed[np.arange(0, 2 * (p.nx + 1) * (p.ny + 1), 2 * (p.nx + 1))] = True   # u_x on left edge
    fixed[2 * p.nx + 1] = True                                                 # u_y at (nx, 0)
    r[fixed] = 0.0
    lam, info = f.adjoint(r, rtol=1e-12)
    res = r - oracle.spmv(rp, col, val, lam)
    assert np.linalg.norm(res) <= 1e-11 * np.linalg.norm(r)
    f.destroy()


def test_spmv_bench_entry():
    from nbots_amd import synth
    from nbots_amd.fem import Fem
    p = synth.cantilever_traction(64, 32)
    f = Fem.from_problem(p)
    assert f.spmv_ms(64, 5) > 0 and f.spmv_ms(32, 5) > 0 and f.spmv_ms(16, 5) > 0
    assert f.spmv_ms(65, 5) > 0     # symmetric FP64 layout
    f.destroy()


def check_against_reference_solve(u, F, g, rtol):
    """u against the displacement the compiled reference's FP64 Jacobi-PCG
    computed (tests/golden/c2|c3_reference_solve.npz): every 101st / 401st
    dof, ||u||, the compliance F.u and the values w_k.u for the 8 vectors of
    synth.projections.  A difference e moves each w_k.u by a random amount
    of standard deviation ||e|| / sqrt(3), so the projections estimate
    ||e|| / ||u|| over the whole vector; all within rtol."""
    from nbots_amd import synth
    un = float(g["u_norm"])
    us = u[g["idx"]]
    assert np.linalg.norm(us - g["u"]) <= rtol * np.linalg.norm(g["u"])
    assert abs(np.linalg.norm(u) - un) <= rtol * un
    assert abs(float(F @ u) - float(g["F_dot_u"])) <= rtol * abs(float(g["F_dot_u"]))
    d = synth.projections(u.size) @ u - g["proj"]
    est = float(np.sqrt(3.0 * np.mean(d * d))) / un
    print(f"samples rel {np.linalg.norm(us - g['u']) / np.linalg.norm(g['u']):.2e}, "
          f"projection estimate of ||u - u_ref|| / ||u_ref||: {est:.2e}")
    assert est <= rtol
    assert np.abs(d).max() <= 2.5 * rtol * un


@pytest.mark.parametrize("name,nx,ny", [("c2_reference_solve", 1024, 512),
                                        ("c3_reference_solve", 2048, 1024)])
def test_fp64_against_reference_solve(name, nx, ny):
    """Lognormal cantilever, right-edge shear traction, 1024x512 (1,051,650
    DOF) and 2048x1024 (4,200,450 DOF, the bench workload): the device's K
    and F equal the compiled reference's (SHA-256), and its FP64 Jacobi-PCG
    with the reference's stop test (tol 1e-8 ||F||), on the symmetric
    layout, takes the reference's iteration count within 0.5 % (the dot
    products are summed in another order) and reaches the reference's
    displacement within 1e-7 (sampled dofs and ||u||)."""
    from nbots_amd import synth
    from nbots_amd.fem import Fem, SOLVE_REFERENCE
    g = load_golden(name)
    p = synth.cantilever_traction(nx, ny)
    f = Fem.from_problem(p)
    rp, col, val, F = f.system()
    assert sha(val) == str(g["K_sha"]) and sha(F) == str(g["F_sha"])
    del rp, col, val
    info = f.solve(SOLVE_REFERENCE, tol=float(g["tol"]), max_iter=p.N)
    assert info["status"] == int(g["status"]) == 0
    assert info["layout"] == 1
    n_ref = int(g["niter"])
    assert abs(info["iters"] - n_ref) <= 0.005 * n_ref, (info["iters"], n_ref)
    u = f.displacement()
    check_against_reference_solve(u, F, g, 1e-7)
    f.destroy()
