"""Host-side behaviour of the drop-in API that needs no GPU: the
reference's exit(1) on an entry outside the pattern (sparse.c:207-219),
argument checks of the nbx extensions that run before any device call, and
the object accessors (material.h, static_damage2D.h)."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _diag2():
    import nbots_amd as nb
    return nb.Sparse.from_csr(np.array([0, 1, 2], np.uint32), np.array([0, 1], np.uint32),
                              np.array([1.0, 2.0]), vars_per_node=1)


def test_add_outside_pattern_exits_like_reference():
    code = ("import numpy as np, nbots_amd as nb\n"
            "A = nb.Sparse.from_csr(np.array([0, 1, 2], np.uint32), np.array([0, 1], np.uint32),"
            " np.array([1.0, 2.0]), vars_per_node=1)\n"
            "A.L.nb_sparse_add(A.h, 0, 1, 3.5)\n"
            "print('not reached')\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 1, (r.returncode, r.stdout, r.stderr)
    lines = r.stdout.splitlines()
    assert lines[0] == "ERROR: Entry of sparse matrix is not allocated."
    assert lines[1].startswith("    -> nb_sparse_add(*A=") and lines[1].endswith(
        ", i=0, j=1, value=3.500000)")
    assert "not reached" not in r.stdout


def test_get_outside_pattern_is_zero():
    A = _diag2()
    assert A.L.nb_sparse_get(A.h, 1, 1) == 2.0 and A.L.nb_sparse_get(A.h, 0, 1) == 0.0
    A.destroy()


def test_solve_options_checked_before_device_use():
    import nbots_amd as nb
    from nbots_amd.fem import solve_opts
    from nbots_amd.lib import SolveInfo
    A = _diag2()
    x = np.zeros(2)
    info = SolveInfo()
    for field, bad in (("mixed_variant", 7), ("precond_bits", 8)):
        o = solve_opts()
        setattr(o, field, bad)
        rc = A.L.nbx_sparse_solve_ex(A.h, np.ones(2), x, 2, 1e-8, 10, C.byref(o), C.byref(info))
        assert rc == 4, rc
        assert ("variant" if field == "mixed_variant" else "precond_bits") in nb.lib.last_error()
    A.destroy()


def test_solve_options_defaults():
    from nbots_amd.fem import solve_opts, MIXED_CHEB_FCG
    o = solve_opts()
    assert (o.mixed_variant, o.cheb_degree, o.cheb_ratio, o.precond_bits) == (MIXED_CHEB_FCG, 40,
                                                                               3e4, 16)


def test_fem_create_rejects_empty_mesh():
    from nbots_amd import lib
    L = lib.load()
    h = C.c_void_p()
    rc = L.nbx_fem_create(C.byref(h), 0, np.zeros(2), 0, np.zeros(3, np.uint32))
    assert rc == 4 and not h.value and "empty" in lib.last_error()


def test_material_and_implicit_accessors():
    from nbots_amd import lib
    L = lib.load()
    m = L.nb_material_create()
    for name, v in (("elasticity_module", 210.0), ("poisson_module", 0.25), ("density", 7.8),
                    ("fracture_energy", 0.05), ("traction_limit_stress", 3.0),
                    ("compression_limit_stress", 30.0), ("energy_release_rate", 0.1),
                    ("damage_length_scale", 0.02)):
        getattr(L, "nb_material_set_" + name)(m, v)
    c = L.nb_material_clone(m)
    for name, v in (("elasticity_module", 210.0), ("poisson_module", 0.25), ("density", 7.8),
                    ("fracture_energy", 0.05), ("traction_limit_stress", 3.0),
                    ("compression_limit_stress", 30.0), ("energy_release_rate", 0.1),
                    ("damage_length_scale", 0.02)):
        assert getattr(L, "nb_material_get_" + name)(c) == v
    L.nb_material_destroy(c)
    L.nb_material_destroy(m)
    p = L.nb_fem_implicit_create()
    L.nb_fem_implicit_set_N_steps(p, 7)
    L.nb_fem_implicit_set_N_max_iter(p, 30)
    L.nb_fem_implicit_set_N_max_iter_without_enhance(p, 4)
    L.nb_fem_implicit_set_residual_tolerance(p, 1e-5)
    assert (L.nb_fem_implicit_get_N_steps(p), L.nb_fem_implicit_get_N_max_iter(p),
            L.nb_fem_implicit_get_N_max_iter_without_enhance(p),
            L.nb_fem_implicit_get_residual_tolerance(p)) == (7, 30, 4, 1e-5)
    L.nb_fem_implicit_destroy(p)
