"""Python view of the device-resident FEM system (nbx_fem_* in
include/nbots_amd.h) and of the drop-in driver nb_fem_compute_2D_Solid_Mechanics."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .lib import SolveInfo, SolveOpts, check, last_error, load

SOLVE_REFERENCE, SOLVE_FP64, SOLVE_MIXED = 0, 1, 2
MIXED_CG_RESTART, MIXED_CG_RELIABLE, MIXED_CHEB_FCG, MIXED_AMG_FCG = 0, 1, 2, 3


def solve_opts(variant=MIXED_CHEB_FCG, cheb_degree=None, cheb_ratio=None, inner_rtol=None,
               reliable_delta=None, precond_bits=None, simp=False, true_residual=None,
               replay=None):
    """nbx_solve_opts_t with the library's defaults (the SIMP loop's with
    simp=True), then the given fields."""
    o = SolveOpts()
    if simp:
        load().nbx_simp_opts_default(C.byref(o))
    else:
        load().nbx_solve_opts_default(C.byref(o))
    o.mixed_variant = variant
    if cheb_degree is not None:
        o.cheb_degree = cheb_degree
    if cheb_ratio is not None:
        o.cheb_ratio = cheb_ratio
    if inner_rtol is not None:
        o.inner_rtol = inner_rtol
    if reliable_delta is not None:
        o.reliable_delta = reliable_delta
    if precond_bits is not None:
        o.precond_bits = precond_bits
    if true_residual is not None:
        o.true_residual = int(bool(true_residual))
    if replay is not None:
        o.replay = int(bool(replay))
    return o


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class Fem:
    """A T3 plane-stress system held on the GPU."""

    def __init__(self, nod, conn):
        self.L = load()
        nod = _c(nod, np.float64)
        conn = _c(conn, np.uint32)
        h = C.c_void_p()
        check(self.L.nbx_fem_create(C.byref(h), nod.size // 2, nod, conn.size // 3, conn),
              "nbx_fem_create")
        self.h = h.value
        self.N_nod = nod.size // 2
        self.N_elem = conn.size // 3

    @classmethod
    def from_problem(cls, p, assemble=True, bc=True):
        f = cls(p.nod, p.conn)
        if assemble:
            f.assemble(p.E, p.nu, p.thickness, p.elem_E)
        if bc:
            f.set_bconditions(p)
        return f

    def destroy(self):
        if getattr(self, "h", None):
            self.L.nbx_fem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    @property
    def N(self):
        return 2 * self.N_nod

    def sizes(self):
        n, nnz, blk = C.c_uint32(), C.c_uint64(), C.c_uint64()
        self.L.nbx_fem_sizes(self.h, C.byref(n), C.byref(nnz), C.byref(blk))
        return n.value, nnz.value, blk.value

    def assemble(self, E=1.0, nu=0.3, thickness=1.0, elem_E=None, enabled=None,
                 self_weight=False, density=0.0, gravity=None):
        rc = self.L.nbx_fem_assemble(self.h, E, nu, thickness,
                                     None if elem_E is None else _c(elem_E, np.float64),
                                     None if enabled is None else _c(enabled, np.uint8),
                                     int(self_weight), density,
                                     None if gravity is None else _c(gravity, np.float64))
        return check(rc, "nbx_fem_assemble", ok=(0, 2))

    def set_bconditions(self, p, factor=1.0, ns_fptr=None, ns_fval=None):
        args = []
        for k in ("dv", "nv", "ds", "ns"):
            ids, m, v = getattr(p, k)
            args += [ids.size, _c(ids, np.uint32), _c(m, np.uint8), _c(v, np.float64)]
        rc = self.L.nbx_fem_set_bconditions(
            self.h, _c(p.vtx, np.uint32), _c(p.sgm_ptr, np.uint32), _c(p.sgm_nodes, np.uint32),
            *args, None if ns_fptr is None else _c(ns_fptr, np.int32),
            None if ns_fval is None else _c(ns_fval, np.float64), factor)
        check(rc, "nbx_fem_set_bconditions")

    def solve(self, mode=SOLVE_MIXED, tol=1e-8, max_iter=None, inner_rtol=0.0, u0=None,
              opts=None):
        """opts (solve_opts(...)) overrides inner_rtol's variant selection."""
        info = SolveInfo()
        mi = 20 * self.N if max_iter is None else max_iter
        u0p = None if u0 is None else _c(u0, np.float64)
        if opts is not None:
            rc = self.L.nbx_fem_solve_ex(self.h, mode, tol, mi, C.byref(opts), u0p, C.byref(info))
        else:
            rc = self.L.nbx_fem_solve(self.h, mode, tol, mi, inner_rtol, u0p, C.byref(info))
        check(rc, "nbx_fem_solve", ok=(0, 1))
        return info.as_dict()

    def displacement(self):
        u = np.zeros(self.N)
        check(self.L.nbx_fem_get_displacement(self.h, u), "nbx_fem_get_displacement")
        return u

    def true_residual(self, compensated=True):
        """||F - K u|| / ||F|| of the current displacement, on the device:
        each row in compensated (Dot2) arithmetic, or plain FP64."""
        r = C.c_double()
        check(self.L.nbx_fem_true_residual(self.h, int(compensated), C.byref(r)),
              "nbx_fem_true_residual")
        return r.value

    def system(self):
        """K (scalar CSR, the reference's layout) and F."""
        _, nnz, _ = self.sizes()
        rp = np.zeros(self.N + 1, np.uint32)
        col = np.zeros(nnz, np.uint32)
        val = np.zeros(nnz)
        F = np.zeros(self.N)
        check(self.L.nbx_fem_get_system(self.h, rp, col, val, F), "nbx_fem_get_system")
        return rp, col, val, F

    def load_vector(self):
        """F (after the boundary conditions) without downloading K."""
        F = np.zeros(self.N)
        check(self.L.nbx_fem_get_system(self.h, None, None, None, F), "nbx_fem_get_system")
        return F

    def strain_stress(self):
        e = np.zeros(3 * self.N_elem)
        s = np.zeros(3 * self.N_elem)
        check(self.L.nbx_fem_strain_stress(self.h, e, s), "nbx_fem_strain_stress")
        return e, s

    def compliance(self, rho=None, penal=3.0, E0=1.0, Emin=1e-9):
        Cv = C.c_double()
        dc = np.zeros(self.N_elem)
        ce = np.zeros(self.N_elem)
        check(self.L.nbx_fem_compliance(self.h, None if rho is None else _c(rho, np.float64),
                                        penal, E0, Emin, C.byref(Cv), dc, ce),
              "nbx_fem_compliance")
        return Cv.value, dc, ce

    def adjoint(self, rhs, mode=SOLVE_FP64, rtol=1e-10, max_iter=None):
        lam = np.zeros(self.N)
        info = SolveInfo()
        rc = self.L.nbx_fem_adjoint(self.h, _c(rhs, np.float64), lam, mode, rtol,
                                    20 * self.N if max_iter is None else max_iter, C.byref(info))
        check(rc, "nbx_fem_adjoint", ok=(0, 1))
        return lam, info.as_dict()

    def elem_cross(self, a=None, b=None):
        """a_e' K0_e b_e per element (None = the current displacement)."""
        out = np.zeros(self.N_elem)
        check(self.L.nbx_fem_elem_cross(self.h, None if a is None else _c(a, np.float64),
                                        None if b is None else _c(b, np.float64), out),
              "nbx_fem_elem_cross")
        return out

    def spmv_ms(self, precision=64, reps=20):
        ms = C.c_double()
        check(self.L.nbx_fem_spmv_bench(self.h, precision, reps, C.byref(ms)), "nbx_fem_spmv_bench")
        return ms.value


class Simp:
    """SIMP topology optimisation on a device system (nbx_simp_*): density
    filter of radius rmin, E = Emin + xPhys^penal (E0 - Emin), OC update."""

    def __init__(self, fem, rmin=1.5, volfrac=0.5, penal=3.0, E0=1.0, Emin=1e-9, move=0.2,
                 x0=None):
        self.L = fem.L
        self.fem = fem
        self.N_elem = fem.N_elem
        h = C.c_void_p()
        check(self.L.nbx_simp_create(C.byref(h), fem.h, rmin, volfrac, penal, E0, Emin, move,
                                     None if x0 is None else _c(x0, np.float64)),
              "nbx_simp_create")
        self.h = h.value

    def destroy(self):
        if self.h:
            self.L.nbx_simp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def iterate(self, mode=SOLVE_MIXED, tol=1e-8, max_iter=0, opts=None):
        """One SIMP iteration; returns (compliance, change, solve info)."""
        Cv, ch, info = C.c_double(), C.c_double(), SolveInfo()
        rc = self.L.nbx_simp_iterate(self.h, mode, tol, max_iter,
                                     None if opts is None else C.addressof(opts),
                                     C.byref(Cv), C.byref(ch), C.byref(info))
        check(rc, "nbx_simp_iterate", ok=(0, 1))
        return Cv.value, ch.value, info.as_dict()

    def state(self):
        """(x, xPhys, filtered dc of the last iteration, OC bisection steps)"""
        x, xp, dc = (np.zeros(self.N_elem) for _ in range(3))
        steps = C.c_uint32()
        check(self.L.nbx_simp_get(self.h, x, xp, dc, C.byref(steps)), "nbx_simp_get")
        return x, xp, dc, steps.value

    def filter(self):
        """H as CSR (Hptr, Hidx, Hval) and Hs."""
        n = C.c_uint64()
        check(self.L.nbx_simp_filter(self.h, C.byref(n), None, None, None, None), "nbx_simp_filter")
        ptr = np.zeros(self.N_elem + 1, np.uint32)
        idx = np.zeros(n.value, np.uint32)
        val = np.zeros(n.value)
        hs = np.zeros(self.N_elem)
        check(self.L.nbx_simp_filter(self.h, None, ptr, idx, val, hs), "nbx_simp_filter")
        return ptr, idx, val, hs


class _Params2D(C.Structure):
    """nb_analysis2D_params (analysis2D.h:10-14)."""
    _fields_ = [("thickness", C.c_double), ("revolution_axe", C.c_double),
                ("revolution_coordinate", C.c_char)]


_BC_FN = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_double))


def _bc_eval(fn, x, out):
    out[0], out[1] = fn(x[0], x[1])


class _Objects:
    """The reference's object API for a Problem: mesh, element type,
    material and boundary conditions (freed on exit)."""

    def __init__(self, p, density=0.0, Gf=0.0, ft=0.0):
        L = self.L = load()
        self.mesh = L.nb_mesh2D_create(0)
        check(L.nbx_mesh2D_set_trg(self.mesh, p.N_nod, _c(p.nod, np.float64), p.N_elem,
                                   _c(p.conn, np.uint32), p.vtx.size, _c(p.vtx, np.uint32),
                                   p.sgm_ptr.size - 1, _c(p.sgm_ptr, np.uint32),
                                   _c(p.sgm_nodes, np.uint32)), "nbx_mesh2D_set_trg")
        self.elem = L.nb_fem_elem_create(0)
        self.mat = L.nb_material_create()
        L.nb_material_set_elasticity_module(self.mat, p.E)
        L.nb_material_set_poisson_module(self.mat, p.nu)
        L.nb_material_set_density(self.mat, density)
        L.nb_material_set_fracture_energy(self.mat, Gf)
        L.nb_material_set_traction_limit_stress(self.mat, ft)
        self.bc = L.nb_bcond_create(2)
        for k, (t, w) in (("dv", (0, 0)), ("nv", (1, 0)), ("ds", (0, 1)), ("ns", (1, 1))):
            ids, m, v = getattr(p, k)
            for c in range(ids.size):
                mask = (C.c_bool * 2)(bool(m[2 * c]), bool(m[2 * c + 1]))
                val = (C.c_double * 2)(v[2 * c], v[2 * c + 1])
                L.nb_bcond_push(self.bc, t, w, int(ids[c]), mask, val)
        self._fns = []              # the callbacks live as long as the bcond
        for sid, m, fn in p.ns_fn:
            cb = _BC_FN(lambda x, t, out, fn=fn: _bc_eval(fn, x, out))
            self._fns.append(cb)
            mask = (C.c_bool * 2)(bool(m[0]), bool(m[1]))
            L.nb_bcond_push_function(self.bc, 1, 1, int(sid), mask, C.cast(cb, C.c_void_p))
        self.params = _Params2D(p.thickness, 0.0, b"\0")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        L = self.L
        L.nb_bcond_destroy(self.bc)
        L.nb_material_destroy(self.mat)
        L.nb_fem_elem_destroy(self.elem)
        L.nb_mesh2D_destroy(self.mesh)


def compute_2D_solid_mechanics(p, enabled=None, self_weight=False, gravity=(0.0, 0.0),
                               density=0.0):
    """Run the drop-in nb_fem_compute_2D_Solid_Mechanics on a Problem through
    the reference's object API (mesh, element, material, bcond)."""
    with _Objects(p, density) as o:
        L = o.L
        en = None
        if enabled is not None:
            en = (C.c_bool * p.N_elem)(*[bool(x) for x in enabled])
        disp = np.zeros(p.N)
        strain = np.zeros(3 * p.N_elem)
        en_p = None if en is None else C.cast(en, C.c_void_p)
        st = L.nb_fem_compute_2D_Solid_Mechanics(o.mesh, o.elem, o.mat, o.bc, bool(self_weight),
                                                 _c(gravity, np.float64), 0,
                                                 C.cast(C.pointer(o.params), C.c_void_p),
                                                 en_p, disp, strain)
        stress = np.zeros(3 * p.N_elem)
        if st == 0:
            L.nb_fem_compute_stress_from_strain(p.N_elem, o.elem, o.mat, 0, strain, en_p, stress)
    if st < 0:
        raise RuntimeError(f"nb_fem_compute_2D_Solid_Mechanics: {last_error()}")
    return st, disp, strain, stress


def compute_2D_damage(p, Gf, ft, n_steps, max_iter, max_iter_wo_enhance, res_tol, logfile=None,
                      drop_in=False):
    """The quasi-static damage solver on a Problem through the object API:
    drop_in=True calls nb_fem_compute_2D_Non_Linear_Solid_Mechanics (its
    only output is the log); otherwise nbx_fem_damage_solve, which also
    returns the final displacement and damage."""
    with _Objects(p, 0.0, Gf, ft) as o:
        L = o.L
        imp = L.nb_fem_implicit_create()
        L.nb_fem_implicit_set_N_steps(imp, n_steps)
        L.nb_fem_implicit_set_N_max_iter(imp, max_iter)
        L.nb_fem_implicit_set_N_max_iter_without_enhance(imp, max_iter_wo_enhance)
        L.nb_fem_implicit_set_residual_tolerance(imp, res_tol)
        g = np.zeros(2)
        lf = None if logfile is None else logfile.encode()
        u = np.zeros(p.N)
        d = np.zeros(p.N_elem)
        if drop_in:
            L.nb_fem_compute_2D_Non_Linear_Solid_Mechanics(
                o.mesh, o.elem, o.mat, o.bc, False, g, False, 0,
                C.cast(C.pointer(o.params), C.c_void_p), imp, lf)
            rc = 0
        else:
            rc = L.nbx_fem_damage_solve(o.mesh, o.elem, o.mat, o.bc, 0, g,
                                        C.cast(C.pointer(o.params), C.c_void_p), imp, lf, u, d)
        L.nb_fem_implicit_destroy(imp)
    check(rc, "nbx_fem_damage_solve")
    return u, d
