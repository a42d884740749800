"""CPU oracle for the nbots FEM solve path -- TEST INFRASTRUCTURE ONLY.

Python bindings (ctypes) for

* ``oracle/librefcpu.so``: this repo's C99 restatement of the reference's
  algorithms (``oracle/refcpu.c``), and
* ``oracle/_ref/libnbref.so``: the real reference compiled from its own
  sources plus ``oracle/ref_harness.c`` (present only where it was built).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package; the product never does.
"""
from __future__ import annotations

import copy
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class _Opt:
    """ndpointer that also accepts None."""

    def __init__(self, base):
        self.base = base

    def from_param(self, obj):
        if obj is None:
            return None
        return self.base.from_param(obj)


u32o, u8o, f64o, i32o = _Opt(u32p), _Opt(u8p), _Opt(f64p), _Opt(i32p)
U32, F64, I32 = C.c_uint32, C.c_double, C.c_int


def build(ref: bool = True) -> None:
    """Compile the restatement, and the reference checker when the
    reference sources are present."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref and os.path.isdir("/root/reference/sources"):
        subprocess.run(["make", "-s", "-C", HERE, "ref", "-j8"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "librefcpu.so")
        if not os.path.exists(path):
            build(ref=False)
        L = C.CDLL(path)
        L.orc_trg_edges.restype = U32
        L.orc_trg_edges.argtypes = [U32, U32, u32p, u32p]
        L.orc_graph_from_edges.restype = None
        L.orc_graph_from_edges.argtypes = [U32, U32, u32p, u32p, u32p, u32p]
        L.orc_sparse_create.restype = U32
        L.orc_sparse_create.argtypes = [U32, u32p, u32p, u32p, u32o, U32, u32p, u32o]
        L.orc_plane_stress_D.argtypes = [F64, F64, f64p]
        L.orc_t3_derivatives.restype = I32
        L.orc_t3_derivatives.argtypes = [f64p, f64p, C.POINTER(F64), f64p, f64p]
        L.orc_t3_ke.argtypes = [f64p, F64, F64, f64p, f64p, f64p]
        L.orc_assemble_t3.restype = I32
        L.orc_assemble_t3.argtypes = [U32, f64p, U32, u32p, F64, F64, f64o, F64, u8o,
                                      I32, F64, f64o, u32p, u32p, f64p, f64p]
        L.orc_sparse_get.restype = F64
        L.orc_sparse_get.argtypes = [u32p, u32p, f64p, U32, U32]
        L.orc_set_dirichlet.argtypes = [U32, u32p, u32p, f64p, f64p, U32, F64]
        L.orc_spmv.argtypes = [U32, u32p, u32p, f64p, f64p, f64p]
        L.orc_pcg_jacobi.restype = I32
        L.orc_pcg_jacobi.argtypes = [U32, u32p, u32p, f64p, f64p, f64p, U32, F64,
                                     C.POINTER(U32), C.POINTER(F64)]
        L.orc_gauss_seidel.restype = I32
        L.orc_gauss_seidel.argtypes = [U32, u32p, u32p, f64p, f64p, f64p, U32, F64,
                                       C.POINTER(U32), C.POINTER(F64)]
        L.orc_assemble_t3_damage.restype = I32
        L.orc_assemble_t3_damage.argtypes = [U32, f64p, U32, u32p, F64, F64, f64p, F64, I32, F64,
                                             f64o, u32p, u32p, f64o, f64o]
        L.orc_residual.restype = F64
        L.orc_residual.argtypes = [U32, u32p, u32p, f64p, f64p, f64p, f64o]
        L.orc_damage_update.restype = I32
        L.orc_damage_update.argtypes = [U32, f64p, u32p, f64p, F64, F64, F64, F64, f64p, f64o,
                                        f64o, f64o]
        L.orc_cg.restype = I32
        L.orc_cg.argtypes = [U32, u32p, u32p, f64p, f64p, f64p, U32, F64,
                             C.POINTER(U32), C.POINTER(F64)]
        L.orc_set_bconditions.argtypes = (
            [f64p, u32p, u32p, u32p]
            + [U32, u32p, u8p, f64p] * 3
            + [U32, u32p, u8p, f64p, i32o, f64o, F64, U32, u32p, u32p, f64p, f64p])
        L.orc_strain.restype = I32
        L.orc_strain.argtypes = [U32, f64p, u32p, f64p, f64p]
        L.orc_stress.argtypes = [U32, F64, F64, f64o, u8o, f64p, f64p]
        L.orc_elem_energy.argtypes = [U32, f64p, u32p, F64, F64, f64p, f64p]
        L.orc_simp_compliance_sens.restype = F64
        L.orc_simp_compliance_sens.argtypes = [U32, f64p, f64p, F64, F64, F64, f64o]
        L.orc_filter_build.restype = U32
        L.orc_filter_build.argtypes = [U32, f64p, f64p, F64, u32p, u32o, f64o, f64o]
        L.orc_filter_apply.argtypes = [U32, u32p, u32p, f64p, f64p, f64p, f64p]
        L.orc_filter_sens.argtypes = [U32, u32p, u32p, f64p, f64p, f64p, f64p]
        L.orc_oc_update.restype = F64
        L.orc_oc_update.argtypes = [U32, u32p, u32p, f64p, f64p, f64p, f64p, f64p, f64p,
                                    F64, F64, f64p, f64p]
        _LIB = L
    return _LIB


def ref_available() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libnbref.so"))


def ref():
    """The compiled reference (raises if it was not built here)."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libnbref.so")
        if not os.path.exists(path):
            raise FileNotFoundError("oracle/_ref/libnbref.so is not built")
        R = C.CDLL(path)
        mesh = [U32, f64p, U32, u32p, U32, u32p, U32, u32p, u32p]
        bcs = [U32, u32p, u8p, f64p] * 4
        R.ref_assemble.restype = U32
        R.ref_assemble.argtypes = mesh + bcs + [F64, F64, f64o, F64, u8o, I32, F64, f64o,
                                                I32, u32p, u32o, f64o, f64o]
        R.ref_fem_compute.restype = I32
        R.ref_fem_compute.argtypes = mesh + bcs + [F64, F64, F64, u8o, f64p, f64p, f64p]
        R.ref_damage_compute.restype = None
        R.ref_damage_compute.argtypes = mesh + bcs + [F64, F64, F64, F64, F64, U32, U32, U32, F64,
                                                      C.c_char_p]
        R.ref_pcg.restype = I32
        R.ref_pcg.argtypes = [U32, u32p, u32p, f64p, f64p, f64p, U32, F64, U32,
                              C.POINTER(U32), C.POINTER(F64)]
        R.ref_sparse_save.argtypes = [U32, u32p, u32p, f64p, C.c_char_p]
        R.ref_sparse_save_mat4.argtypes = [U32, u32p, u32p, f64p, C.c_char_p, C.c_char_p]
        R.ref_mat4_save_vec.argtypes = [C.c_char_p, C.c_char_p, f64p, U32]
        R.ref_gauss_seidel.restype = I32
        R.ref_gauss_seidel.argtypes = [U32, u32p, u32p, f64p, f64p, f64p, U32, F64, U32,
                                       C.POINTER(U32), C.POINTER(F64)]
        R.ref_cg.restype = I32
        R.ref_cg.argtypes = [U32, u32p, u32p, f64p, f64p, f64p, U32, F64, U32,
                             C.POINTER(U32), C.POINTER(F64)]
        R.ref_sparse_import.restype = C.c_void_p
        R.ref_sparse_import.argtypes = [U32, u32p, u32p, f64p]
        R.ref_sparse_free.argtypes = [C.c_void_p]
        R.ref_pcg_handle.restype = I32
        R.ref_pcg_handle.argtypes = [C.c_void_p, f64p, f64p, U32, F64, U32,
                                     C.POINTER(U32), C.POINTER(F64)]
        R.ref_spmv.argtypes = [U32, u32p, u32p, f64p, f64p, f64p, U32]
        R.ref_testcase_run.restype = I32
        R.ref_testcase_run.argtypes = [C.c_char_p, U32, u32p]
        R.ref_testcase_run_ex.restype = I32
        R.ref_testcase_run_ex.argtypes = [C.c_char_p, U32, I32, u32p]
        R.ref_testcase_fetch.argtypes = [f64p, u32p, u32p, u32p, u32p, f64p, f64p, f64p, f64p]
        R.ref_testcase_free.argtypes = []
        _REF = R
    return _REF


# ---------------------------------------------------------------------------
# numpy-level wrappers around the restatement
# ---------------------------------------------------------------------------

def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def pattern(N_nod, conn, perm=None, vpn=2):
    """rowptr, col of the reference's 2-dof pattern of a triangle list."""
    L = lib()
    conn = _c(conn, np.uint32)
    ne = conn.size // 3
    edg = np.zeros(6 * ne + 2, np.uint32)
    n_edg = L.orc_trg_edges(N_nod, ne, conn, edg)
    N_adj = np.zeros(N_nod, np.uint32)
    adj_ptr = np.zeros(N_nod + 1, np.uint32)
    adj = np.zeros(2 * n_edg + 1, np.uint32)
    L.orc_graph_from_edges(N_nod, n_edg, edg, N_adj, adj_ptr, adj)
    rowptr = np.zeros(N_nod * vpn + 1, np.uint32)
    permv = None if perm is None else _c(perm, np.uint32)
    nnz = L.orc_sparse_create(N_nod, N_adj, adj_ptr, adj, permv, vpn, rowptr, None)
    col = np.zeros(nnz, np.uint32)
    L.orc_sparse_create(N_nod, N_adj, adj_ptr, adj, permv, vpn, rowptr, col)
    return rowptr, col


def assemble(p, rowptr, col, enabled=None, self_weight=False, density=0.0,
             gravity=(0.0, 0.0)):
    L = lib()
    val = np.zeros(col.size, np.float64)
    F = np.zeros(p.N, np.float64)
    en = None if enabled is None else _c(enabled, np.uint8)
    st = L.orc_assemble_t3(p.N_nod, _c(p.nod, np.float64), p.N_elem, _c(p.conn, np.uint32),
                           p.E, p.nu, None if p.elem_E is None else _c(p.elem_E, np.float64),
                           p.thickness, en, int(self_weight), density,
                           np.array(gravity, np.float64), rowptr, col, val, F)
    return val, F, st


def _bc_args(p):
    out = []
    for lst in (p.dv, p.nv, p.ds):
        ids, m, v = lst
        out += [ids.size, _c(ids, np.uint32), _c(m, np.uint8), _c(v, np.float64)]
    ids, m, v = p.ns
    out += [ids.size, _c(ids, np.uint32), _c(m, np.uint8), _c(v, np.float64)]
    return out


def set_bconditions(p, rowptr, col, val, F, factor=1.0, ns_fptr=None, ns_fval=None):
    L = lib()
    L.orc_set_bconditions(_c(p.nod, np.float64), _c(p.vtx, np.uint32),
                          _c(p.sgm_ptr, np.uint32), _c(p.sgm_nodes, np.uint32),
                          *_bc_args(p),
                          None if ns_fptr is None else _c(ns_fptr, np.int32),
                          None if ns_fval is None else _c(ns_fval, np.float64),
                          factor, p.N, rowptr, col, val, F)


def spmv(rowptr, col, val, x):
    y = np.zeros(rowptr.size - 1, np.float64)
    lib().orc_spmv(rowptr.size - 1, rowptr, col, val, _c(x, np.float64), y)
    return y


def pcg(rowptr, col, val, b, x0=None, max_iter=None, tol=1e-8):
    N = rowptr.size - 1
    x = np.zeros(N) if x0 is None else np.array(x0, np.float64)
    it, tr = U32(0), F64(0)
    st = lib().orc_pcg_jacobi(N, rowptr, col, val, _c(b, np.float64), x,
                              N if max_iter is None else max_iter, tol,
                              C.byref(it), C.byref(tr))
    return x, it.value, tr.value, st


def cg(rowptr, col, val, b, x0=None, max_iter=None, tol=1e-8):
    """Unpreconditioned CG restated from conjugate_gradient.c:13-77."""
    N = rowptr.size - 1
    x = np.zeros(N) if x0 is None else np.array(x0, np.float64)
    it, tr = U32(0), F64(0)
    st = lib().orc_cg(N, rowptr, col, val, _c(b, np.float64), x,
                      N if max_iter is None else max_iter, tol, C.byref(it), C.byref(tr))
    return x, it.value, tr.value, st


def gauss_seidel(rowptr, col, val, b, x0=None, max_iter=None, tol=1e-8):
    """Gauss-Seidel of gauss_seidel.c:13-63: returns x, sweeps, sqrt(error),
    status."""
    N = rowptr.size - 1
    x = np.zeros(N) if x0 is None else np.array(x0, np.float64)
    it, tr = U32(0), F64(0)
    st = lib().orc_gauss_seidel(N, rowptr, col, val, _c(b, np.float64), x,
                                60 * N if max_iter is None else max_iter, tol,
                                C.byref(it), C.byref(tr))
    return x, it.value, tr.value, st


def damage_solve(p, Gf, ft, n_steps, max_iter, max_iter_wo_enhance, res_tol, logfile=None):
    """nb_fem_compute_2D_Non_Linear_Solid_Mechanics (static_damage2D.c:239-472)
    on the restatement, with the conjugate-gradient path
    (enable_Cholesky_solver false), plane stress, no self weight: load
    factor (n+1)/n_steps on the boundary conditions, residual F - K u of the
    damaged K, Jacobi-PCG (tol 1e-8, max_iter N) for du from the previous du
    (Gauss-Seidel from there when the PCG does not converge), u += du,
    strain and damage update.  Writes the reference's log lines
    (to logfile if given) and returns u, damage, r and the log text."""
    L = lib()
    N = p.N
    rp, col = pattern(p.N_nod, p.conn)
    nod, conn = _c(p.nod, np.float64), _c(p.conn, np.uint32)
    u = np.zeros(N)
    du = np.zeros(N)
    dmg = np.zeros(p.N_elem)
    r = np.full(p.N_elem, ft / np.sqrt(p.E))
    strain_ = np.zeros(3 * p.N_elem)
    res = np.zeros(N)
    lines = ["FEM: Damage Model"]
    for n in range(n_steps):
        lines.append(f"  [ Load step {n + 1}]")
        r_prev = r.copy()
        du[:] = 0.0
        it = 0
        wo = 0
        best = 0.0
        while True:
            val = np.zeros(rp[-1])
            F = np.zeros(N)
            st = L.orc_assemble_t3_damage(p.N_nod, nod, p.N_elem, conn, p.E, p.nu, dmg,
                                          p.thickness, 0, 0.0, None, rp, col, val, F)
            if st:
                raise RuntimeError("distorted element")
            set_bconditions(p, rp, col, val, F, factor=(n + 1.0) / float(n_steps))
            nrm = L.orc_residual(N, rp, col, val, F, u, res)
            if it == 0:
                best = nrm
            elif nrm < best:
                best = nrm
                wo = 0
            else:
                wo += 1
            if (nrm < res_tol and it > 0) or it >= max_iter or wo >= max_iter_wo_enhance:
                lines.append("    [%i: Residual: %s ]" % (it, _fmt_e(nrm)))
                break
            du, _, _, st = pcg(rp, col, val, res, x0=du, max_iter=N, tol=1e-8)
            if st != 0:
                # static_damage2D.c:414-427: Gauss-Seidel from the CG's du,
                # 60 N sweeps, the same tolerance; logged only if it fails too
                du, gi, ge, gst = gauss_seidel(rp, col, val, res, x0=du, max_iter=60 * N,
                                               tol=1e-8)
                if gst != 0:
                    lines.append("    [Conjugate Gradient (p. Jacobi) and Gauss-Seidel fails.]")
                    lines.append("      [Gauss-Seidel iterations: %u]" % gi)
                    lines.append("      [Gauss-Seidel error: %s ]" % _fmt_e(ge))
            lines.append("    [%i: Residual: %s]" % (it, _fmt_e(nrm)))
            u += du
            L.orc_damage_update(p.N_elem, nod, conn, u, p.E, p.nu, Gf, ft, r_prev, r, dmg,
                                strain_)
            it += 1
    text = "\n".join(lines) + "\n"
    if logfile:
        with open(logfile, "a") as f:
            f.write(text)
    return dict(u=u, damage=dmg, r=r, strain=strain_, log=text)


def _fmt_e(v):
    """C's %e."""
    return "%e" % v


def strain(p, u):
    s = np.zeros(3 * p.N_elem)
    lib().orc_strain(p.N_elem, _c(p.nod, np.float64), _c(p.conn, np.uint32),
                     _c(u, np.float64), s)
    return s


def stress(p, strain_, enabled=None):
    s = np.zeros(3 * p.N_elem)
    lib().orc_stress(p.N_elem, p.E, p.nu,
                     None if p.elem_E is None else _c(p.elem_E, np.float64),
                     None if enabled is None else _c(enabled, np.uint8),
                     _c(strain_, np.float64), s)
    return s


def fn_neumann(p):
    """p with its function-valued Neumann segments (p.ns_fn) appended to the
    constant ones, and the (ns_fptr, ns_fval) sample arrays of
    orc_set_bconditions: each function evaluated at its segment's nodes in
    segment order, at time 0 (set_bconditions.c:87-131)."""
    ns_fn = getattr(p, "ns_fn", ())
    if not ns_fn:
        return p, None, None
    ids, m, v = p.ns
    fptr = [-1] * ids.size
    fval = []
    for sid, mask, fn in ns_fn:
        fptr.append(len(fval) // 2)
        for n in p.sgm_nodes[p.sgm_ptr[sid]:p.sgm_ptr[sid + 1]]:
            fval += list(fn(float(p.nod[2 * n]), float(p.nod[2 * n + 1])))
    q = copy.copy(p)
    q.ns = (np.concatenate([ids, np.array([s for s, _, _ in ns_fn], np.uint32)]),
            np.concatenate([m, np.array([b for _, mk, _ in ns_fn for b in mk], np.uint8)]),
            np.concatenate([v, np.zeros(2 * len(ns_fn))]))
    q.ns_fn = ()
    return q, np.array(fptr, np.int32), np.array(fval, np.float64)


def fem_solve(p, tol=1e-8, max_iter=None, enabled=None):
    """The reference driver flow (static_elasticity2D.c:31-97) on the
    restatement: pattern, assembly, BCs, x=0, PCG(max_iter=N, tol), strain."""
    rowptr, col = pattern(p.N_nod, p.conn)
    val, F, st = assemble(p, rowptr, col, enabled=enabled)
    if st:
        return None
    q, fptr, fval = fn_neumann(p)
    set_bconditions(q, rowptr, col, val, F, ns_fptr=fptr, ns_fval=fval)
    u, it, tr, cst = pcg(rowptr, col, val, F, max_iter=max_iter, tol=tol)
    eps = strain(p, u)
    return dict(rowptr=rowptr, col=col, val=val, F=F, u=u, niter=it,
                tol_reached=tr, status=cst, strain=eps, stress=stress(p, eps, enabled))


def elem_energy(p, u):
    ce = np.zeros(p.N_elem)
    lib().orc_elem_energy(p.N_elem, _c(p.nod, np.float64), _c(p.conn, np.uint32),
                          p.nu, p.thickness, _c(u, np.float64), ce)
    return ce


def simp_sens(rho, ce, penal=3.0, E0=1.0, Emin=1e-9):
    dc = np.zeros(rho.size)
    C_ = lib().orc_simp_compliance_sens(rho.size, _c(rho, np.float64), _c(ce, np.float64),
                                        penal, E0, Emin, dc)
    return C_, dc


def filter_build(cx, cy, rmin):
    L = lib()
    n = cx.size
    Hptr = np.zeros(n + 1, np.uint32)
    cnt = L.orc_filter_build(n, _c(cx, np.float64), _c(cy, np.float64), rmin, Hptr,
                             None, None, None)
    Hidx = np.zeros(cnt, np.uint32)
    Hval = np.zeros(cnt)
    Hs = np.zeros(n)
    L.orc_filter_build(n, _c(cx, np.float64), _c(cy, np.float64), rmin, Hptr, Hidx, Hval, Hs)
    return Hptr, Hidx, Hval, Hs


def filter_apply(H, x):
    Hptr, Hidx, Hval, Hs = H
    out = np.zeros(Hs.size)
    lib().orc_filter_apply(Hs.size, Hptr, Hidx, Hval, Hs, _c(x, np.float64), out)
    return out


def filter_sens(H, dc):
    Hptr, Hidx, Hval, Hs = H
    out = np.zeros(Hs.size)
    lib().orc_filter_sens(Hs.size, Hptr, Hidx, Hval, Hs, _c(dc, np.float64), out)
    return out


def oc_update(H, area, x, dc, dv, volfrac, move=0.2):
    Hptr, Hidx, Hval, Hs = H
    xnew = np.zeros(Hs.size)
    xphys = np.zeros(Hs.size)
    ch = lib().orc_oc_update(Hs.size, Hptr, Hidx, Hval, Hs, _c(area, np.float64),
                             _c(x, np.float64), _c(dc, np.float64), _c(dv, np.float64),
                             volfrac, move, xnew, xphys)
    return xnew, xphys, ch


# ---------------------------------------------------------------------------
# wrappers around the compiled reference
# ---------------------------------------------------------------------------

def _mesh_args(p):
    return [p.N_nod, _c(p.nod, np.float64), p.N_elem, _c(p.conn, np.uint32),
            p.vtx.size, _c(p.vtx, np.uint32), p.sgm_ptr.size - 1,
            _c(p.sgm_ptr, np.uint32), _c(p.sgm_nodes, np.uint32)]


def ref_assemble(p, apply_bc=True, enabled=None, self_weight=False, density=0.0,
                 gravity=(0.0, 0.0)):
    R = ref()
    rowptr = np.zeros(p.N + 1, np.uint32)
    args = _mesh_args(p) + _bc_args(p)
    tail = [p.E, p.nu, None if p.elem_E is None else _c(p.elem_E, np.float64), p.thickness,
            None if enabled is None else _c(enabled, np.uint8), int(self_weight), density,
            np.array(gravity, np.float64), int(apply_bc)]
    nnz = R.ref_assemble(*args, *tail, rowptr, None, None, None)
    col = np.zeros(nnz, np.uint32)
    val = np.zeros(nnz)
    F = np.zeros(p.N)
    st = R.ref_assemble(*args, *tail, rowptr, col, val, F)
    return rowptr, col, val, F, (1 if st == 0xFFFFFFFF else 0)


def ref_assemble_timed(p):
    """ref_assemble with BCs, timed: (rowptr, col, val, F, status,
    seconds of the pattern call, seconds of the pattern + assembly + BCs
    call)."""
    import time
    R = ref()
    rowptr = np.zeros(p.N + 1, np.uint32)
    args = _mesh_args(p) + _bc_args(p)
    tail = [p.E, p.nu, None if p.elem_E is None else _c(p.elem_E, np.float64), p.thickness,
            None, 0, 0.0, np.zeros(2), 1]
    t0 = time.perf_counter()
    nnz = R.ref_assemble(*args, *tail, rowptr, None, None, None)
    t_pat = time.perf_counter() - t0
    col = np.zeros(nnz, np.uint32)
    val = np.zeros(nnz)
    F = np.zeros(p.N)
    t0 = time.perf_counter()
    st = R.ref_assemble(*args, *tail, rowptr, col, val, F)
    t_all = time.perf_counter() - t0
    return rowptr, col, val, F, (1 if st == 0xFFFFFFFF else 0), t_pat, t_all


def ref_fem_compute(p, enabled=None):
    disp = np.zeros(p.N)
    eps = np.zeros(3 * p.N_elem)
    sig = np.zeros(3 * p.N_elem)
    st = ref().ref_fem_compute(*_mesh_args(p), *_bc_args(p), p.E, p.nu, p.thickness,
                               None if enabled is None else _c(enabled, np.uint8),
                               disp, eps, sig)
    return st, disp, eps, sig


def ref_damage_compute(p, Gf, ft, n_steps, max_iter, max_iter_wo_enhance, res_tol, logfile):
    """The compiled reference's nb_fem_compute_2D_Non_Linear_Solid_Mechanics;
    appends its log to logfile."""
    ref().ref_damage_compute(*_mesh_args(p), *_bc_args(p), p.E, p.nu, p.thickness, Gf, ft,
                             n_steps, max_iter, max_iter_wo_enhance, res_tol, logfile.encode())


def ref_pcg(rowptr, col, val, b, x0=None, max_iter=None, tol=1e-8, threads=1):
    N = rowptr.size - 1
    x = np.zeros(N) if x0 is None else np.array(x0, np.float64)
    it, tr = U32(0), F64(0)
    st = ref().ref_pcg(N, rowptr, col, val, _c(b, np.float64), x,
                       N if max_iter is None else max_iter, tol, threads,
                       C.byref(it), C.byref(tr))
    return x, it.value, tr.value, st


def ref_gauss_seidel(rowptr, col, val, b, x0=None, max_iter=None, tol=1e-8, threads=1):
    """The compiled reference's nb_sparse_solve_Gauss_Seidel."""
    N = rowptr.size - 1
    x = np.zeros(N) if x0 is None else np.array(x0, np.float64)
    it, tr = U32(0), F64(0)
    st = ref().ref_gauss_seidel(N, rowptr, col, val, _c(b, np.float64), x,
                                60 * N if max_iter is None else max_iter, tol, threads,
                                C.byref(it), C.byref(tr))
    return x, it.value, tr.value, st


def ref_cg(rowptr, col, val, b, x0=None, max_iter=None, tol=1e-8, threads=1):
    """The compiled reference's nb_sparse_solve_conjugate_gradient."""
    N = rowptr.size - 1
    x = np.zeros(N) if x0 is None else np.array(x0, np.float64)
    it, tr = U32(0), F64(0)
    st = ref().ref_cg(N, rowptr, col, val, _c(b, np.float64), x,
                      N if max_iter is None else max_iter, tol, threads,
                      C.byref(it), C.byref(tr))
    return x, it.value, tr.value, st


def ref_sparse_save(rowptr, col, val, path):
    ref().ref_sparse_save(rowptr.size - 1, rowptr, col, val, path.encode())


def ref_sparse_save_mat4(rowptr, col, val, path, label):
    ref().ref_sparse_save_mat4(rowptr.size - 1, rowptr, col, val, path.encode(), label.encode())


def ref_mat4_save_vec(path, label, x):
    x = _c(x, np.float64)
    ref().ref_mat4_save_vec(path.encode(), label.encode(), x, x.size)


def ref_spmv(rowptr, col, val, x, threads=1):
    y = np.zeros(rowptr.size - 1)
    ref().ref_spmv(rowptr.size - 1, rowptr, col, val, _c(x, np.float64), y, threads)
    return y
