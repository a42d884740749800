"""nbots_amd -- the nbots FEM solve path (T3 plane-stress assembly,
Jacobi-PCG on nb_sparse_t, SIMP sensitivities) for one MI355X (gfx950).

The product is the C-ABI library ``libnbots_amd.so`` (``include/nbots_amd.h``);
this package holds its sources (``csrc/``), its build (``build.py``), thin
ctypes wrappers (``lib.py``, this module, ``fem.py``) and the synthetic
workload generator (``synth.py``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as _lib
from .lib import SolveInfo, check, last_error, load

SOLVE_REFERENCE, SOLVE_FP64, SOLVE_MIXED = 0, 1, 2

__all__ = ["Sparse", "SolveInfo", "load", "last_error", "check",
           "SOLVE_REFERENCE", "SOLVE_FP64", "SOLVE_MIXED"]


def _graph_struct(N_adj, adj_lists):
    """An nb_graph_t built in Python memory (kept alive by the returned tuple)."""
    N = len(N_adj)
    rows = [np.ascontiguousarray(a, np.uint32) for a in adj_lists]
    nadj = np.ascontiguousarray(N_adj, np.uint32)
    ptrs = (C.POINTER(C.c_uint32) * max(N, 1))()
    for i, r in enumerate(rows):
        ptrs[i] = r.ctypes.data_as(C.POINTER(C.c_uint32)) if r.size else C.POINTER(C.c_uint32)()
    g = _lib.Graph(N, nadj.ctypes.data_as(C.POINTER(C.c_uint32)), ptrs, None, None)
    return g, (rows, nadj, ptrs)


class Sparse:
    """Drop-in nb_sparse_t (headers/nb/solver_bot/sparse/sparse.h)."""

    def __init__(self, handle):
        self.L = load()
        self.h = handle

    @classmethod
    def from_graph(cls, N_adj, adj_lists, vars_per_node=2, perm=None):
        L = load()
        g, keep = _graph_struct(N_adj, adj_lists)
        h = L.nb_sparse_create(C.byref(g), None if perm is None else
                               np.ascontiguousarray(perm, np.uint32), vars_per_node)
        del keep
        return cls(h)

    @classmethod
    def from_csr(cls, rowptr, col, val, vars_per_node=1):
        """Pattern via a graph whose adjacency is the CSR's off-diagonal
        node columns (rows must carry their diagonal), then bulk values."""
        rowptr = np.asarray(rowptr, np.uint32)
        col = np.asarray(col, np.uint32)
        vpn = vars_per_node
        n_nod = (rowptr.size - 1) // vpn
        adj = []
        for i in range(n_nod):
            cols = col[rowptr[i * vpn]:rowptr[i * vpn + 1]] // vpn
            nb = np.unique(cols)
            adj.append(nb[nb != i])
        A = cls.from_graph([a.size for a in adj], adj, vpn)
        rp2, c2, _ = A.export()
        if not (np.array_equal(rp2, rowptr) and np.array_equal(c2, col)):
            A.destroy()
            raise ValueError("CSR pattern is not a graph pattern with diagonal")
        A.import_values(val)
        return A

    def destroy(self):
        if self.h:
            self.L.nb_sparse_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    @property
    def N(self):
        return self.L.nb_sparse_get_size(self.h)

    @property
    def nnz(self):
        return self.L.nb_sparse_get_nnz(self.h)

    def export(self):
        rp = np.zeros(self.N + 1, np.uint32)
        col = np.zeros(self.nnz, np.uint32)
        val = np.zeros(self.nnz)
        self.L.nbx_sparse_export_csr(self.h, rp, col, val)
        return rp, col, val

    def import_values(self, val):
        self.L.nbx_sparse_import_values(self.h, np.ascontiguousarray(val, np.float64))

    def add(self, i, j, v):
        self.L.nb_sparse_add(self.h, i, j, v)

    def get(self, i, j):
        return self.L.nb_sparse_get(self.h, i, j)

    def set_dirichlet(self, rhs, idx, value):
        self.L.nb_sparse_set_Dirichlet_condition(self.h, rhs, idx, value)

    def matvec(self, x):
        y = np.zeros(self.N)
        self.L.nb_sparse_multiply_vector(self.h, np.ascontiguousarray(x, np.float64), y, 1)
        return y

    def is_blocked(self):
        return self.L.nbx_sparse_is_blocked(self.h)

    def solve_cg_jacobi(self, b, x0=None, max_iter=None, tol=1e-8):
        """nb_sparse_solve_CG_precond_Jacobi: returns x, status, niter, tol_reached."""
        x = np.zeros(self.N) if x0 is None else np.array(x0, np.float64)
        it, tr = C.c_uint32(0), C.c_double(0)
        st = self.L.nb_sparse_solve_CG_precond_Jacobi(
            self.h, np.ascontiguousarray(b, np.float64), x,
            self.N if max_iter is None else max_iter, tol, C.byref(it), C.byref(tr), 1)
        if st < 0:
            raise RuntimeError(f"GPU solve failed: {last_error()}")
        return x, st, it.value, tr.value

    def solve_cg(self, b, x0=None, max_iter=None, tol=1e-8):
        """nb_sparse_solve_conjugate_gradient: returns x, status, niter, tol_reached."""
        x = np.zeros(self.N) if x0 is None else np.array(x0, np.float64)
        it, tr = C.c_uint32(0), C.c_double(0)
        st = self.L.nb_sparse_solve_conjugate_gradient(
            self.h, np.ascontiguousarray(b, np.float64), x,
            self.N if max_iter is None else max_iter, tol, C.byref(it), C.byref(tr), 1)
        if st < 0:
            raise RuntimeError(f"GPU solve failed: {last_error()}")
        return x, st, it.value, tr.value

    def solve_gauss_seidel(self, b, x0=None, max_iter=None, tol=1e-8):
        """nb_sparse_solve_Gauss_Seidel: returns x, status, sweeps, error."""
        x = np.zeros(self.N) if x0 is None else np.array(x0, np.float64)
        it, tr = C.c_uint32(0), C.c_double(0)
        st = self.L.nb_sparse_solve_Gauss_Seidel(
            self.h, np.ascontiguousarray(b, np.float64), x,
            60 * self.N if max_iter is None else max_iter, tol, C.byref(it), C.byref(tr), 1)
        if st < 0:
            raise RuntimeError(f"GPU solve failed: {last_error()}")
        return x, st, it.value, tr.value

    def solve(self, b, mode=SOLVE_FP64, tol=1e-8, max_iter=None, x0=None, inner_rtol=0.0,
              opts=None):
        """nbx_sparse_solve, or nbx_sparse_solve_ex when opts (a SolveOpts,
        nbots_amd.fem.solve_opts(...)) is given."""
        x = np.zeros(self.N) if x0 is None else np.array(x0, np.float64)
        info = SolveInfo()
        mi = 10 * self.N if max_iter is None else max_iter
        bb = np.ascontiguousarray(b, np.float64)
        if opts is not None:
            rc = self.L.nbx_sparse_solve_ex(self.h, bb, x, mode, tol, mi, C.byref(opts),
                                            C.byref(info))
        else:
            rc = self.L.nbx_sparse_solve(self.h, bb, x, mode, tol, mi, inner_rtol, C.byref(info))
        check(rc, "nbx_sparse_solve", ok=(0, 1))
        return x, info.as_dict()
