"""nbx_sparse_solve_CG_precond_Jacobi_noalloc: the drop-in Jacobi-PCG on a
CSR matrix and vectors in device memory, with its work in the library's
fixed device array, against the drop-in nb_sparse_solve_CG_precond_Jacobi
on the same systems: calls in a row, the later ones smaller (the array
keeps the earlier calls' contents), to the bar of the other PCG parity
tests (the two sum their dot products in different orders): converged,
the same status, iterations within 2, x within rel-L2 1e-10 and
tol_reached under the tolerance; cut at 7 iterations, the same status and
count, tol_reached within 1e-9 relative and x within 1e-11.
A size beyond the maximum is refused with NBX_ERR_ARG."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(case):
    from nbots_amd import synth
    import oracle
    p = (synth.cantilever_point(64, 32, jitter=0.1) if case == "c1_64x32"
         else synth.cantilever_traction(16, 8))
    rp, col = oracle.pattern(p.N_nod, p.conn)
    val, F, _ = oracle.assemble(p, rp, col)
    oracle.set_bconditions(p, rp, col, val, F)
    return rp, col, val, F


def _hip():
    """The HIP runtime the library itself links, for the test's device
    buffers: symbols looked up through the library's own handle resolve in
    its dependencies (another HIP runtime in the process, such as one a
    Python package bundles, could not open the GPU)."""
    from nbots_amd import lib, load
    load()
    H = C.CDLL(lib.LIB_PATH)
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    H.hipFree.argtypes = [C.c_void_p]
    return H


class _Dev:
    """A device copy of a numpy array (hipMalloc + hipMemcpy)."""
    H2D, D2H = 1, 2

    def __init__(self, H, a):
        self.H, self.a = H, np.ascontiguousarray(a)
        self.p = C.c_void_p()
        assert H.hipMalloc(C.byref(self.p), max(self.a.nbytes, 8)) == 0
        assert H.hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, self.H2D) == 0

    def get(self):
        out = np.empty_like(self.a)
        assert self.H.hipMemcpy(out.ctypes.data, self.p, out.nbytes, self.D2H) == 0
        return out

    def free(self):
        self.H.hipFree(self.p)


def _noalloc(L, rp, col, val, b, tol, max_iter):
    H = _hip()
    bufs = [_Dev(H, a) for a in (rp.astype(np.uint32), col.astype(np.uint32), val, b,
                                 np.zeros(b.size))]
    it, tr = C.c_uint32(0), C.c_double(0.0)
    st = L.nbx_sparse_solve_CG_precond_Jacobi_noalloc(
        b.size, *(d.p.value for d in bufs), max_iter, tol, C.byref(it), C.byref(tr), None)
    x = bufs[4].get()
    for d in bufs:
        d.free()
    return st, x, it.value, tr.value


def test_noalloc_matches_drop_in_twice():
    """Calls in a row on the 64x32 cantilever (4,290 rows) and then the
    smaller 16x8 one (306 rows), each solved to convergence and cut at 7
    iterations."""
    from nbots_amd import Sparse, load
    L = load()
    for case in ("c1_64x32", "traction_16x8"):
        rp, col, val, F = _system(case)
        tol = 1e-8 * float(np.linalg.norm(F))
        A = Sparse.from_csr(rp, col, val, vars_per_node=2)
        x0, st0, it0, tr0 = A.solve_cg_jacobi(F, tol=tol)
        x70, st70, it70, tr70 = A.solve_cg_jacobi(F, tol=tol, max_iter=7)
        A.destroy()
        st, x, it, tr = _noalloc(L, rp, col, val, F, tol, F.size)
        print(case, "drop-in", st0, it0, tr0, "no-alloc", st, it, tr)
        assert st == st0 == 0 and abs(it - it0) <= 2 and tr <= tol
        assert np.linalg.norm(x - x0) <= 1e-10 * np.linalg.norm(x0)
        st7, x7, it7, tr7 = _noalloc(L, rp, col, val, F, tol, 7)
        assert (st7, it7) == (st70, it70) == (1, 7)
        assert abs(tr7 - tr70) <= 1e-9 * tr70
        assert np.linalg.norm(x7 - x70) <= 1e-11 * np.linalg.norm(x70)


def test_noalloc_refuses_oversize():
    from nbots_amd import load, lib
    L = load()
    n = 262145
    it, tr = C.c_uint32(0), C.c_double(0.0)
    st = L.nbx_sparse_solve_CG_precond_Jacobi_noalloc(n, 1, 1, 1, 1, 1, 10, 1e-8, C.byref(it),
                                                      C.byref(tr), None)
    assert st == 4 and "no-allocation maximum" in lib.last_error()
