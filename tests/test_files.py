"""On-disk matrix formats of the drop-in (host code, no GPU): nb_sparse_save
(sparse.c:97-112) and the MATLAB v4 records of matlab_v4.c, against files the
compiled reference's own writers produced (tests/golden/cant_4x2_K.txt and
cant_4x2.mat: K after BCs as "K", F as "F", K before BCs as "K0", for the
4x2 jittered cantilever)."""
import os

import numpy as np

from conftest import GOLDEN

import oracle


def _system():
    from nbots_amd import synth
    p = synth.cantilever_point(4, 2, jitter=0.1)
    rp, col = oracle.pattern(p.N_nod, p.conn)
    val0, F0, _ = oracle.assemble(p, rp, col)
    val, F = val0.copy(), F0.copy()
    oracle.set_bconditions(p, rp, col, val, F)
    return rp, col, val0, val, F


def _read(path):
    with open(path, "rb") as f:
        return f.read()


def test_text_save_byte_identical(tmp_path):
    import nbots_amd as nb
    rp, col, _, val, _ = _system()
    A = nb.Sparse.from_csr(rp, col, val, vars_per_node=2)
    out = str(tmp_path / "K.txt")
    A.L.nb_sparse_save(A.h, out.encode())
    assert _read(out) == _read(os.path.join(GOLDEN, "cant_4x2_K.txt"))
    A.destroy()


def test_mat4_save_byte_identical(tmp_path):
    import nbots_amd as nb
    rp, col, val0, val, F = _system()
    A = nb.Sparse.from_csr(rp, col, val, vars_per_node=2)
    A0 = nb.Sparse.from_csr(rp, col, val0, vars_per_node=2)
    out = str(tmp_path / "sys.mat").encode()
    L = A.L
    L.nb_mat4_clear(out)
    L.nb_sparse_save_mat4(A.h, out, b"K")
    L.nb_mat4_save_vec(out, b"F", np.ascontiguousarray(F), F.size)
    L.nb_sparse_save_mat4(A0.h, out, b"K0")
    assert _read(out.decode()) == _read(os.path.join(GOLDEN, "cant_4x2.mat"))
    A.destroy()
    A0.destroy()


def test_mat4_read_reference_file():
    """Records found by label in a file the reference wrote; the matrix comes
    back in the reference's row layout (sorted columns) with its values."""
    import nbots_amd as nb
    rp, col, val0, val, F = _system()
    path = os.path.join(GOLDEN, "cant_4x2.mat").encode()
    L = nb.lib.load()
    assert L.nb_mat4_exist(path, b"K") == 1 and L.nb_mat4_exist(path, b"K0") == 1
    assert L.nb_mat4_exist(path, b"F") == 1 and L.nb_mat4_exist(path, b"G") == 0
    for label, want in ((b"K", val), (b"K0", val0)):
        A = nb.Sparse.from_csr(np.array([0, 1], np.uint32), np.array([0], np.uint32),
                               np.array([1.0]), vars_per_node=1)
        L.nb_sparse_read_mat4(A.h, path, label)
        rp2, col2, val2 = A.export()
        assert np.array_equal(rp2, rp) and np.array_equal(col2, col) and np.array_equal(val2, want)
        A.destroy()
    x = np.zeros(F.size)
    L.nb_mat4_read_vec(path, b"F", x)
    assert np.array_equal(x, F)


def test_mat4_round_trip_unsorted_rows(tmp_path):
    """A record whose entries are not in column order within a row (written
    by hand): the reader sorts each row's columns and moves the values with
    them."""
    import struct
    import nbots_amd as nb
    path = str(tmp_path / "u.mat")
    rows = [1, 2, 1, 2]          # 1-based, entries given column-major but swapped
    cols = [2, 2, 1, 1]
    vals = [3.0, 4.0, 1.0, 2.0]  # A = [[1, 3], [2, 4]]
    with open(path, "wb") as f:
        f.write(struct.pack("<5i", 2, 5, 3, 0, 2) + b"A\0")
        f.write(struct.pack("<15d", *(rows + [2] + cols + [2] + vals + [0])))
    A = nb.Sparse.from_csr(np.array([0, 1], np.uint32), np.array([0], np.uint32),
                           np.array([1.0]), vars_per_node=1)
    A.L.nb_sparse_read_mat4(A.h, path.encode(), b"A")
    rp, col, val = A.export()
    assert rp.tolist() == [0, 2, 4] and col.tolist() == [0, 1, 0, 1]
    assert val.tolist() == [1.0, 3.0, 2.0, 4.0]
    A.destroy()
