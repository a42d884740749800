"""ctypes bindings of libnbots_amd.so (include/nbots_amd.h) and of
libnbx_objects.so (include/nbx_objects.h).

libnbots_amd.so reads the FEM input objects (mesh, element type, material,
boundary conditions, implicit-scheme parameters) through accessors it
imports; in a drop-in link libnbots provides them.  Here, without libnbots,
libnbx_objects.so provides them: it is loaded first, into the global symbol
scope, and the product library after it.  ``load()`` returns one namespace
holding the functions of both.

Both libraries are built in-tree by ``python -m nbots_amd.build``.  Loading
fails loudly when either is missing: there is no CPU fallback for any entry
point.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnbots_amd.so")
OBJECTS_PATH = os.path.join(HERE, "libnbx_objects.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "nbots_amd.h")
OBJECTS_HEADER = os.path.join(os.path.dirname(HERE), "include", "nbx_objects.h")

u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class _Opt:
    def __init__(self, base):
        self.base = base

    def from_param(self, obj):
        return None if obj is None else self.base.from_param(obj)


u32o, u8o, f64o, i32o = _Opt(u32p), _Opt(u8p), _Opt(f64p), _Opt(i32p)
U32, U64, F64, I32, VP = C.c_uint32, C.c_uint64, C.c_double, C.c_int, C.c_void_p


class SolveInfo(C.Structure):
    _fields_ = [("iters", C.c_uint32), ("outer", C.c_uint32), ("fp64_iters", C.c_uint32),
                ("status", C.c_int32), ("rel_res", C.c_double), ("tol_reached", C.c_double),
                ("ms", C.c_double), ("fallback", C.c_int32), ("layout", C.c_int32),
                ("ms_spmv", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SolveOpts(C.Structure):
    """nbx_solve_opts_t"""
    _fields_ = [("mixed_variant", C.c_int32), ("cheb_degree", C.c_uint32),
                ("cheb_ratio", C.c_double), ("inner_rtol", C.c_double),
                ("reliable_delta", C.c_double), ("precond_bits", C.c_int32),
                ("true_residual", C.c_int32), ("replay", C.c_int32),
                ("reserved", C.c_int32)]


class Graph(C.Structure):
    """nb_graph_t (headers/nb/graph_bot/graph.h:25-31)."""
    _fields_ = [("N", C.c_uint32), ("N_adj", C.POINTER(C.c_uint32)),
                ("adj", C.POINTER(C.POINTER(C.c_uint32))),
                ("wi", C.POINTER(C.c_double)), ("wij", C.POINTER(C.POINTER(C.c_double)))]


_LIB = None

# name -> (restype, argtypes)
SIGS = {
    "nb_graph_get_memsize": (U32, []),
    "nb_graph_init": (None, [VP]),
    "nb_graph_finish": (None, [VP]),
    "nb_graph_clear": (None, [VP]),
    "nb_sparse_create": (VP, [VP, u32o, U32]),
    "nb_sparse_clone": (VP, [VP]),
    "nb_sparse_save": (None, [VP, C.c_char_p]),
    "nb_sparse_read_mat4": (None, [VP, C.c_char_p, C.c_char_p]),
    "nb_sparse_save_mat4": (None, [VP, C.c_char_p, C.c_char_p]),
    "nb_mat4_exist": (C.c_short, [C.c_char_p, C.c_char_p]),
    "nb_mat4_clear": (None, [C.c_char_p]),
    "nb_mat4_read_vec": (None, [C.c_char_p, C.c_char_p, f64p]),
    "nb_mat4_save_vec": (None, [C.c_char_p, C.c_char_p, f64p, U32]),
    "nb_sparse_destroy": (None, [VP]),
    "nb_sparse_reset": (None, [VP]),
    "nb_sparse_set": (None, [VP, U32, U32, F64]),
    "nb_sparse_set_identity_row": (None, [VP, U32]),
    "nb_sparse_make_diagonal": (None, [VP, F64]),
    "nb_sparse_get": (F64, [VP, U32, U32]),
    "nb_sparse_get_and_set": (F64, [VP, U32, U32, F64]),
    "nb_sparse_is_non_zero": (C.c_bool, [VP, U32, U32]),
    "nb_sparse_add": (None, [VP, U32, U32, F64]),
    "nb_sparse_scale": (None, [VP, F64]),
    "nb_sparse_get_size": (U32, [VP]),
    "nb_sparse_get_nnz": (U32, [VP]),
    "nb_sparse_get_asym": (F64, [VP]),
    "nb_sparse_get_frobenius_norm": (F64, [VP]),
    "nb_sparse_multiply_scalar": (None, [VP, F64, U32]),
    "nb_sparse_multiply_vector": (None, [VP, f64p, f64p, U32]),
    "nb_sparse_set_Dirichlet_condition": (None, [VP, f64p, U32, F64]),
    "nb_sparse_solve_CG_precond_Jacobi": (I32, [VP, f64p, f64p, U32, F64, C.POINTER(U32),
                                                C.POINTER(F64), U32]),
    "nb_sparse_solve_conjugate_gradient": (I32, [VP, f64p, f64p, U32, F64, C.POINTER(U32),
                                                 C.POINTER(F64), U32]),
    "nb_sparse_solve_Gauss_Seidel": (I32, [VP, f64p, f64p, U32, F64, C.POINTER(U32),
                                           C.POINTER(F64), U32]),
    "nb_mesh2D_create": (VP, [I32]),
    "nb_mesh2D_destroy": (None, [VP]),
    "nb_mesh2D_get_N_invtx": (U32, [VP]),
    "nb_mesh2D_get_N_insgm": (U32, [VP]),
    "nb_mesh2D_get_N_nodes": (U32, [VP]),
    "nb_mesh2D_get_N_elems": (U32, [VP]),
    "nb_mesh2D_node_get_x": (F64, [VP, U32]),
    "nb_mesh2D_node_get_y": (F64, [VP, U32]),
    "nb_mesh2D_elem_get_adj": (U32, [VP, U32, C.c_uint8]),
    "nb_mesh2D_get_invtx": (U32, [VP, U32]),
    "nb_mesh2D_insgm_get_N_nodes": (U32, [VP, U32]),
    "nb_mesh2D_insgm_get_node": (U32, [VP, U32, U32]),
    "nb_mesh2D_elem_get_N_adj": (U32, [VP, U32]),
    "nb_mesh2D_insgm_get_length": (F64, [VP, U32]),
    "nb_mesh2D_insgm_subsgm_get_length": (F64, [VP, U32, U32]),
    "nbx_mesh2D_set_trg": (I32, [VP, U32, f64p, U32, u32p, U32, u32p, U32, u32p, u32p]),
    "nb_fem_elem_create": (VP, [I32]),
    "nb_fem_elem_destroy": (None, [VP]),
    "nb_fem_elem_get_N_gpoints": (C.c_uint8, [VP]),
    "nb_fem_elem_get_N_nodes": (C.c_uint8, [VP]),
    "nb_material_create": (VP, []),
    "nb_material_destroy": (None, [VP]),
    "nb_material_set_poisson_module": (None, [VP, F64]),
    "nb_material_set_elasticity_module": (None, [VP, F64]),
    "nb_material_set_density": (None, [VP, F64]),
    "nb_material_get_poisson_module": (F64, [VP]),
    "nb_material_get_elasticity_module": (F64, [VP]),
    "nb_material_get_density": (F64, [VP]),
    "nb_material_clone": (VP, [VP]),
    "nb_material_set_fracture_energy": (None, [VP, F64]),
    "nb_material_set_traction_limit_stress": (None, [VP, F64]),
    "nb_material_set_compression_limit_stress": (None, [VP, F64]),
    "nb_material_set_energy_release_rate": (None, [VP, F64]),
    "nb_material_set_damage_length_scale": (None, [VP, F64]),
    "nb_material_get_fracture_energy": (F64, [VP]),
    "nb_material_get_traction_limit_stress": (F64, [VP]),
    "nb_material_get_compression_limit_stress": (F64, [VP]),
    "nb_material_get_energy_release_rate": (F64, [VP]),
    "nb_material_get_damage_length_scale": (F64, [VP]),
    "nb_fem_implicit_create": (VP, []),
    "nb_fem_implicit_destroy": (None, [VP]),
    "nb_fem_implicit_set_N_steps": (None, [VP, U32]),
    "nb_fem_implicit_set_N_max_iter": (None, [VP, U32]),
    "nb_fem_implicit_set_N_max_iter_without_enhance": (None, [VP, U32]),
    "nb_fem_implicit_set_residual_tolerance": (None, [VP, F64]),
    "nb_fem_implicit_get_N_steps": (U32, [VP]),
    "nb_fem_implicit_get_N_max_iter": (U32, [VP]),
    "nb_fem_implicit_get_N_max_iter_without_enhance": (U32, [VP]),
    "nb_fem_implicit_get_residual_tolerance": (F64, [VP]),
    "nb_fem_compute_2D_Non_Linear_Solid_Mechanics": (None, [VP, VP, VP, VP, C.c_bool, f64p,
                                                            C.c_bool, I32, VP, VP, C.c_char_p]),
    "nbx_fem_damage_solve": (I32, [VP, VP, VP, VP, I32, f64o, VP, VP, C.c_char_p, f64o, f64o]),
    "nb_bcond_get_memsize": (C.c_uint16, [C.c_uint8]),
    "nb_bcond_init": (None, [VP, C.c_uint8]),
    "nb_bcond_finish": (None, [VP]),
    "nb_bcond_create": (VP, [C.c_uint8]),
    "nb_bcond_destroy": (None, [VP]),
    "nb_bcond_clear": (None, [VP]),
    "nb_bcond_get_N_dof": (C.c_uint8, [VP]),
    "nb_bcond_push": (None, [VP, I32, I32, U32, C.POINTER(C.c_bool), C.POINTER(C.c_double)]),
    "nb_bcond_push_function": (None, [VP, I32, I32, U32, C.POINTER(C.c_bool), VP]),
    "nb_bcond_iter_get_memsize": (C.c_uint16, []),
    "nb_bcond_iter_init": (None, [VP]),
    "nb_bcond_iter_finish": (None, [VP]),
    "nb_bcond_iter_create": (VP, []),
    "nb_bcond_iter_destroy": (None, [VP]),
    "nb_bcond_iter_set_conditions": (None, [VP, VP, I32, I32]),
    "nb_bcond_iter_has_more": (C.c_bool, [VP]),
    "nb_bcond_iter_go_next": (None, [VP]),
    "nb_bcond_iter_get_id": (U32, [VP]),
    "nb_bcond_iter_get_mask": (C.c_bool, [VP, C.c_uint8]),
    "nb_bcond_iter_val_is_function": (C.c_bool, [VP]),
    "nb_bcond_iter_get_val": (None, [VP, C.c_uint8, f64p, F64, f64p]),
    "nb_fem_compute_2D_Solid_Mechanics": (I32, [VP, VP, VP, VP, C.c_bool, f64o, I32, VP, VP,
                                                f64p, f64p]),
    "nb_fem_compute_stress_from_strain": (None, [U32, VP, VP, I32, f64p, VP, f64p]),
    "nbx_last_error": (C.c_char_p, []),
    "nbx_device_info": (I32, [C.c_char_p, U32, C.POINTER(U32), C.POINTER(F64)]),
    "nbx_sparse_solve": (I32, [VP, f64p, f64p, I32, F64, U32, F64, C.POINTER(SolveInfo)]),
    "nbx_sparse_solve_ex": (I32, [VP, f64p, f64p, I32, F64, U32, C.POINTER(SolveOpts),
                                  C.POINTER(SolveInfo)]),
    "nbx_solve_opts_default": (None, [C.POINTER(SolveOpts)]),
    "nbx_simp_opts_default": (None, [C.POINTER(SolveOpts)]),
    "nbx_sparse_spmv_dev": (I32, [VP, VP, VP]),
    "nbx_sparse_is_blocked": (I32, [VP]),
    "nbx_sparse_solve_CG_precond_Jacobi_noalloc": (I32, [U32, VP, VP, VP, VP, VP, U32, F64,
                                                         C.POINTER(U32), C.POINTER(F64), VP]),
    "nbx_sparse_export_csr": (I32, [VP, u32o, u32o, f64o]),
    "nbx_sparse_import_values": (I32, [VP, f64p]),
    "nbx_fem_create": (I32, [C.POINTER(VP), U32, f64p, U32, u32p]),
    "nbx_fem_destroy": (None, [VP]),
    "nbx_fem_sizes": (I32, [VP, C.POINTER(U32), C.POINTER(U64), C.POINTER(U64)]),
    "nbx_fem_assemble": (I32, [VP, F64, F64, F64, f64o, u8o, I32, F64, f64o]),
    "nbx_fem_set_bconditions": (I32, [VP, u32p, u32p, u32p]
                                + [U32, u32p, u8p, f64p] * 4 + [i32o, f64o, F64]),
    "nbx_fem_solve": (I32, [VP, I32, F64, U32, F64, f64o, C.POINTER(SolveInfo)]),
    "nbx_fem_solve_ex": (I32, [VP, I32, F64, U32, C.POINTER(SolveOpts), f64o,
                               C.POINTER(SolveInfo)]),
    "nbx_fem_get_displacement": (I32, [VP, f64p]),
    "nbx_fem_true_residual": (I32, [VP, I32, C.POINTER(F64)]),
    "nbx_fem_get_system": (I32, [VP, u32o, u32o, f64o, f64o]),
    "nbx_fem_strain_stress": (I32, [VP, f64p, f64p]),
    "nbx_fem_compliance": (I32, [VP, f64o, F64, F64, F64, C.POINTER(F64), f64o, f64o]),
    "nbx_fem_adjoint": (I32, [VP, f64p, f64p, I32, F64, U32, C.POINTER(SolveInfo)]),
    "nbx_fem_elem_cross": (I32, [VP, f64o, f64o, f64p]),
    "nbx_fem_spmv_bench": (I32, [VP, I32, U32, C.POINTER(F64)]),
    "nbx_simp_create": (I32, [C.POINTER(VP), VP, F64, F64, F64, F64, F64, F64, f64o]),
    "nbx_simp_destroy": (None, [VP]),
    "nbx_simp_iterate": (I32, [VP, I32, F64, U32, C.c_void_p, C.POINTER(F64), C.POINTER(F64),
                               C.POINTER(SolveInfo)]),
    "nbx_simp_get": (I32, [VP, f64o, f64o, f64o, C.POINTER(U32)]),
    "nbx_simp_filter": (I32, [VP, C.POINTER(U64), u32o, u32o, f64o, f64o]),
}


def _declared(path):
    import re
    txt = open(path).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nbx?_[A-Za-z0-9_]+)\s*\(", txt)))


def import_symbols():
    """The accessors libnbots_amd.so imports: include/nbots_amd.h between
    NBX_IMPORTS_BEGIN and NBX_IMPORTS_END."""
    import re
    txt = open(HEADER).read()
    sec = txt[txt.index("NBX_IMPORTS_BEGIN"):txt.index("NBX_IMPORTS_END")]
    sec = re.sub(r"/\*.*?\*/", "", "/*" + sec + "*/", flags=re.S)
    return sorted(set(re.findall(r"\b(nb_[A-Za-z0-9_]+)\s*\(", sec)))


def header_symbols():
    """Functions libnbots_amd.so defines: those declared in
    include/nbots_amd.h outside the imports section."""
    imports = set(import_symbols())
    return [n for n in _declared(HEADER) if n not in imports]


def objects_symbols():
    """Functions libnbx_objects.so defines: include/nbx_objects.h and the
    imports section of include/nbots_amd.h."""
    own = [n for n in _declared(OBJECTS_HEADER) if n not in _declared(HEADER)]
    return sorted(set(own) | set(import_symbols()))


class _Namespace:
    """The functions of both libraries under one name (product first)."""

    def __init__(self, product, objects):
        self._product = product
        self._objects = objects

    def __getattr__(self, name):
        try:
            return getattr(self._product, name)
        except AttributeError:
            return getattr(self._objects, name)


def load(path: str = LIB_PATH):
    global _LIB
    if _LIB is not None:
        return _LIB
    for p in (OBJECTS_PATH, path):
        if not os.path.exists(p):
            raise RuntimeError(f"{p} is missing: build it with `python -m nbots_amd.build` "
                               "(the HIP library is required; there is no CPU fallback)")
    objects = C.CDLL(OBJECTS_PATH, mode=C.RTLD_GLOBAL)
    product = C.CDLL(path)
    L = _Namespace(product, objects)
    for name, (res, args) in SIGS.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def last_error() -> str:
    e = load().nbx_last_error()
    return e.decode() if e else ""


def check(rc: int, what: str, ok=(0,)):
    if rc not in ok:
        raise RuntimeError(f"{what} failed with status {rc}: {last_error()}")
    return rc
