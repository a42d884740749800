"""The C-ABI libraries load without a GPU and export every symbol their
headers declare (no compute calls); libnbots_amd.so defines none of the
reference's object types and imports exactly the accessors its header
lists, so that in a drop-in link those objects stay libnbots' own."""
import ctypes
import os
import subprocess

from nbots_amd import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_built_in_tree():
    assert os.path.exists(lib.LIB_PATH), "run python -m nbots_amd.build"


def _nm(path, *flags):
    r = subprocess.run(["nm", "-D", *flags, path], capture_output=True, text=True, check=True)
    return {ln.split()[-1] for ln in r.stdout.splitlines() if ln.strip()}


def test_exports_every_declared_symbol():
    names = lib.header_symbols()
    assert len(names) > 50
    objects = ctypes.CDLL(lib.OBJECTS_PATH, mode=ctypes.RTLD_GLOBAL)
    L = ctypes.CDLL(lib.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    missing = [n for n in lib.objects_symbols() if not hasattr(objects, n)]
    assert not missing, missing
    # and every declared symbol has a ctypes signature in the bindings
    assert not [n for n in names + lib.objects_symbols() if n not in lib.SIGS]


OBJECT_APIS = ("nb_mesh2D_", "nb_bcond_", "nb_material_", "nb_fem_elem_", "nb_fem_implicit_",
               "nb_graph_")


def test_product_defines_no_reference_object():
    """libnbots_amd.so exports no function of the reference's object types
    (their layouts are libnbots' own) and leaves undefined exactly the
    accessors listed between NBX_IMPORTS_BEGIN and NBX_IMPORTS_END."""
    defined = _nm(lib.LIB_PATH, "--defined-only")
    assert not sorted(n for n in defined if n.startswith(OBJECT_APIS))
    undefined = {n for n in _nm(lib.LIB_PATH, "--undefined-only") if n.startswith("nb")}
    assert undefined == set(lib.import_symbols())
    # the objects library, for callers without libnbots, defines them all
    assert set(lib.import_symbols()) <= _nm(lib.OBJECTS_PATH, "--defined-only")


def test_device_code_is_gfx950():
    """The embedded device code object targets gfx950 (and nothing else)."""
    blob = open(lib.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob
    assert b"--gfx942" not in blob and b"--gfx90a" not in blob


def test_header_compiles_as_c99():
    src = "#include \"nbx_objects.h\"\nint main(void){return 0;}\n"
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only",
                        f"-I{os.path.join(ROOT, 'include')}", "-x", "c", "-"],
                       input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_no_cpu_fallback_in_product():
    """The product never loads the oracle: no source under nbots_amd/ names it."""
    for dp, _, files in os.walk(os.path.join(ROOT, "nbots_amd")):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h")):
                txt = open(os.path.join(dp, f)).read()
                assert "import oracle" not in txt and "refcpu" not in txt, f
