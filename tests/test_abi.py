"""The C-ABI library loads without a GPU and exports every symbol that
include/nbots_amd.h declares (no compute calls)."""
import ctypes
import os
import subprocess

from nbots_amd import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_built_in_tree():
    assert os.path.exists(lib.LIB_PATH), "run python -m nbots_amd.build"


def test_exports_every_declared_symbol():
    names = lib.header_symbols()
    assert len(names) > 60
    L = ctypes.CDLL(lib.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and every declared symbol has a ctypes signature in the bindings
    assert not [n for n in names if n not in lib.SIGS]


def test_device_code_is_gfx950():
    """The embedded device code object targets gfx950 (and nothing else)."""
    blob = open(lib.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob
    assert b"--gfx942" not in blob and b"--gfx90a" not in blob


def test_header_compiles_as_c99():
    src = "#include \"nbots_amd.h\"\nint main(void){return 0;}\n"
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
