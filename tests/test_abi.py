"""The C-ABI library loads without a GPU and exports every entry point that
include/glacx.h declares (no compute calls)."""
import ctypes
import os
import re

import glac_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"\b(glacx_\w+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ["glacx_create", "glacx_flow", "glacx_metropolis", "glacx_measure", "glacx_load_bin"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(glac_amd.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_abi():
    glac_amd.library()  # sets argtypes for every entry point; raises if one is absent


def test_create_rejects_bad_shapes_before_touching_a_device():
    """Extent and size checks come before any HIP call, so they hold on a
    machine without a GPU: odd or too small extents, unknown precision, and
    lattices whose direction region would exceed a 32-bit buffer range."""
    lib = glac_amd.library()
    h = ctypes.c_void_p()
    for dims, prec in [((3, 4, 4, 4), 0), ((4, 4, 4, 0), 0), ((4, 4, 4, 4), 7),
                       ((80, 80, 80, 80), 0), ((96, 96, 96, 96), 1), ((514, 2, 2, 2), 0)]:
        rc = lib.glacx_create((ctypes.c_int * 4)(*dims), prec, 0, ctypes.byref(h))
        assert rc == -1, (dims, prec, rc)
