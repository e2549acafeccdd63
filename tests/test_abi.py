"""The C-ABI library loads and exports every entry point include/gfmc.h
declares (no GPU calls)."""
import ctypes
import os
import re

from pygfmc_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "gfmc.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gfmc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    names = declared_functions()
    assert len(names) >= 15
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert set(declared_functions()) <= bound


def test_version_and_error_strings():
    L = _lib.load()
    assert b"gfx950" in L.gfmc_version()
    assert isinstance(L.gfmc_last_error(), bytes)


def test_problem_create_rejects_bad_descriptors_without_a_gpu():
    """Validation happens before any device call."""
    L = _lib.load()
    d = _lib.ProblemDesc()
    d.n_sites = 0
    h = ctypes.c_void_p()
    assert L.gfmc_problem_create(ctypes.byref(d), 0, ctypes.byref(h)) == -1
    assert b"n_sites" in L.gfmc_last_error()
    d.n_sites = 4
    assert L.gfmc_problem_create(ctypes.byref(d), 7, ctypes.byref(h)) == -1
    assert b"precision" in L.gfmc_last_error()
