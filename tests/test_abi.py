"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import glob
import os
import re

import cola_amd
from cola_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        # C-ABI declarations: a cola_status / const char* return type at line start
        names |= set(re.findall(r"^(?:cola_status|const char \*)\s*(cola_[a-z0-9_]+)\s*\(", text, flags=re.M))
    return names


def test_library_loads_and_exports_every_declared_symbol():
    L = cola_amd.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert declared_functions() == set(_lib.SIGNATURES)


def test_exports_are_c_symbols():
    out = os.popen(f"nm -D --defined-only {_lib.LIB_PATH}").read()
    exported = set(re.findall(r" T (cola_[a-z0-9_]+)$", out, flags=re.M))
    assert declared_functions() <= exported


def test_version_and_error_text():
    L = cola_amd.lib()
    assert L.cola_version().startswith(b"cola")
    # An invalid call returns a status and sets the error text (no exit()).
    h = ctypes.c_void_p()
    st = L.cola_rmat_graph(0, 1.0, 0.5, 0.2, 0.2, 1, 0, 1, ctypes.byref(h))
    assert st == _lib.COLA_ERR_INVALID
    assert b"log_n" in L.cola_last_error()


def test_cpp_drop_in_driver_compiles_and_links(tmp_path):
    """Reference-style host code (vector-taking d_csr_zip / d_csc_zip
    constructors, nnz_split, single-stage wrapper) builds against
    include/sparse_data.h and the library."""
    from tests.cpp_driver import build_driver
    exe = build_driver(tmp_path)
    assert os.path.exists(exe)


def test_cpp_reference_driver_compiles_and_links(tmp_path):
    """A driver whose load, struct and call lines are those of the reference's
    src/distSpMM_ppl_batch.cu (bin2vec, nnz_split, d_csc_zip / d_csr_zip from
    std::vectors, run_distSpMM_ppl_batch<true, ...>) builds against
    include/dist_spmm_ppl_batch.h and the library."""
    from tests.cpp_driver import build_driver
    exe = build_driver(tmp_path, "reference_driver")
    assert os.path.exists(exe)
