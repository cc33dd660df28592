"""CPU checks of the C ABI: libsmat.so loads without a GPU and exports every
function include/smat.h declares; the Python binding covers them all; the
product library links no part of the oracle."""
import ctypes
import os
import re
import subprocess

import pytest

import smat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "smat.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:[A-Za-z_][\w\s\*]*?[\s\*])([A-Za-z_]\w*)\s*\(", src, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while", "return")))


def test_header_parses():
    names = declared_functions()
    assert "smat_spmm_execute" in names and "mmaOBTKernel" in names
    assert len(names) >= 18


def test_library_exports_every_declared_symbol():
    if not os.path.exists(smat.LIB_PATH):
        pytest.skip("libsmat.so not built")
    lib = ctypes.CDLL(smat.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_functions()) == set(smat.SIGNATURES)


def test_error_state_without_gpu():
    if not os.path.exists(smat.LIB_PATH):
        pytest.skip("libsmat.so not built")
    l = smat.lib()
    l.smat_clear_error()
    assert l.smat_last_status() == 0
    assert l.smat_version().startswith(b"smat")
    # argument validation happens before any device work
    h = ctypes.c_void_p()
    st = l.smat_bcsr_from_csr(16, 16, 0, None, None, None, 12, 0, None, ctypes.byref(h))
    assert st == smat.SMAT_ERR_INVALID
    assert b"bk" in l.smat_last_error()


def test_product_does_not_link_oracle():
    if not os.path.exists(smat.LIB_PATH):
        pytest.skip("libsmat.so not built")
    out = subprocess.run(["nm", "-D", smat.LIB_PATH], capture_output=True, text=True).stdout
    assert " or_" not in out and "ref_keeper" not in out
    deps = subprocess.run(["ldd", smat.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in deps and "smat_ref" not in deps
