"""The C-ABI library loads and exports exactly what include/*.h declares
(no compute calls: this runs without a GPU)."""
import glob
import os
import re
import subprocess

import gpu_spmv_amd._lib as L
from conftest import ROOT


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(axspmv_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_declares_functions():
    assert len(declared()) > 30


def test_library_exports_every_declared_symbol():
    lib = L.load()
    for name in sorted(declared()):
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (axspmv_\w+)", out))
    assert declared() <= exported
    assert exported <= declared(), exported - declared()


def test_python_binding_covers_header():
    assert declared() == set(L.SIGNATURES)


def test_abi_version_and_error_path():
    lib = L.load()
    assert lib.axspmv_abi_version() == 1
    import ctypes as C
    rc = lib.axspmv_csr_check(None)
    assert rc == L.E_INVALID
    assert b"NULL" in lib.axspmv_last_error()


def test_wave_selftest_rejects_bad_arguments():
    """Checked before any device call: partial waves, unknown ops, NULL."""
    lib = L.load()
    assert lib.axspmv_wave_selftest(0, None, None, None, None, 65, None) == L.E_INVALID
    assert b"whole waves" in lib.axspmv_last_error()
    assert lib.axspmv_wave_selftest(11, None, None, None, None, 64, None) == L.E_INVALID
    assert b"unknown op" in lib.axspmv_last_error()
    assert lib.axspmv_wave_selftest(0, None, None, None, None, 64, None) == L.E_INVALID
    assert b"NULL" in lib.axspmv_last_error()
