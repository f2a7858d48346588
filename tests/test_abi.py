"""The C-ABI library loads and exports every function include/ufvm.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from matlab_ufvm_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "ufvm.h")).read()
    txt = re.sub(r"/\*.*?\*/", " ", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ufvm_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    assert "ufvm_pcg_solve" in names and "ufvm_apply_amg" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(_lib.EXPORTED) == declared()


def test_no_cpu_fallback_when_library_missing(tmp_path):
    import pytest
    with pytest.raises(RuntimeError):
        _lib._lib = None
        try:
            _lib.load(str(tmp_path / "absent.so"))
        finally:
            _lib._lib = None
