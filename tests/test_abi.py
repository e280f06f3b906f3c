"""The C-ABI library loads and exports every symbol include/tsmi_hip.h declares
(no GPU work), and the host-side argument checks reject bad input."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "tsmi_hip.h")).read()
    return sorted(set(re.findall(r"\b(tsmi_[a-z_0-9]+)\s*\(", text)))


def test_header_lists_match_binding():
    from tsmi import _lib
    assert declared_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_all_symbols():
    from tsmi import _lib
    L = _lib.load()
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_invalid_descriptor_rejected_without_gpu():
    from tsmi import _lib
    L = _lib.load()
    h = C.c_void_p()
    d = _lib.Desc(0, 10, 0, 0, 4, 4, 0, 64, 0, 0, 1e-3, 1.0, 0.0, 0.0, 0.1, 0.1, 1.0, 1.0, 0.0)
    st = L.tsmi_create(C.byref(d), C.byref(h))
    assert st != 0 and b"positive" in L.tsmi_last_error()
    d = _lib.Desc(10, 10, 0, 0, 4, 4, 0, 16, 0, 0, 1e-3, 1.0, 0.0, 0.0, 0.1, 0.1, 1.0, 1.0, 0.0)
    assert L.tsmi_create(C.byref(d), C.byref(h)) != 0 and b"precision" in L.tsmi_last_error()
    d = _lib.Desc(10, 10, 0, 0, 8, 8, 0, 64, 0, 0, 1e-3, 1.0, 0.0, 0.0, 0.1, 0.1, 1.0, 1.0, 0.0)
    assert L.tsmi_create(C.byref(d), C.byref(h)) != 0 and b"halo" in L.tsmi_last_error()
    assert L.tsmi_evolve(None, 1) != 0


def test_pair_descriptor_checks_without_gpu():
    from tsmi import _lib
    L = _lib.load()
    h = C.c_void_p()
    a = _lib.Desc(10, 10, 0, 0, 4, 4, 0, 64, 0, 0, 1e-3, 1.0, 0.0, 0.0, 0.1, 0.1, 1.0, 1.0, 0.0)
    b = _lib.Desc(12, 10, 0, 0, 4, 4, 0, 64, 0, 0, 1e-3, 1.0, 0.0, 0.0, 0.1, 0.1, 1.0, 1.0, 0.0)
    d = _lib.PairDesc(a, b, 0.0, 1.0, 0.0)
    assert L.tsmi_pair_create(C.byref(d), C.byref(h)) == 1 and b"same lattice" in L.tsmi_last_error()
    b = _lib.Desc(10, 10, 0, 0, 4, 4, 0, 64, 0, 0, 1e-3, 1.0, 0.0, 0.5, 0.1, 0.1, 1.0, 1.0, 0.0)
    d = _lib.PairDesc(a, b, 0.0, 1.0, 0.0)
    assert L.tsmi_pair_create(C.byref(d), C.byref(h)) == 4 and b"LHY" in L.tsmi_last_error()
    assert L.tsmi_pair_evolve(None, 1, 1) != 0
    assert L.tsmi_pair_set_state(None, 0, None, None) != 0


def test_no_cpu_fallback_when_library_missing(monkeypatch, tmp_path):
    from tsmi import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(_lib.TsmiError):
        _lib.load()
