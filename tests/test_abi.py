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


def _plan(dim_x, dim_y, periodic, precision, tile_rows=0, sweeps=0, imag=0):
    from tsmi import _lib
    L = _lib.load()
    d = _lib.Desc(dim_x, dim_y, periodic, periodic, 4, 4, imag, precision, tile_rows, sweeps, 1e-3, 1.0, 1.0, 0.0,
                  0.1, 0.1, 0.1 * dim_x, 0.1 * dim_y, 0.0)
    out = (C.c_int32 * 12)()
    st = L.tsmi_plan_geometry(C.byref(d), out)
    return st, list(out), L


# One FP64 segment plus its apron rows must span less than 2^31 bytes: the
# out-of-range offset 0x80000000 that drops a store or zeroes a load has to lie
# outside every segment's buffer range.  Lattices whose planes fit an MI355X
# (40000^2 FP64: ~141 GB of planes) and a caller-given segment height of 2^20
# rows both reach that limit.
@pytest.mark.parametrize("dim,periodic,precision,tile_rows", [
    (40000, 1, 64, 0), (40000, 0, 64, 0), (40000, 1, 32, 0), (60000, 1, 64, 0),
    (40000, 1, 64, 1 << 20), (20000, 0, 32, 1 << 20), (8192, 1, 32, 0), (257, 0, 64, 0),
])
def test_plan_geometry_segments_stay_addressable(dim, periodic, precision, tile_rows):
    st, g, L = _plan(dim, dim, periodic, precision, tile_rows=tile_rows)
    assert st == 0, L.tsmi_last_error()
    W, H, pitch, strips, segs, rows, sweeps, strips2, segs2, rows2, Wd, Hd = g
    assert (W, H) == (dim + 8 * periodic, dim + 8 * periodic)
    assert rows >= 2 and rows % 2 == 0 or rows == H
    assert segs * rows >= H and (segs - 1) * rows < H
    assert (rows + 8) * pitch * 8 < 2 ** 31
    assert sweeps == 2   # real time, both axes alike
    assert (Wd, Hd) == ((W + 8, H + 8) if periodic else (W, H))
    assert pitch % 64 == 0 and pitch >= Wd
    assert rows2 >= 2 and (rows2 + 16) * pitch * 8 < 2 ** 31
    assert segs2 * rows2 >= Hd and (segs2 - 1) * rows2 < Hd
    if tile_rows:
        # a caller's height is kept where it fits and cut where it does not
        fits = (tile_rows + 8) * pitch * 8 < 2 ** 31
        assert rows == min(tile_rows, H) if fits else rows < tile_rows


def test_plan_geometry_one_step_per_pass_where_two_are_not_exact():
    st, g, _ = _plan(300, 300, 0, 64, imag=1)
    assert st == 0 and g[6] == 1 and g[7:10] == [0, 0, 0]
    st, g, _ = _plan(300, 300, 1, 32, sweeps=1)
    assert st == 0 and g[6] == 1 and g[10:12] == [308, 308]


def test_plan_geometry_rejects_lattices_past_the_offset_limit():
    st, _, L = _plan(20_000_000, 4, 0, 64)
    assert st == 1 and b"too large" in L.tsmi_last_error()
