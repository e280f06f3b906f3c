"""tsmi_energy_device_noalloc (the entry point that allocates nothing) against
tsmi_energy on the same states: two calls in a row, the second on a smaller
lattice, in both precisions, through the C ABI.  The state and potential live
in caller-owned device memory (torch tensors, row pitch = tile width); the
results are written to device memory and read back after a synchronisation.

Both entry points run the same reduction kernels over the same inner region
with the same grid, so the bar is FP64's 1e-12 relative (equal bits expected).
"""
import ctypes as C
import sys

import numpy as np
import pytest

from tsmi import workloads as W

pytestmark = pytest.mark.gpu

CASES = [  # (n_x, n_y, periodic_x, periodic_y, g, g_lhy): the second call is smaller
    (520, 200, True, False, 3.0, 0.5),
    (96, 72, False, True, 2.0, 0.0),
]


def _desc(lat, wl, precision):
    from tsmi import _lib
    return _lib.Desc(lat.dim_x, lat.dim_y, int(lat.periodic_x), int(lat.periodic_y), lat.halo_x, lat.halo_y,
                     0, precision, 0, 0, wl.dt, wl.mass, wl.g, wl.g_lhy, lat.delta_x, lat.delta_y,
                     lat.length_x, lat.length_y, 0.0)


@pytest.mark.parametrize("precision", [64, 32])
def test_noalloc_energy_matches_tsmi_energy(precision):
    import torch
    from tsmi import _lib
    from tsmi.hip import kernel_for
    L = _lib.load()
    torch.cuda.set_device(0)
    dt = torch.float64 if precision == 64 else torch.float32
    for (nx, ny, px, py, g, g_lhy) in CASES:
        print(f"[noalloc] FP{precision} {nx}x{ny} periodic ({px}, {py})", file=sys.stderr, flush=True)
        wl = W.small_case(nx, ny, px, py, steps=0, g=g, g_lhy=g_lhy)
        lat = wl.lattice
        k = kernel_for(wl, precision=precision)
        try:
            ref = k.energy()
        finally:
            k.close()
        re = torch.from_numpy(np.ascontiguousarray(wl.psi.real)).to("cuda", dt)
        im = torch.from_numpy(np.ascontiguousarray(wl.psi.imag)).to("cuda", dt)
        v = torch.from_numpy(np.ascontiguousarray(wl.v, dtype=np.float64)).to("cuda")
        out = torch.full((6,), float("nan"), dtype=torch.float64, device="cuda")
        assert re.shape == (lat.H, lat.W) and v.shape == (lat.H, lat.W)
        d = _desc(lat, wl, precision)
        torch.cuda.synchronize()
        st = L.tsmi_energy_device_noalloc(C.byref(d), re.data_ptr(), im.data_ptr(), lat.W, v.data_ptr(),
                                          out.data_ptr(), torch.cuda.current_stream().cuda_stream)
        assert st == 0, L.tsmi_last_error()
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        print(f"[noalloc] tsmi_energy {ref}\n[noalloc] noalloc     {got}", file=sys.stderr, flush=True)
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0.0)


def test_noalloc_rejects_bad_arguments():
    import torch
    from tsmi import _lib
    L = _lib.load()
    torch.cuda.set_device(0)
    wl = W.small_case(64, 48, False, False, steps=0, g=1.0)
    lat = wl.lattice
    d = _desc(lat, wl, 64)
    buf = torch.zeros((lat.H, lat.W), dtype=torch.float64, device="cuda")
    out = torch.zeros(6, dtype=torch.float64, device="cuda")
    # row pitch below the tile width
    assert L.tsmi_energy_device_noalloc(C.byref(d), buf.data_ptr(), buf.data_ptr(), lat.W - 1, None,
                                        out.data_ptr(), None) == 1
    assert b"pitch" in L.tsmi_last_error()
    # an inner region wider than the scratch array covers (the check comes before any device work)
    wide = _lib.Desc(4096 * 240 + 1, 2, 0, 0, 4, 4, 0, 64, 0, 0, 1e-3, 1.0, 0.0, 0.0, 0.1, 0.1, 1.0, 1.0, 0.0)
    assert L.tsmi_energy_device_noalloc(C.byref(wide), buf.data_ptr(), buf.data_ptr(), 4096 * 240 + 1, None,
                                        out.data_ptr(), None) == 1
    assert b"wider" in L.tsmi_last_error()
    assert L.tsmi_energy_device_noalloc(None, None, None, 0, None, None, None) == 1
