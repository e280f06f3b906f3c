"""The batched step (tsmi_batch_*): every lattice of a batch gets exactly the
single-lattice evolution.  FP64 against the oracle within 1e-12 relative L2
over the whole tile; FP32 against the single-context FP32 kernel within 1e-6;
graph replays, uneven sizes, mixed boundaries, imaginary time, the autotuner."""
import numpy as np
import pytest

from tsmi import workloads as W

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


MIXED = [
    dict(n_x=64, n_y=64, periodic_x=False, periodic_y=False, g=2.0),
    dict(n_x=96, n_y=80, periodic_x=True, periodic_y=True, g=1.0),
    dict(n_x=33, n_y=17, periodic_x=False, periodic_y=False),
    dict(n_x=300, n_y=40, periodic_x=True, periodic_y=False, g=0.5),
    dict(n_x=40, n_y=130, periodic_x=False, periodic_y=True, g=0.5),
    dict(n_x=8, n_y=8, periodic_x=True, periodic_y=True),
    dict(n_x=37, n_y=1, periodic_x=False, periodic_y=False),
    dict(n_x=520, n_y=70, periodic_x=False, periodic_y=False, g=3.0),
]


def workloads(specs, **kw):
    out = []
    for i, s in enumerate(specs):
        out.append(W.small_case(s["n_x"], s["n_y"], s["periodic_x"], s["periodic_y"], g=s.get("g", 0.0),
                                seed=W.SEED_BASE + 200 + i, **kw))
    return out


def oracle_ref(oracle, wl, calls):
    psi = wl.psi
    for n in calls:
        psi = oracle.evolve(wl, steps=n, psi=psi)
    return psi


def oracle_e(oracle, wls):
    return [oracle.exp_potential(w.v, w.dt, w.imag_time) for w in wls]


@pytest.mark.parametrize("calls", [[9], [7, 5], [4, 4, 4]], ids=lambda c: "calls" + "_".join(map(str, c)))
def test_batch_fp64_matches_oracle(oracle, calls):
    from tsmi.batch import HipBatch
    wls = workloads(MIXED)
    with HipBatch(wls, precision=64, e_planes=oracle_e(oracle, wls)) as b:
        for n in calls:
            b.evolve(n)
        for i, wl in enumerate(wls):
            assert rel_l2(b.state(i), oracle_ref(oracle, wl, calls)) < 1e-12, i


def test_batch_imaginary_time_matches_oracle(oracle):
    from tsmi.batch import HipBatch
    wls = workloads(MIXED[:5], imag_time=True)
    nts = [oracle.make_desc(w).norm_target for w in wls]
    with HipBatch(wls, precision=64, norm_targets=nts, e_planes=oracle_e(oracle, wls)) as b:
        b.evolve(6)
        b.evolve(5)
        for i, wl in enumerate(wls):
            assert rel_l2(b.state(i), oracle_ref(oracle, wl, [6, 5])) < 1e-12, i


def test_batch_fp32_matches_single_context(oracle):
    from tsmi.batch import HipBatch
    from tsmi.hip import kernel_for
    wls = workloads(MIXED)
    es = oracle_e(oracle, wls)
    with HipBatch(wls, precision=32, tile_rows=16, e_planes=es) as b:
        b.evolve(10)
        for i, wl in enumerate(wls):
            k = kernel_for(wl, precision=32, tile_rows=16, e_planes=es[i])
            try:
                k.evolve(10)
                single = k.sample()
            finally:
                k.close()
            assert rel_l2(b.state(i), single) < 1e-6, i
            assert rel_l2(b.state(i), oracle_ref(oracle, wl, [10])) < 2e-5, i


def test_batch_tune_keeps_states(oracle):
    """The autotuner (segment heights and, where free, one or two steps per pass)
    leaves every state where it was: the evolution continues exactly."""
    from tsmi.batch import HipBatch
    wls = workloads(MIXED)
    with HipBatch(wls, precision=64, e_planes=oracle_e(oracle, wls)) as b:
        b.evolve(3)
        ms = b.tune(steps=2)
        assert ms > 0
        info = [b.info(i) for i in range(len(wls))]
        assert all(r["rows_per_segment"] >= 2 and r["rows_per_segment"] % 2 == 0 for r in info)
        assert all(r["steps_per_pass"] in (1, 2) for r in info)
        b.evolve(4)
        for i, wl in enumerate(wls):
            assert rel_l2(b.state(i), oracle_ref(oracle, wl, [3, 4])) < 1e-12, i


def test_batch_rejects_mixed_precision():
    from tsmi import _lib
    from tsmi.batch import desc_of
    import ctypes as C
    L = _lib.load()
    wls = workloads(MIXED[:2])
    descs = (_lib.Desc * 2)(desc_of(wls[0], 32), desc_of(wls[1], 64))
    h = C.c_void_p()
    assert L.tsmi_batch_create(descs, 2, C.byref(h)) == 4
    assert b"same precision" in L.tsmi_last_error()


def test_batch_non_unit_e_keeps_e_form(oracle):
    """One lattice with an E table that is not a pure phase puts the whole batch
    on the E form; every lattice still matches the oracle."""
    import ctypes as C
    from tsmi.batch import HipBatch
    wls = workloads(MIXED[:4])
    es = oracle_e(oracle, wls)
    er, ei = es[1]
    damp = 0.999 - 0.01 * np.linspace(0.0, 1.0, er.shape[1])[None, :]
    es[1] = (er * damp, ei * damp)
    refs = []
    for wl, (a, b) in zip(wls, es):
        d = oracle.make_desc(wl)
        re = np.ascontiguousarray(wl.psi.real).copy()
        im = np.ascontiguousarray(wl.psi.imag).copy()
        oracle.lib().ts_oracle_evolve(C.byref(d), 9, oracle._p(re), oracle._p(im),
                                      oracle._p(np.ascontiguousarray(a)), oracle._p(np.ascontiguousarray(b)))
        refs.append(re + 1j * im)
    with HipBatch(wls, precision=64, e_planes=es) as b:
        b.evolve(9)
        for i in range(len(wls)):
            assert rel_l2(b.state(i), refs[i]) < 1e-12, i
