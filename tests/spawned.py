"""The C++ programs the tests start, all run once when the test session starts.

A process that has initialised the GPU must not start another program: on the
MI355X hosts that spawn ended the pytest process itself (a segmentation fault
inside the fork + exec of the child) after GPU tests had run in it.  So every
program the test suite needs -- tests/cpp/test_solver (this project's C++ API
on the HIP kernel types), oracle/_ref/dropin_check (HipKernel inside the
reference library) and the reference driver's text I/O -- is started here,
from pytest_sessionstart, before any test runs and before anything in the
pytest process touches the GPU; the tests then read the recorded output.

Programs that need a GPU are started only when the session may run GPU tests
(the marker expression does not exclude them) and the host has a GPU device
node; on other hosts their tests are deselected or skipped anyway.
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
TEST_SOLVER = os.path.join(ROOT, "trotter-suzuki-mpi_amd", "lib", "test_solver")
DROPIN = os.path.join(ROOT, "oracle", "_ref", "dropin_check")

# single-component cases of HipKernel inside the reference library.  calls:
# one Solver::evolve call per entry; update_at: an update_potential call before
# that step (the time-dependent branch of Solver::evolve) with the same table.
DROPIN_CASES = [
    dict(n_x=64, n_y=64, periodic_x=False, periodic_y=False, steps=40, g=5.0, g_lhy=1.0),
    dict(n_x=96, n_y=64, periodic_x=True, periodic_y=True, steps=30, g=2.0),
    dict(n_x=64, n_y=48, periodic_x=True, periodic_y=False, steps=30, g=3.0),
    dict(n_x=64, n_y=64, periodic_x=False, periodic_y=False, imag_time=True, steps=50, g=10.0),
    dict(n_x=300, n_y=140, periodic_x=False, periodic_y=True, steps=12, g=1.0),
    # several calls and a mid-run potential update on the lattices that take two
    # steps per pass
    dict(n_x=300, n_y=200, periodic_x=True, periodic_y=True, steps=24, g=2.0, calls=(11, 7, 6), update_at=15),
    dict(n_x=260, n_y=136, periodic_x=False, periodic_y=False, steps=21, g=4.0, calls=(9, 12), update_at=4),
]

PAIR_DROPIN = [
    dict(n_x=40, n_y=36, periodic_x=False, periodic_y=False, omega_r=2.0, omega_i=0.7, mass_b=1.7, calls=[7, 5]),
    dict(n_x=96, n_y=64, periodic_x=True, periodic_y=True, omega_r=1.0, calls=[9]),
    dict(n_x=64, n_y=48, periodic_x=True, periodic_y=False, calls=[6, 1]),
    dict(n_x=40, n_y=36, periodic_x=False, periodic_y=False, imag_time=True, omega_r=2.0, omega_i=0.7, calls=[7, 5]),
    dict(n_x=300, n_y=140, periodic_x=False, periodic_y=True, omega_r=0.5, calls=[8]),
]

IO_CASES = [
    dict(n_x=24, n_y=16, periodic_x=True, periodic_y=True),
    dict(n_x=20, n_y=12, periodic_x=False, periodic_y=False),
    dict(n_x=18, n_y=14, periodic_x=True, periodic_y=False),
    dict(n_x=16, n_y=18, periodic_x=False, periodic_y=True),
]


def case_id(c: dict) -> str:
    return "_".join(f"{k}{v}" for k, v in c.items())


def dropin_key(case: dict, precision: int) -> str:
    return f"dropin-{precision}-{case_id(case)}"


def pair_dropin_key(case: dict) -> str:
    return "pair-dropin-" + case_id(case)


def io_key(case: dict) -> str:
    return "io-" + case_id(case)


_results: dict = {}
_workdir: str | None = None
_started = False


def _log(msg: str) -> None:
    print(f"[spawned] {msg}", file=sys.stderr, flush=True)


def _run(key: str, args: list, timeout: int = 300) -> None:
    t = time.perf_counter()
    try:
        out = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
        _results[key] = dict(returncode=out.returncode, stdout=out.stdout, stderr=out.stderr)
    except subprocess.TimeoutExpired as e:
        _results[key] = dict(returncode=-1, stdout=e.stdout or "", stderr=f"timed out after {timeout} s")
    _log(f"{key}: rc {_results[key]['returncode']} in {time.perf_counter() - t:.1f} s")


def rows(key: str) -> list:
    """The JSON lines a program printed (asserts that it ran and succeeded)."""
    r = _results.get(key)
    assert r is not None, f"{key} was not run at session start"
    assert r["returncode"] == 0, r["stderr"][-2000:]
    return [json.loads(l) for l in r["stdout"].splitlines() if l.startswith("{")]


def path(*parts: str) -> str:
    assert _workdir is not None
    return os.path.join(_workdir, *parts)


def has(key: str) -> bool:
    return key in _results


def start(gpu: bool) -> None:
    """Runs every program the selected tests may need (pytest_sessionstart)."""
    global _workdir, _started
    if _started:
        return
    _started = True
    sys.path[:0] = [p for p in (ROOT, os.path.join(ROOT, "trotter-suzuki-mpi_amd")) if p not in sys.path]
    from oracle import pyoracle
    from tsmi import workloads as W

    _workdir = tempfile.mkdtemp(prefix="tsmi-spawned-")
    if os.path.exists(TEST_SOLVER):
        _run("errors", [TEST_SOLVER, "hip", "errors"], timeout=120)
        if pyoracle.ref_available("strict"):
            for case in IO_CASES:
                wl = W.small_case(steps=0, g=1.0, **case)
                d = path(io_key(case))
                os.makedirs(d)
                cpath = os.path.join(d, "case.bin")
                pyoracle.write_case(cpath, wl, 0)
                _run(io_key(case) + "-ours", [TEST_SOLVER, "io", cpath, os.path.join(d, "ours")], timeout=60)
                _run(io_key(case) + "-ref", [pyoracle.ref_driver("strict"), cpath, "--io", os.path.join(d, "ref")],
                     timeout=60)
    if not gpu:
        return
    _log("starting the GPU programs before any test runs")
    if os.path.exists(TEST_SOLVER):
        for kernel in ("hip", "hip-fp32"):
            _run(f"scenarios-{kernel}", [TEST_SOLVER, kernel])
            for name in ("mixed_BEC", "imaginary_mixed_BEC"):
                _run(f"pair-scenario-{kernel}-{name}", [TEST_SOLVER, kernel, name])
    if os.access(DROPIN, os.X_OK):
        for case in DROPIN_CASES:
            c = dict(case)
            calls = c.pop("calls", None)
            update_at = c.pop("update_at", None)
            wl = W.small_case(**c)
            cpath = path(dropin_key(case, 64) + ".bin")
            pyoracle.write_case(cpath, wl, wl.steps)
            for prec in (64, 32):
                args = [DROPIN, cpath, str(prec)]
                if calls:
                    args.append(",".join(str(n) for n in calls))
                    if update_at is not None:
                        args.append(str(update_at))
                _run(dropin_key(case, prec), args)
        for case in PAIR_DROPIN:
            pw = W.pair_case(**case, g_a=3.0, g_b=2.0, g_ab=1.5)
            cpath = path(pair_dropin_key(case) + ".bin")
            pyoracle.write_case_pair(cpath, pw)
            _run(pair_dropin_key(case), [DROPIN, cpath, "64"])


def finish() -> None:
    if _workdir and os.path.isdir(_workdir):
        shutil.rmtree(_workdir, ignore_errors=True)
