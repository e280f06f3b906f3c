#!/usr/bin/env python3
"""Trotter-Suzuki step throughput on one MI355X.

Workload: the 8192 x 8192 real-time GPE vortex lattice (periodic, side 64,
g = 50, harmonic trap of weight 0.01, dt = 5e-4, 16 seeded vortices), built on
the device.  One step = one full second-order Trotter-Suzuki step of the whole
lattice = one launch of the fused step kernel.  The FP32-bulk and the FP64 forms
are timed on the same GPU from the same initial state; the FP32 result after K
steps is compared with the FP64 one (relative L2 over the inner lattice).

Prints one JSON line: value = site-updates/s of the FP32 step (inner sites x
steps / time), with the roofline of the step kernel, the FP64 figures, and the
reference CPU kernel timed on this host's cores on a bounded sample.

roofline.achieved is the HBM bandwidth the dominant kernel would need for the
bytes it must move: per launch, the state read once and written once and E read
once per inner site (FP32 with E as one phase plane: 20 B; FP64: 40 B), whatever
the number of steps the launch advances; frac = achieved / 8 TB/s.  The
single-sweep figure (24 B per site and step, times the steps per launch) is
reported beside it as single_sweep_equivalent_GBs: it exceeds the peak when a
launch advances two steps.

The CPU baseline runs first, before anything in this process touches the GPU
(a process that has initialised the GPU must not start another program).

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "trotter-suzuki-mpi_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0        # MI355X HBM3E
BYTES_PER_SU = {32: 24, 64: 48}   # psi read + psi write + E read (two planes) per inner site and step
# What a launch must move per inner site when E is a pure phase (the case here):
# the state read and written once, E read as one phase plane -- per launch,
# whether the launch advances one step or two
KERNEL_BYTES_PER_SITE = {32: 20, 64: 40}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "step_8192_traffic.json")
# the sources of the step kernels' device code: a traffic record measured on
# other sources is not reported
KERNEL_SOURCES = ("trotter-suzuki-mpi_amd/csrc/step_kernel.hpp", "trotter-suzuki-mpi_amd/csrc/fused_kernel.hpp")
ACCURACY_STEPS = 1000


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=8192, help="lattice side (default: the 8192 throughput config)")
    ap.add_argument("--tile-rows", type=int, default=0, help="rows per wavefront segment (0: library choice)")
    ap.add_argument("--precision", choices=["both", "32", "64"], default="both")
    ap.add_argument("--cpu-steps", type=int, default=12, help="timed steps of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the FP32-vs-FP64 comparison over 1000 steps (when --steps is not 1000)")
    ap.add_argument("--no-api", action="store_true", help="skip timing Solver::evolve through the C++ API")
    ap.add_argument("--no-studies", action="store_true",
                    help="skip the precision study, the batch, adjoint and two-component measurements")
    ap.add_argument("--only-study", choices=["precision_study", "batch", "adjoint", "pair", "observables"],
                    help="only run that study and print its JSON object (for profiling)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy (for counter calibration) and exit")
    ap.add_argument("--stream-ceiling", action="store_true",
                    help="only time plain elementwise kernels over planes of the 8192^2 FP32 tile and exit")
    return ap.parse_args()


def calibrate_copy():
    import torch
    n = 1 << 28   # 2^28 float32 = 1 GiB read + 1 GiB written per launch
    a = torch.rand(n, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    for _ in range(5):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_kernel_bytes_read": 4 * n, "copy_kernel_bytes_written": 4 * n, "launches": 5}))


def stream_ceiling(launches: int = 20):
    """What plain streaming kernels reach on this GPU over planes the size of the
    8192^2 FP32 tile with its 8-wide halo (8256 x 8208 floats): a copy (one
    plane read, one written), an add (two read, one written) and addcmul (three
    read, one written), each timed with HIP events over `launches` launches."""
    import torch
    n = 8256 * 8208
    planes = [torch.rand(n, device="cuda", dtype=torch.float32) for _ in range(3)]
    out = torch.empty_like(planes[0])
    a, b, c = planes
    kernels = {"copy": (lambda: torch.mul(a, 1.0, out=out), 2),
               "add": (lambda: torch.add(a, b, out=out), 3),
               "addcmul": (lambda: torch.addcmul(c, a, b, out=out), 4)}
    res = {"plane_bytes": 4 * n, "launches": launches}
    for name, (fn, nplanes) in kernels.items():
        for _ in range(3):
            fn()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0.record()
        for _ in range(launches):
            fn()
        t1.record()
        torch.cuda.synchronize()
        us = t0.elapsed_time(t1) * 1e3 / launches
        res[name] = {"us": us, "bytes": nplanes * 4 * n, "TB_per_s": nplanes * 4 * n / us / 1e6}
    print(json.dumps({"stream_ceiling": res}), flush=True)


def time_precision(prec, lat, psi_re, psi_im, e_re, e_im, v, params, steps, warmup, tile_rows):
    import torch
    from tsmi.hip import HipKernel
    k = HipKernel(lat, delta_t=params["dt"], mass=params["mass"], coupling=params["g"], precision=prec,
                  tile_rows=tile_rows)
    stream = torch.cuda.current_stream()
    k.set_stream(stream.cuda_stream)
    k.set_exp_potential_device(e_re.data_ptr(), e_im.data_ptr())
    k.set_potential_device(v.data_ptr())
    k.set_state_device(psi_re.data_ptr(), psi_im.data_ptr())
    if warmup:
        k.evolve(warmup)
    k.set_state_device(psi_re.data_ptr(), psi_im.data_ptr())
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    k.evolve(steps)
    t1.record(stream)
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1)
    out_re, out_im = torch.empty_like(psi_re), torch.empty_like(psi_im)
    k.get_state_device(out_re.data_ptr(), out_im.data_ptr())
    torch.cuda.synchronize()
    norm = k.squared_norm()
    geo = k.geometry()
    # the same run again with every step-kernel launch bracketed by HIP events on
    # this stream: the per-launch duration of the dominant kernel
    k.set_state_device(psi_re.data_ptr(), psi_im.data_ptr())
    torch.cuda.synchronize()
    k.kernel_timing(True)
    k.evolve(steps)
    kt = k.kernel_time()
    k.kernel_timing(False)
    k.close()
    return ms, out_re, out_im, norm, geo, kt


def host_cpu() -> dict:
    """This host's CPU: model, logical CPUs, the ones this process may use, and
    whether it has AVX-512 (x86-64-v4)."""
    model, flags = "unknown", set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("flags") and not flags:
                    flags = set(line.split(":", 1)[1].split())
    except OSError:
        pass
    v4 = {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"} <= flags
    return {"model": model, "logical_cpus": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "avx512": v4}


def cpu_baseline(lat, psi, v, params, steps):
    """The reference CPU kernel (oracle/_ref) on this host, on the full lattice:
    its OpenMP CPUBlock built from its sources with its release flags, for
    AVX-512 (x86-64-v4) where the CPU has it, else for x86-64-v3.  Threads: the
    CPUs this process may use, at most 16 (the host share of one GPU)."""
    from oracle import pyoracle
    from tsmi.workloads import Workload
    cpu = host_cpu()
    threads = min(16, cpu["affinity_cpus"])
    wl = Workload("config2", lat, psi, v, params["dt"], steps, g=params["g"], mass=params["mass"])
    sample = f"{lat.dim_x}x{lat.dim_y} periodic GPE (the benchmark lattice and state), 1 warm-up + {steps} timed steps"
    host = {"cpu_model": cpu["model"], "host_logical_cpus": cpu["logical_cpus"],
            "process_cpus": cpu["affinity_cpus"]}
    kind = "fast4" if cpu["avx512"] and pyoracle.ref_available("fast4") else "fast"
    if pyoracle.ref_available(kind):
        r = pyoracle.time_reference(wl, steps=steps, n_warmup=1, kind=kind, threads=threads)
        su = lat.dim_x * lat.dim_y * r["steps"] / r["seconds"]
        march = "x86-64-v4" if kind == "fast4" else "x86-64-v3"
        return {"value": su, "unit": "site-updates/s", "cores": r["threads"], "kind": "reference",
                "sample": f"reference CPUBlock built from its sources (-Ofast -march={march} -fopenmp), " + sample,
                **host}
    # the oracle restatement (OpenMP over rows) when the reference build is absent
    os.environ["OMP_NUM_THREADS"] = str(threads)
    import numpy as np
    d = pyoracle.make_desc(wl, norm_target=1.0)
    er, ei = pyoracle.exp_potential(v, params["dt"], False)
    re, im = np.ascontiguousarray(psi.real), np.ascontiguousarray(psi.imag)
    pyoracle.lib().ts_oracle_evolve(d, 1, pyoracle._p(re), pyoracle._p(im), pyoracle._p(er), pyoracle._p(ei))
    t = time.perf_counter()
    pyoracle.lib().ts_oracle_evolve(d, steps, pyoracle._p(re), pyoracle._p(im), pyoracle._p(er), pyoracle._p(ei))
    sec = time.perf_counter() - t
    return {"value": lat.dim_x * lat.dim_y * steps / sec, "unit": "site-updates/s", "cores": threads, "kind": "port",
            "sample": "oracle restatement (OpenMP over rows), " + sample, **host}


def sources_sha16() -> str:
    import hashlib
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def measured_traffic(prec: int, n: int):
    """HBM bytes per launch of the dominant kernel from the committed counter
    record, if it was measured on these kernel sources at this size."""
    if not os.path.exists(TRAFFIC_FILE):
        return None, "no counter record"
    with open(TRAFFIC_FILE) as f:
        rec = json.load(f)
    if rec.get("source_sha16") != sources_sha16():
        return None, "counter record measured on other kernel sources"
    tr = rec.get(f"fp{prec}")
    if not tr or tr.get("n") != n:
        return None, "no counter record at this size"
    return tr.get("bytes_per_launch"), TRAFFIC_FILE[len(ROOT) + 1:]


def api_leg(lat, psi_h, v_h, params, prec: int, steps: int, warmup: int = 4):
    """Solver::evolve(steps) through the C++ API (Lattice2D, State, Potential,
    Hamiltonian, Solver with kernel type "hip-fp32"/"hip"), in this process."""
    import ctypes as C
    import numpy as np
    from tsmi import _lib
    _lib.load()   # the HIP library first, with PyTorch's runtime (see tsmi._lib)
    path = os.path.join(ROOT, "trotter-suzuki-mpi_amd", "lib", "libtsmi_api_timing.so")
    L = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    L.tsmi_api_time_evolve.argtypes = [C.c_int32, C.c_double, C.c_int32, dp, dp, dp, C.c_double, C.c_double,
                                       C.c_double, C.c_int32, C.c_int32, C.c_int32, dp]
    L.tsmi_api_last_error.restype = C.c_char_p
    re, im = np.ascontiguousarray(psi_h.real), np.ascontiguousarray(psi_h.imag)
    v = np.ascontiguousarray(v_h)
    out = np.zeros(5)
    st = L.tsmi_api_time_evolve(lat.dim_x, lat.length_x, 1, re.ctypes.data_as(dp), im.ctypes.data_as(dp),
                                v.ctypes.data_as(dp), params["dt"], params["mass"], params["g"], prec, warmup, steps,
                                out.ctypes.data_as(dp))
    if st != 0:
        raise RuntimeError(L.tsmi_api_last_error().decode())
    per_step = (out[1] - out[2]) / (steps - 2)
    kind = "hip-fp32" if prec == 32 else "hip"
    return {"route": f"Solver(..., \"{kind}\").evolve(-{steps}) (soft update) + get_squared_norm() through the "
                     f"C++ API, less the same with 2 steps",
            "ms_per_step": per_step * 1e3, "site_updates_per_s": lat.dim_x * lat.dim_y / per_step,
            "evolve_soft_s": out[1], "evolve_soft_2_s": out[2], "kernel_build_s": out[0], "norm": out[3],
            "evolve_with_get_sample_s": out[4],
            "evolve_with_get_sample_note": f"evolve({steps}) as the reference examples call it: the steps plus the "
                                           "copy of the whole state into the State's host buffers (PCIe, FP64)"}


def precision_study(steps: int = 10000):
    """Imaginary-time GPE ground state, 2048^2 closed, g=100, dt=1e-4, FP32 bulk
    against FP64 on the GPU (the FP64 path agrees with the oracle to 1e-12)."""
    import numpy as np
    import torch
    from tsmi import workloads as W
    from tsmi.hip import kernel_for
    wl = W.config1(steps=steps)
    lat = wl.lattice
    nt = W.squared_norm(lat, wl.psi)
    out = {}
    for prec in (64, 32):
        k = kernel_for(wl, precision=prec, norm_target=nt)
        try:
            k.synchronize()
            t = time.perf_counter()
            k.evolve(steps)
            k.synchronize()
            sec = time.perf_counter() - t
            out[prec] = dict(psi=k.sample(), energy=k.energy(), sec=sec)
        finally:
            k.close()
    e64, e32 = out[64]["energy"], out[32]["energy"]
    iy, ix = lat.inner_y, lat.inner_x
    a32, a64 = out[32]["psi"][iy, ix], out[64]["psi"][iy, ix]
    # plain pairwise sums, not BLAS dot: under rocprofv3's preloaded library the
    # BLAS path returned NaN for these finite (partly subnormal) arrays
    rel = float(np.sqrt(np.sum(np.abs(a32 - a64) ** 2) / np.sum(np.abs(a64) ** 2)))
    sites = lat.dim_x * lat.dim_y
    return {"workload": f"imaginary-time GPE ground state, {lat.dim_x}x{lat.dim_y} closed, side 24, g=100, "
                        f"dt=1e-4, {steps} steps",
            "fp64": {"energy": e64[0], "norm": e64[5], "seconds": out[64]["sec"],
                     "site_updates_per_s": sites * steps / out[64]["sec"]},
            "fp32": {"energy": e32[0], "norm": e32[5], "seconds": out[32]["sec"],
                     "site_updates_per_s": sites * steps / out[32]["sec"],
                     "energy_rel_err_vs_fp64": abs(e32[0] - e64[0]) / abs(e64[0]), "rel_l2_vs_fp64": rel}}


def batch_study(steps: int = 500, count: int = 512):
    """512 independent lattices of seeded sides in {64..1024}, mixed closed and
    periodic, advanced together in FP32 (one persistent launch per step,
    graph-captured), after the autotuner."""
    from tsmi import workloads as W
    from tsmi.batch import HipBatch
    t = time.perf_counter()
    wls = W.config3(count=count, steps=steps)
    build = time.perf_counter() - t
    sites = sum(w.lattice.dim_x * w.lattice.dim_y for w in wls)
    with HipBatch(wls, precision=32) as b:
        b.synchronize()
        ms_tuned = b.tune(steps=3)
        b.evolve(steps)          # records the graph
        b.synchronize()
        t = time.perf_counter()
        b.evolve(steps)
        b.synchronize()
        sec = time.perf_counter() - t
        rows = sorted({(b.info(i)["H"], b.info(i)["steps_per_pass"], b.info(i)["rows_per_segment"])
                       for i in range(len(wls))})
    return {"workload": f"{count} real-time lattices, sides 64..1024 (seeded), mixed closed/periodic, FP32, "
                        f"{steps} graph-captured steps", "inner_sites": sites,
            "site_updates_per_s": sites * steps / sec, "ms_per_step": sec * 1e3 / steps,
            "tuned_ms_per_step": ms_tuned, "tile_height_steps_per_pass_rows": rows, "host_build_s": build}


def adjoint_study():
    """dL/dV of L = 1 - |<target|psi(T)>|^2 on the 1024^2 closed lattice, 2000
    real-time steps, checkpoints every 100, FP32 and FP64, checked against the
    reference solver's central differences (tests/golden/adjoint_config4.npz)."""
    import numpy as np
    from tsmi import workloads as W
    from tsmi.hip import kernel_for
    from tsmi.losses import infidelity
    wl, target = W.config4()
    fx = os.path.join(ROOT, "tests", "golden", "adjoint_config4.npz")
    fd = np.load(fx) if os.path.exists(fx) else None
    out = {"workload": "1024x1024 closed, 16 Gaussians, 2000 steps dt=1e-3, checkpoints every 100"}
    grads = {}
    for prec in (64, 32):
        k = kernel_for(wl, precision=prec)
        try:
            k.synchronize()
            t = time.perf_counter()
            k.evolve_checkpointed(wl.steps, 100)
            k.synchronize()
            fwd = time.perf_counter() - t
            L, g = infidelity(wl.lattice, target, k.sample())
            t = time.perf_counter()
            gv, _ = k.adjoint(g)
            rev = time.perf_counter() - t
        finally:
            k.close()
        grads[prec] = gv
        o = {"loss": L, "forward_s": fwd, "reverse_s": rev}
        if fd is not None:
            pts = fd["points"]
            ref = fd["fd"]
            o["max_rel_err_vs_reference_fd"] = float(np.max(np.abs(gv.ravel()[pts] - ref)) / np.max(np.abs(ref)))
        out[f"fp{prec}"] = o
    out["fp32_rel_err_vs_fp64"] = float(np.max(np.abs(grads[32] - grads[64])) / np.max(np.abs(grads[64])))
    return out


def observables_study(n: int = 8192, calls: int = 6):
    """The read-outs (tsmi_energy, tsmi_expected_values, tsmi_squared_norm) on
    the vortex lattice in both precisions: host-clock time per call (device
    passes, the copy of the results and the synchronisation)."""
    import torch
    from tsmi import workloads_torch as WT
    from tsmi.hip import HipKernel
    lat, psi, v, params = WT.config2(n=n, device="cuda")
    psi_re, psi_im = psi.real.contiguous(), psi.imag.contiguous()
    e_re, e_im = WT.exp_potential(v, params["dt"], False)
    out = {"lattice": [lat.dim_x, lat.dim_y], "calls": calls}
    for prec in (32, 64):
        k = HipKernel(lat, delta_t=params["dt"], mass=params["mass"], coupling=params["g"], precision=prec)
        k.set_exp_potential_device(e_re.data_ptr(), e_im.data_ptr())
        k.set_potential_device(v.data_ptr())
        k.set_state_device(psi_re.data_ptr(), psi_im.data_ptr())
        torch.cuda.synchronize()
        for name, fn in (("energy", k.energy), ("expected_values", k.expected_values),
                         ("squared_norm", k.squared_norm)):
            fn()
            t = time.perf_counter()
            for _ in range(calls):
                fn()
            out[f"fp{prec}_{name}_ms_per_call"] = (time.perf_counter() - t) / calls * 1e3
        k.close()
    return out


def pair_study(n: int = 4096, steps: int = 200):
    """Two-component real-time GPE with Rabi coupling on an n x n periodic
    lattice (two seeded Thomas-Fermi clouds, g_a = g_b = 50, g_ab = 40,
    omega = 1).  FP32: two steps of both components per pass (pair2_kernel, the
    Rabi rotation applied between and after them; the first and last step of
    the call run alone), also timed at one step per pass; FP64: one pair_kernel
    pass per step.  Each pass is followed by the halo exchange."""
    import numpy as np
    from tsmi import workloads as W
    from tsmi.lattice import Lattice2D
    from tsmi.pair import HipPair
    lat = Lattice2D(n, 64.0, periodic_x=True, periodic_y=True)
    X, Y = lat.mesh()
    v_a = 0.5 * 0.01 * (X * X + Y * Y)
    v_b = 0.5 * 0.012 * (X * X + Y * Y)
    amp = np.sqrt(np.maximum(10.0 - v_a, 0.0) / 50.0)
    psi_a = (amp * np.exp(1j * 0.3 * X)).astype(np.complex128)
    psi_b = (0.5 * amp * np.exp(-1j * 0.2 * Y)).astype(np.complex128)
    pw = W.PairWorkload(f"pair_{n}", lat, psi_a, psi_b, v_a, v_b, 5e-4, (steps,), g_a=50.0, g_b=50.0,
                        g_ab=40.0, omega_r=1.0)
    sites = n * n
    out = {"workload": f"two-component real-time GPE with Rabi coupling, {n}x{n} periodic, g_a = g_b = 50, "
                       f"g_ab = 40, omega = 1, dt = 5e-4, {steps} steps", "inner_sites": sites}
    for key, prec, sweeps in (("fp32", 32, 0), ("fp32_one_step_per_pass", 32, 1), ("fp64", 64, 0)):
        with HipPair(pw, precision=prec, sweeps=sweeps) as h:
            h.evolve(4)
            h.synchronize()
            t = time.perf_counter()
            h.evolve(steps)
            h.synchronize()
            sec = time.perf_counter() - t
            e = h.energy()
            geo = h.geometry()
        alg = 2 * BYTES_PER_SU[prec] * sites * steps
        out[key] = {"ms_per_step": sec * 1e3 / steps,
                            "component_site_updates_per_s": 2 * sites * steps / sec,
                            "algorithmic_GB_per_s": alg / sec / 1e9, "norm": e[11], "geometry": geo}
    return out


def evolve_precision(prec, lat, psi_re, psi_im, e_re, e_im, params, steps):
    """steps steps from the initial state; returns the final planes and the
    squared norms before and after (FP64 sums on the device)."""
    import torch
    from tsmi.hip import HipKernel
    k = HipKernel(lat, delta_t=params["dt"], mass=params["mass"], coupling=params["g"], precision=prec)
    try:
        k.set_stream(torch.cuda.current_stream().cuda_stream)
        k.set_exp_potential_device(e_re.data_ptr(), e_im.data_ptr())
        k.set_state_device(psi_re.data_ptr(), psi_im.data_ptr())
        n0 = k.squared_norm()
        k.evolve(steps)
        out_re, out_im = torch.empty_like(psi_re), torch.empty_like(psi_im)
        k.get_state_device(out_re.data_ptr(), out_im.data_ptr())
        torch.cuda.synchronize()
        return out_re, out_im, n0, k.squared_norm()
    finally:
        k.close()


def rel_l2_inner(lat, a, b):
    import torch
    iy, ix = lat.inner_y, lat.inner_x
    x = torch.complex(*[t[iy, ix] for t in a])
    y = torch.complex(*[t[iy, ix] for t in b])
    return float(torch.linalg.vector_norm(x - y) / torch.linalg.vector_norm(y))


def roofline_of(r, prec, sites, n):
    """The dominant kernel's roofline: the bytes a launch must move per inner
    site (state in, state out, E as one phase plane) over its mean duration."""
    t = r["kernel_us"] * 1e-6
    need = KERNEL_BYTES_PER_SITE[prec] * sites
    achieved = need / t / 1e9
    traffic, src = measured_traffic(prec, n)
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": r["kernel"], "kernel_us_per_launch": r["kernel_us"], "launches": r["launches"],
            "steps_per_launch": r["sweeps"], "bytes_per_site_per_launch": KERNEL_BYTES_PER_SITE[prec],
            "bytes_per_launch": need, "e_form": "phase plane (E of a real potential as one plane)",
            "traffic_source": src,
            "traffic_over_required": traffic / need if traffic else None,
            "traffic_rate_GBs": traffic / t / 1e9 if traffic else None,
            "single_sweep_equivalent_GBs": r["sweeps"] * BYTES_PER_SU[prec] * sites / t / 1e9}


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build runs the step on a single MI355X "
              "(multi-GPU sharding is a design note, not part of this build); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    if args.calibrate_copy or args.stream_ceiling or args.only_study:
        if not torch.cuda.is_available():
            print("bench.py needs a GPU", file=sys.stderr)
            sys.exit(2)
        if args.calibrate_copy:
            calibrate_copy()
            return
        if args.stream_ceiling:
            stream_ceiling()
            return
        fn = {"precision_study": precision_study, "batch": batch_study, "adjoint": adjoint_study,
              "pair": pair_study, "observables": observables_study}[args.only_study]
        print(json.dumps({args.only_study: fn()}), flush=True)
        return
    from tsmi import workloads_torch as WT

    # The CPU baseline first: it starts the reference program, which this process
    # may do only before it has touched the GPU.
    cpu = None
    psi_h = v_h = None
    if not args.no_cpu_baseline:
        log(f"building the {args.n}^2 vortex lattice on the host")
        lat, psi_c, v_c, params = WT.config2(n=args.n, device="cpu")
        psi_h, v_h = psi_c.numpy(), v_c.numpy()
        del psi_c, v_c
        log(f"CPU baseline: reference kernel, {args.cpu_steps} steps of the full lattice")
        cpu = cpu_baseline(lat, psi_h, v_h, params, args.cpu_steps)
        log(f"CPU baseline: {cpu['value']:.4g} SU/s on {cpu['cores']} threads ({cpu['kind']}, {cpu['cpu_model']}, "
            f"{cpu['host_logical_cpus']} logical CPUs)")

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(2)
    torch.cuda.set_device(0)
    if psi_h is not None:
        psi = torch.from_numpy(psi_h).to("cuda")
        v = torch.from_numpy(v_h).to("cuda")
    else:
        log(f"building the {args.n}^2 vortex lattice on the device")
        lat, psi, v, params = WT.config2(n=args.n, device="cuda")
    psi_re, psi_im = psi.real.contiguous(), psi.imag.contiguous()
    e_re, e_im = WT.exp_potential(v, params["dt"], False)
    sites = lat.dim_x * lat.dim_y

    results = {}
    precs = [32, 64] if args.precision == "both" else [int(args.precision)]
    for prec in precs:
        log(f"FP{prec}: {args.warmup} warm-up + {args.steps} timed steps")
        ms, o_re, o_im, norm, geo, kt = time_precision(prec, lat, psi_re, psi_im, e_re, e_im, v, params,
                                                       args.steps, args.warmup, args.tile_rows)
        per_step = ms / args.steps
        su = sites * args.steps / (ms * 1e-3)
        # dominant kernel: the fused two-step pass when the run used it
        fused = kt["fused_launches"] > 0
        launches = kt["fused_launches"] if fused else kt["single_launches"]
        kms = (kt["fused_ms"] if fused else kt["single_ms"]) / max(1, launches)
        results[prec] = dict(ms=ms, per_step=per_step, su=su, norm=norm, geo=geo, out=(o_re, o_im),
                             kernel="step2_kernel" if fused else "step_kernel", kernel_us=kms * 1e3,
                             launches=launches, sweeps=2 if fused else 1, kt=kt)
        rf = roofline_of(results[prec], prec, sites, args.n)
        results[prec]["roofline"] = rf
        log(f"FP{prec}: {per_step * 1e3:.1f} us/step, {su:.4g} SU/s; {rf['kernel']} {kms * 1e3:.1f} us/launch x "
            f"{launches} ({rf['steps_per_launch']} steps each): {rf['achieved']:.0f} GB/s of the bytes a launch "
            f"must move = {rf['frac']:.3f} of HBM peak; norm {norm:.15f}")

    accuracy = None
    if 32 in results and 64 in results and not args.no_accuracy:
        if args.steps == ACCURACY_STEPS:
            outs = {p: (results[p]["out"], None, results[p]["norm"]) for p in (32, 64)}
            n0 = {}
        else:
            log(f"accuracy: FP32 and FP64, {ACCURACY_STEPS} steps each from the same state")
            outs, n0 = {}, {}
            for p in (32, 64):
                o_re, o_im, a, b = evolve_precision(p, lat, psi_re, psi_im, e_re, e_im, params, ACCURACY_STEPS)
                outs[p] = ((o_re, o_im), a, b)
                n0[p] = a
        rel = rel_l2_inner(lat, outs[32][0], outs[64][0])
        accuracy = {"steps": ACCURACY_STEPS, "fp32_rel_l2_vs_fp64": rel,
                    "norm_drift_fp32": abs(outs[32][2] - (n0.get(32) or 1.0)),
                    "norm_drift_fp64": abs(outs[64][2] - (n0.get(64) or 1.0))}
        log(f"FP32 vs FP64 after {ACCURACY_STEPS} steps: relative L2 {rel:.3e}; norm drift FP32 "
            f"{accuracy['norm_drift_fp32']:.2e}, FP64 {accuracy['norm_drift_fp64']:.2e}")

    api = None
    if not args.no_api and 32 in results:
        log("C++ API: Solver::evolve through HipKernel (hip-fp32)")
        if psi_h is None:
            psi_h, v_h = psi.cpu().numpy(), v.cpu().numpy()
        torch.cuda.synchronize()
        api = api_leg(lat, psi_h, v_h, params, 32, max(args.steps, 100))
        log(f"C++ API: {api['ms_per_step'] * 1e3:.1f} us/step ({api['site_updates_per_s']:.4g} SU/s); "
            f"evolve({max(args.steps, 100)}) with the copy to the host {api['evolve_with_get_sample_s']:.3f} s")

    main_prec = precs[0]
    r = results[main_prec]
    line = {
        "metric": "site-updates/s per Trotter step & % HBM peak; FP32 vs FP64 at fixed L2 error",
        "value": r["su"],
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"f{main_prec}",
        "data": "synthetic (seeded 16-vortex Thomas-Fermi state)",
        "config": {"workload": f"real-time GPE vortex dynamics, {lat.dim_x}x{lat.dim_y} periodic, side 64, "
                               "g=50, trap 0.01, dt=5e-4",
                   "lattice": [lat.dim_x, lat.dim_y], "tile": [lat.W, lat.H], "geometry": r["geo"],
                   "parallelism": "single GPU"},
        "roofline": r["roofline"],
        "cpu_baseline": cpu,
    }
    if 64 in results and main_prec != 64:
        r64 = results[64]
        line["fp64"] = {"value": r64["su"], "ms_per_step": r64["per_step"], "roofline": r64["roofline"]}
        line["fp32_over_fp64"] = r["su"] / r64["su"]
    if accuracy:
        line["fp32_rel_l2_vs_fp64"] = accuracy["fp32_rel_l2_vs_fp64"]
        line["accuracy"] = accuracy
    if api:
        line["api"] = api
    if not args.no_studies:
        for name, fn in (("precision_study", precision_study), ("batch", batch_study), ("adjoint", adjoint_study),
                         ("pair", pair_study)):
            log(f"{name} ...")
            t = time.perf_counter()
            line[name] = fn()
            log(f"{name}: done in {time.perf_counter() - t:.1f} s")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
