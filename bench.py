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

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "trotter-suzuki-mpi_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0        # MI355X HBM3E
BYTES_PER_SU = {32: 24, 64: 48}   # psi read + psi write + E read per inner site (real time)
# What the kernel reads and writes per inner site and step when E is a pure phase
# (the case here): E as one phase plane instead of two planes
KERNEL_BYTES_PER_SU = {32: 20, 64: 40}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "step_8192_traffic.json")


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
    ap.add_argument("--no-studies", action="store_true",
                    help="skip the precision study, the batch, adjoint and two-component measurements")
    ap.add_argument("--only-study", choices=["precision_study", "batch", "adjoint", "pair"],
                    help="only run that study and print its JSON object (for profiling)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy (for counter calibration) and exit")
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


def cpu_baseline(lat, psi, v, params, steps):
    """The reference CPU kernel (oracle/_ref) on this host, on the full lattice."""
    from oracle import pyoracle
    from tsmi.workloads import Workload
    threads = min(16, len(os.sched_getaffinity(0)))
    wl = Workload("config2", lat, psi, v, params["dt"], steps, g=params["g"], mass=params["mass"])
    sample = f"{lat.dim_x}x{lat.dim_y} periodic GPE (the benchmark lattice), 1 warm-up + {steps} timed steps"
    if pyoracle.ref_available("fast"):
        r = pyoracle.time_reference(wl, steps=steps, n_warmup=1, kind="fast", threads=threads)
        su = lat.dim_x * lat.dim_y * r["steps"] / r["seconds"]
        return {"value": su, "unit": "site-updates/s", "cores": r["threads"], "kind": "reference",
                "sample": "reference CPUBlock built from its sources (-Ofast -march=x86-64-v3 -fopenmp), " + sample}
    # the oracle restatement (OpenMP over rows) when the reference build is absent
    os.environ["OMP_NUM_THREADS"] = str(threads)
    import numpy as np
    d = pyoracle.make_desc(wl, norm_target=1.0)
    er, ei = pyoracle.exp_potential(v, params["dt"], False)
    re, im = np.ascontiguousarray(psi.real), np.ascontiguousarray(psi.imag)
    pyoracle.lib().ts_oracle_evolve(d, 1, pyoracle._p(re), pyoracle._p(im), pyoracle._p(er), pyoracle._p(ei))
    t = time.perf_counter()
    pyoracle.lib().ts_oracle_evolve(d, steps, pyoracle._p(re), pyoracle._p(im), pyoracle._p(er), pyoracle._p(ei))
    s = time.perf_counter() - t
    return {"value": lat.dim_x * lat.dim_y * steps / s, "unit": "site-updates/s", "cores": threads, "kind": "port",
            "sample": "oracle restatement (OpenMP over rows), " + sample}


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


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build runs the step on a single MI355X "
              "(multi-GPU sharding is a design note, not part of this build); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.only_study:
        fn = {"precision_study": precision_study, "batch": batch_study, "adjoint": adjoint_study,
              "pair": pair_study}[args.only_study]
        print(json.dumps({args.only_study: fn()}), flush=True)
        return
    from tsmi import workloads_torch as WT

    torch.cuda.set_device(0)
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
        sweeps = 2 if fused else 1
        ach = sweeps * BYTES_PER_SU[prec] * sites / (kms * 1e-3) / 1e9
        results[prec] = dict(ms=ms, per_step=per_step, su=su, achieved=ach, norm=norm, geo=geo, out=(o_re, o_im),
                             kernel="step2_kernel" if fused else "step_kernel", kernel_us=kms * 1e3,
                             launches=launches, sweeps=sweeps, kt=kt)
        log(f"FP{prec}: {per_step * 1e3:.1f} us/step, {su:.4g} SU/s; {results[prec]['kernel']} "
            f"{kms * 1e3:.1f} us/launch x {launches} ({sweeps} steps each): {ach:.0f} GB/s algorithmic; "
            f"norm {norm:.15f}")

    rel = None
    if 32 in results and 64 in results:
        iy, ix = lat.inner_y, lat.inner_x
        a32 = torch.complex(*[t[iy, ix] for t in results[32]["out"]])
        a64 = torch.complex(*[t[iy, ix] for t in results[64]["out"]])
        rel = float(torch.linalg.vector_norm(a32 - a64) / torch.linalg.vector_norm(a64))
        log(f"FP32 vs FP64 after {args.steps} steps: relative L2 {rel:.3e}")

    main_prec = precs[0]
    r = results[main_prec]
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f).get(f"fp{main_prec}")
        if tr and tr.get("n") == args.n:
            traffic = tr.get("bytes_per_launch")

    cpu = None
    if not args.no_cpu_baseline:
        log(f"CPU baseline: reference kernel, {args.cpu_steps} steps of the full lattice")
        psi_h = psi.cpu().numpy()
        v_h = v.cpu().numpy()
        cpu = cpu_baseline(lat, psi_h, v_h, params, args.cpu_steps)
        log(f"CPU baseline: {cpu['value']:.4g} SU/s on {cpu['cores']} threads ({cpu['kind']})")

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
        "data": "synthetic (seeded 16-vortex Thomas-Fermi state built on the device)",
        "config": {"workload": f"real-time GPE vortex dynamics, {lat.dim_x}x{lat.dim_y} periodic, side 64, "
                               "g=50, trap 0.01, dt=5e-4",
                   "lattice": [lat.dim_x, lat.dim_y], "tile": [lat.W, lat.H], "geometry": r["geo"],
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": r["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": r["achieved"] / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": r["kernel"], "kernel_us_per_launch": r["kernel_us"], "launches": r["launches"],
                     "bytes_per_site_update": BYTES_PER_SU[main_prec], "sweeps_per_launch": r["sweeps"],
                     "algorithmic_bytes_per_launch": r["sweeps"] * BYTES_PER_SU[main_prec] * sites,
                     "e_form": "phase plane (one plane of E's two)",
                     "kernel_bytes_per_site_update": KERNEL_BYTES_PER_SU[main_prec],
                     "kernel_bytes_achieved": r["achieved"] * KERNEL_BYTES_PER_SU[main_prec] / BYTES_PER_SU[main_prec]},
        "cpu_baseline": cpu,
    }
    if 64 in results and main_prec != 64:
        r64 = results[64]
        line["fp64"] = {"value": r64["su"], "ms_per_step": r64["per_step"],
                        "roofline": {"bound": "hbm", "achieved": r64["achieved"], "peak": HBM_PEAK_GBS,
                                     "unit": "GB/s", "frac": r64["achieved"] / HBM_PEAK_GBS,
                                     "kernel": r64["kernel"], "kernel_us_per_launch": r64["kernel_us"],
                                     "bytes_per_site_update": BYTES_PER_SU[64], "sweeps_per_launch": r64["sweeps"]}}
        line["fp32_over_fp64"] = r["su"] / r64["su"]
        line["fp32_rel_l2_vs_fp64"] = rel
        line["norm_drift"] = {"fp32": abs(results[32]["norm"] - 1.0), "fp64": abs(r64["norm"] - 1.0)}
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
