#!/usr/bin/env python3
"""Benchmark: cell-updates/s of the MEUMAPPS-SS semi-implicit spectral step.

Workload: 256^3 periodic grid, one composition c + 3 tetragonal variants
(eigenstrain diag(0.02,-0.01,-0.01) permuted per variant), cubic
C11/C12/C44 = 250/150/100, dt = 0.005, initial fields c = 0.4 + U[-0.01,0.01],
eta_v = U[-0.01,0.01] from ran_2 with seed 7.  One "step" is one full
semi-implicit step (gradients, KKS chemistry, elastic fixed-point solve,
elastic energy, Cahn-Hilliard and Allen-Cahn updates, clamp and
diagnostics) on inputs resident in HBM.  The elastep0 step (zeroth-order
elastic solve) runs before the timed region.

Reported: FP64 cell-updates/s as `value` (timed steps with no event records
between the kernels), the mixed FP32/FP64 rate and its relative L2 error
against FP64 on the same steps, the live HIP-event timing of the dominant
kernel against the HBM roofline (the same number of steps run again with that
kernel's launches bracketed by events on the step's stream), and the CPU
oracle timed on this host's cores on a bounded sample.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


def step_bytes_per_cell(V, E, n, w):
    """Algorithmic bytes per cell-update: one read + write per minimal 3D FFT
    and one traversal per fused pointwise pass (N_FFT = 7V+5E+15n,
    N_pw = 8V+5E+6+n(V+18))."""
    nfft = 7 * V + 5 * E + 15 * n
    npw = 8 * V + 5 * E + 6 + n * (V + 18)
    return w * (2 * nfft + npw)


# kernel classes (mm_probe kinds); the one named by --probe is timed inside the
# timed region, the others one after another after it
KINDS = {"x_r2c": 0, "y": 1, "z": 2, "x_c2r": 3, "chem": 4, "polar": 5, "elener": 6, "ks_elastic": 7,
         "zf_elastic": 8, "zf_other": 9, "x_polar": 10}
PASS_KINDS = ("x_r2c", "y", "x_c2r", "chem", "elener", "zf_elastic", "zf_other", "x_polar")
# kernel symbol of each headline class (for the rocprofv3 counter runs)
KERNEL_NAME = {"x_polar": "k_x_polar", "zf_elastic": "k_z_fused<OpElastic>"}


def host_cpu():
    """CPU model, logical CPUs this process may run on and the physical cores
    among them (distinct (package, core) pairs)."""
    cpus = sorted(os.sched_getaffinity(0))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    cores = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            cores.add((open(base + "physical_package_id").read().strip(), open(base + "core_id").read().strip()))
        except OSError:
            cores.add(("?", str(c)))
    return {"model": model, "logical_cpus": len(cpus), "physical_cores": len(cores)}


def run_gpu(params, phi, cons, precision, steps, warmup, probe_kind, graph=False, pass_steps=0):
    import numpy as np
    import torch

    from meumapps_ss_amd import _lib
    from meumapps_ss_amd.sim import Simulation

    lib = _lib.load()
    shape = phi.shape[1:]
    sim = Simulation(params, precision)
    sim.upload(phi=phi, cons=cons, u=np.zeros((3, *shape)), gradu=np.zeros((9, *shape)))
    if graph:
        sim.set_graph(True)
    var = (ctypes.c_int32 * 5)()
    tms = (ctypes.c_float * 15)()
    _lib.check(lib.mm_get_tuning(sim.ctx, var, tms), "mm_get_tuning")
    tuning = {k: {"variant": var[i], "ms": [round(tms[3 * i + v], 4) for v in range(3)]}
              for i, k in enumerate(["x_r2c", "x_c2r", "y", "z", "y_tiled"])}
    sim.step(1 + warmup)  # elastep0 step + warm-up
    # timed region: no event records between the kernels
    _lib.check(lib.mm_probe(sim.ctx, -1), "mm_probe")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rows = sim.step(steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    state = sim.download()
    # the same number of steps again with the roofline kernel's launches
    # bracketed by HIP events on the step's stream (live, inside the step)
    _lib.check(lib.mm_probe(sim.ctx, probe_kind), "mm_probe")
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    sim.step(steps, diag=False)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    ms = ctypes.c_double()
    nl = ctypes.c_int64()
    by = ctypes.c_double()
    _lib.check(lib.mm_probe_read(sim.ctx, ctypes.byref(ms), ctypes.byref(nl), ctypes.byref(by)), "probe")
    # after the timed region: live HIP-event timing of the other kernel classes,
    # pass_steps steps each (the state keeps advancing; it is not compared)
    kernels = {}
    for name in (PASS_KINDS if pass_steps > 0 else ()):
        _lib.check(lib.mm_probe(sim.ctx, KINDS[name]), "mm_probe")
        sim.step(pass_steps, diag=False)
        kms, knl, kby = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
        _lib.check(lib.mm_probe_read(sim.ctx, ctypes.byref(kms), ctypes.byref(knl), ctypes.byref(kby)), "probe")
        if knl.value:
            avg = kms.value / knl.value
            gbs = kby.value / knl.value / (avg * 1e-3) / 1e9
            kernels[name] = {"avg_launch_ms": avg, "launches_per_step": knl.value / pass_steps,
                             "ms_per_step": kms.value / pass_steps,
                             "bytes_per_launch": kby.value / knl.value,
                             "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS}
    sim.close()
    return dict(seconds=t1 - t0, rows=rows, probe_ms=ms.value, probe_launches=nl.value,
                probe_bytes=by.value, probe_seconds=t3 - t2, state=state, tuning=tuning, kernels=kernels)


def run_ensemble(steps, warmup, graph=True, per_size=16):
    """The 48-member ensemble (16 each at 64^3, 96^3, 128^3, MIXED): one graph
    launch per step for all members, or (graph=False) the members stepped one
    after another through the eager path."""
    import numpy as np
    import torch

    from meumapps_ss_amd.configs import ensemble, initial_fields
    from meumapps_ss_amd.params import MM_MIXED
    from meumapps_ss_amd.sim import Ensemble, Simulation

    ws = ensemble(per_size=per_size)
    params = [w.params() for w in ws]
    cells = sum(p.nx * p.ny * p.nz for p in params)
    t_create = time.perf_counter()
    if graph:
        ens = Ensemble(params, MM_MIXED)
        members = ens.members
    else:
        ens = None
        members = [Simulation(p, MM_MIXED) for p in params]
    t_create = time.perf_counter() - t_create
    for m, w in zip(members, ws):
        phi, cons = initial_fields(w)
        shape = phi.shape[1:]
        m.upload(phi=phi, cons=cons, u=np.zeros((3, *shape)), gradu=np.zeros((9, *shape)))

    def adv(k):
        if graph:
            ens.step(k)
        else:
            for _ in range(k):
                for m in members:
                    m.step(1)

    adv(1 + warmup)  # elastep0 step (eager) and warm-up; the first regular step records the graph
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    adv(steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    if graph:
        ens.close()
    else:
        for m in members:
            m.close()
    return dict(seconds=t1 - t0, cells=cells, members=len(params), create_s=t_create)


def cpu_baseline(params, phi, cons, threads, steps=2):
    """Seconds per regular step of the CPU oracle over `steps` steps."""
    import numpy as np

    from oracle import Oracle

    o = Oracle(params, threads)
    o.set_state(phi=phi, cons=cons, u=np.zeros((3, *phi.shape[1:])), gradu=np.zeros((9, *phi.shape[1:])))
    o.set_step(params.elastep0 + 1)  # skip the zeroth-order solve: time regular steps
    t0 = time.perf_counter()
    o.step(steps)
    t1 = time.perf_counter()
    o.close()
    return (t1 - t0) / steps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--ensemble-steps", type=int, default=10,
                    help="steps of the 48-member ensemble (0 skips it)")
    ap.add_argument("--graph", action="store_true", help="run the 256^3 step in graph mode")
    ap.add_argument("--calib-copy", action="store_true",
                    help="end with three 1 GiB device copies (torch.neg, float64) that calibrate HBM "
                         "byte counters of a profiler run")
    ap.add_argument("--pass-steps", type=int, default=3,
                    help="after the timed region, steps per kernel class timed live for the per-kernel "
                         "HBM rates (0 skips them)")
    ap.add_argument("--probe", default="zf_elastic", choices=sorted(KINDS),
                    help="kernel class timed live inside the timed region for the roofline line "
                         "(default: the fused elastic z kernel k_z_fused<OpElastic>, the kernel with "
                         "the largest time per step: 4 launches of ~0.47 ms at 256^3 FP64, against "
                         "~0.43 ms for the polarization pass k_x_polar)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build keeps the whole grid on a single MI355X "
              "and has no multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from meumapps_ss_amd.configs import initial_fields, workload
    from meumapps_ss_amd.params import MM_FP64, MM_MIXED

    torch.cuda.set_device(0)
    w = workload(2, n=args.size)
    p = w.params()
    phi, cons = initial_fields(w)
    n3 = args.size ** 3
    kind = KINDS[args.probe]

    r64 = run_gpu(p, phi, cons, MM_FP64, args.steps, args.warmup, kind, args.graph, args.pass_steps)
    cups64 = n3 * args.steps / r64["seconds"]
    n_el = [r["n_elastic"] for r in r64["rows"]]
    n_avg = sum(n_el) / len(n_el)
    out = {
        "metric": "cell-updates/s at 256^3 3-variant: FP64 vs mixed FP32 at 1e-4; % HBM peak",
        "value": cups64,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r64["seconds"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (ran_2 uniform noise, seed 7)",
        "config": {"workload": f"{args.size}^3 c + 3 tetragonal variants, semi-implicit spectral step, FP64",
                   "grid": args.size, "variants": 3, "compositions": 1, "dt": 0.005, "seed": 7,
                   "elastic_iterations_per_step": n_avg, "parallelism": "single GPU"},
        "fft_autotune": r64["tuning"],
    }
    if r64["probe_launches"]:
        avg_ms = r64["probe_ms"] / r64["probe_launches"]
        bytes_per_launch = r64["probe_bytes"] / r64["probe_launches"]
        ach = bytes_per_launch / (avg_ms * 1e-3) / 1e9
        traffic = None
        tj = os.path.join(ROOT, "profiles", "pmc_traffic.json")
        key = f"{KERNEL_NAME.get(args.probe, args.probe)} {args.size}^3 FP64"
        if os.path.exists(tj):
            # HBM bytes per launch of this kernel on this workload from the committed
            # rocprofv3 FETCH_SIZE / WRITE_SIZE runs (copy-calibrated)
            t = json.load(open(tj)).get(key)
            if t:
                traffic = t["fetch_bytes"] + t["write_bytes"]
        out["roofline"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                           "kernel": KERNEL_NAME.get(args.probe, args.probe),
                           "avg_launch_ms": avg_ms, "bytes_per_launch": bytes_per_launch,
                           "launches_per_step": r64["probe_launches"] / args.steps,
                           "ms_per_step": r64["probe_ms"] / args.steps,
                           "share_of_step": r64["probe_ms"] / (1e3 * r64["probe_seconds"]),
                           "timed_steps_ms_per_step": 1e3 * r64["probe_seconds"] / args.steps}
    if r64["kernels"]:
        out["kernels"] = r64["kernels"]
    bpc = step_bytes_per_cell(3, 1, n_avg, 8)
    out["step_roofline"] = {"bytes_per_cell_update": bpc,
                            "achieved_GBs": bpc * cups64 / 1e9,
                            "frac": bpc * cups64 / 1e9 / HBM_PEAK_GBS}
    if not args.no_mixed:
        r32 = run_gpu(p, phi, cons, MM_MIXED, args.steps, args.warmup, kind, args.graph, args.pass_steps)
        cups32 = n3 * args.steps / r32["seconds"]
        l2 = {}
        for k in ("phi", "cons"):
            a, b = r64["state"][k], r32["state"][k]
            l2[k] = float(np.linalg.norm(b - a) / np.linalg.norm(a))
            l2[k + "_maxabs"] = float(np.abs(b - a).max())
        bpc32 = step_bytes_per_cell(3, 1, n_avg, 4)
        out["mixed"] = {"value": cups32, "speedup_vs_fp64": cups32 / cups64, "rel_l2_vs_fp64": l2,
                        "ms_per_step": 1e3 * r32["seconds"] / args.steps,
                        "probe_kernel": args.probe,
                        "probe_avg_launch_ms": (r32["probe_ms"] / r32["probe_launches"]
                                                if r32["probe_launches"] else None),
                        "step_roofline_frac": bpc32 * cups32 / 1e9 / HBM_PEAK_GBS}
        if r32["kernels"]:
            out["mixed"]["kernels"] = r32["kernels"]
    if args.ensemble_steps > 0:
        eg = run_ensemble(args.ensemble_steps, 2, graph=True)
        ee = run_ensemble(args.ensemble_steps, 2, graph=False)
        out["ensemble"] = {
            "workload": "48 members (16 each at 64^3, 96^3, 128^3), 3 tetragonal variants, eigenstrain "
                        "U[0.005,0.03], kappa U[0.5,2], seeds 100-147, mixed FP32/FP64",
            "steps": args.ensemble_steps,
            "graph_cell_updates_per_s": eg["cells"] * args.ensemble_steps / eg["seconds"],
            "graph_ms_per_step": 1e3 * eg["seconds"] / args.ensemble_steps,
            "eager_sequential_cell_updates_per_s": ee["cells"] * args.ensemble_steps / ee["seconds"],
            "eager_ms_per_step": 1e3 * ee["seconds"] / args.ensemble_steps,
            "create_s": eg["create_s"]}
    if args.calib_copy:
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.neg(a, out=b)
        torch.cuda.synchronize()
        out["calib_copy_bytes"] = {"read": a.numel() * 8, "write": b.numel() * 8}
    if not args.no_cpu:
        hc = host_cpu()
        # one OpenMP thread per physical core this process may use
        threads = args.cpu_threads or hc["physical_cores"]
        secs = cpu_baseline(p, phi, cons, threads)
        out["cpu_baseline"] = {"value": n3 / secs, "unit": "cell-updates/s", "cores": threads,
                               "kind": "port",
                               "sample": f"2 steps of the {args.size}^3 3-variant workload (after the "
                                         f"elastep0 step) by the CPU binary64 oracle, {threads} OpenMP threads",
                               "seconds_per_step": secs, "cpu_model": hc["model"],
                               "physical_cores": hc["physical_cores"], "logical_cpus": hc["logical_cpus"]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
