"""Mixed FP32/FP64 against FP64 over a whole run of the 256^3 workload.

Workload: 256^3 grid, c + 3 tetragonal variants (eigenstrain diag(0.02,
-0.01, -0.01) permuted per variant), cubic C11/C12/C44 = 250/150/100,
dt = 0.005, initial fields c = 0.4 + U[-0.01, 0.01], eta_v = U[-0.01, 0.01]
from ran_2 with seed 7, 2000 steps after the zeroth-order elastic step, the
fields compared every 500 steps (the dump interval of the workload).

Both precisions advance side by side from the same uploaded fields; at each
comparison point the fields come back to the host and are compared per
field: relative L2 and max-abs of phi (all variants) and c, the mean
composition of each run against the initial mean; at the end the elastic
iterations of each run and the number of steps where they differ.

With the initial amplitude of the workload the order parameters decay
towards the homogeneous state.  ``--init smooth`` starts instead from smooth
random fields (spectral content below a quarter of Nyquist, seed 7 of
numpy's PCG64): eta_v in [-0.15, 1.1] and c in [0.3, 0.6], a microstructure
that coarsens, hits the clamp and carries elastic interactions.

    python -m meumapps_ss_amd.mixed_check [--n 256] [--steps 2000] [--every 500] [--init noise|smooth]

prints one progress line per comparison point and one JSON line at the end.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np


def smooth_fields(shape, V, E, seed=7, cut=0.25):
    """Random fields with spectral content below cut * Nyquist, scaled to
    [-0.15, 1.1] (eta) and [0.3, 0.6] (c)."""
    rng = np.random.default_rng(seed)
    kz = np.fft.fftfreq(shape[0])[:, None, None]
    ky = np.fft.fftfreq(shape[1])[None, :, None]
    kx = np.fft.rfftfreq(shape[2])[None, None, :]
    keep = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2) <= cut * 0.5

    def one(lo, hi):
        f = np.fft.rfftn(rng.standard_normal(shape)) * keep
        b = np.fft.irfftn(f, s=shape, axes=(0, 1, 2))
        return lo + (hi - lo) * (b - b.min()) / (b.max() - b.min())

    return np.stack([one(-0.15, 1.1) for _ in range(V)]), np.stack([one(0.3, 0.6) for _ in range(E)])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--every", type=int, default=500)
    ap.add_argument("--init", default="noise", choices=["noise", "smooth"])
    args = ap.parse_args()

    from .configs import initial_fields, workload
    from .params import MM_FP64, MM_MIXED
    from .sim import Simulation

    w = workload(2, n=args.n)
    p = w.params()
    phi, cons = initial_fields(w)
    shape = phi.shape[1:]
    if args.init == "smooth":
        phi, cons = smooth_fields(shape, p.vartot, p.numel)
    c0 = cons.mean(axis=(1, 2, 3))
    sims = {}
    for name, prec in (("fp64", MM_FP64), ("mixed", MM_MIXED)):
        s = Simulation(p, prec)
        s.upload(phi=phi, cons=cons, u=np.zeros((3, *shape)), gradu=np.zeros((9, *shape)))
        s.step(1)  # the elastep0 step
        sims[name] = s
    points, done = [], 0
    iters = {"fp64": [], "mixed": []}
    t0 = time.perf_counter()
    while done < args.steps:
        k = min(args.every, args.steps - done)
        rows = {name: s.step(k) for name, s in sims.items()}
        done += k
        for name in iters:
            iters[name] += [r["n_elastic"] for r in rows[name]]
        a, b = sims["fp64"].download(), sims["mixed"].download()
        pt = {"step": done}
        for key in ("phi", "cons"):
            d = b[key] - a[key]
            pt[f"{key}_rel_l2"] = float(np.linalg.norm(d) / np.linalg.norm(a[key]))
            pt[f"{key}_max_abs"] = float(np.abs(d).max())
        pt["mean_c_fp64_minus_c0"] = float(rows["fp64"][-1]["caviter"][0] - c0[0])
        pt["mean_c_mixed_minus_c0"] = float(rows["mixed"][-1]["caviter"][0] - c0[0])
        pt["phi_max"] = float(a["phi"].max())
        points.append(pt)
        print(json.dumps(pt), flush=True)
    for s in sims.values():
        s.close()
    out = {"workload": w.name, "init": args.init, "steps": args.steps, "every": args.every, "points": points,
           "elastic_iterations": {k: int(sum(v)) for k, v in iters.items()},
           "steps_with_other_iteration_count": int(sum(x != y for x, y in zip(iters["fp64"], iters["mixed"]))),
           "seconds": time.perf_counter() - t0}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
