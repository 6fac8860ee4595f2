"""The differentiation workload's gradient against the oracle's tangent-linear
step (FP64 reference derivative), at the configured size.

Workload as in meumapps_ss_amd.adjoint_check: 128^3 grid, c + 2 order
parameters (two precipitate phases, eigenstrain m (1, 1/2, 1/2), m = 0.02),
dt = 0.005, 200 steps after a restart with u = 0, checkpoints every 20;
eta = U[-0.01, 0.01], c = 0.4 + U[-0.01, 0.01] (ran_2 seed 3); loss = mean
squared difference of the final eta from a run with kappa x1.1, M x0.9,
L x1.1, m x1.1.  For each direction delta of (eta_0, kappa, M, L, m) it
compares <dL/d(inputs), delta> from the GPU adjoint (chained through the
set-up Jacobian) with <dL/d eta_T, tangent(delta)> from the oracle's tangent
run; both take their own forward run's decisions, and the forward runs are
compared (final eta to 1e-10, the clamp pattern, the elastic control).

    python tests/adjoint_tangent_check.py [--n 128] [--steps 200] [--dirs 2] [--threads 16]

prints a progress line per 20 oracle steps and one JSON line at the end.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--every", type=int, default=20)
    ap.add_argument("--dirs", type=int, default=2)
    ap.add_argument("--threads", type=int, default=16)
    args = ap.parse_args()

    import torch

    from meumapps_ss_amd.adjoint import AdjointRun, build_jacobian, chain
    from meumapps_ss_amd.adjoint_check import build_for
    from meumapps_ss_amd.configs import Workload, initial_fields, model_inputs
    from meumapps_ss_amd.params import MMAdjGrads
    from oracle import Oracle

    torch.cuda.set_device(0)
    n, ns, ev = args.n, args.steps, args.every
    theta = {"kappa": 1.0, "M": 1.0, "L": 1.0, "eps": 0.02}
    names = list(theta)
    build = build_for(n, 2)
    w = Workload("adjoint", model_inputs(n, 1, (0.02, 0.01, 0.01), 0.005, numph=3, seed=3), 3, 0.01, (0.4,), ns)
    phi0, cons0 = initial_fields(w)
    pert = {"kappa": 1.1, "M": 0.9, "L": 1.1, "eps": 1.1}
    r = AdjointRun(build({k: theta[k] * pert[k] for k in theta}))
    target, _, _ = r.forward(phi0, cons0, ns, ns)
    r.close()

    p = build(theta)
    run = AdjointRun(p)
    t0 = time.perf_counter()
    phiT, ck, rows = run.forward(phi0, cons0, ns, ev)
    lam = 2.0 * (phiT - target) / phiT.size
    gphi, _, g = run.backward(lam, ck, ns, ev)
    torch.cuda.synchronize()
    t_gpu = time.perf_counter() - t0
    run.close()
    jac = build_jacobian(build, theta)
    dth = chain(g, jac)
    print(f"GPU forward + adjoint {t_gpu:.1f} s; loss {float(np.mean((phiT - target) ** 2)):.6e}", flush=True)

    rng = np.random.default_rng(2024)
    checks = []
    for i in range(args.dirs):
        dphi = rng.standard_normal(phi0.shape) * 0.01
        dt_ = rng.standard_normal(len(names))
        ad = float(np.sum(gphi * dphi)) + sum(dth[k] * dt_[j] * abs(theta[k]) for j, k in enumerate(names))
        # the same direction in the raw parameter arrays
        dpar = MMAdjGrads()
        for f in ("grad_coeff", "wellph", "cmobs", "trans_ppt"):
            arr = getattr(dpar, f)
            v = sum(jac[k][f] * dt_[j] * abs(theta[k]) for j, k in enumerate(names))
            for q in range(len(v)):
                arr[q] = float(v[q])
        dpar.b_mob = float(sum(jac[k]["b_mob"][0] * dt_[j] * abs(theta[k]) for j, k in enumerate(names)))
        o = Oracle(p, args.threads)
        shape = phi0.shape[1:]
        o.set_state(phi=phi0, cons=cons0, u=np.zeros((3, *shape)), gradu=np.zeros((9, *shape)))
        o.set_step(p.elastep0 + 1)
        o.tlm_set(dphi, None, None, dpar)
        t1 = time.perf_counter()
        ctl = []
        for s in range(ns):
            rr = o.step(1)[0]
            ctl.append((rr["n_elastic"], rr["cfac"]))
            if (s + 1) % 20 == 0:
                print(f"direction {i}: oracle step {s + 1}/{ns} ({time.perf_counter() - t1:.0f} s)", flush=True)
        st = o.get_state()
        tl = float(np.sum(lam * o.tlm_get()["phi"]))
        o.close()
        same = {"phi_max_abs_diff": float(np.abs(st["phi"] - phiT).max()),
                "clamp_pattern_equal": bool(np.array_equal((st["phi"] == 0) | (st["phi"] == 1),
                                                           (phiT == 0) | (phiT == 1))),
                "elastic_control_equal": ctl == [(q["n_elastic"], q["cfac"]) for q in rows]}
        rel = abs(ad - tl) / max(abs(tl), 1e-300)
        checks.append({"adjoint": ad, "tangent": tl, "rel_err": rel, "forward": same,
                       "oracle_s": time.perf_counter() - t1})
        print(f"direction {i}: adjoint {ad:.15e} tangent {tl:.15e} rel {rel:.2e} {same}", flush=True)
    print(json.dumps({"workload": f"{n}^3 c + 2 variants, FP64, {ns} steps, checkpoints every {ev}",
                      "gpu_forward_adjoint_s": t_gpu, "grad": dth, "checks": checks,
                      "max_rel_err": max(c["rel_err"] for c in checks)}))


if __name__ == "__main__":
    main()
