"""Gradient check of the differentiation workload on one GPU.

Workload: 128^3 grid, c + 2 order parameters (two precipitate phases, one
variant each, eigenstrain m (1, 1/2, 1/2) with m = 0.02), FP64, dt = 0.005,
200 steps after the zeroth-order elastic solve (restart with u = 0),
checkpoints every 20 steps with recomputation.  Initial fields: c = 0.4 +
U[-0.01, 0.01], eta_v = eta0 + U[-0.01, 0.01] from ran_2 with seed 3 (eta0 = 0).

The step is not differentiable where an order parameter reaches 0 while
another is positive: the f_diff correction of Multi_MEUMAPPS.f:824-841
switches on phi > 0, a jump of the driving force.  Competing variants drive
each other to 0 (and the clamp holds them there), so along a run of this
workload central differences do not converge to the derivative; the adjoint
is the derivative of the piecewise-smooth map with the branches of the run.
``--variants 1 --eta0 0.5`` gives a control run (one variant growing towards 1,
no correction term) where the map is smooth along the trajectory.  Target: the
final eta fields of a run from the same initial fields with the parameters
perturbed (kappa x1.1, M x0.9, L x1.1, m x1.1).  Loss: mean squared
difference of the final eta fields.  Gradients with respect to kappa, M, L, m
and the initial eta, checked along 8 random directions of the joint space
against central differences with step 1e-6.

    python -m meumapps_ss_amd.adjoint_check [--n 128] [--steps 200] [--every 20] [--dirs 8]

prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np


def build_for(n, variants=2):
    from .configs import model_inputs
    from .params import setup

    numph = 3 if variants == 2 else 2

    def build(th):
        m = th["eps"]
        inp = model_inputs(n, 1, (m, 0.5 * m, 0.5 * m), 0.005, kappa=th["kappa"], mob_c=th["M"],
                           mob_eta=th["L"], numph=numph, seed=3)
        return setup(inp)
    return build


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--every", type=int, default=20)
    ap.add_argument("--dirs", type=int, default=8)
    ap.add_argument("--h", type=float, default=1e-6)
    ap.add_argument("--eta0", type=float, default=0.0)
    ap.add_argument("--variants", type=int, default=2, choices=[1, 2])
    args = ap.parse_args()

    import torch

    from .adjoint import AdjointRun, build_jacobian, chain
    from .configs import Workload, initial_fields, model_inputs

    torch.cuda.set_device(0)
    n, ns, ev = args.n, args.steps, args.every
    theta = {"kappa": 1.0, "M": 1.0, "L": 1.0, "eps": 0.02}
    names = list(theta)
    build = build_for(n, args.variants)
    w = Workload("adjoint", model_inputs(n, 1, (0.02, 0.01, 0.01), 0.005,
                                              numph=3 if args.variants == 2 else 2, seed=3),
                 3, 0.01, (0.4,), ns)
    phi0, cons0 = initial_fields(w)
    phi0 = phi0 + args.eta0

    def run_loss(th, phi, target=None):
        run = AdjointRun(build(th))
        phiT, _, _ = run.forward(phi, cons0, ns, ns)
        run.close()
        return phiT if target is None else float(np.mean((phiT - target) ** 2))

    pert = {"kappa": 1.1, "M": 0.9, "L": 1.1, "eps": 1.1}
    target = run_loss({k: theta[k] * pert[k] for k in theta}, phi0)

    run = AdjointRun(build(theta))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    phiT, ck, rows = run.forward(phi0, cons0, ns, ev)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    L0 = float(np.mean((phiT - target) ** 2))
    gphi, _, g = run.backward(2.0 * (phiT - target) / phiT.size, ck, ns, ev)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    run.close()
    dth = chain(g, build_jacobian(build, theta))

    rng = np.random.default_rng(2024)
    checks = []
    for _ in range(args.dirs):
        dphi = rng.standard_normal(phi0.shape) * 0.01
        dt_ = rng.standard_normal(len(names))
        ad = float(np.sum(gphi * dphi)) + sum(dth[k] * dt_[i] * abs(theta[k]) for i, k in enumerate(names))
        h = args.h
        up = {k: theta[k] + h * dt_[i] * abs(theta[k]) for i, k in enumerate(names)}
        dn = {k: theta[k] - h * dt_[i] * abs(theta[k]) for i, k in enumerate(names)}
        fd = (run_loss(up, phi0 + h * dphi, target) - run_loss(dn, phi0 - h * dphi, target)) / (2 * h)
        checks.append({"adjoint": ad, "fd": fd, "rel_err": abs(ad - fd) / max(abs(fd), 1e-300)})
    out = {
        "workload": f"{n}^3 c + {args.variants} variant(s), FP64, {ns} steps, checkpoints every {ev}, "
                    f"eta0 = {args.eta0} + U[-0.01,0.01]",
        "loss": L0,
        "grad": {k: dth[k] for k in names},
        "grad_phi0_norm": float(np.linalg.norm(gphi)),
        "forward_s": t1 - t0,
        "backward_s": t2 - t1,
        "elastic_iterations": sorted(set(r["n_elastic"] for r in rows)),
        "fd_step": args.h,
        "checks": checks,
        "max_rel_err": max(c["rel_err"] for c in checks),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
