"""Gradient check of the differentiation workload on one GPU.

Workload: 128^3 grid, c + 2 order parameters (two precipitate phases, one
variant each, eigenstrain m (1, 1/2, 1/2) with m = 0.02), FP64, dt = 0.005,
200 steps after the zeroth-order elastic solve (restart with u = 0),
checkpoints every 20 steps with recomputation.  Initial fields: c = 0.4 +
U[-0.01, 0.01], eta_v = eta0 + U[-0.01, 0.01] from ran_2 with seed 3 (eta0 = 0).
Target: the final eta fields of a run from the same initial fields with the
parameters perturbed (kappa x1.1, M x0.9, L x1.1, m x1.1).  Loss: mean squared
difference of the final eta fields.  Gradients with respect to kappa, M, L, m
and the initial eta, checked along 8 random directions of the joint space.

The step takes two data-dependent branches per point and variant: the
phi > 0 tests of the f_diff correction (Multi_MEUMAPPS.f:824-841) and the
clamp to [0, 1] (:1034-1035).  From eta = U[-0.01, 0.01] about half the
points sit at the clamp from the first step on, so a perturbation of the
inputs changes decisions at some points, and the map is not differentiable
there.  Each direction therefore reports
* the central difference of the plain run (step --h) and the number of
  branch decisions the +h and -h runs take differently from the base run
  (mm_branch_tape in count mode);
* central differences of the frozen-branch map (the base run's decisions
  replayed, mode 3) for each step in --hs: the map whose derivative the
  adjoint computes, smooth in the inputs.
``--variants 1 --eta0 0.5`` gives a control run (one variant growing towards 1,
no correction term) where the map is smooth along the trajectory.

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


MM_BRANCH_RECORD, MM_BRANCH_COUNT, MM_BRANCH_REPLAY = 1, 2, 3


def check(n=128, steps=200, every=20, dirs=8, h=1e-6, hs=(1e-4, 1e-5, 1e-6, 1e-7), eta0=0.0, variants=2,
          seed=2024):
    """Run the check; returns the result dictionary (see the module text)."""
    import torch

    from .adjoint import AdjointRun, build_jacobian, chain
    from .configs import Workload, initial_fields, model_inputs

    ns, ev = steps, every
    theta = {"kappa": 1.0, "M": 1.0, "L": 1.0, "eps": 0.02}
    names = list(theta)
    build = build_for(n, variants)
    w = Workload("adjoint", model_inputs(n, 1, (0.02, 0.01, 0.01), 0.005,
                                              numph=3 if variants == 2 else 2, seed=3),
                 3, 0.01, (0.4,), ns)
    phi0, cons0 = initial_fields(w)
    phi0 = phi0 + eta0

    tape = {}

    def run_loss(th, phi, target=None, mode=0):
        run = AdjointRun(build(th))
        if mode:
            run.set_branch_tape(mode, tape["buf"], ns)
        phiT, _, _ = run.forward(phi, cons0, ns, ns)
        flips = run.branch_flips() if mode == MM_BRANCH_COUNT else None
        run.close()
        val = phiT if target is None else float(np.mean((phiT - target) ** 2))
        return val, flips

    pert = {"kappa": 1.1, "M": 0.9, "L": 1.1, "eps": 1.1}
    target, _ = run_loss({k: theta[k] * pert[k] for k in theta}, phi0)

    run = AdjointRun(build(theta))
    tape["buf"] = run.new_branch_tape(ns)
    run.set_branch_tape(MM_BRANCH_RECORD, tape["buf"], ns)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    phiT, ck, rows = run.forward(phi0, cons0, ns, ev)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    run.set_branch_tape(0)
    L0 = float(np.mean((phiT - target) ** 2))
    gphi, _, g = run.backward(2.0 * (phiT - target) / phiT.size, ck, ns, ev)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    run.close()
    dth = chain(g, build_jacobian(build, theta))
    # decisions at the branch points over the run: how many sit at the clamp
    # or at phi = 0
    tb = tape["buf"].cpu().numpy()
    clamped = int(np.count_nonzero((tb >> 1) & 3))
    positive = int(np.count_nonzero(tb & 1))

    rng = np.random.default_rng(seed)
    checks = []
    for _ in range(dirs):
        dphi = rng.standard_normal(phi0.shape) * 0.01
        dt_ = rng.standard_normal(len(names))
        ad = float(np.sum(gphi * dphi)) + sum(dth[k] * dt_[i] * abs(theta[k]) for i, k in enumerate(names))

        def fd(hh, mode):
            up = {k: theta[k] + hh * dt_[i] * abs(theta[k]) for i, k in enumerate(names)}
            dn = {k: theta[k] - hh * dt_[i] * abs(theta[k]) for i, k in enumerate(names)}
            lu, fu = run_loss(up, phi0 + hh * dphi, target, mode)
            ld, fdn = run_loss(dn, phi0 - hh * dphi, target, mode)
            return (lu - ld) / (2 * hh), fu, fdn

        plain, fp, fm = fd(h, MM_BRANCH_COUNT)
        frozen = {f"{hh:g}": fd(hh, MM_BRANCH_REPLAY)[0] for hh in hs}
        rel = lambda x: abs(ad - x) / max(abs(x), 1e-300)
        checks.append({"adjoint": ad, "fd": plain, "rel_err": rel(plain), "flips_plus": fp, "flips_minus": fm,
                       "fd_frozen": frozen, "rel_err_frozen": {k: rel(v) for k, v in frozen.items()}})
    best_h = min(frozen, key=lambda k: max(c["rel_err_frozen"][k] for c in checks))
    return {
        "workload": f"{n}^3 c + {variants} variant(s), FP64, {ns} steps, checkpoints every {ev}, "
                    f"eta0 = {eta0} + U[-0.01,0.01]",
        "loss": L0,
        "grad": {k: dth[k] for k in names},
        "grad_phi0_norm": float(np.linalg.norm(gphi)),
        "forward_s": t1 - t0,
        "backward_s": t2 - t1,
        "elastic_iterations": sorted(set(r["n_elastic"] for r in rows)),
        "branch_decisions": {"points_x_variants_x_steps": int(tb.size), "clamped": clamped,
                             "phi_positive": positive},
        "fd_step": h,
        "checks": checks,
        "max_rel_err": max(c["rel_err"] for c in checks),
        "frozen_best_step": float(best_h),
        "max_rel_err_frozen": max(c["rel_err_frozen"][best_h] for c in checks),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--every", type=int, default=20)
    ap.add_argument("--dirs", type=int, default=8)
    ap.add_argument("--h", type=float, default=1e-6)
    ap.add_argument("--hs", default="1e-4,1e-5,1e-6,1e-7")
    ap.add_argument("--eta0", type=float, default=0.0)
    ap.add_argument("--variants", type=int, default=2, choices=[1, 2])
    args = ap.parse_args()

    import torch

    torch.cuda.set_device(0)
    out = check(args.n, args.steps, args.every, args.dirs, args.h, tuple(float(x) for x in args.hs.split(",")),
                args.eta0, args.variants)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
