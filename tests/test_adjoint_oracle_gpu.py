"""The GPU adjoint against an independent FP64 reference derivative: the
oracle's tangent-linear step (oracle_tlm_*, pinned by central differences in
tests/test_tlm_cpu.py).

Dot-product test: for a linear read-out L = <w_phi, phi_T> + <w_c, c_T> of a
run and a direction delta of (phi_0, c_0, parameter arrays), the adjoint
seeded with lambda = (w_phi, w_c) gives <adjoint, delta>, the tangent run
gives <w, tangent(delta)>, and the two must agree.  Both take the decisions
of their own forward run (phi > 0 in the f_diff correction,
Multi_MEUMAPPS.f:824-841; the clamp, :1034-1035; the elastic iteration count
and cfac); the forward runs are compared first (fields to 1e-10, the clamp
pattern exactly, the elastic control exactly).  Bar: 1e-11 relative (both
are exact derivatives of the same map; rounding alone separates them)."""
import numpy as np
import pytest

from meumapps_ss_amd.adjoint import AdjointRun, raw_vector
from meumapps_ss_amd.params import MMAdjGrads
from parity_cases import smooth_field
from tangent_dirs import direction, run

pytestmark = pytest.mark.gpu

TOL = 1e-11


def _workload_start(n, seed=3):
    """The differentiation workload's parameters and initial fields (two
    precipitate phases, eigenstrain m (1, 1/2, 1/2), m = 0.02;
    eta = U[-0.01, 0.01], c = 0.4 + U[-0.01, 0.01] from ran_2 seed 3)."""
    from meumapps_ss_amd.adjoint_check import build_for
    from meumapps_ss_amd.configs import Workload, initial_fields, model_inputs
    p = build_for(n, 2)({"kappa": 1.0, "M": 1.0, "L": 1.0, "eps": 0.02})
    inp = model_inputs(n, 1, (0.02, 0.01, 0.01), 0.005, numph=3, seed=seed)
    phi, cons = initial_fields(Workload("adjoint", inp, seed, 0.01, (0.4,), 1))
    return p, phi, cons


def _three_variant_start(n, seed=11):
    from meumapps_ss_amd.configs import model_inputs
    from meumapps_ss_amd.params import setup
    p = setup(model_inputs(n, 3, (0.02, 0.01, 0.01), 0.005))
    rng = np.random.default_rng(seed)
    shape = (n, n, n)
    phi = np.stack([smooth_field(rng, shape, -0.1, 1.0) for _ in range(3)])
    cons = smooth_field(rng, shape, 0.35, 0.55)[None]
    return p, phi, cons


def _dot(a, b):
    ra, rb = raw_vector(a), raw_vector(b)
    return float(sum(np.dot(ra[k], rb[k]) for k in ra))


@pytest.mark.parametrize("case,n,nsteps,every", [("workload", 32, 20, 5), ("workload", 64, 50, 10), ("three_variants", 16, 8, 3)])
def test_adjoint_matches_oracle_tangent(gpu_lib, case, n, nsteps, every):
    p, phi0, cons0 = _workload_start(n) if case == "workload" else _three_variant_start(n)
    print(f"case {case} n={n} steps={nsteps}", flush=True)
    rng = np.random.default_rng(7)
    w_phi = rng.standard_normal(phi0.shape)
    w_c = rng.standard_normal(cons0.shape)
    # three directions: all inputs, the initial order parameters only, the
    # parameter arrays only
    dirs = []
    dphi = rng.standard_normal(phi0.shape) * 0.01
    dcons = rng.standard_normal(cons0.shape) * 0.001
    dpar = direction(p, rng)
    dirs.append((dphi, dcons, dpar))
    dirs.append((rng.standard_normal(phi0.shape) * 0.01, np.zeros_like(cons0), MMAdjGrads()))
    dirs.append((np.zeros_like(phi0), np.zeros_like(cons0), direction(p, rng)))

    ar = AdjointRun(p)
    phiT, ck, rows = ar.forward(phi0, cons0, nsteps, every)
    consT = ar.sim.download()["cons"]
    gphi, gcons, g = ar.backward(w_phi, ck, nsteps, every, lam_cons=w_c)
    ar.close()
    adj = [float(np.sum(gphi * d[0]) + np.sum(gcons * d[1])) + _dot(g, d[2]) for d in dirs]

    for i, d in enumerate(dirs):
        st, t, pats, ctl = run(p, phi0, cons0, nsteps, (d[0], d[1], None, d[2]), nthreads=4)
        if i == 0:
            # the same forward run: fields, clamp pattern, elastic control
            assert np.abs(st["phi"] - phiT).max() < 1e-10
            assert np.abs(st["cons"] - consT).max() < 1e-10
            assert np.array_equal(pats[-1], np.packbits((phiT == 0.0) | (phiT == 1.0)))
            assert ctl == [(r["n_elastic"], r["cfac"]) for r in rows]
        tl = float(np.sum(w_phi * t["phi"]) + np.sum(w_c * t["cons"]))
        rel = abs(adj[i] - tl) / max(abs(tl), 1e-300)
        print(f"direction {i}: adjoint {adj[i]:.15e} tangent {tl:.15e} rel {rel:.2e}", flush=True)
        assert rel <= TOL, (i, adj[i], tl)
