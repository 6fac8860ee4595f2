"""The configured workloads at their own sizes against the oracle (FP64 to the
parity bar) and, for the mixed form, against FP64 on the GPU.

* 64^3, c + 1 variant with a dilatational eigenstrain, dt = 0.01, ran_2 seed 1:
  the elastep0 step and 10 steps;
* 256^3, c + 3 tetragonal variants, dt = 0.005, seed 7 (the benchmark
  workload): the elastep0 step and one regular step (the oracle takes several
  seconds per step at this size on 16 threads);
* 128^3, c + 2 variants in two precipitate phases, seed 3: the elastep0 step
  and 4 steps;
* 512^3 mixed (seed 11): the elastep0 step and 3 steps against FP64 on the GPU,
  relative L2 <= 1e-4 per field.
"""
import numpy as np
import pytest

from meumapps_ss_amd.configs import initial_fields, workload
from meumapps_ss_amd.params import MM_FP64, MM_MIXED
from meumapps_ss_amd.sim import Simulation
from oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _gpu(p, phi, cons, prec, nsteps):
    shape = phi.shape[1:]
    with Simulation(p, prec) as s:
        s.upload(phi=phi, cons=cons, u=np.zeros((3, *shape)), gradu=np.zeros((9, *shape)))
        rows = s.step(nsteps)
        return s.download(), rows


@pytest.mark.parametrize("k,n,nsteps", [(1, 64, 11), (2, 256, 2), (5, 128, 5)])
def test_workload_fp64_matches_oracle(gpu_lib, k, n, nsteps):
    w = workload(k, n=n)
    p = w.params()
    phi, cons = initial_fields(w)
    st, rows = _gpu(p, phi, cons, MM_FP64, nsteps)
    o = Oracle(p, 16)
    o.set_state(phi=phi, cons=cons)
    ref_rows = o.step(nsteps)
    ro = o.get_state()
    o.close()
    for key in ("phi", "cons"):
        err = np.abs(st[key] - ro[key]).max()
        assert err < TOL, f"{key} max-abs {err:.3e}"
    assert [r["n_elastic"] for r in rows] == [q["n_elastic"] for q in ref_rows]
    for r, q in zip(rows, ref_rows):
        assert abs(r["caviter"][0] - q["caviter"][0]) < 1e-13


def test_workload_512_mixed_matches_fp64(gpu_lib):
    w = workload(3, n=512)
    p = w.params()
    phi, cons = initial_fields(w)
    a, _ = _gpu(p, phi, cons, MM_FP64, 4)
    b, _ = _gpu(p, phi, cons, MM_MIXED, 4)
    for key in ("phi", "cons"):
        l2 = np.linalg.norm(b[key] - a[key]) / np.linalg.norm(a[key])
        assert l2 < 1e-4, f"{key} rel L2 {l2:.2e}"
