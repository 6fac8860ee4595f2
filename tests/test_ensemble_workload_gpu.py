"""The ensemble workload at its own size: 48 members of the three-variant
tetragonal model, 16 each at 64^3, 96^3 and 128^3, ran_2 seeds 100-147,
eigenstrain magnitude U[0.005, 0.03] and kappa U[0.5, 2], mixed precision,
every regular step one graph launch for all members (configs.ensemble).

Each member is one run of the loop body Multi_MEUMAPPS.f:548-1106, so
* every member's fields and fort.98 rows after the elastep0 step and 5 graph
  steps equal those of its own single eager run, bit for bit;
* one member of each size is within 1e-4 relative L2 (phi and c) of the FP64
  oracle run from the same initial fields, with the same elastic iteration
  counts.
"""
import numpy as np
import pytest

from meumapps_ss_amd.configs import ensemble, initial_fields
from meumapps_ss_amd.params import MM_MIXED
from meumapps_ss_amd.sim import Ensemble, Simulation
from oracle import Oracle
from test_graph_gpu import _rows_same, _same, _zeros

pytestmark = pytest.mark.gpu

NSTEPS = 6  # the elastep0 step and 5 regular (graph) steps


def test_ensemble_workload_members_equal_single_runs_and_oracle(gpu_lib):
    ws = ensemble()
    assert len(ws) == 48
    params = [w.params() for w in ws]
    assert sorted({p.nx for p in params}) == [64, 96, 128]
    fields = [initial_fields(w) for w in ws]
    with Ensemble(params, MM_MIXED) as ens:
        for m, (phi, cons) in zip(ens.members, fields):
            m.upload(phi=phi, cons=cons, **_zeros(phi.shape[1:]))
        rows = ens.step(1)
        for _ in range(NSTEPS - 1):
            rows += ens.step(1)
        got = [m.download() for m in ens.members]
    print(f"ensemble: {len(ws)} members, {NSTEPS} steps", flush=True)
    checked = set()
    for i, (p, (phi, cons)) in enumerate(zip(params, fields)):
        with Simulation(p, MM_MIXED) as s:
            s.upload(phi=phi, cons=cons, **_zeros(phi.shape[1:]))
            ref = s.step(NSTEPS)
            _same(got[i], s.download(), f"member {i} ({p.nx}^3)")
        for k in range(NSTEPS):
            _rows_same(rows[k][i], ref[k], f"member {i} step {k + 1}")
        if p.nx in checked:
            continue
        checked.add(p.nx)
        o = Oracle(p, 16)
        o.set_state(phi=phi, cons=cons)
        orows = o.step(NSTEPS)
        ro = o.get_state()
        o.close()
        assert [r["n_elastic"] for r in orows] == [rows[k][i]["n_elastic"] for k in range(NSTEPS)]
        for key in ("phi", "cons"):
            l2 = np.linalg.norm(got[i][key] - ro[key]) / np.linalg.norm(ro[key])
            print(f"member {i} ({p.nx}^3) {key}: rel L2 vs FP64 oracle {l2:.2e}", flush=True)
            assert l2 < 1e-4, f"member {i} {key} rel L2 {l2:.2e}"
    assert checked == {64, 96, 128}
