"""The FP64 GPU step against the oracle at the configured horizons.

The fixtures (tests/golden/long_*.npz, written by
tests/golden/make_long_fixtures.py from the oracle) hold, for each dump of a
run (every ``ifreq`` steps, Multi_MEUMAPPS.f:1074-1100), a strided subsample
of every phi and c field and per field the sum, sum of squares, minimum and
maximum, plus the fort.98 / fort.99 row of every step, including the lagged
check-loop columns glob_min, glob_max, phi_max and solid_tot
(Multi_MEUMAPPS.f:905-906, fed by :1016-1062).  The GPU runs the same
workload from the same ran_2 initial fields and must match at every dump:

* the 64^3 one-variant workload, 1000 steps, dumps every 100;
* the 256^3 three-variant workload (the benchmark workload), dumps every 100;
* a 256^3 three-variant run from amplitude-0.3 fields whose solid fraction
  grows over the run (the workload itself decays towards phi = 0).
"""
import importlib.util
import os

import numpy as np
import pytest

from meumapps_ss_amd.configs import initial_fields
from meumapps_ss_amd.params import MM_FP64
from meumapps_ss_amd.sim import Simulation

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
_spec = importlib.util.spec_from_file_location("make_long_fixtures",
                                               os.path.join(HERE, "golden", "make_long_fixtures.py"))
mlf = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(mlf)

TOL = 1e-10


@pytest.mark.parametrize("name", ["workload1_64", "grow_256", "workload2_256"])
def test_long_horizon_fp64_matches_oracle(gpu_lib, name):
    f = mlf.path(name)
    if not os.path.exists(f):
        pytest.skip(f"no fixture {os.path.basename(f)}")
    d = np.load(f)
    w, _, every, _ = mlf.case(name)
    stride = int(d["stride"])
    p = w.params()
    phi, cons = initial_fields(w)
    n3 = p.nx * p.ny * p.nz
    sub = (slice(None), slice(None, None, stride), slice(None, None, stride), slice(None, None, stride))
    rows = []
    worst = 0.0
    with Simulation(p, MM_FP64) as s:
        s.upload(phi=phi, cons=cons, u=np.zeros((3, *phi.shape[1:])), gradu=np.zeros((9, *phi.shape[1:])))
        for i, dstep in enumerate(d["dump_steps"]):
            rows += s.step(every)
            assert rows[-1]["step"] == dstep
            st = s.download()
            for key in ("phi", "cons"):
                e = np.abs(st[key][sub] - d[f"{key}_sub"][i]).max()
                worst = max(worst, e)
                assert e < TOL, f"{name} step {dstep} {key} subsample max-abs {e:.3e}"
                ref = d[f"{key}_stats"][i]
                a = st[key]
                got = np.stack([a.sum(axis=(1, 2, 3)), (a * a).sum(axis=(1, 2, 3)), a.min(axis=(1, 2, 3)),
                                a.max(axis=(1, 2, 3))], axis=1)
                # sums to the bar per point, extrema to the bar
                assert np.all(np.abs(got[:, 0] - ref[:, 0]) < TOL * n3), (name, dstep, key, "sum")
                assert np.all(np.abs(got[:, 1] - ref[:, 1]) < TOL * n3 * 2 * np.abs(ref[:, 3]).max()), \
                    (name, dstep, key, "sum of squares")
                assert np.all(np.abs(got[:, 2:] - ref[:, 2:]) < TOL), (name, dstep, key, "min/max")
            print(f"{name} step {dstep}: phi_max {st['phi'].max():.4e}, mean phi "
                  f"{st['phi'].mean():.4e}, worst max-abs so far {worst:.2e}", flush=True)
    nd = len(rows)
    diag = np.array([[r[k] for k in mlf.DIAG_KEYS] for r in rows])
    ref = d["diag"][:nd]
    assert [r["n_elastic"] for r in rows] == list(d["n_elastic"][:nd])
    # energies (fort.98 columns 2-6) relative, lagged check-loop columns absolute
    for j in range(5):
        sc = np.abs(ref[:, j]).max() + 1e-300
        assert np.abs(diag[:, j] - ref[:, j]).max() < 1e-9 * sc, (name, mlf.DIAG_KEYS[j])
    lag = ref[:, 5:]
    ok = ~np.isnan(lag)
    assert np.abs(diag[:, 5:][ok] - lag[ok]).max() < TOL, (name, "lagged columns")
    cav = np.array([r["caviter"] for r in rows])
    assert np.abs(cav - d["caviter"][:nd]).max() < 1e-12
