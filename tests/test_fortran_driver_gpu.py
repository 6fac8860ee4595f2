"""The step driven from Fortran through the ISO_C_BINDING module
(meumapps_ss_amd/fortran/mm_step_driver.f90 over mm_abi.f90, built by
build() with flang): the parameter block and the fields go in as files, and
the fields after the steps and the fort.98 rows come back.  Against the same
step driven from Python through ctypes: identical bits and identical rows;
against the oracle: the parity bar."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from meumapps_ss_amd import build as mbuild
from meumapps_ss_amd.params import MM_FP64, MM_MIXED
from meumapps_ss_amd.records import fort98_row
from meumapps_ss_amd.sim import Simulation, soa_to_aos
from oracle import Oracle
from parity_cases import make_case

DRIVER = mbuild.DRIVER
pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran step driver not built (no flang)")]


@pytest.mark.parametrize("name,prec", [("v3_tetra", MM_FP64), ("v1_e2", MM_FP64), ("v3_tetra", MM_MIXED)])
def test_fortran_driver_matches_python_and_oracle(gpu_lib, tmp_path, name, prec):
    p, phi, cons = make_case(name, seed=9)
    nsteps = 6
    V, E = p.vartot, p.numel
    shape = phi.shape[1:]
    (tmp_path / "params.bin").write_bytes(bytes(p))
    flds = np.concatenate([soa_to_aos(phi).ravel(), soa_to_aos(cons).ravel()])
    flds.tofile(tmp_path / "fields.bin")
    r = subprocess.run([DRIVER, "params.bin", "fields.bin", str(nsteps), str(prec), "out.bin"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    out = np.fromfile(tmp_path / "out.bin")
    n3 = int(np.prod(shape))
    f_phi, f_cons = out[:V * n3], out[V * n3:]
    with Simulation(p, prec) as s:
        s.upload(phi=phi, cons=cons, u=np.zeros((3, *shape)), gradu=np.zeros((9, *shape)),
                 el_en=np.zeros(shape))
        rows = s.step(nsteps)
        rec = s.download_records()
    assert np.array_equal(f_phi, rec["phi"]) and np.array_equal(f_cons, rec["cons"])
    assert lines == [fort98_row(d) for d in rows]
    if prec == MM_FP64:
        o = Oracle(p, 4)
        o.set_state(phi=phi, cons=cons)
        o.step(nsteps)
        ro = o.get_state()
        assert np.abs(f_phi - soa_to_aos(ro["phi"]).ravel()).max() < 1e-10
        assert np.abs(f_cons - soa_to_aos(ro["cons"]).ravel()).max() < 1e-10
