"""Langevin noise steps on the GPU against the CPU oracle.

Steps with step <= noise_step add the reference's ran_2 / langevin stream
(host-generated, Multi_MEUMAPPS.f:552-571) to phi_noise, whose transform joins
the Allen-Cahn update (:963-978); later steps reset it.  Bar: phi and c within
1e-10 max-abs of the oracle over noise steps and the steps after; the
generator state after the noise steps equal; MIXED within 1e-4 relative L2
of FP64.
"""
import numpy as np
import pytest

from meumapps_ss_amd._lib import MMError
from meumapps_ss_amd.params import MM_FP64, MM_MIXED
from meumapps_ss_amd.sim import Simulation
from oracle import Oracle
from parity_cases import make_case

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _zero(phi):
    return np.zeros((3, *phi.shape[1:])), np.zeros((9, *phi.shape[1:]))


@pytest.mark.parametrize("name,seed", [("v3_tetra", -11), ("v2_nonsq", 12345), ("v3_inhomog", -3)])
def test_noise_steps_match_oracle(gpu_lib, name, seed):
    p, phi, cons = make_case(name, seed=7)
    p.noise_step = p.step + 2  # three noise steps, then two without
    sd = 0.02
    o = Oracle(p, 4)
    o.set_state(phi=phi, cons=cons)
    o.set_noise(sd, seed)
    ref_rows = o.step(5)
    ro = o.get_state()
    r_idum, r_state, _ = o.get_noise()
    with Simulation(p, MM_FP64) as s:
        u0, g0 = _zero(phi)
        s.upload(phi=phi, cons=cons, u=u0, gradu=g0)
        s.set_noise(sd, seed)
        rows = s.step(5)
        st = s.download()
        idum, state = s.noise_state()
    for k in ("phi", "cons"):
        err = np.abs(st[k] - ro[k]).max()
        assert err < TOL, f"{k} max-abs {err:.3e}"
    assert (idum, state) == (r_idum, r_state)
    assert [r["n_elastic"] for r in rows] == [q["n_elastic"] for q in ref_rows]


def test_noise_graph_and_mixed(gpu_lib):
    """Graph mode (noise steps run eagerly, the regular steps after them from
    the recorded graph) equals eager mode; MIXED within 1e-4 of FP64."""
    p, phi, cons = make_case("v3_noise", seed=8)
    p.noise_step = p.step + 3
    out = {}
    for key, prec, graph in (("eager", MM_FP64, False), ("graph", MM_FP64, True), ("mixed", MM_MIXED, False)):
        with Simulation(p, prec) as s:
            u0, g0 = _zero(phi)
            s.upload(phi=phi, cons=cons, u=u0, gradu=g0)
            s.set_noise(0.01, -21)
            if graph:
                s.set_graph(True)
            s.step(12)
            out[key] = s.download()
    for k in ("phi", "cons"):
        assert np.array_equal(out["graph"][k], out["eager"][k]), k
        l2 = np.linalg.norm(out["mixed"][k] - out["eager"][k]) / np.linalg.norm(out["eager"][k])
        assert l2 < 1e-4, f"{k} rel L2 {l2:.2e}"


def test_noise_step_needs_a_generator(gpu_lib):
    p, phi, cons = make_case("v1_dilat")
    p.noise_step = p.step
    with Simulation(p, MM_FP64) as s:
        u0, g0 = _zero(phi)
        s.upload(phi=phi, cons=cons, u=u0, gradu=g0)
        with pytest.raises(MMError):
            s.step(1)
        s.set_noise(0.01, -1)
        s.step(2)
