"""The reference's fatal error flags as return codes of mm_step.

The reference stops (ModFunc.f:52-67, ModElastic.f:446-450) when a small
inverse fails; the C ABI returns the same err_flag value from mm_step instead
(include/meumapps_ss.h MM_ERR_*).  Each case makes the matrix exactly
singular (or the backtracking give up) in the first step, runs the oracle and
the FP64 GPU step on the same inputs and expects the same code from both:

* 2  KKS Zintf (Multi_MEUMAPPS.f:690): matrix phase with A = 0 and phi = 1 at
     one point, so the first column of Zintf is zero there;
* 3  acoustic tensor (ModElastic.f:187-189): precipitate moduli with C11 <= 0,
     which InitElasticMat replaces by zero (ModInitials.f:84-87), so Q(k) = 0;
* 6  compliance mix of elener (ModElastic.f:597-601): two variants with
     p(phi) = 1 at one point and precipitate moduli exactly twice the matrix
     ones, so sum p S_p + (1 - sum p) S_1 = 2 S_p - S_1 = 0 there;
* 7  elastic backtracking (ModElastic.f:446-450): precipitates at 0.15 x the
     matrix moduli make every relaxation factor the ramp returns to diverge
     again until newfac < 1e-7.

After an error the context takes a valid upload and steps again normally.
"""
import ctypes

import numpy as np
import pytest

from meumapps_ss_amd.configs import model_inputs
from meumapps_ss_amd.params import MM_FP64, MMDiag, setup
from meumapps_ss_amd.sim import Simulation
from oracle import Oracle
from parity_cases import smooth_field

pytestmark = pytest.mark.gpu


def _case(name):
    n = 8
    if name == "zsolve":
        inp = model_inputs(n, 1, (0.01, 0.01, 0.01), 0.005)
        inp.phases[0].gparams[0] = 0.0
        p = setup(inp)
        phi = np.full((1, n, n, n), 0.2)
        phi[0, 3, 4, 5] = 1.0
        return p, phi, np.full((1, n, n, n), 0.4), 2
    if name == "qinv":
        p = setup(model_inputs(n, 1, (0.01, 0.01, 0.01), 0.005,
                               cela2={(1, 1): -1.0, (1, 2): 150.0, (4, 4): 100.0}))
        return p, np.full((1, n, n, n), 0.2), np.full((1, n, n, n), 0.4), 3
    if name == "cinv_ener":
        p = setup(model_inputs(n, 2, (0.01, 0.01, 0.01), 0.005,
                               cela2={(1, 1): 500.0, (1, 2): 300.0, (4, 4): 200.0}))
        assert np.array_equal(2 * p.sela_array()[1], p.sela_array()[0])
        phi = np.full((2, n, n, n), 0.1)
        phi[:, 3, 4, 5] = 1.0
        return p, phi, np.full((1, n, n, n), 0.4), 6
    if name == "newfac":
        n, f = 16, 0.15
        p = setup(model_inputs(n, 3, (0.02, -0.01, 0.01), 0.005,
                               cela2={(1, 1): 250 * f, (1, 2): 150 * f, (4, 4): 100 * f}))
        rng = np.random.default_rng(3)
        shape = (n, n, n)
        phi = np.stack([smooth_field(rng, shape, -0.15, 1.1) for _ in range(3)])
        cons = np.stack([smooth_field(rng, shape, 0.3, 0.6)])
        return p, phi, cons, 7
    raise ValueError(name)


@pytest.mark.parametrize("name", ["zsolve", "qinv", "cinv_ener", "newfac"])
def test_error_flag_returned_like_the_oracle(gpu_lib, name):
    p, phi, cons, code = _case(name)
    o = Oracle(p, 4)
    o.set_state(phi=phi, cons=cons)
    rc_o = o.lib.oracle_step(o.h, ctypes.byref(MMDiag()))
    o.close()
    assert rc_o == code
    shape = phi.shape[1:]
    zeros = dict(u=np.zeros((3, *shape)), gradu=np.zeros((9, *shape)))
    with Simulation(p, MM_FP64) as s:
        s.upload(phi=phi, cons=cons, **zeros)
        rc = s.lib.mm_step(s.ctx, 1, None)
        assert rc == code, f"{name}: mm_step returned {rc}, the oracle {code}"
        # the context recovers: valid fields (phi in [0, 0.3], away from the
        # singular points) step normally, unless the inputs themselves are
        # singular for every field (the acoustic tensor)
        if name != "qinv":
            s.upload(phi=np.full_like(phi, 0.05), cons=np.full_like(cons, 0.4), **zeros)
            s.set_step(p.elastep0 + 1)
            if name == "newfac":
                return
            rc = s.lib.mm_step(s.ctx, 1, None)
            assert rc == 0, f"{name}: step after the error returned {rc}"
        else:
            rc = s.lib.mm_step(s.ctx, 1, None)
            assert rc == code
