"""The oracle's tangent-linear step (oracle_tlm_*) against central differences
of the oracle's own step.

The tangent runs in lockstep with the primal and takes the primal's
decisions (the phi > 0 tests of the f_diff correction,
Multi_MEUMAPPS.f:824-841; the clamp, :1034-1035; the elastic iteration count
and cfac sequence, ModElastic.f:433-488), so it is the derivative of the
frozen-decision map.  Each check runs the base, +h and -h runs, asserts that
they took the same decisions at every step (the clamp pattern of phi after
each step, which also fixes the next step's phi > 0 tests, and the elastic
iteration counts and cfac), and then compares the tangent of phi, c and
grad u after the run with the central difference.  The direction moves the
state and every differentiated parameter array (grad_coeff, wellph, cmobs,
b_mob, trans_ppt); for variants whose eigenstrain components sum to zero
(where elener's ``transsum > 1e-9`` guard, ModElastic.f:640, switches el_grad
off) the trans_ppt direction keeps that sum zero.
"""
import numpy as np
import pytest

from oracle import Oracle
from parity_cases import make_case, smooth_field
from tangent_dirs import direction, moved, run

NSTEPS = 4
H = 1e-5
TOL = 1e-7


def check(p, phi0, cons0, seed=1, nsteps=NSTEPS, h=H):
    rng = np.random.default_rng(seed)
    dphi = rng.standard_normal(phi0.shape) * 0.01
    dcons = rng.standard_normal(cons0.shape) * 0.001
    dp = direction(p, rng)
    _, t, pat0, ctl0 = run(p, phi0, cons0, nsteps, (dphi, dcons, None, dp))
    su, _, patu, ctlu = run(moved(p, dp, h), phi0 + h * dphi, cons0 + h * dcons, nsteps)
    sd, _, patd, ctld = run(moved(p, dp, -h), phi0 - h * dphi, cons0 - h * dcons, nsteps)
    assert ctlu == ctl0 and ctld == ctl0, "elastic control differs between the +-h runs"
    for a, b, c in zip(pat0, patu, patd):
        assert np.array_equal(a, b) and np.array_equal(a, c), "a clamp decision differs between the +-h runs"
    errs = {}
    for k in ("phi", "cons", "gradu"):
        fd = (su[k] - sd[k]) / (2 * h)
        scale = np.abs(t[k]).max()
        assert scale > 0, k
        errs[k] = np.abs(fd - t[k]).max() / scale
    assert max(errs.values()) < TOL, errs
    return ctl0, errs


@pytest.mark.parametrize("name", ["v2_twophase", "v3_tetra", "v3_inhomog", "v1_e2", "v1_trigC", "v12_rot"])
def test_tangent_matches_frozen_differences_on_parity_cases(name):
    """Smooth initial fields over [-0.15, 1.1]: the correction term, clamped
    points, nv = 3 / 12 tetragonal and rotated eigenstrains (zero eigenstrain
    sum: el_grad off), inhomogeneous moduli (cubic and trigonal precipitate),
    two compositions."""
    p, phi, cons = make_case(name)
    check(p, phi, cons)


def _two_variant(n, eta0, amp, seed=3):
    from meumapps_ss_amd.configs import Workload, initial_fields, model_inputs
    from meumapps_ss_amd.params import setup
    inp = model_inputs(n, 1, (0.02, 0.01, 0.01), 0.005, numph=3, seed=seed)
    phi, cons = initial_fields(Workload("two precipitate phases", inp, seed, amp, (0.4,), NSTEPS))
    return setup(inp), phi + eta0, cons


def test_tangent_at_the_clamp():
    """The differentiation workload's start (two precipitate phases,
    eta = U[-0.01, 0.01]): about half of the points sit at the clamp and the
    phi > 0 tests decide the correction term at every point."""
    p, phi, cons = _two_variant(16, 0.0, 0.01)
    check(p, phi, cons, nsteps=6)


def test_tangent_where_the_mixture_exceeds_one():
    """eta = 0.4 + U[-0.3, 0.3] in two variants: sum_v P(phi_v) > 1 at about a
    fifth of the points, where the polarization mixture drops its matrix term
    (ModElastic.f:262-270) while elener keeps (1 - sum P) S (:590-597)."""
    p, phi, cons = _two_variant(16, 0.4, 0.3)
    P = phi ** 3 * (6 * phi ** 2 - 15 * phi + 10)
    assert (P.sum(0) > 1).mean() > 0.1
    check(p, phi, cons)


def test_tangent_through_backtracking():
    """Soft precipitates (0.24 x the matrix moduli): the elastic loop of the
    first step diverges, cfac is cut and ramped back (ModElastic.f:438-481);
    the tangent follows every iteration with that iteration's cfac."""
    from meumapps_ss_amd.configs import model_inputs
    from meumapps_ss_amd.params import setup
    n, f = 16, 0.24
    p = setup(model_inputs(n, 3, (0.02, -0.01, 0.01), 0.005,
                           cela2={(1, 1): 250 * f, (1, 2): 150 * f, (4, 4): 100 * f}))
    rng = np.random.default_rng(3)
    phi = np.stack([smooth_field(rng, (n, n, n), -0.15, 1.1) for _ in range(3)])
    cons = np.stack([smooth_field(rng, (n, n, n), 0.3, 0.6)])
    ctl, _ = check(p, phi, cons, nsteps=2)
    assert ctl[0][0] > 10 and ctl[0][1] == 1.0


def test_tangent_is_linear_and_refuses_unsupported_steps():
    """Tangent of a sum of directions = sum of tangents (to rounding); a noise
    step or the zeroth-order step is refused while a tangent is set."""
    p, phi, cons = make_case("v2_twophase")
    rng = np.random.default_rng(5)
    a = [rng.standard_normal(phi.shape) * 0.01 for _ in range(2)]
    outs = []
    for dphi in (a[0], a[1], a[0] + a[1]):
        _, t, _, _ = run(p, phi, cons, 2, (dphi, None, None, None))
        outs.append(t["phi"])
    assert np.abs(outs[0] + outs[1] - outs[2]).max() < 1e-12 * np.abs(outs[2]).max()
    o = Oracle(p, 1)
    o.set_state(phi=phi, cons=cons)
    o.set_step(p.elastep0)
    o.tlm_set(a[0])
    with pytest.raises(RuntimeError, match="20001"):
        o.step(1)
    o.close()
