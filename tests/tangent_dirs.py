"""Directions and runs for the tangent-linear checks: a random direction in
the differentiated parameter arrays, parameters moved along it, and an oracle
run from a restart with u = 0 that can carry a tangent."""
import numpy as np

from meumapps_ss_amd.params import MMAdjGrads
from oracle import Oracle


def direction(p, rng):
    """Random parameter direction, scaled to each entry's magnitude."""
    d = MMAdjGrads()
    V = p.vartot
    for i in range(9 * V):
        d.grad_coeff[i] = rng.standard_normal() * abs(p.grad_coeff[i])
    for i in range(p.numph):
        d.wellph[i] = rng.standard_normal() * abs(p.wellph[i])
    for i in range(p.numel):
        d.cmobs[i] = rng.standard_normal() * abs(p.cmobs[i])
    d.b_mob = rng.standard_normal() * abs(p.b_mob)
    for v in range(V):
        tp = np.array(p.trans_ppt[6 * v:6 * v + 6])
        dt = rng.standard_normal(6) * np.abs(tp)
        if not tp.sum() > 1e-9:
            dt -= dt.mean()
        for m in range(6):
            d.trans_ppt[6 * v + m] = dt[m]
    return d


def moved(p, d, h):
    q = type(p).from_buffer_copy(p)
    for f in ("grad_coeff", "wellph", "cmobs", "trans_ppt"):
        a, b = getattr(q, f), getattr(d, f)
        for i in range(len(b)):
            a[i] = a[i] + h * b[i]
    q.b_mob = q.b_mob + h * d.b_mob
    return q


def run(p, phi0, cons0, nsteps, tangent=None, nthreads=1):
    """nsteps regular steps from a restart with u = 0 at elastep0 + 1;
    returns (final state, tangent or None, clamp patterns, (n_elastic, cfac))."""
    o = Oracle(p, nthreads)
    shape = phi0.shape[1:]
    o.set_state(phi=phi0, cons=cons0, u=np.zeros((3, *shape)), gradu=np.zeros((9, *shape)))
    o.set_step(p.elastep0 + 1)
    if tangent is not None:
        o.tlm_set(*tangent)
    pats, ctl = [], []
    for _ in range(nsteps):
        r = o.step(1)[0]
        ctl.append((r["n_elastic"], r["cfac"]))
        ph = o.get_state()["phi"]
        pats.append(np.packbits((ph == 0.0) | (ph == 1.0)))
    st = o.get_state()
    t = o.tlm_get() if tangent is not None else None
    o.close()
    return st, t, pats, ctl
