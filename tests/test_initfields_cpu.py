"""InitFields / NucAdd on the host (ModInitials.f:415-826), one process.

Pins: the single nucleus against the lattice-point count of a ball and the
exact rational ellipsoid test; the random field against the exact ran_2
stream; the multiple and added nuclei against their invariants (disjoint
nuclei, the requested count, variants at 1, the clearance box, the last
variant cleared) and the composition switch of 'cppt'."""
import numpy as np
import pytest
from scipy import ndimage

from meumapps_ss_amd.initfields import Ran2, init_fields, nuc_add
from meumapps_ss_amd.params import Inputs
from test_abi_cpu import _ran2_bigint


def _inp(n=16, V=2, **kw):
    inp = Inputs()
    inp.nx = inp.ny = inp.nz = n
    inp.vartot = V
    inp.numel_in = 2
    inp.conc = [(1, 0.4, 0.9, 0.0)]
    for k, v in kw.items():
        setattr(inp, k, v)
    return inp


def test_single_sphere_lattice_count():
    inp = _inp(16, 2, nucstat="single", targv=2, radppt=(4, 4, 4), phi_init=0.0)
    phi, cons = init_fields(inp, Ran2(-1))
    assert (phi[1] == 1.0).sum() == 257  # integer points with i^2 + j^2 + k^2 <= 16
    assert not phi[0].any() and cons.min() == cons.max() == 0.4
    # centred on the reference's mnx = Nx/2 (1-based) -> index 7
    assert phi[1, 7, 7, 7] == 1.0 and phi[1, 7, 7, 11] == 1.0 and phi[1, 7, 7, 12] == 0.0


def test_single_ellipsoid_matches_exact_test():
    r = (3, 5, 2)
    inp = _inp(16, 1, nucstat="single", targv=0, radppt=r, phi_init=0.05)
    phi, _ = init_fields(inp, Ran2(-1))
    k, j, i = np.meshgrid(*(np.arange(16) - 7,) * 3, indexing="ij")  # centre mnx = 8 (1-based)
    a, b, c = r
    exact = i * i * b * b * c * c + j * j * a * a * c * c + k * k * a * a * b * b <= (a * b * c) ** 2
    assert np.array_equal(phi[0] == 1.0, exact)
    assert np.all(phi[0][~exact] == 0.05)


def test_random_field_follows_the_stream():
    n, V = 8, 3
    inp = _inp(n, V, nucstat="random", phi_init=0.0)
    rng = Ran2(-9)
    phi, _ = init_fields(inp, rng)
    u = np.array(_ran2_bigint(-9, n ** 3)).reshape(n, n, n)  # x fastest
    var = (u * V + 1.0).astype(int)
    for v in range(V):
        assert np.array_equal(phi[v] == 0.1, var == v + 1)
    assert np.all((phi == 0.1).sum(axis=0) == 1)
    ref = Ran2(-9)
    ref.fill(n ** 3)
    assert rng.snapshot() == ref.snapshot()


def test_multiple_nuclei_and_cppt():
    inp = _inp(24, 3, nucstat="multiple", targv=0, nmaxppt=6, var_lo=0, var_hi=3, radppt=(2, 2, 1),
               phi_init=0.0, constat="none")
    phi, _ = init_fields(inp, Ran2(-4))
    assert set(np.unique(phi)) <= {0.0, 1.0}
    assert np.all((phi > 0).sum(axis=0) <= 1)  # no point in two nuclei
    lab, count = ndimage.label((phi > 0).any(axis=0))
    assert count == 6
    # 'cppt' compositions with a target variant
    inp2 = _inp(24, 3, nucstat="multiple", targv=2, nmaxppt=3, radppt=(2, 2, 2), phi_init=0.0, constat="cppt")
    phi2, cons2 = init_fields(inp2, Ran2(-5))
    assert not phi2[0].any() and not phi2[2].any()
    inside = phi2[1] > 0.9
    assert np.all(cons2[0][inside] == 0.9) and np.all(cons2[0][~inside] == 0.4)


def test_nucadd_addmult():
    n, V = 32, 3
    inp = _inp(n, V, nucstat="addmult", targv=0, nmaxppt=4, radppt=(2, 2, 2), addrad=(2, 1, 3))
    phi = np.zeros((V, n, n, n))
    phi[V - 1] = 0.005  # below the 0.01 test: does not block
    nuc_add(inp, phi, Ran2(-13))
    assert not (phi[V - 1] == 1.0).any()  # randvar in 1..V-1, the last variant cleared inside
    lab, count = ndimage.label((phi[: V - 1] > 0).any(axis=0))
    assert count == 4
    ball = (phi[: V - 1] == 1.0).any(axis=0)
    assert not (phi[V - 1][ball]).any()
    # centres stay inside the clearance box: no nucleus touches the faces
    zz, yy, xx = np.nonzero(ball)
    assert xx.min() >= 1 and yy.min() >= 1 and zz.min() >= 1
    assert xx.max() <= n - 2 and yy.max() <= n - 2 and zz.max() <= n - 2


def test_nucadd_other_status_is_a_no_op():
    inp = _inp(8, 2, nucstat="single")
    phi = np.full((2, 8, 8, 8), 0.3)
    rng = Ran2(-3)
    nuc_add(inp, phi, rng)
    assert np.all(phi == 0.3) and rng.snapshot() == Ran2(-3).snapshot()
