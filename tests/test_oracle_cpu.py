"""Pins of the CPU oracle (the parity reference of the tests) against things the oracle's own
formulas do not restate: numpy's pocketfft and brute-force DFTs, closed-form
single-mode solutions, thermodynamic and mechanical identities, and
invariants.  Each check is chosen so that a plausible mistake (a dropped term,
a sign, an index, a transposed operand, a Nyquist convention) fails it.

The reference ships no tests, golden vectors or fixtures (its only checks are
fatal runtime guards), so these are the pins.
"""
import numpy as np
import pytest

from meumapps_ss_amd.configs import model_inputs
from meumapps_ss_amd.params import setup
from oracle import Oracle
from oracle import pyoracle


def _oracle(n=16, nvar=1, tstr=(0.01, 0.01, 0.01), dt=0.005, **kw):
    p = setup(model_inputs(n, nvar, tstr, dt, **kw))
    return p, Oracle(p, 2)


# ------------------------------------------------------------------- FFT
@pytest.mark.parametrize("shape", [(8, 8, 8), (12, 8, 16), (6, 10, 4), (16, 24, 12)])
def test_f_trans_matches_numpy(shape):
    nz, ny, nx = shape
    p = setup(model_inputs(8, 1, (0, 0, 0), 0.01))
    p.nx, p.ny, p.nz = nx, ny, nz
    o = Oracle(p, 2)
    x = np.random.default_rng(1).standard_normal(shape)
    got = o.f_trans(x)
    ref = np.fft.rfftn(x) / x.size
    assert np.abs(got - ref).max() < 1e-15 * max(1, np.abs(ref).max()) * 50


def test_f_trans_brute_force_dft():
    nz, ny, nx = 4, 6, 8
    p = setup(model_inputs(8, 1, (0, 0, 0), 0.01))
    p.nx, p.ny, p.nz = nx, ny, nz
    o = Oracle(p, 1)
    x = np.random.default_rng(2).standard_normal((nz, ny, nx))
    got = o.f_trans(x)
    z, y, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    for kz in range(nz):
        for ky in range(ny):
            for kx in range(nx // 2 + 1):
                ph = np.exp(-2j * np.pi * (kx * xx / nx + ky * y / ny + kz * z / nz))
                assert abs(got[kz, ky, kx] - (x * ph).sum() / x.size) < 1e-14


def test_inv_trans_c2r_semantics_and_roundtrip():
    nz, ny, nx = 8, 12, 16
    p = setup(model_inputs(8, 1, (0, 0, 0), 0.01))
    p.nx, p.ny, p.nz = nx, ny, nz
    o = Oracle(p, 2)
    rng = np.random.default_rng(3)
    c = rng.standard_normal((nz, ny, nx // 2 + 1)) + 1j * rng.standard_normal((nz, ny, nx // 2 + 1))
    got = o.inv_trans(c)
    t = np.fft.ifft(np.fft.ifft(c, axis=0), axis=1) * ny * nz
    t[..., 0] = t[..., 0].real
    t[..., -1] = t[..., -1].real
    ref = np.fft.irfft(t, n=nx, axis=2) * nx
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    x = rng.standard_normal((nz, ny, nx))
    assert np.abs(o.inv_trans(o.f_trans(x)) - x).max() < 1e-14


# --------------------------------------------------------------- LINPACK
def test_invert_against_numpy_and_singular_flag():
    rng = np.random.default_rng(4)
    for n in (2, 3, 6):
        a = rng.standard_normal((n, n)) + n * np.eye(n)
        rc, inv = pyoracle.invert(a, 9)
        assert rc == 0
        assert np.abs(inv - np.linalg.inv(a)).max() < 1e-13
    rc, _ = pyoracle.invert(np.ones((3, 3)), 9)
    assert rc == 9


# ------------------------------------------------------------- k-space
def test_gradient_single_mode_and_nyquist():
    """grad of cos(2 pi m x/N) is the exact derivative; at m = N/2 kf is zeroed
    (grad = 0) but kf_sq is kept, so div(G grad) = -kappa pi^2 phi
    (ModInitials.f:395-401)."""
    n = 16
    p, o = _oracle(n)
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    for m in (1, 3, n // 2):
        k = 2 * np.pi * m / n
        phi = 0.3 + 0.01 * np.cos(k * x) + 0.02 * np.sin(2 * np.pi * 2 * y / n)
        o = Oracle(p, 2)
        o.set_state(phi=phi[None], cons=np.full((1, n, n, n), 0.4))
        o.step(1)
        g = o.get_work(0)
        kd = o.get_work(1)[0]
        ky = 2 * np.pi * 2 / n
        dx = -0.01 * k * np.sin(k * x) if m != n // 2 else 0.0 * x
        dy = 0.02 * ky * np.cos(ky * y)
        assert np.abs(g[0] - dx).max() < 1e-14
        assert np.abs(g[1] - dy).max() < 1e-14
        assert np.abs(g[2]).max() < 1e-14
        kappa = p.grad_coeff[0]
        lap = -kappa * (k * k * 0.01 * np.cos(k * x) + ky * ky * 0.02 * np.sin(ky * y))
        assert np.abs(kd - lap).max() < 1e-13


def test_gradient_energy_closed_form():
    n = 16
    p, o = _oracle(n)
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    k = 2 * np.pi * 2 / n
    a = 0.05
    o.set_state(phi=(0.4 + a * np.cos(k * z))[None], cons=np.full((1, n, n, n), 0.4))
    d = o.step(1)[0]
    kappa = p.grad_coeff[8]
    assert abs(d["fint"] - 0.5 * kappa * a * a * k * k * n ** 3 / 2) < 1e-10


# ------------------------------------------------------ Cahn-Hilliard
def test_ch_single_mode_decay_factor():
    """phi = 0 makes the flux F vanish, so one step multiplies a Fourier mode of
    c by 1/(1 + dt M |k|^2) and keeps the mean exactly."""
    n = 16
    p, o = _oracle(n, dt=0.01, mob_c=0.7)
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    kv = 2 * np.pi * np.array([1, 2, 3]) / n
    c = 0.4 + 0.01 * np.cos(kv[0] * x + kv[1] * y + kv[2] * z)
    o.set_state(phi=np.zeros((1, n, n, n)), cons=c[None])
    rows = o.step(1)
    cn = o.get_state()["cons"][0]
    fac = 1.0 / (1.0 + 0.01 * 0.7 * (kv ** 2).sum())
    assert np.abs(cn - (0.4 + 0.01 * fac * np.cos(kv[0] * x + kv[1] * y + kv[2] * z))).max() < 1e-15
    assert abs(rows[0]["caviter"][0] - 0.4) < 1e-15


def test_ch_kks_equilibrium_profile_is_stationary():
    """c = x1 + (x2 - x1) p(phi) with a uniform diffusion potential (x1, x2 in
    KKS equilibrium) has zero flux: dc/dt = M div(grad c + F) = 0 up to the
    pointwise chain rule on a smooth grid function.  A flipped sign of F would
    give |dc| ~ 2 dt M |lap p| 0.4."""
    n = 32
    p, o = _oracle(n, dt=0.005)
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    phi = 0.5 + 0.2 * np.cos(2 * np.pi * x / n)
    pp = phi ** 3 * (6 * phi ** 2 - 15 * phi + 10)
    x1 = 0.35
    x2 = x1 + 0.4  # A(x1 - 0.3) = A(x2 - 0.7) with equal curvatures
    c = x1 + (x2 - x1) * pp
    o.set_state(phi=phi[None], cons=c[None])
    o.step(1)
    dc = np.abs(o.get_state()["cons"][0] - c).max()
    lap_p = np.abs(np.fft.irfftn(-(2 * np.pi / n) ** 2 * np.fft.rfftn(pp), s=pp.shape, axes=(0, 1, 2))).max()
    assert dc < 1e-3 * 2 * 0.005 * 1.0 * lap_p * 0.4


# ------------------------------------------------------------- chemistry
def _uniform_energy_and_force(p, phis, c):
    n = p.nx
    o = Oracle(p, 2)
    o.set_state(phi=np.stack([np.full((n, n, n), f) for f in phis]), cons=np.full((1, n, n, n), c))
    d = o.step(1)[0]
    fg = o.get_work(2)[:, 0, 0, 0]
    return (d["fch"] + d["fw"]) / n ** 3, fg


@pytest.mark.parametrize("nvar,numph", [(1, 2), (3, 2), (1, 3)])
def test_driving_force_is_derivative_of_free_energy(nvar, numph):
    """KKS identity: f_grad_v = d f(phi, c)/d phi_v at fixed c, with
    f = sum_p P_p G_p(x_p) + W-wells and x_p from the KKS conditions.
    Central differences of the oracle's own fch + fw densities."""
    p = setup(model_inputs(8, nvar, (0.0, 0.0, 0.0), 0.005, numph=numph))
    V = p.vartot
    base = [0.3, 0.15, 0.6][:V]
    c = 0.45
    h = 1e-5
    _, fg = _uniform_energy_and_force(p, base, c)
    for v in range(V):
        up = list(base); up[v] += h
        dn = list(base); dn[v] -= h
        fu, _ = _uniform_energy_and_force(p, up, c)
        fd, _ = _uniform_energy_and_force(p, dn, c)
        fdiff = (fu - fd) / (2 * h)
        assert abs(fg[v] - fdiff) < 1e-6 * max(1.0, abs(fdiff)), (v, fg[v], fdiff)


def test_f_diff_correction_brute_force():
    """f_diff_v = f_grad_v - sum_{u != v, phi_u > 0}(f_grad_u - divGgrad phi_u)/(1 + count)
    when phi_v > 0, else f_grad_v (Multi_MEUMAPPS.f:824-841), plus el_grad."""
    n = 8
    p = setup(model_inputs(n, 3, (0.02, -0.01, -0.01), 0.005))
    rng = np.random.default_rng(5)
    phi = rng.uniform(-0.3, 1.0, (3, n, n, n))
    o = Oracle(p, 1)
    o.set_state(phi=phi, cons=0.4 + 0.05 * rng.standard_normal((1, n, n, n)))
    o.step(1)
    fg, kd, fdf, eg = o.get_work(2), o.get_work(1), o.get_work(3), o.get_work(4)
    pos = phi > 0
    ref = fg.copy()
    for v in range(3):
        others = [u for u in range(3) if u != v]
        s = sum(np.where(pos[u], fg[u] - kd[u], 0.0) for u in others)
        cnt = 1 + sum(pos[u].astype(int) for u in others)
        ref[v] = np.where(pos[v], fg[v] - s / cnt, fg[v])
    assert np.abs(fdf - eg - ref).max() < 1e-12 * max(1, np.abs(ref).max())


def test_allen_cahn_linear_mode():
    """phi = phi0 + d cos(k.x), d = 1e-7: one step multiplies the mode by
    (1 - dt L f'(phi0)) / (1 + dt L kappa |k|^2), with f' the derivative of
    the uniform driving force (pinned above)."""
    n = 16
    p = setup(model_inputs(n, 1, (0.0, 0.0, 0.0), 0.005, mob_eta=0.8))
    phi0, c, d = 0.5, 0.45, 1e-7
    h = 1e-5
    _, fu = _uniform_energy_and_force(p, [phi0 + h], c)
    _, fd = _uniform_energy_and_force(p, [phi0 - h], c)
    fprime = (fu[0] - fd[0]) / (2 * h)
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    k = 2 * np.pi * np.array([2, 1, 0]) / n
    mode = np.cos(k[0] * x + k[1] * y)
    o = Oracle(p, 2)
    o.set_state(phi=(phi0 + d * mode)[None], cons=np.full((1, n, n, n), c))
    o.step(1)
    new = o.get_state()["phi"][0]
    amp = (new * mode).sum() / (mode * mode).sum()
    kappa = p.grad_coeff[0]
    dtl = 0.005 * 0.8
    ref = d * (1 - dtl * fprime) / (1 + dtl * kappa * (k ** 2).sum())
    assert abs(amp - ref) < 1e-4 * abs(ref)


# ------------------------------------------------------------ elasticity
def _stress_and_divergence(p, st, phi, P=None):
    """sigma = C (eps(u) - sum P eps0) for homogeneous C (P = p(phi) unless
    given); returns max |k_j sigma_ij(k)| over non-Nyquist k relative to
    max |k||sigma(k)|."""
    C = p.cela_array()[1]
    V = p.vartot
    if P is None:
        P = phi ** 3 * (6 * phi ** 2 - 15 * phi + 10)
    t = p.trans_ppt_array()
    g = st["gradu"]  # g[d*3 + i] = d_d u_i
    eps = np.empty((6,) + phi.shape[1:])
    eps[0], eps[1], eps[2] = g[0], g[4], g[8]
    eps[3] = g[5] + g[7]
    eps[4] = g[2] + g[6]
    eps[5] = g[1] + g[3]
    w = np.array([1, 1, 1, 2, 2, 2.0])
    for m in range(6):
        eps[m] -= w[m] * sum(P[v] * t[v, m] for v in range(V))
    sig = np.einsum("ab,b...->a...", C, eps)
    n = phi.shape[1]
    S = np.fft.fftn(sig, axes=(1, 2, 3))
    kk = np.fft.fftfreq(n) * 2 * np.pi
    kz, ky, kx = np.meshgrid(kk, kk, kk, indexing="ij")
    nyq = (np.abs(kx) == np.pi) | (np.abs(ky) == np.pi) | (np.abs(kz) == np.pi)
    full = [[0, 5, 4], [5, 1, 3], [4, 3, 2]]
    kv = [kx, ky, kz]
    worst, scale = 0.0, 0.0
    for i in range(3):
        div = sum(kv[j] * S[full[i][j]] for j in range(3))
        worst = max(worst, np.abs(np.where(nyq, 0, div)).max())
        scale = max(scale, np.abs(S[full[i][0]] * kx).max())
    return worst / scale


@pytest.mark.parametrize("nvar,tstr,hi", [(1, (0.01, 0.01, 0.01), 1.0), (1, (0.02, -0.01, 0.005), 1.0),
                                          (3, (0.02, -0.01, -0.01), 0.3)])
def test_elastic_solution_is_in_equilibrium(nvar, tstr, hi):
    """With homogeneous moduli and sum p(phi) < 1 the compliance mix is exactly
    the matrix compliance, and after eldis the stress C (eps(u) - eps0 p) is
    divergence-free in every non-Nyquist mode (Khachaturyan microelasticity):
    pins Q(k), R = kf.x, the Voigt maps and the signs."""
    from parity_cases import smooth_field

    n = 16
    p = setup(model_inputs(n, nvar, tstr, 0.005))
    rng = np.random.default_rng(6)
    phi = np.stack([smooth_field(rng, (n, n, n), 0.0, hi) for _ in range(nvar)])
    o = Oracle(p, 2)
    o.set_state(phi=phi, cons=np.full((1, n, n, n), 0.4))
    rows = o.step(1)
    assert rows[0]["n_elastic"] == 4  # converged at once, minimum of 4 iterations
    st = o.get_state()
    assert _stress_and_divergence(p, st, phi) < 1e-12


@pytest.mark.parametrize("nvar,tstr", [(1, (0.02, -0.01, 0.005)), (3, (0.02, -0.01, -0.01))])
def test_zeroth_order_guess_is_the_equilibrium_of_the_even_part(nvar, tstr):
    """The zeroth-order guess of step == elastep0 (ModElastic.f:105-236) sums
    only Re(P_hat) (sum_0 is real*8, :29,151) and takes R = i kf . sum_0.  Re
    P_hat is the transform of the even part P_e(x) = (P(x) + P(-x))/2, so u0
    is the equilibrium displacement of P_e: C(eps(u0) - eps0 P_e) is
    divergence-free in every non-Nyquist mode, and against P itself it is
    not (the odd part is dropped).  Pins the truncation, the sign of R and
    Q(k) of the guess."""
    from parity_cases import smooth_field

    n = 16
    p = setup(model_inputs(n, nvar, tstr, 0.005))
    rng = np.random.default_rng(21)
    phi = np.stack([smooth_field(rng, (n, n, n), 0.0, 0.9) for _ in range(nvar)])
    P = phi ** 3 * (6 * phi ** 2 - 15 * phi + 10)
    o = Oracle(p, 2)
    o.set_state(phi=phi, cons=np.full((1, n, n, n), 0.4))
    _, gu = o.elastic_zeroth(P)
    r = (-np.arange(n)) % n
    Pe = 0.5 * (P + P[:, r][:, :, r][:, :, :, r])
    assert _stress_and_divergence(p, dict(gradu=gu), phi, Pe) < 1e-12
    assert _stress_and_divergence(p, dict(gradu=gu), phi, P) > 1e-3


def test_heterogeneous_fixed_point_is_in_equilibrium():
    """Inhomogeneous moduli: the converged fixed point of ModElastic.f:241-481
    satisfies div[C(x)(eps(u) - eps0 p)] = 0 with C(x) = (sum P S_p + (1-sum P) S_1)^-1,
    to the 1e-7 convergence tolerance of the iteration."""
    from parity_cases import smooth_field

    n = 16
    p = setup(model_inputs(n, 1, (0.01, 0.005, 0.01), 0.005,
                           cela2={(1, 1): 300.0, (1, 2): 140.0, (4, 4): 120.0}))
    rng = np.random.default_rng(9)
    phi = smooth_field(rng, (n, n, n), 0.0, 1.0)[None]
    o = Oracle(p, 2)
    o.set_state(phi=phi, cons=np.full((1, n, n, n), 0.4))
    rows = o.step(1)
    assert 4 <= rows[0]["n_elastic"] <= 10
    st = o.get_state()
    P = phi[0] ** 3 * (6 * phi[0] ** 2 - 15 * phi[0] + 10)
    S1, S2 = p.sela_array()[0], p.sela_array()[1]
    M = P[..., None, None] * S2 + (1 - P)[..., None, None] * S1
    Cx = np.linalg.inv(M)
    g = st["gradu"]
    t = p.trans_ppt_array()[0]
    eps = np.stack([g[0] - P * t[0], g[4] - P * t[1], g[8] - P * t[2], g[5] + g[7] - 2 * P * t[3],
                    g[2] + g[6] - 2 * P * t[4], g[1] + g[3] - 2 * P * t[5]], axis=-1)
    sig = np.moveaxis(np.einsum("...ab,...b->...a", Cx, eps), -1, 0)
    Sg = np.fft.fftn(sig, axes=(1, 2, 3))
    kk = np.fft.fftfreq(n) * 2 * np.pi
    kz, ky, kx = np.meshgrid(kk, kk, kk, indexing="ij")
    nyq = (np.abs(kx) == np.pi) | (np.abs(ky) == np.pi) | (np.abs(kz) == np.pi)
    full = [[0, 5, 4], [5, 1, 3], [4, 3, 2]]
    kv = [kx, ky, kz]
    worst = max(np.abs(np.where(nyq, 0, sum(kv[j] * Sg[full[i][j]] for j in range(3)))).max() for i in range(3))
    scale = max(np.abs(Sg[full[i][0]] * kx).max() for i in range(3))
    assert worst / scale < 1e-3


def test_elastic_energy_hooke_closed_form():
    """Homogeneous moduli: el_en = 1/2 eps:C:eps (matrix and precipitate terms
    add up to one Hooke energy when every variant has trans_ppt sum > 1e-9)."""
    n = 8
    p = setup(model_inputs(n, 1, (0.01, 0.02, 0.005), 0.005))
    rng = np.random.default_rng(7)
    phi = rng.uniform(0.0, 1.0, (1, n, n, n))
    o = Oracle(p, 1)
    o.set_state(phi=phi, cons=np.full((1, n, n, n), 0.4))
    o.step(1)
    st = o.get_state()
    C = p.cela_array()[1]
    P = phi[0] ** 3 * (6 * phi[0] ** 2 - 15 * phi[0] + 10)
    g = st["gradu"]
    t = p.trans_ppt_array()[0]
    eps = np.stack([g[0] - P * t[0], g[4] - P * t[1], g[8] - P * t[2], g[5] + g[7], g[2] + g[6], g[1] + g[3]])
    ref = 0.5 * np.einsum("a...,ab,b...->...", eps, C, eps)
    assert np.abs(st["el_en"] - ref).max() < 1e-12 * np.abs(ref).max()


def test_traceless_eigenstrain_switches_el_grad_off():
    """trans_ppt summing to 0 fails the > 1e-9 guard (ModElastic.f:653)."""
    n = 8
    p = setup(model_inputs(n, 3, (0.02, -0.01, -0.01), 0.005))
    o = Oracle(p, 1)
    o.set_state(phi=np.random.default_rng(8).uniform(0, 1, (3, n, n, n)), cons=np.full((1, n, n, n), 0.4))
    o.step(1)
    assert np.abs(o.get_work(4)).max() == 0.0


# eigenstrains with every Voigt component non-zero (tensor shear), trace > 0
# so the transsum > 1e-9 guard of ModElastic.f:653 keeps el_grad on
_SHEAR_EPS0 = ((0.012, -0.004, 0.006, 0.003, -0.002, 0.005),
               (0.004, 0.010, -0.003, -0.004, 0.006, 0.002))


def _elener_params(kind, n):
    """dilat: one variant, dilatational, homogeneous moduli; shear: two
    variants of one phase with full (shear-carrying) eigenstrains; twophase:
    two precipitate phases, the first with its own cubic moduli, both with
    shear-carrying eigenstrains."""
    if kind == "dilat":
        return setup(model_inputs(n, 1, (0.01, 0.01, 0.01), 0.005))
    if kind == "shear":
        p = setup(model_inputs(n, 2, (0.01, 0.01, 0.01), 0.005))
    else:
        p = setup(model_inputs(n, 1, (0.01, 0.01, 0.01), 0.005, numph=3,
                               cela2={(1, 1): 320.0, (1, 2): 130.0, (4, 4): 70.0}))
    for v in range(2):
        for m in range(6):
            p.trans_ppt[v * 6 + m] = _SHEAR_EPS0[v][m]
    return p


@pytest.mark.parametrize("kind", ["dilat", "shear", "twophase"])
def test_el_grad_is_the_phi_derivative_of_el_en(kind):
    """elener's el_grad (ModElastic.f:653-678) is the derivative of the local
    elastic energy density el_en = 1/2 sigma : S(phi) : sigma with respect to
    phi_v at fixed grad u: p'(phi_v) [(el_mat - el_ppt_v) - sigma : eps0_v]
    (engineering shear, so the shear terms of sigma : eps0 count twice).
    Checked against the central difference of the oracle's own el_en, point
    by point, for a dilatational eigenstrain, shear-carrying eigenstrains and
    a second precipitate phase with other moduli (el_mat != el_ppt).  A
    flipped sign of the sigma : eps0 term, a dropped factor 2 on its shear
    terms or a flipped (el_mat - el_ppt) fails by orders of magnitude."""
    n = 8
    p = _elener_params(kind, n)
    V = p.vartot
    rng = np.random.default_rng(31)
    phi = rng.uniform(0.05, 0.45, (V, n, n, n))
    gu = rng.uniform(-0.02, 0.02, (9, n, n, n))
    cons = np.full((p.numel, n, n, n), 0.4)
    o = Oracle(p, 1)

    def run(ph):
        o.set_state(phi=ph, cons=cons, gradu=gu)
        o.set_pphi()
        rc, en, eg = o.elener()
        assert rc == 0
        return en, eg

    en0, eg0 = run(phi)
    # el_en is the Hooke energy 1/2 eps_el : C(phi) : eps_el with the compliance mix
    P = phi ** 3 * (6 * phi ** 2 - 15 * phi + 10)
    S = p.sela_array()
    Smix = (1 - P.sum(0))[..., None, None] * S[0]
    ph_of_v = [ph for ph in range(1, p.numph) for _ in range(p.nvar[ph])]
    for v in range(V):
        Smix = Smix + P[v][..., None, None] * S[ph_of_v[v]]
    t = p.trans_ppt_array()
    w = np.array([1, 1, 1, 2, 2, 2.0])
    g = gu
    eps = np.stack([g[0], g[4], g[8], g[5] + g[7], g[2] + g[6], g[1] + g[3]], axis=-1)
    eps = eps - np.einsum("v...,vm->...m", P, t) * w
    sig = np.linalg.solve(Smix, eps[..., None])[..., 0]
    assert np.abs(en0 - 0.5 * (sig * eps).sum(-1)).max() < 1e-12 * np.abs(en0).max()
    h = 1e-5
    for v in range(V):
        dp = np.zeros_like(phi)
        dp[v] = h
        fd = (run(phi + dp)[0] - run(phi - dp)[0]) / (2 * h)
        scale = np.abs(eg0[v]).max()
        assert scale > 1e-3
        assert np.abs(fd - eg0[v]).max() < 1e-7 * scale, (kind, v)


@pytest.mark.parametrize("kind", ["dilat", "shear"])
def test_el_grad_gives_the_equilibrium_energy_change(kind):
    """Envelope theorem: at elastic equilibrium (homogeneous moduli, where
    eldis converges to the exact equilibrium) the change of the total elastic
    energy sum el_en under a local change delta phi equals
    sum el_grad . delta phi, because the energy is stationary in u.  The
    equilibrium is recomputed by eldis (ModElastic.f:241-481) at phi +- h
    delta phi; el_grad comes from the base state only."""
    from parity_cases import smooth_field

    n = 16
    p = _elener_params(kind, n)
    V = p.vartot
    rng = np.random.default_rng(33)
    phi = np.stack([smooth_field(rng, (n, n, n), 0.0, 0.45) for _ in range(V)])
    cons = np.full((p.numel, n, n, n), 0.4)
    z = np.arange(n)
    zz, yy, xx = np.meshgrid(z, z, z, indexing="ij")
    o = Oracle(p, 2)

    def solve(ph):
        o.set_state(phi=ph, cons=cons, u=np.zeros((3, n, n, n)), gradu=np.zeros((9, n, n, n)))
        o.set_step(p.elastep0 + 1)
        o.set_pphi()
        rc, nit, _ = o.eldis()
        assert rc == 0
        return o.get_state()["el_en"], o.get_work(4)

    en0, eg0 = solve(phi)
    for v in range(V):
        c = (3 + 5 * v, 7, 10 - 2 * v)
        r2 = (zz - c[0]) ** 2 + (yy - c[1]) ** 2 + (xx - c[2]) ** 2
        bump = np.zeros_like(phi)
        bump[v] = np.exp(-r2 / 8.0)
        lin = float((eg0 * bump).sum())
        h = 1e-5
        fd = (solve(phi + h * bump)[0].sum() - solve(phi - h * bump)[0].sum()) / (2 * h)
        assert abs(lin) > 1e-3
        assert abs(fd - lin) < 1e-6 * abs(lin), (kind, v, fd, lin)


# ------------------------------------------------------------ invariants
def test_mass_clamp_and_lagged_diagnostics():
    from parity_cases import make_case

    p, phi, cons = make_case("v3_tetra")
    o = Oracle(p, 2)
    o.set_state(phi=phi, cons=cons)
    rows = o.step(4)
    st = o.get_state()
    assert st["phi"].min() >= 0.0 and st["phi"].max() <= 1.0
    for r in rows:
        assert abs(r["caviter"][0] - cons.mean()) < 1e-14
    # fort.98 columns glob_min .. solid_tot lag one step behind the check loop
    assert rows[0]["glob_min"] == 1.0 and rows[0]["phi_max"] == 0.0
    o2 = Oracle(p, 2)
    o2.set_state(phi=phi, cons=cons)
    o2.step(1)
    s1 = o2.get_state()
    assert abs(rows[1]["phi_max"] - s1["phi"].max()) == 0.0
    assert abs(rows[1]["glob_max"] - max(0.0, s1["cons"].max())) == 0.0
