"""Langevin noise (Multi_MEUMAPPS.f:552-571, 963-978) on the CPU.

The uniform stream ran_2 (ran_2.f) of the oracle against exact modular
recurrences; the product's host draws (mm_langevin) against the oracle's; the
polar Box-Muller law of langevin (ModFunc.f:77-100); and the oracle's noise
step from phi = 0, where the step reduces to a closed form: f_diff = 0, so
phi = clamp(IFFT(dt F(phi_noise) / (1 - dt L k.G.k))).
"""
import ctypes
import math

import numpy as np
import pytest

from meumapps_ss_amd import _lib
from meumapps_ss_amd.configs import model_inputs
from meumapps_ss_amd.params import setup
from oracle import Oracle, pyoracle
from test_abi_cpu import _ran2_bigint


def _stream(fn, seed, n, *extra):
    idum = ctypes.c_int32(seed)
    st = (ctypes.c_int32 * 35)()
    out = [fn(ctypes.byref(idum), st, *extra) for _ in range(n)]
    return out, idum.value, list(st)


@pytest.mark.parametrize("seed", [-7, -1, -123456])
def test_oracle_ran2_matches_exact_recurrences(seed):
    got, _, _ = _stream(pyoracle.load().oracle_ran2, seed, 3000)
    assert got == _ran2_bigint(seed, 3000)


@pytest.mark.parametrize("seed", [-7, 0, 12345])  # 12345 > 0: no table initialisation
def test_langevin_product_matches_oracle(seed):
    a = _stream(pyoracle.load().oracle_langevin, seed, 20000, 0.3)
    b = _stream(_lib.load().mm_langevin, seed, 20000, 0.3)
    assert a == b


def test_langevin_polar_law():
    """Outputs from the uniform stream: pairs mapped to (-1, 1)^2, rejected
    outside the open unit disc, v2 sqrt(-2 ln r / r) sd kept; acceptance
    pi/4; mean 0 and standard deviation sd."""
    sd, n = 0.25, 100000
    lib = pyoracle.load()
    got, _, _ = _stream(lib.oracle_langevin, -5, n, sd)
    uni = _ran2_bigint(-5, int(2.6 * n) + 1000)
    out, used, i = [], 0, 0
    while len(out) < n:
        v1, v2 = 2.0 * uni[i] - 1.0, 2.0 * uni[i + 1] - 1.0
        i += 2
        used += 1
        r = v1 * v1 + v2 * v2
        if 0.0 < r < 1.0:
            out.append(v2 * math.sqrt(-2.0 * math.log(r) / r) * sd)
    assert got == out
    assert abs(n / used - math.pi / 4) < 0.01
    g = np.array(got)
    assert abs(g.mean()) < 4 * sd / math.sqrt(n)
    assert abs(g.std() / sd - 1) < 0.01


def test_noise_steps_from_zero_phi_closed_form():
    n, V, seed, sd = 16, 3, -11, 0.05
    p = setup(model_inputs(n, V, (0.02, -0.01, -0.01), 0.005))
    p.noise_step = p.step + 1  # two noise steps
    o = Oracle(p, 2)
    o.set_state(phi=np.zeros((V, n, n, n)), cons=np.full((1, n, n, n), 0.4))
    o.set_noise(sd, seed)
    o.step(1)
    _, _, pn = o.get_noise()
    draws, _, _ = _stream(pyoracle.load().oracle_langevin, seed, 2 * V * n ** 3, sd)
    d = np.array(draws).reshape(2, n, n, n, V).transpose(0, 4, 1, 2, 3)  # z, y, x, variant order
    assert np.array_equal(pn, d[0])
    # Allen-Cahn from phi = 0 (f_diff = 0: no driving force, no stress)
    kfx, kqx = pyoracle.kaxis(n, n // 2 + 1)
    kfy, kqy = pyoracle.kaxis(n, n)
    kfz, kqz = kfy, kqy
    KQZ, KQY, KQX = np.meshgrid(kqz, kqy, kqx, indexing="ij")
    AZ, AY, AX = np.meshgrid(kfz, kfy, kfx, indexing="ij")
    G = p.grad_coeff_array()
    dt, b = p.t_step, p.b_mob
    phi = o.get_state()["phi"]
    for v in range(V):
        g = G[v]
        kgk = (g[0, 0] * KQX + g[1, 1] * KQY + g[2, 2] * KQZ + (g[0, 1] + g[1, 0]) * -(AX * AY)
               + (g[1, 2] + g[2, 1]) * -(AY * AZ) + (g[2, 0] + g[0, 2]) * -(AZ * AX))
        ref = np.fft.irfftn(dt * np.fft.rfftn(d[0, v]) / (1.0 - dt * b * kgk), s=(n, n, n), axes=(0, 1, 2))
        ref = np.clip(ref, 0.0, 1.0)
        assert np.abs(phi[v] - ref).max() < 1e-14
    assert np.abs(o.get_state()["cons"] - 0.4).max() < 1e-15
    # the second noise step accumulates; the third (past noise_step) resets
    o.step(1)
    _, _, pn = o.get_noise()
    assert np.array_equal(pn, d[0] + d[1])
    o.step(1)
    _, _, pn = o.get_noise()
    assert not pn.any()


def test_noise_step_without_generator_is_refused():
    p = setup(model_inputs(8, 1, (0.01, 0.01, 0.01), 0.005))
    p.noise_step = p.step
    o = Oracle(p, 1)
    o.set_state(phi=np.zeros((1, 8, 8, 8)), cons=np.full((1, 8, 8, 8), 0.4))
    with pytest.raises(RuntimeError):
        o.step(1)


@pytest.mark.parametrize("seed,V,n3", [(-7, 3, 1500000), (12345, 1, 70000), (-2, 2, 40000)])
def test_threaded_noise_loop_matches_sequential(seed, V, n3):
    """mm_noise_accumulate (jump-ahead streams, threaded Box-Muller around one
    in-order shuffle pass) against the oracle's sequential loop: the same
    draws and generator over two consecutive steps; 1.5e6 points x 3 spans
    several blocks of uniforms."""
    lib, olib = _lib.load(), pyoracle.load()
    D = ctypes.POINTER(ctypes.c_double)
    ia, ib = ctypes.c_int32(seed), ctypes.c_int32(seed)
    sa, sb = (ctypes.c_int32 * 35)(), (ctypes.c_int32 * 35)()
    a, b = np.zeros(V * n3), np.zeros(V * n3)
    for _ in range(2):
        assert lib.mm_noise_accumulate(ctypes.byref(ia), sa, 0.2, V, n3, a.ctypes.data_as(D)) == 0
        olib.oracle_noise_accumulate(ctypes.byref(ib), sb, 0.2, V, n3, b.ctypes.data_as(D))
    assert np.array_equal(a, b)
    assert ia.value == ib.value and list(sa) == list(sb)
