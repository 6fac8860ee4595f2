"""Shared parity cases: model inputs and seeded initial fields that exercise
the branches of the step (clamp at 0 and 1, phi > 0 tests of the f_diff
correction, overlapping variants, inhomogeneous moduli with a real elastic
fixed-point iteration, a second precipitate phase, two compositions)."""
from __future__ import annotations

import numpy as np

from meumapps_ss_amd.configs import model_inputs
from meumapps_ss_amd.params import setup


def smooth_field(rng, shape, lo, hi, cut=0.25):
    """Random field with spectral content below cut*Nyquist, scaled to [lo, hi]."""
    a = rng.standard_normal(shape)
    f = np.fft.rfftn(a)
    kz = np.fft.fftfreq(shape[0])[:, None, None]
    ky = np.fft.fftfreq(shape[1])[None, :, None]
    kx = np.fft.rfftfreq(shape[2])[None, None, :]
    k = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    f[k > cut * 0.5] = 0.0
    b = np.fft.irfftn(f, s=shape, axes=(0, 1, 2))
    b = (b - b.min()) / (b.max() - b.min())
    return lo + (hi - lo) * b


CASES = {
    # name: (n, nvar, tstr, dt, extra model kwargs, init kind, numel_in)
    "v1_dilat": (16, 1, (0.01, 0.01, 0.01), 0.01, {}, "rich"),
    "v3_tetra": (16, 3, (0.02, -0.01, -0.01), 0.005, {}, "rich"),
    "v3_noise": (32, 3, (0.02, -0.01, -0.01), 0.005, {}, "noise"),
    "v2_twophase": (16, 1, (0.02, 0.01, 0.01), 0.005, {"numph": 3}, "rich"),
    "v3_inhomog": (16, 3, (0.02, 0.01, -0.01), 0.005,
                   {"cela2": {(1, 1): 300.0, (1, 2): 140.0, (4, 4): 120.0}}, "rich"),
    # nv = 2 has no eigenstrain variants in the reference set-up (ModInitials.f:200-340
    # fills nv = 1, 3, 12), so this case runs the elastic loop on zero eigenstrain
    "v2_nonsq": (24, 2, (0.015, 0.015, -0.01), 0.005, {}, "rich"),
    # radix-3 lengths with elasticity: 24^3 and 96^3 (the ensemble's middle size)
    "v3_n24": (24, 3, (0.02, -0.01, -0.01), 0.005, {}, "rich"),
    "v3_n96": (96, 3, (0.02, -0.01, -0.01), 0.005, {}, "rich"),
    "v1_e2": (16, 1, (0.01, 0.01, 0.01), 0.005, {"numel_in": 3}, "rich"),
    # 16 x 16 x 512: the length-512 z kernels (radix-4 plan of the fused elastic pass)
    "v3_z512": (16, 3, (0.02, -0.01, -0.01), 0.005, {"nz": 512}, "rich"),
    # 16 x 16 x 192 and x 384: radix-3 column lengths of the fused elastic pass
    # with 12 points per thread (the 512-thread launch shape)
    "v3_z192": (16, 3, (0.02, -0.01, -0.01), 0.005, {"nz": 192}, "rich"),
    "v3_z384": (16, 3, (0.02, -0.01, -0.01), 0.005, {"nz": 384}, "rich"),
    # nv = 12 from sym_var rotations (InitTransGrad, ModInitials.f:200-340): an
    # orthorhombic eigenstrain turned about z in steps of 15 degrees
    "v12_rot": (16, 12, (0.02, -0.01, 0.005), 0.005, {"sym_var": "z15"}, "rich"),
    # non-cubic precipitate moduli completed by eltensor (ModInitials.f:16-72):
    # tetragonal (Ctype 2), hexagonal (Ctype 3) and trigonal (Ctype 4) against a
    # cubic matrix, so the moduli are inhomogeneous and the acoustic tensor of
    # the precipitate phase is not cubic
    "v1_tetC": (16, 1, (0.01, 0.012, 0.006), 0.005,
                {"ctype2": 2, "cela2": {(1, 1): 260.0, (1, 2): 140.0, (1, 3): 120.0, (3, 3): 230.0,
                                        (4, 4): 90.0, (6, 6): 110.0}}, "rich"),
    "v1_hexC": (16, 1, (0.01, 0.012, 0.006), 0.005,
                {"ctype2": 3, "cela2": {(1, 1): 260.0, (1, 2): 140.0, (1, 3): 120.0, (3, 3): 230.0,
                                        (4, 4): 90.0}}, "rich"),
    "v1_trigC": (16, 1, (0.01, 0.012, 0.006), 0.005,
                 {"ctype2": 4, "cela2": {(1, 1): 260.0, (1, 2): 140.0, (1, 3): 120.0, (3, 3): 230.0,
                                         (4, 4): 90.0, (1, 5): 15.0}}, "rich"),
}


def z_rotations(n=12, step_deg=15.0):
    out = []
    for k in range(n):
        a = np.deg2rad(step_deg * k)
        c, s = np.cos(a), np.sin(a)
        out.append(np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]]))
    return np.stack(out)


def make_case(name, seed=0):
    n, nvar, tstr, dt, kw, init = CASES[name]
    kw = dict(kw)
    numel_in = kw.pop("numel_in", 2)
    nz = kw.pop("nz", n)
    sym = kw.pop("sym_var", None)
    inp = model_inputs(n, nvar, tstr, dt, **kw)
    inp.nz = nz
    if sym == "z15":
        inp.sym_var = z_rotations()
    if numel_in != 2:
        # a second composition: A, G0, c0 rows extend to 2*numel+1
        inp.numel_in = numel_in
        for ph in inp.phases:
            a, g0, c0 = ph.gparams
            ph.gparams = [a, 0.5 * a, g0, c0, c0 - 0.1]
        inp.conc = [(1, 0.4, 1.0, 0.0), (2, 0.25, 1.0, 0.0)]
    p = setup(inp)
    V, E = p.vartot, p.numel
    shape = (nz, n, n)
    rng = np.random.default_rng(seed + 17 * n + 5 * V)
    if init == "noise":
        phi = rng.uniform(-0.01, 0.01, (V, *shape))
        cons = 0.4 + rng.uniform(-0.01, 0.01, (E, *shape))
    else:
        phi = np.stack([smooth_field(rng, shape, -0.15, 1.1) for _ in range(V)])
        cons = np.stack([smooth_field(rng, shape, 0.3, 0.6) for _ in range(E)])
    return p, phi, cons
