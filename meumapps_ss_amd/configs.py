"""The benchmark workloads mapped onto the reference's Input.txt model.

Each workload is expressed through the inputs the reference reads
(Multi_MEUMAPPS.f:117-206) and turned into ``mm_params`` by
:func:`params.setup`, so the numbers the kernels see are the ones the
reference's set-up would produce.  Mapping choices (the reference cannot
express some of them literally):

* kappa enters as ``grad_coeff = 5*del_N*EIntf/del_N**2`` (ModInitials.f:163-192),
  so kappa = 1 means EIntf = (0.2, 0.2, 0.2) with del_N = 1; that also fixes the
  well depth W = 6*sum(EIntf)/(5*del_N) (Multi_MEUMAPPS.f:364).
* M = diff_coeff/del_N**2, L = b_mob.
* The anti-phase well between variants uses intfantiph = sum(EIntf), so
  W_1 = W (Multi_MEUMAPPS.f:375).
* Phase free energies are the reference's quadratic KKS form with A = 10,
  G0 = 0 (matrix, c0 = 0.3) and G0 = -6 (precipitate, c0 = 0.7) in reduced
  units; Input.txt carries them in "J/mol", i.e. multiplied by 1.055e-5.
* Eigenstrain interpolation is p(eta) (the reference has no eta**2 form).
* Initial fields: c = c0 + U[-a, a], eta_v = U[-a, a], drawn from one ran_2
  stream with iseed = -seed (mm_ic_uniform), read back as restart records.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Sequence, Tuple

import numpy as np

from .params import Inputs, MMParams, PhaseInput, setup

JMOL = 1.055e-5


@dataclass
class Workload:
    name: str
    inputs: Inputs
    seed: int
    amp: float
    cav: Tuple[float, ...]
    nsteps: int

    def params(self) -> MMParams:
        return setup(self.inputs)


def model_inputs(n: int, nvar: int, tstr: Sequence[float], dt: float, *, kappa: float = 1.0,
                 mob_c: float = 1.0, mob_eta: float = 1.0, numph: int = 2, step0: int = 1,
                 nsteps: int = 10, c11=250.0, c12=150.0, c44=100.0, seed: int = 1,
                 ctype2: int = 1, cela2=None, eintf=None) -> Inputs:
    e = kappa / 5.0
    eintf = tuple(eintf) if eintf is not None else (e, e, e)
    cub = {(1, 1): c11, (1, 2): c12, (4, 4): c44}
    matrix = PhaseInput(nvar=1, ctype=1, cela=dict(cub), gparams=[10.0 * JMOL, 0.0 * JMOL, 0.3])
    phases = [matrix]
    for ph in range(2, numph + 1):
        g0, c0 = (-6.0, 0.7) if ph == 2 else (-5.0, 0.6)
        phases.append(PhaseInput(nvar=nvar if ph == 2 else 1, ctype=ctype2 if ph == 2 else 1,
                                 cela=dict(cela2 if (cela2 is not None and ph == 2) else cub),
                                 gparams=[10.0 * JMOL, g0 * JMOL, c0], tstr=tuple(tstr), eintf=eintf))
    vartot = sum(p.nvar for p in phases) - 1
    return Inputs(nx=n, ny=n, nz=n, del_n=1.0, nrun=2, step=step0, elastep0=step0, nt=step0 + nsteps,
                  t_step=dt, noise_step=0, b_mob=mob_eta, diff_coeff=mob_c, iseed=-seed,
                  vartot=vartot, numph=numph, numel_in=2, phases=phases, intfantiph=sum(eintf),
                  conc=[(1, 0.4, 1.0, 0.0)])


def workload(k: int, n: int = None) -> Workload:
    """Workload k = 1..5 of the benchmark definition (sizes overridable)."""
    if k == 1:
        inp = model_inputs(n or 64, 1, (0.01, 0.01, 0.01), 0.01, nsteps=1000, seed=1)
        return Workload("64^3 c + 1 variant, dilatational 0.01, FP64", inp, 1, 0.01, (0.4,), 1000)
    if k == 2:
        inp = model_inputs(n or 256, 3, (0.02, -0.01, -0.01), 0.005, nsteps=2000, seed=7)
        return Workload("256^3 c + 3 tetragonal variants, FP64", inp, 7, 0.01, (0.4,), 2000)
    if k == 3:
        inp = model_inputs(n or 512, 3, (0.02, -0.01, -0.01), 0.005, nsteps=1000, seed=11)
        return Workload("512^3 c + 3 tetragonal variants, mixed FP32/FP64", inp, 11, 0.01, (0.4,), 1000)
    if k == 5:
        inp = model_inputs(n or 128, 1, (0.02, 0.01, 0.01), 0.005, numph=3, nsteps=200, seed=3)
        return Workload("128^3 c + 2 variants (2 precipitate phases), FP64", inp, 3, 0.01, (0.4,), 200)
    raise ValueError(k)


def ensemble(sizes=(64, 96, 128), per_size: int = 16, seed0: int = 100, nsteps: int = 500,
             draw_seed: int = 4):
    """The ensemble workload: per_size members at each grid size of the
    3-variant tetragonal model, member i with ran_2 seed seed0 + i, eigenstrain
    magnitude m ~ U[0.005, 0.03] (diag(m, -m/2, -m/2) permuted per variant) and
    kappa ~ U[0.5, 2.0], drawn from numpy's PCG64 with draw_seed."""
    rng = np.random.default_rng(draw_seed)
    out = []
    i = 0
    for n in sizes:
        for _ in range(per_size):
            m = float(rng.uniform(0.005, 0.03))
            kappa = float(rng.uniform(0.5, 2.0))
            seed = seed0 + i
            inp = model_inputs(n, 3, (m, -0.5 * m, -0.5 * m), 0.005, kappa=kappa, nsteps=nsteps, seed=seed)
            out.append(Workload(f"{n}^3 member {i}: eigenstrain {m:.4f}, kappa {kappa:.3f}", inp, seed, 0.01,
                                (0.4,), nsteps))
            i += 1
    return out


def initial_fields(w: Workload):
    """Uniform-noise initial fields from ran_2 (host side, through the C ABI)."""
    import ctypes

    from . import _lib

    lib = _lib.load()
    inp = w.inputs
    n3 = inp.nx * inp.ny * inp.nz
    V, E = inp.vartot, inp.numel
    phi = np.empty(V * n3)
    cons = np.empty(E * n3)
    cav = np.asarray(w.cav, dtype=np.float64)
    D = ctypes.POINTER(ctypes.c_double)
    rc = lib.mm_ic_uniform(inp.nx, inp.ny, inp.nz, V, E, -w.seed, w.amp, cav.ctypes.data_as(D),
                           phi.ctypes.data_as(D), cons.ctypes.data_as(D))
    if rc:
        raise RuntimeError(f"mm_ic_uniform {rc}")
    shp = (inp.nz, inp.ny, inp.nx)
    return phi.reshape(V, *shp), cons.reshape(E, *shp)
