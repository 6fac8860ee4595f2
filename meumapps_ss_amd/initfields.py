"""Initial fields and added nuclei on the host, for one process.

Mirrors ``InitFields`` (ModInitials.f:415-690, the fresh-start branch
``nrun = 1`` of Multi_MEUMAPPS.f:455-459) and ``NucAdd`` (ModInitials.f:695-826,
called after a restart read when ``nucstat /= 'none'``, :483-487).  Both draw
from the same ran_2 stream (seed ``iseed - myid``, :229) that the Langevin
noise of the first steps continues, so they take and advance a :class:`Ran2`
whose state is later handed to ``Simulation.set_noise``.

Fields are structure-of-arrays with shape (ncomp, Nz, Ny, Nx); the reference's
1-based point (i, j, k) is ``[k-1, j-1, i-1]``.  One process owns the whole
grid (ist = 1, ien = N), so the MPI broadcasts and reductions of the
reference are identities here.  The reference's integer conversions of
``ran_2 * n + 1.0`` truncate toward zero, and the ellipsoid test of the
single nucleus is evaluated in binary32 (``float(i) / float(radppt)``).
"""
from __future__ import annotations

import ctypes
from typing import Tuple

import numpy as np

from . import _lib

MAX_TRIES = 1_000_000


class Ran2:
    """The reference's ran_2 stream (mm_ran2): idum and its saved state."""

    def __init__(self, seed: int, state=None):
        self.lib = _lib.load()
        self.idum = ctypes.c_int32(int(seed))
        self.state = (ctypes.c_int32 * 35)(*(state or [0] * 35))

    def __call__(self) -> float:
        return self.lib.mm_ran2(ctypes.byref(self.idum), self.state)

    def fill(self, n: int) -> np.ndarray:
        out = np.empty(n)
        _lib.check(self.lib.mm_ran2_fill(ctypes.byref(self.idum), self.state, n,
                                         out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "mm_ran2_fill")
        return out

    def snapshot(self) -> Tuple[int, list]:
        return self.idum.value, list(self.state)


def _ball_mask(shape, center, r0sq):
    """rdist = (i - ix)^2 + (j - jy)^2 + (k - kz)^2 <= r0ppt over the 1-based
    grid (no periodic wrap, as in the reference)."""
    nz, ny, nx = shape
    k = np.arange(1, nz + 1)[:, None, None]
    j = np.arange(1, ny + 1)[None, :, None]
    i = np.arange(1, nx + 1)[None, None, :]
    ix, jy, kz = center
    return (i - ix) ** 2 + (j - jy) ** 2 + (k - kz) ** 2 <= r0sq


def init_fields(inp, rng: Ran2):
    """InitFields (ModInitials.f:415-690) for one process.  Returns (phi, cons)
    as (V, Nz, Ny, Nx) and (E, Nz, Ny, Nx) arrays; rng advances by the draws
    the reference makes."""
    nx, ny, nz = inp.nx, inp.ny, inp.nz
    V, E = inp.vartot, inp.numel
    shape = (nz, ny, nx)
    phi = np.full((V, *shape), float(inp.phi_init))
    cav = {int(c[0]): float(c[1]) for c in inp.conc}
    cimax = {int(c[0]): float(c[2]) for c in inp.conc}
    cons = np.stack([np.full(shape, cav[e + 1]) for e in range(E)])
    mnx, mny, mnz = nx // 2, ny // 2, nz // 2
    r1, r2, r3 = (int(r) for r in inp.radppt)
    r0ppt = r1 * r1 + r2 * r2 + r3 * r3
    targv = int(inp.targv)
    nuc = inp.nucstat
    if nuc == "single":  # :466-515
        if targv == 0:
            targv = 1
        f32 = np.float32
        for k in range(-r3, r3 + 1):
            for j in range(-r2, r2 + 1):
                for i in range(-r1, r1 + 1):
                    t1 = f32(f32(i) / f32(r1)) ** 2
                    t2 = f32(f32(j) / f32(r2)) ** 2
                    t3 = f32(f32(k) / f32(r3)) ** 2
                    s = float(t1) + float(t2) + float(t3)
                    if s <= 1.0:
                        ii, jj, kk = i + mnx, j + mny, k + mnz
                        if 1 <= ii <= nx and 1 <= jj <= ny and 1 <= kk <= nz:
                            phi[targv - 1, kk - 1, jj - 1, ii - 1] = 1.0
        if inp.constat == "cppt":
            # the reference measures all three offsets from mnx (:500)
            kk = np.arange(1, nz + 1)[:, None, None]
            jj = np.arange(1, ny + 1)[None, :, None]
            ii = np.arange(1, nx + 1)[None, None, :]
            rd = (ii - mnx) ** 2 + (jj - mnx) ** 2 + (kk - mnx) ** 2
            phi[targv - 1][rd <= r0ppt] = 1.0
    elif nuc == "multiple":  # :521-611
        nppt = 1
        while nppt <= inp.nmaxppt:
            for _ in range(MAX_TRIES):
                ix = int(rng() * nx + 1.0)
                jy = int(rng() * ny + 1.0)
                kz = int(rng() * nz + 1.0)
                randvar = int(inp.var_lo + rng() * (inp.var_hi - inp.var_lo) + 1.0) if targv == 0 else targv
                ball = _ball_mask(shape, (ix, jy, kz), r0ppt)
                if not np.any(ball[None] & (phi > 0.0)):
                    break
            else:
                raise RuntimeError("no room for another nucleus")
            phi[randvar - 1][ball] = 1.0
            nppt += 1
    elif nuc == "random":  # :616-629
        u = rng.fill(nx * ny * nz).reshape(shape)  # x fastest, as the i loop
        var = (u * V + 1.0).astype(np.int64)
        for v in range(V):
            phi[v][var == v + 1] = 0.1
    elif nuc not in ("none", "addmult"):
        raise ValueError(f"nucstat {nuc!r}")
    if inp.constat == "cppt":  # :634-661
        if targv < 1:
            raise ValueError("constat 'cppt' needs a target variant")
        inside = phi[targv - 1] > 0.9
        for e in range(E):
            cons[e][inside] = cimax[e + 1]
    return phi, cons


def nuc_add(inp, phi: np.ndarray, rng: Ran2) -> np.ndarray:
    """NucAdd (ModInitials.f:695-826) for one process: with nucstat 'addmult',
    nmaxppt nuclei of random size and variant in the box kept clear of the
    faces, each placed where no variant exceeds 0.01.  phi is (V, Nz, Ny, Nx)
    and is changed in place (and returned)."""
    if inp.nucstat != "addmult":
        return phi
    nx, ny, nz = inp.nx, inp.ny, inp.nz
    V = inp.vartot
    shape = (nz, ny, nx)
    rp, ad = [int(r) for r in inp.radppt], [int(a) for a in inp.addrad]
    box = [n - 2 * (r + a + 1) for n, r, a in zip((nx, ny, nz), rp, ad)]
    targv = int(inp.targv)
    nppt = 1
    while nppt <= inp.nmaxppt:
        for _ in range(MAX_TRIES):
            ix = int(rng() * box[0] + 1.0 + rp[0] + ad[0])
            jy = int(rng() * box[1] + 1.0 + rp[1] + ad[1])
            kz = int(rng() * box[2] + 1.0 + rp[2] + ad[2])
            randvar = int(rng() * (V - 1) + 1.0) if targv == 0 else targv
            r1 = int(rng() * ad[0])
            r2 = int(rng() * ad[1])
            r3 = int(rng() * ad[2])
            r0ppt = (rp[0] + r1) ** 2 + (rp[1] + r2) ** 2 + (rp[2] + r3) ** 2
            ball = _ball_mask(shape, (ix, jy, kz), r0ppt)
            if not np.any(ball[None] & (phi > 0.01)):
                break
        else:
            raise RuntimeError("no room for another nucleus")
        phi[randvar - 1][ball] = 1.0
        phi[V - 1][ball] = 0.0
        nppt += 1
    return phi
