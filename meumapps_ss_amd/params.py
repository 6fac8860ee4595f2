"""Host-side set-up: Input.txt parsing and the ModGparams state the step reads.

This mirrors the part of the reference that stays on the host and runs once:

* the positional, list-directed ``Input.txt`` read (Multi_MEUMAPPS.f:117-206),
* the J/mol -> J/m^3 conversion by the float32 literal ``1.055e-5`` and the
  dislocation energy (Multi_MEUMAPPS.f:218-236),
* ``InitElasticMat`` / ``eltensor`` (ModInitials.f:16-116),
* ``InitTransGrad`` for nv = 1, 3 and 12 (ModInitials.f:122-343),
* interface width, well depths, gradient coefficients and mobilities
  (Multi_MEUMAPPS.f:353-405),

and packs the result into :class:`MMParams`, the ``mm_params`` struct of
``include/meumapps_ss.h``.
"""
from __future__ import annotations

import ctypes
import os
import re
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

MM_MAX_VAR = 16
MM_MAX_PH = 4
MM_MAX_EL = 4
MM_GP_STRIDE = 2 * MM_MAX_EL + 1

MM_FP64 = 0
MM_MIXED = 1

# The reference divides by the default-real literal 1.055e-5 (Multi_MEUMAPPS.f:221,225);
# binary32 rounding of that literal, widened to binary64.
F32_1055E5 = float(np.float32(1.055e-5))


class MMParams(ctypes.Structure):
    """ctypes mirror of ``mm_params`` (include/meumapps_ss.h)."""

    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("vartot", ctypes.c_int32), ("numel", ctypes.c_int32), ("numph", ctypes.c_int32),
        ("nvar", ctypes.c_int32 * MM_MAX_PH),
        ("step", ctypes.c_int32), ("elastep0", ctypes.c_int32), ("noise_step", ctypes.c_int32),
        ("pad0", ctypes.c_int32),
        ("t_step", ctypes.c_double), ("b_mob", ctypes.c_double),
        ("cmobs", ctypes.c_double * MM_MAX_EL),
        ("grad_coeff", ctypes.c_double * (MM_MAX_VAR * 9)),
        ("trans_ppt", ctypes.c_double * (MM_MAX_VAR * 6)),
        ("cela", ctypes.c_double * (MM_MAX_PH * 36)),
        ("sela", ctypes.c_double * (MM_MAX_PH * 36)),
        ("gparams", ctypes.c_double * (MM_MAX_PH * MM_GP_STRIDE)),
        ("wellph", ctypes.c_double * MM_MAX_PH),
        ("e_mean", ctypes.c_double * 6),
    ]

    # convenient numpy views (Fortran index order documented in the header)
    def grad_coeff_array(self) -> np.ndarray:
        a = np.ctypeslib.as_array(self.grad_coeff).reshape(MM_MAX_VAR, 3, 3)
        return a[: self.vartot].transpose(0, 2, 1)  # [v, i, j] = grad_coeff(i,j,v)

    def trans_ppt_array(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.trans_ppt).reshape(MM_MAX_VAR, 6)[: self.vartot]

    def cela_array(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.cela).reshape(MM_MAX_PH, 6, 6)[: self.numph].transpose(0, 2, 1)

    def sela_array(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.sela).reshape(MM_MAX_PH, 6, 6)[: self.numph].transpose(0, 2, 1)

    def gparams_array(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.gparams).reshape(MM_MAX_PH, MM_GP_STRIDE)[
            : self.numph, : 2 * self.numel + 1]


class MMDiag(ctypes.Structure):
    """ctypes mirror of ``mm_diag``: one fort.98 + fort.99 row."""

    _fields_ = [
        ("step", ctypes.c_int32), ("n_elastic", ctypes.c_int32),
        ("fint", ctypes.c_double), ("fw", ctypes.c_double), ("fch", ctypes.c_double),
        ("elen", ctypes.c_double), ("etot", ctypes.c_double),
        ("glob_min", ctypes.c_double), ("glob_max", ctypes.c_double),
        ("phi_max", ctypes.c_double), ("solid_tot", ctypes.c_double),
        ("caviter", ctypes.c_double * MM_MAX_EL),
        ("cfac", ctypes.c_double),
    ]

    def as_dict(self, nel: int) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "caviter"}
        d["caviter"] = [self.caviter[i] for i in range(nel)]
        return d


class MMAdjGrads(ctypes.Structure):
    """ctypes mirror of ``mm_adj_grads``: dL/d of the mm_params arrays."""

    _fields_ = [
        ("grad_coeff", ctypes.c_double * (MM_MAX_VAR * 9)),
        ("wellph", ctypes.c_double * MM_MAX_PH),
        ("cmobs", ctypes.c_double * MM_MAX_EL),
        ("b_mob", ctypes.c_double),
        ("trans_ppt", ctypes.c_double * (MM_MAX_VAR * 6)),
    ]


@dataclass
class PhaseInput:
    """One phase block of Input.txt (Multi_MEUMAPPS.f:178-204)."""

    nvar: int
    ctype: int
    cela: dict                    # {(i,j): value}, 1-based Voigt indices as read
    gparams: List[float]          # 2*numel_in - 1 values (A_e, G0, c0_e) in J/mol
    tstr: Sequence[float] = (0.0, 0.0, 0.0)
    eintf: Sequence[float] = (0.0, 0.0, 0.0)


@dataclass
class Inputs:
    """Every value the reference reads from Input.txt, by name."""

    ndim: int = 1
    dims: Sequence[int] = (1, 1)
    nx: int = 64
    ny: int = 64
    nz: int = 64
    del_n: float = 1.0
    nrun: int = 2
    step: int = 1
    elastep0: int = 1
    nt: int = 2
    t_step: float = 0.01
    noise_step: int = 0
    b_mob: float = 1.0
    diff_coeff: float = 1.0
    m_star: float = 0.0
    sd: float = 0.0
    noise: float = 0.0
    ifreq: int = 500
    nucstat: str = "none"
    targv: int = 0
    vartot: int = 1
    nmaxppt: int = 0
    iseed: int = -1
    var_lo: int = 1
    var_hi: int = 1
    radppt: Sequence[int] = (0, 0, 0)
    addrad: Sequence[int] = (0, 0, 0)
    disc_ener: float = 0.0
    numph: int = 2
    numel_in: int = 2               # as read; the reference uses numel_in - 1
    intfantiph: float = 0.0
    e_mean: Sequence[float] = (0.0, 0.0, 0.0)
    constat: str = "none"
    conc: List[Sequence[float]] = field(default_factory=lambda: [(1, 0.4, 1.0, 0.0)])
    phi_init: float = 0.0
    phases: List[PhaseInput] = field(default_factory=list)
    sym_var: Optional[np.ndarray] = None  # (12,3,3) rotations for nv = 12

    @property
    def numel(self) -> int:
        return self.numel_in - 1


def eltensor(c: np.ndarray, ctype: int) -> np.ndarray:
    """Complete a 6x6 stiffness from its independent constants (ModInitials.f:16-72)."""
    c = c.copy()
    if ctype == 1:
        c[0, 2] = c[0, 1]; c[1, 1] = c[0, 0]; c[1, 2] = c[0, 2]
        c[2, 2] = c[0, 0]; c[4, 4] = c[3, 3]; c[5, 5] = c[3, 3]
    elif ctype == 2:
        c[1, 1] = c[0, 0]; c[4, 4] = c[3, 3]; c[1, 2] = c[0, 2]
    elif ctype == 3:
        c[1, 1] = c[0, 0]; c[1, 2] = c[0, 2]; c[4, 4] = c[3, 3]
        c[5, 5] = 0.5 * (c[0, 0] - c[0, 1])
    elif ctype == 4:
        c[1, 1] = c[0, 0]; c[1, 2] = c[0, 2]; c[1, 4] = -c[0, 4]; c[3, 5] = -c[0, 4]
        c[4, 4] = c[3, 3]; c[5, 5] = 0.5 * (c[0, 0] - c[0, 1])
    for j in range(1, 6):
        for i in range(j):
            c[j, i] = c[i, j]
    return c


class SetupError(ValueError):
    """Raised for the reference's initerr / err_flag 20 conditions."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


def _rotations_delta() -> np.ndarray:
    """Habit-plane rotation of InitTransGrad nv=12 (ModInitials.f:219-229).

    The reference evaluates ``1.d0/dsqrt(2.0)`` with a default-real argument;
    sqrt of an exactly representable small integer is the same in both kinds.
    """
    s2, s3, s6 = np.sqrt(2.0), np.sqrt(3.0), np.sqrt(6.0)
    q = np.zeros((3, 3))
    q[0, 0] = 1.0 / s2; q[1, 0] = 0.0; q[2, 0] = -1.0 / s2
    q[0, 1] = 1.0 / s3; q[1, 1] = 1.0 / s3; q[2, 1] = 1.0 / s3
    q[0, 2] = 1.0 / s6; q[1, 2] = -2.0 / s6; q[2, 2] = 1.0 / s6
    return q


def init_trans_grad(inp: Inputs, nphc: int, i0: int, b_width: float,
                    trans: np.ndarray, gcoef: np.ndarray) -> None:
    """InitTransGrad (ModInitials.f:122-343); nphc and i0 are 1-based."""
    ph = inp.phases[nphc - 1]
    nv, eintf, tstr = ph.nvar, list(ph.eintf), list(ph.tstr)
    fvn = i0 + nv - 1
    if nv == 1:
        trans[i0 - 1, 0:3] = tstr
        for d in range(3):
            gcoef[i0 - 1, d, d] = b_width * eintf[d]
    if nv == 3:
        trans[i0 - 1, 0:3] = (tstr[0], tstr[1], tstr[2])
        trans[i0, 0:3] = (tstr[1], tstr[0], tstr[2])
        trans[i0 + 1, 0:3] = (tstr[1], tstr[2], tstr[0])
        for v in range(i0 - 1, fvn):
            for d in range(3):
                gcoef[v, d, d] = b_width * eintf[d]
    if nv == 12:
        if inp.sym_var is None:
            raise SetupError(2, "Check input inp file: Initial error 2")
        q = _rotations_delta()
        a = np.diag([b_width * e for e in eintf])
        a = (q @ a) @ q.T
        b = np.diag(tstr)
        b = (q @ b) @ q.T
        for k, v in enumerate(range(i0 - 1, fvn)):
            s = inp.sym_var[k]
            gcoef[v] = (s.T @ a) @ s
            bp = (s.T @ b) @ s
            trans[v] = (bp[0, 0], bp[1, 1], bp[2, 2], bp[1, 2], bp[0, 2], bp[0, 1])


def setup(inp: Inputs) -> MMParams:
    """Multi_MEUMAPPS.f:153-448 restricted to what the step reads."""
    numel = inp.numel
    numph = inp.numph
    if numel < 1 or numel > MM_MAX_EL or numph < 2 or numph > MM_MAX_PH:
        raise SetupError(20001, "unsupported numph/numel")
    if inp.vartot < 1 or inp.vartot > MM_MAX_VAR:
        raise SetupError(20001, "unsupported vartot")
    nvar = [ph.nvar for ph in inp.phases]
    if len(nvar) != numph:
        raise SetupError(4, "Check variant number: Initial error 4")
    if inp.vartot + 1 - sum(nvar) != 0:
        raise SetupError(1, "Check variant numbers: Initial error 1")
    pmax = 2 * numel + 1
    gp = np.zeros((numph, pmax))
    cela = np.zeros((numph, 6, 6))
    for p, ph in enumerate(inp.phases):
        if len(ph.gparams) != pmax:
            raise SetupError(5, "Check element number: Initial error 5")
        gp[p] = ph.gparams
        for (i, j), val in ph.cela.items():
            cela[p, i - 1, j - 1] = val
    # J/mol -> J/m^3 (:218-227): rows 1..numel and numel+1
    gp[:, : numel + 1] = gp[:, : numel + 1] / F32_1055E5
    gp[0, :numel] = gp[0, :numel] + inp.disc_ener
    # InitElasticMat (ModInitials.f:76-116)
    sela = np.zeros_like(cela)
    for p, ph in enumerate(inp.phases):
        if cela[p, 0, 0] <= 0:
            cela[p] = 0.0
            sela[p] = 0.0
            continue
        cela[p] = eltensor(cela[p], ph.ctype)
        try:
            s = np.linalg.inv(cela[p])
        except np.linalg.LinAlgError:
            raise SetupError(20, "Poor cond. no; halting program (flag 20)")
        if 1.0 / (np.abs(cela[p]).sum(0).max() * np.abs(s).sum(0).max()) <= 1e-16:
            raise SetupError(20, "Poor cond. no; halting program (flag 20)")
        sela[p] = s
    b_width = 5.0 * inp.del_n
    trans = np.zeros((inp.vartot, 6))
    gcoef = np.zeros((inp.vartot, 3, 3))
    wellph = np.zeros(numph)
    ivn = 0
    for nphc in range(2, numph + 1):
        sume = sum(inp.phases[nphc - 1].eintf)
        wellph[nphc - 1] = 6.0 * sume / b_width
        ivn += nvar[nphc - 2]
        init_trans_grad(inp, nphc, ivn, b_width, trans, gcoef)
    wellph[0] = 6.0 * inp.intfantiph / b_width
    dn2 = inp.del_n * inp.del_n
    gcoef = gcoef / dn2
    cmobs = [inp.diff_coeff / dn2] * numel

    p = MMParams()
    p.nx, p.ny, p.nz = inp.nx, inp.ny, inp.nz
    p.vartot, p.numel, p.numph = inp.vartot, numel, numph
    for i, n in enumerate(nvar):
        p.nvar[i] = n
    p.step, p.elastep0, p.noise_step = inp.step, inp.elastep0, inp.noise_step
    p.t_step, p.b_mob = inp.t_step, inp.b_mob
    for e in range(numel):
        p.cmobs[e] = cmobs[e]
    for v in range(inp.vartot):
        for i in range(3):
            for j in range(3):
                p.grad_coeff[v * 9 + j * 3 + i] = gcoef[v, i, j]
        for m in range(6):
            p.trans_ppt[v * 6 + m] = trans[v, m]
    for ph in range(numph):
        for i in range(6):
            for j in range(6):
                p.cela[ph * 36 + j * 6 + i] = cela[ph, i, j]
                p.sela[ph * 36 + j * 6 + i] = sela[ph, i, j]
        for j in range(pmax):
            p.gparams[ph * MM_GP_STRIDE + j] = gp[ph, j]
        p.wellph[ph] = wellph[ph]
    em = list(inp.e_mean) + [0.0] * (6 - len(inp.e_mean))
    for m in range(6):
        p.e_mean[m] = em[m]
    return p


# --------------------------------------------------------------- Input.txt
_NUM = re.compile(r"[+-]?(\d+\.?\d*|\.\d+)([eEdD][+-]?\d+)?")


def _tokens(line: str) -> List[str]:
    line = line.split("!")[0]
    return [t for t in re.split(r"[,\s]+", line.strip()) if t]


def _num(tok: str) -> float:
    return float(tok.replace("d", "e").replace("D", "e"))


def read_input_txt(path: str, sym_var_path: Optional[str] = None) -> Inputs:
    """Read Input.txt in the reference's record order (Multi_MEUMAPPS.f:123-201).

    Each Fortran list-directed ``read(2,*)`` consumes one record; the header
    records (``buffchar``) are skipped.
    """
    with open(path) as f:
        recs = [l for l in f.read().splitlines()]
    it = iter(recs)

    def rec() -> List[str]:
        return _tokens(next(it))

    inp = Inputs()
    rec()  # buffchar
    t = rec(); inp.ndim = int(t[0]); inp.dims = (int(t[1]), int(t[2]))
    rec()
    t = rec(); inp.nx, inp.ny, inp.nz = int(t[0]), int(t[1]), int(t[2]); inp.del_n = _num(t[3])
    t = rec()
    inp.nrun, inp.step, inp.elastep0, inp.nt = int(t[0]), int(t[1]), int(t[2]), int(t[3])
    inp.t_step, inp.noise_step = _num(t[4]), int(t[5])
    t = rec(); inp.b_mob, inp.diff_coeff, inp.m_star = _num(t[0]), _num(t[1]), _num(t[2])
    t = rec(); inp.sd, inp.noise = _num(t[0]), _num(t[1])
    t = rec(); inp.ifreq = int(t[0])
    rec()
    t = rec()
    inp.nucstat = t[0].strip("'\""); inp.targv = int(t[1]); inp.vartot = int(t[2])
    inp.nmaxppt = int(t[3]); inp.iseed = int(t[4])
    t = rec(); inp.var_lo, inp.var_hi = int(t[0]), int(t[1])
    t = rec(); inp.radppt = tuple(int(x) for x in t[:3])
    t = rec(); inp.addrad = tuple(int(x) for x in t[:3])
    t = rec(); inp.disc_ener = _num(t[0])
    rec()
    t = rec(); inp.numph, inp.numel_in = int(t[0]), int(t[1])
    t = rec(); inp.intfantiph = _num(t[0])
    t = rec(); inp.e_mean = tuple(_num(x) for x in t[:3])
    numel = inp.numel_in - 1
    pmax = 2 * numel + 1
    rec()
    t = rec(); inp.constat = t[0].strip("'\"")
    inp.conc = []
    for _ in range(numel):
        t = rec(); inp.conc.append((int(t[0]), _num(t[1]), _num(t[2]), _num(t[3])))
    rec()
    t = rec(); inp.phi_init = _num(t[0])
    t = rec(); nphc, nv, ct = int(t[0]), int(t[1]), int(t[2])
    t = rec(); cel = {(1, 1): _num(t[0]), (1, 2): _num(t[1]), (4, 4): _num(t[2])}
    t = rec(); g = [_num(x) for x in t[:pmax]]
    inp.phases = [PhaseInput(nvar=nv, ctype=ct, cela=cel, gparams=g)]
    for i in range(2, inp.numph + 1):
        rec()
        t = rec(); nphc, nv, ct = int(t[0]), int(t[1]), int(t[2])
        if ct == 1:
            t = rec(); cel = {(1, 1): _num(t[0]), (1, 2): _num(t[1]), (4, 4): _num(t[2])}
        else:
            t = rec()
            cel = {(1, 1): _num(t[0]), (1, 2): _num(t[1]), (1, 3): _num(t[2]),
                   (2, 2): _num(t[3]), (2, 3): _num(t[4]), (3, 3): _num(t[5])}
            t = rec()
            cel.update({(4, 4): _num(t[0]), (4, 5): _num(t[1]), (4, 6): _num(t[2]),
                        (5, 5): _num(t[3]), (5, 6): _num(t[4]), (6, 6): _num(t[5])})
        t = rec(); ts = tuple(_num(x) for x in t[:3])
        t = rec(); ei = tuple(_num(x) for x in t[:3])
        t = rec(); g = [_num(x) for x in t[:pmax]]
        if i != nphc:
            raise SetupError(4, "Check variant number: Initial error 4")
        inp.phases.append(PhaseInput(nvar=nv, ctype=ct, cela=cel, gparams=g, tstr=ts, eintf=ei))
    if sym_var_path is None:
        cand = os.path.join(os.path.dirname(os.path.abspath(path)), "sym_var.inp")
        sym_var_path = cand if os.path.exists(cand) else None
    if sym_var_path:
        inp.sym_var = read_sym_var(sym_var_path)
    return inp


def read_sym_var(path: str, nvar: int = 12) -> np.ndarray:
    """The variant rotations of sym_var.inp as InitTransGrad reads them
    (ModInitials.f:307-310): per variant three list-directed reads
    ``read(21,*) (SYM(i,j), j=1,3)``, one per row i.  Each read starts at a new
    record, takes three values (continuing onto the next records if one holds
    fewer) and discards the rest of the record it ends in."""
    recs = [_tokens(l) for l in open(path).read().splitlines()]
    out = np.zeros((nvar, 3, 3))
    r = 0
    for v in range(nvar):
        for i in range(3):
            vals = []
            while len(vals) < 3:
                if r >= len(recs):
                    raise SetupError(2, "Check input inp file: Initial error 2")
                vals += recs[r]
                r += 1
            out[v, i] = [_num(x) for x in vals[:3]]
    return out


def write_input_txt(inp: Inputs, path: str) -> None:
    """Write an Input.txt the reference (and :func:`read_input_txt`) can read."""
    L = []
    L.append("'MPI'")
    L.append(f"{inp.ndim} {inp.dims[0]} {inp.dims[1]}")
    L.append("'SIM'")
    L.append(f"{inp.nx} {inp.ny} {inp.nz} {inp.del_n!r}")
    L.append(f"{inp.nrun} {inp.step} {inp.elastep0} {inp.nt} {inp.t_step!r} {inp.noise_step}")
    L.append(f"{inp.b_mob!r} {inp.diff_coeff!r} {inp.m_star!r}")
    L.append(f"{inp.sd!r} {inp.noise!r}")
    L.append(f"{inp.ifreq}")
    L.append("'NUC'")
    L.append(f"'{inp.nucstat}' {inp.targv} {inp.vartot} {inp.nmaxppt} {inp.iseed}")
    L.append(f"{inp.var_lo} {inp.var_hi}")
    L.append(" ".join(str(x) for x in inp.radppt))
    L.append(" ".join(str(x) for x in inp.addrad))
    L.append(f"{inp.disc_ener!r}")
    L.append("'PHASES'")
    L.append(f"{inp.numph} {inp.numel_in}")
    L.append(f"{inp.intfantiph!r}")
    L.append(" ".join(repr(float(x)) for x in inp.e_mean))
    L.append("'CONC'")
    L.append(f"'{inp.constat}'")
    for c in inp.conc:
        L.append(f"{int(c[0])} {c[1]!r} {c[2]!r} {c[3]!r}")
    m = inp.phases[0]
    L.append("'MATRIX'")
    L.append(f"{inp.phi_init!r}")
    L.append(f"1 {m.nvar} {m.ctype}")
    L.append(f"{m.cela.get((1,1),0.0)!r} {m.cela.get((1,2),0.0)!r} {m.cela.get((4,4),0.0)!r}")
    L.append(" ".join(repr(float(x)) for x in m.gparams))
    for i, ph in enumerate(inp.phases[1:], start=2):
        L.append(f"'PHASE{i}'")
        L.append(f"{i} {ph.nvar} {ph.ctype}")
        if ph.ctype == 1:
            L.append(f"{ph.cela.get((1,1),0.0)!r} {ph.cela.get((1,2),0.0)!r} {ph.cela.get((4,4),0.0)!r}")
        else:
            L.append(" ".join(repr(ph.cela.get(k, 0.0)) for k in [(1,1),(1,2),(1,3),(2,2),(2,3),(3,3)]))
            L.append(" ".join(repr(ph.cela.get(k, 0.0)) for k in [(4,4),(4,5),(4,6),(5,5),(5,6),(6,6)]))
        L.append(" ".join(repr(float(x)) for x in ph.tstr))
        L.append(" ".join(repr(float(x)) for x in ph.eintf))
        L.append(" ".join(repr(float(x)) for x in ph.gparams))
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")
