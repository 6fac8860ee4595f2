"""The reference's restart / dump records and run files.

``data_save.NNNN`` (read at Multi_MEUMAPPS.f:466-481, written at :510-520 and
:1087-1095, read by data_ext_modular.f:34-42) is a Fortran unformatted
sequential file.  With one process (ist = (1,1,1), ien = (Nx,Ny,Nz)):

  record 1: ist(3), ien(3)                                   integer*4
  record 2: Cons(E,Nx,Ny,Nz), phi(V,Nx,Ny,Nz), el_en(Nx,Ny,Nz)  real*8
  record 3: u1, u2, u3                                       real*8 (Nx,Ny,Nz) each
  record 4: grad_1_u1, grad_1_u2, grad_1_u3
  record 5: grad_2_u1, grad_2_u2, grad_2_u3
  record 6: grad_3_u1, grad_3_u2, grad_3_u3

These are exactly the arrays ``mm_upload`` takes and ``mm_download`` returns
(component-fastest phi and Cons, then u1..u3 and the nine gradients in record
order).  Records are framed by 4-byte little-endian length markers before and
after the data, as gfortran, flang and nvfortran write them by default (the
reference's data_ext_modular.f built with flang reads these files).  A record
longer than 2**31 - 9 bytes is split into subrecords whose markers carry a
negative sign where the record continues: gfortran's convention, which
flang's runtime does not read; nvfortran's form for such records is not
checked.

``step_NNNNNNN/data_extract_mult.in`` (:500-505, :1077-1083) holds
``Nx Ny Nz nprocs vartot numel`` and ``dims`` (list-directed).  fort.98 rows
use format ``(i8,10e15.7)`` (:905-906, :1113); fort.99 rows are list-directed
``step, Caviter(1..numel)`` (:1064).
"""
from __future__ import annotations

import os
import struct
from typing import Dict, Iterable, List, Optional, Sequence

import numpy as np

MAX_SUBRECORD = 2 ** 31 - 9


def _write_record(f, payload: bytes, max_sub: int = MAX_SUBRECORD) -> None:
    n = len(payload)
    if n <= max_sub:
        f.write(struct.pack("<i", n))
        f.write(payload)
        f.write(struct.pack("<i", n))
        return
    pos, first = 0, True
    while pos < n:
        chunk = payload[pos:pos + max_sub]
        last = pos + len(chunk) >= n
        head = len(chunk) if last else -len(chunk)
        tail = len(chunk) if first else -len(chunk)
        f.write(struct.pack("<i", head))
        f.write(chunk)
        f.write(struct.pack("<i", tail))
        pos += len(chunk)
        first = False


def _read_record(f) -> bytes:
    parts = []
    while True:
        h = f.read(4)
        if len(h) < 4:
            raise EOFError("end of file inside a record")
        n = struct.unpack("<i", h)[0]
        data = f.read(abs(n))
        if len(data) != abs(n):
            raise EOFError("truncated record")
        t = struct.unpack("<i", f.read(4))[0]
        if abs(t) != abs(n):
            raise ValueError("record markers disagree")
        parts.append(data)
        if n >= 0:
            return b"".join(parts)


def _f8(arrays: Iterable[np.ndarray]) -> bytes:
    return b"".join(np.ascontiguousarray(a, dtype="<f8").tobytes() for a in arrays)


def write_data_save(path: str, shape, cons_aos: np.ndarray, phi_aos: np.ndarray, el_en: np.ndarray,
                    u: np.ndarray, gradu: np.ndarray, max_sub: int = MAX_SUBRECORD) -> None:
    """Write one data_save file.  shape = (Nx, Ny, Nz); cons_aos / phi_aos in
    the reference's component-fastest order (flat E*N^3 / V*N^3), el_en N^3,
    u (3, N^3) = u1, u2, u3, gradu (9, N^3) in record order."""
    nx, ny, nz = shape
    n3 = nx * ny * nz
    u = np.asarray(u).reshape(3, n3)
    gradu = np.asarray(gradu).reshape(9, n3)
    with open(path, "wb") as f:
        _write_record(f, struct.pack("<6i", 1, 1, 1, nx, ny, nz), max_sub)
        _write_record(f, _f8([np.ravel(cons_aos), np.ravel(phi_aos), np.ravel(el_en)]), max_sub)
        _write_record(f, _f8([u[0], u[1], u[2]]), max_sub)
        for d in range(3):
            _write_record(f, _f8([gradu[3 * d + i] for i in range(3)]), max_sub)


def read_data_save(path: str, V: int, E: int) -> Dict[str, np.ndarray]:
    """Read one data_save file written for a single process."""
    with open(path, "rb") as f:
        ist_ien = struct.unpack("<6i", _read_record(f))
        ist, ien = ist_ien[:3], ist_ien[3:]
        if tuple(ist) != (1, 1, 1):
            raise ValueError(f"data_save of a decomposed run (ist = {ist}); one process expected")
        nx, ny, nz = ien
        n3 = nx * ny * nz
        r2 = np.frombuffer(_read_record(f), dtype="<f8")
        if r2.size != (E + V + 1) * n3:
            raise ValueError(f"record 2 holds {r2.size} values, expected {(E + V + 1) * n3}")
        cons = r2[:E * n3].copy()
        phi = r2[E * n3:(E + V) * n3].copy()
        el_en = r2[(E + V) * n3:].copy()
        u = np.frombuffer(_read_record(f), dtype="<f8").reshape(3, n3).copy()
        gradu = np.concatenate([np.frombuffer(_read_record(f), dtype="<f8") for _ in range(3)]).reshape(9, n3)
    return dict(shape=(nx, ny, nz), cons=cons, phi=phi, el_en=el_en, u=u, gradu=gradu.copy())


def write_extract_in(path: str, shape, nprocs: int, vartot: int, numel: int, dims: Sequence[int]) -> None:
    nx, ny, nz = shape
    with open(path, "w") as f:
        f.write(f" {nx} {ny} {nz} {nprocs} {vartot} {numel}\n")
        f.write(" " + " ".join(str(int(d)) for d in dims) + "\n")


def fort98_row(d: dict) -> str:
    """One fort.98 row, format (i8,10e15.7)."""
    vals = [d["fint"], d["fw"], d["fch"], d["elen"], d["etot"], d["glob_min"], d["glob_max"], d["phi_max"],
            d["solid_tot"]]
    return f"{d['step']:8d}" + "".join(_e15_7(v) for v in vals)


def _e15_7(v: float) -> str:
    """Fortran E15.7: 0.dddddddE+xx with a leading blank for positives."""
    if v == 0.0:
        return f"{'0.0000000E+00':>15s}"
    m, e = f"{abs(v):.6E}".split("E")
    digits = m.replace(".", "")
    exp = int(e) + 1
    s = f"{'-' if v < 0 else ''}0.{digits}E{'+' if exp >= 0 else '-'}{abs(exp):02d}"
    return f"{s:>15s}"


def fort99_row(d: dict, numel: int) -> str:
    """One fort.99 row (list-directed step, Caviter(1..numel))."""
    return f" {d['step']:11d}" + "".join(f"  {c:.17G}" for c in d["caviter"][:numel])


def dump_step(root: str, step: int, shape, vartot: int, numel: int, sim_state: Dict[str, np.ndarray]) -> str:
    """Write step_NNNNNNN/data_extract_mult.in and data_save.0001 from the
    mm_download arrays of a Simulation (sim.download_records())."""
    d = os.path.join(root, f"step_{step:07d}")
    os.makedirs(d, exist_ok=True)
    write_extract_in(os.path.join(d, "data_extract_mult.in"), shape, 1, vartot, numel, (1, 1))
    write_data_save(os.path.join(d, "data_save.0001"), shape, sim_state["cons"], sim_state["phi"],
                    sim_state["el_en"], sim_state["u"], sim_state["gradu"])
    return d
