"""Host-side mirror of the reference driver around the C ABI.

``Simulation`` plays the part of the Fortran main program's loop
(Multi_MEUMAPPS.f:548-1106): it owns an ``mm_ctx`` (the whole grid resident
in HBM), uploads the restart record set, advances steps and returns the
fort.98 / fort.99 rows.  Arrays are numpy, structure-of-arrays with shape
(ncomp, Nz, Ny, Nx) (x fastest); conversion to the reference's component-
fastest Fortran layout happens here, at the boundary.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional

import numpy as np

from . import _lib
from .params import MM_FP64, MM_MIXED, MMDiag, MMParams

FIELD = {"phi": 0, "cons": 1, "el_en": 2, "f_diff": 3, "strain": 4, "grad_phi": 5,
         "k_del_sq": 6, "el_grad": 7}


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def soa_to_aos(a: np.ndarray) -> np.ndarray:
    """(ncomp, nz, ny, nx) -> Fortran a(ncomp, Nx, Ny, Nz) memory order."""
    return np.ascontiguousarray(np.moveaxis(a, 0, -1), dtype=np.float64)


def aos_to_soa(a: np.ndarray, ncomp: int, shape) -> np.ndarray:
    return np.ascontiguousarray(np.moveaxis(a.reshape(*shape, ncomp), -1, 0))


class Simulation:
    def __init__(self, params: MMParams, precision: int = MM_FP64, _ctx: Optional[ctypes.c_void_p] = None):
        self.lib = _lib.load()
        self.p = params
        self.precision = precision
        self.owned = _ctx is None
        if self.owned:
            self.ctx = ctypes.c_void_p()
            _lib.check(self.lib.mm_create(ctypes.byref(params), precision, ctypes.byref(self.ctx)), "mm_create")
        else:
            self.ctx = _ctx
        self.shape = (params.nz, params.ny, params.nx)
        self.V, self.E = params.vartot, params.numel

    def close(self):
        if self.ctx and self.owned:
            self.lib.mm_destroy(self.ctx)
        self.ctx = ctypes.c_void_p()

    def set_graph(self, on: bool = True, iters: int = 0):
        """Replay the regular step as one recorded graph (results identical)."""
        _lib.check(self.lib.mm_set_graph(self.ctx, int(on), iters), "mm_set_graph")

    def set_noise(self, sd: float, idum: int, state=None):
        """Langevin noise of the steps with step <= noise_step (mm_set_noise):
        amplitude sd, ran_2 generator (idum, 35-word state; None = program
        start).  For a restart run idum = Input.txt iseed (minus the rank)."""
        st = None if state is None else (ctypes.c_int32 * 35)(*state)
        _lib.check(self.lib.mm_set_noise(self.ctx, float(sd), int(idum), st), "mm_set_noise")

    def noise_state(self):
        """(idum, 35-word state) of the generator after the last noise step."""
        idum = ctypes.c_int32(0)
        st = (ctypes.c_int32 * 35)()
        _lib.check(self.lib.mm_get_noise_state(self.ctx, ctypes.byref(idum), st), "mm_get_noise_state")
        return idum.value, list(st)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- state
    def upload(self, phi=None, cons=None, u=None, gradu=None, el_en=None):
        phi_a = None if phi is None else soa_to_aos(np.asarray(phi).reshape(self.V, *self.shape))
        cons_a = None if cons is None else soa_to_aos(np.asarray(cons).reshape(self.E, *self.shape))
        u_a = None if u is None else np.ascontiguousarray(u, dtype=np.float64)
        g_a = None if gradu is None else np.ascontiguousarray(gradu, dtype=np.float64)
        e_a = None if el_en is None else np.ascontiguousarray(el_en, dtype=np.float64)
        _lib.check(self.lib.mm_upload(self.ctx, _ptr(phi_a), _ptr(cons_a), _ptr(e_a), _ptr(u_a), _ptr(g_a)),
                   "mm_upload")

    def download(self):
        n3 = int(np.prod(self.shape))
        phi = np.empty(self.V * n3); cons = np.empty(self.E * n3)
        el = np.empty(self.shape); u = np.empty((3, *self.shape)); gu = np.empty((9, *self.shape))
        _lib.check(self.lib.mm_download(self.ctx, _ptr(phi), _ptr(cons), _ptr(el), _ptr(u), _ptr(gu)),
                   "mm_download")
        return dict(phi=aos_to_soa(phi, self.V, self.shape), cons=aos_to_soa(cons, self.E, self.shape),
                    el_en=el, u=u, gradu=gu)

    def download_records(self) -> dict:
        """The arrays of the reference's data_save records, as mm_download
        returns them: phi / cons component-fastest (flat), el_en, u (3, N^3)
        = u1..u3, gradu (9, N^3) in record order (grad_1_u1, grad_1_u2, ...)."""
        n3 = int(np.prod(self.shape))
        phi = np.empty(self.V * n3); cons = np.empty(self.E * n3)
        el = np.empty(n3); u = np.empty(3 * n3); gu = np.empty(9 * n3)
        _lib.check(self.lib.mm_download(self.ctx, _ptr(phi), _ptr(cons), _ptr(el), _ptr(u), _ptr(gu)),
                   "mm_download")
        return dict(phi=phi, cons=cons, el_en=el, u=u.reshape(3, n3), gradu=gu.reshape(9, n3))

    def upload_records(self, rec: dict):
        """Inverse of download_records (the restart read of Multi_MEUMAPPS.f:470-481)."""
        arrs = [np.ascontiguousarray(rec[k], dtype=np.float64).ravel() for k in ("phi", "cons", "el_en", "u", "gradu")]
        _lib.check(self.lib.mm_upload(self.ctx, *[_ptr(a) for a in arrs]), "mm_upload")

    def field(self, name: str) -> np.ndarray:
        which = FIELD[name]
        ptr = ctypes.c_void_p(); esz = ctypes.c_int32(); nf = ctypes.c_int32()
        _lib.check(self.lib.mm_field(self.ctx, which, ctypes.byref(ptr), ctypes.byref(esz), ctypes.byref(nf)),
                   "mm_field")
        out = np.empty((nf.value, *self.shape))
        _lib.check(self.lib.mm_field_to_host(self.ctx, which, _ptr(out)), "mm_field_to_host")
        return out

    def device_field(self, name: str):
        """(device address, element size, number of fields) of a resident field."""
        ptr = ctypes.c_void_p(); esz = ctypes.c_int32(); nf = ctypes.c_int32()
        _lib.check(self.lib.mm_field(self.ctx, FIELD[name], ctypes.byref(ptr), ctypes.byref(esz),
                                     ctypes.byref(nf)), "mm_field")
        return ptr.value, esz.value, nf.value

    @property
    def step_count(self) -> int:
        return self.lib.mm_get_step(self.ctx)

    def set_step(self, s: int):
        _lib.check(self.lib.mm_set_step(self.ctx, s), "mm_set_step")

    # ----------------------------------------------------------------- step
    def step(self, n: int = 1, diag: bool = True) -> List[dict]:
        if not diag:
            _lib.check(self.lib.mm_step(self.ctx, n, None), "mm_step")
            return []
        rows = (MMDiag * n)()
        _lib.check(self.lib.mm_step(self.ctx, n, rows), "mm_step")
        return [r.as_dict(self.E) for r in rows]


class Ensemble:
    """Independent members advanced together, one graph launch per regular
    step (``mm_batch_*``).  ``members[i]`` is a ``Simulation`` view of member
    i for upload / download / field access; the batch owns the contexts."""

    def __init__(self, params: List[MMParams], precision: int = MM_FP64):
        self.lib = _lib.load()
        n = len(params)
        arr = (MMParams * n)(*params)
        self.b = ctypes.c_void_p()
        _lib.check(self.lib.mm_batch_create(n, ctypes.cast(arr, ctypes.c_void_p), precision,
                                            ctypes.byref(self.b)), "mm_batch_create")
        self.members = []
        for i, p in enumerate(params):
            c = ctypes.c_void_p()
            _lib.check(self.lib.mm_batch_member(self.b, i, ctypes.byref(c)), "mm_batch_member")
            self.members.append(Simulation(p, precision, _ctx=c))

    def step(self, n: int = 1) -> List[List[dict]]:
        """Advance every member n steps; returns rows[k][i] for member i."""
        m = len(self.members)
        rows = (MMDiag * (n * m))()
        _lib.check(self.lib.mm_batch_step(self.b, n, ctypes.cast(rows, ctypes.c_void_p)), "mm_batch_step")
        return [[rows[k * m + i].as_dict(self.members[i].E) for i in range(m)] for k in range(n)]

    def close(self):
        if self.b:
            for s in self.members:
                s.ctx = ctypes.c_void_p()
            self.lib.mm_batch_destroy(self.b)
            self.b = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fort98_line(d: dict) -> str:
    """One fort.98 row, format (i8,10e15.7) (Multi_MEUMAPPS.f:905-906,1113)."""
    from .records import fort98_row

    return fort98_row(d)
