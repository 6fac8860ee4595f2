"""TEST INFRASTRUCTURE ONLY: ctypes binding of ``oracle/liboracle.so``.

Fields are numpy arrays in structure-of-arrays form, shape (ncomp, nz, ny, nx)
(x fastest), matching the oracle's C layout.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int, c_void_p

import numpy as np

from meumapps_ss_amd.params import MMAdjGrads, MMDiag, MMParams

HERE = os.path.dirname(os.path.abspath(__file__))
# MM_ORACLE_LIB points the tests at another build of oracle.c (a copy with a
# planted mistake, to show that the pins catch it)
LIB = os.environ.get("MM_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
_lib = None


def build() -> str:
    r = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stderr)
    return LIB


# The reference's own ran_2.f (self-contained) compiled with flang where the
# reference tree is present: draw-for-draw check of the ran_2 restatements.
REF_SRC = "/root/reference/ran_2.f"
REF_RAN2 = os.path.join(HERE, "_ref", "libran2_ref.so")
# and its reader of the data_save records, data_ext_modular.f
REF_DATA_EXT = os.path.join(HERE, "_ref", "data_ext_ref")


def build_ref():
    """Build oracle/_ref (the reference's ran_2.f as a library and its
    data_ext_modular.f as a program); None where the reference tree is
    absent."""
    if not os.path.exists(REF_SRC):
        return None
    r = subprocess.run(["make", "-s", "-C", HERE, "ref"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("reference ran_2 build failed:\n" + r.stderr)
    return REF_RAN2


def load_ref_ran2():
    """ctypes handle of the reference ran_2 (real*8 function ran_2(idum),
    SAVEd state), or None where it has not been built."""
    if not os.path.exists(REF_RAN2):
        return None
    lib = ctypes.CDLL(REF_RAN2)
    lib.ran_2_.restype = c_double
    lib.ran_2_.argtypes = [POINTER(ctypes.c_int32)]
    return lib


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = ctypes.CDLL(LIB)
    D = POINTER(c_double)
    sig = {
        "oracle_create": (c_void_p, [POINTER(MMParams), c_int]),
        "oracle_destroy": (None, [c_void_p]),
        "oracle_set_threads": (c_int, [c_void_p, c_int]),
        "oracle_set_state": (c_int, [c_void_p, D, D, D, D, D]),
        "oracle_get_state": (c_int, [c_void_p, D, D, D, D, D]),
        "oracle_get_work": (c_int, [c_void_p, c_int, D]),
        "oracle_get_step": (c_int, [c_void_p]),
        "oracle_get_lagged": (None, [c_void_p, D]),
        "oracle_set_lagged": (None, [c_void_p, D]),
        "oracle_set_step": (None, [c_void_p, c_int]),
        "oracle_step": (c_int, [c_void_p, POINTER(MMDiag)]),
        "oracle_f_trans": (c_int, [c_void_p, D, D]),
        "oracle_inv_trans": (c_int, [c_void_p, D, D]),
        "oracle_invert": (c_int, [D, c_int, c_int]),
        "oracle_kaxis": (None, [c_int, c_int, D, D]),
        "oracle_eldis": (c_int, [c_void_p, POINTER(c_int), D]),
        "oracle_elastic_zeroth": (c_int, [c_void_p, D, D, D]),
        "oracle_set_pphi": (c_int, [c_void_p]),
        "oracle_elener": (c_int, [c_void_p]),
        "oracle_ran2": (c_double, [POINTER(ctypes.c_int32), POINTER(ctypes.c_int32)]),
        "oracle_langevin": (c_double, [POINTER(ctypes.c_int32), POINTER(ctypes.c_int32), c_double]),
        "oracle_set_noise": (c_int, [c_void_p, c_double, ctypes.c_int32, POINTER(ctypes.c_int32)]),
        "oracle_noise_accumulate": (None, [POINTER(ctypes.c_int32), POINTER(ctypes.c_int32), c_double, c_int,
                                           ctypes.c_long, D]),
        "oracle_get_noise": (c_int, [c_void_p, POINTER(ctypes.c_int32), POINTER(ctypes.c_int32), D]),
        "oracle_tlm_set": (c_int, [c_void_p, D, D, D, POINTER(MMAdjGrads)]),
        "oracle_tlm_get": (c_int, [c_void_p, D, D, D]),
        "oracle_tlm_off": (None, [c_void_p]),
    }
    for k, (r, a) in sig.items():
        f = getattr(lib, k)
        f.restype = r
        f.argtypes = a
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(POINTER(c_double))


class Oracle:
    def __init__(self, params: MMParams, nthreads: int = 1):
        self.lib = load()
        self.p = params
        self.h = self.lib.oracle_create(ctypes.byref(params), nthreads)
        if not self.h:
            raise ValueError("oracle_create rejected the parameters")
        self.nx, self.ny, self.nz = params.nx, params.ny, params.nz
        self.V, self.E = params.vartot, params.numel
        self.shape = (self.nz, self.ny, self.nx)

    def close(self):
        if self.h:
            self.lib.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_threads(self, n):
        self.lib.oracle_set_threads(self.h, n)

    def set_state(self, phi=None, cons=None, u=None, gradu=None, el_en=None):
        conv = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        phi, cons, u, gradu, el_en = map(conv, (phi, cons, u, gradu, el_en))
        self.lib.oracle_set_state(self.h, _p(phi), _p(cons), _p(u), _p(gradu), _p(el_en))

    def get_state(self):
        V, E = self.V, self.E
        phi = np.empty((V, *self.shape)); cons = np.empty((E, *self.shape))
        u = np.empty((3, *self.shape)); gu = np.empty((9, *self.shape)); el = np.empty(self.shape)
        self.lib.oracle_get_state(self.h, _p(phi), _p(cons), _p(u), _p(gu), _p(el))
        return dict(phi=phi, cons=cons, u=u, gradu=gu, el_en=el)

    def get_work(self, which):
        n = {0: 3 * self.V, 1: self.V, 2: self.V, 3: self.V, 4: self.V, 5: 3 * self.E, 6: self.V}[which]
        out = np.empty((n, *self.shape))
        self.lib.oracle_get_work(self.h, which, _p(out))
        return out

    @property
    def step_count(self):
        return self.lib.oracle_get_step(self.h)

    def get_lagged(self):
        v = np.empty(4)
        self.lib.oracle_get_lagged(self.h, _p(v))
        return v

    def set_lagged(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        self.lib.oracle_set_lagged(self.h, _p(v))

    def set_step(self, s):
        self.lib.oracle_set_step(self.h, s)

    def step(self, n=1):
        rows = []
        for _ in range(n):
            d = MMDiag()
            rc = self.lib.oracle_step(self.h, ctypes.byref(d))
            if rc:
                raise RuntimeError(f"oracle_step error {rc}")
            rows.append(d.as_dict(self.E))
        return rows

    def f_trans(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        H = self.nx // 2 + 1
        out = np.empty((self.nz, self.ny, H), dtype=np.complex128)
        self.lib.oracle_f_trans(self.h, _p(x), out.ctypes.data_as(POINTER(c_double)))
        return out

    def inv_trans(self, s):
        s = np.ascontiguousarray(s, dtype=np.complex128)
        out = np.empty(self.shape)
        self.lib.oracle_inv_trans(self.h, s.ctypes.data_as(POINTER(c_double)), _p(out))
        return out

    def set_noise(self, sd, idum, state=None):
        st = None if state is None else (ctypes.c_int32 * 35)(*state)
        rc = self.lib.oracle_set_noise(self.h, sd, idum, st)
        if rc:
            raise RuntimeError(f"oracle_set_noise error {rc}")

    def get_noise(self):
        idum = ctypes.c_int32(0)
        st = (ctypes.c_int32 * 35)()
        pn = np.empty((self.V, *self.shape))
        self.lib.oracle_get_noise(self.h, ctypes.byref(idum), st, _p(pn))
        return idum.value, list(st), pn

    def elastic_zeroth(self, pphi):
        """The zeroth-order guess alone from P_phi fields (V, nz, ny, nx):
        returns (u (3, ...), gradu (9, ...))."""
        pphi = np.ascontiguousarray(pphi, dtype=np.float64)
        u = np.empty((3, *self.shape)); gu = np.empty((9, *self.shape))
        rc = self.lib.oracle_elastic_zeroth(self.h, _p(pphi), _p(u), _p(gu))
        if rc:
            raise RuntimeError(f"oracle_elastic_zeroth error {rc}")
        return u, gu

    def set_pphi(self):
        """P_phi = p(phi) of the current phi (input of eldis / elener)."""
        self.lib.oracle_set_pphi(self.h)

    def elener(self):
        """elener alone at the current phi, P_phi and grad u; returns
        (rc, el_en, el_grad)."""
        rc = self.lib.oracle_elener(self.h)
        return rc, self.get_state()["el_en"], self.get_work(4)

    def tlm_set(self, dphi=None, dcons=None, dgradu=None, dpar: MMAdjGrads = None):
        """Start the tangent-linear run: every following step also advances
        the tangent of (phi, c, grad u) along this direction (state arrays
        and a parameter direction in the mm_adj_grads layout; None = 0)."""
        conv = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        dphi, dcons, dgradu = map(conv, (dphi, dcons, dgradu))
        rc = self.lib.oracle_tlm_set(self.h, _p(dphi), _p(dcons), _p(dgradu),
                                     None if dpar is None else ctypes.byref(dpar))
        if rc:
            raise RuntimeError(f"oracle_tlm_set error {rc}")

    def tlm_get(self):
        V, E = self.V, self.E
        dphi = np.empty((V, *self.shape)); dcons = np.empty((E, *self.shape)); dgu = np.empty((9, *self.shape))
        rc = self.lib.oracle_tlm_get(self.h, _p(dphi), _p(dcons), _p(dgu))
        if rc:
            raise RuntimeError(f"oracle_tlm_get error {rc}")
        return dict(phi=dphi, cons=dcons, gradu=dgu)

    def tlm_off(self):
        self.lib.oracle_tlm_off(self.h)

    def eldis(self):
        n = c_int(0)
        cf = c_double(0)
        rc = self.lib.oracle_eldis(self.h, ctypes.byref(n), ctypes.byref(cf))
        return rc, n.value, cf.value


def invert(a: np.ndarray, flag: int = 1):
    """oracle_invert on a copy of the (n,n) matrix; returns (rc, inverse)."""
    lib = load()
    n = a.shape[0]
    f = np.asfortranarray(a, dtype=np.float64).copy(order="F")
    rc = lib.oracle_invert(f.ctypes.data_as(POINTER(c_double)), n, flag)
    return rc, np.array(f, order="C")


def kaxis(n: int, count: int):
    lib = load()
    kf = np.empty(count); kq = np.empty(count)
    lib.oracle_kaxis(n, count, _p(kf), _p(kq))
    return kf, kq
