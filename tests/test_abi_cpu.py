"""CPU checks of the C ABI library: it loads, exports every function that
include/meumapps_ss.h declares, and its host-only helpers behave (no GPU
compute is called here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from meumapps_ss_amd import _lib
from meumapps_ss_amd.params import MMDiag, MMParams

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "meumapps_ss.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mm_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 19
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.EXPORTED)


def test_struct_layouts_match_header():
    # offsets fixed by the header: 14 int32 then doubles
    assert MMParams.t_step.offset == 14 * 4 + 0 * 8 or MMParams.t_step.offset == 56
    assert ctypes.sizeof(MMParams) == 56 + 8 * (2 + 4 + 16 * 9 + 16 * 6 + 4 * 36 * 2 + 4 * 9 + 4 + 6)
    assert MMDiag.fint.offset == 8
    assert ctypes.sizeof(MMDiag) == 8 + 8 * (5 + 4 + 4 + 1)


def test_errstr_reference_messages():
    lib = _lib.load()
    assert b"Poor cond" in lib.mm_errstr(2)
    assert b"newfac is too small" in lib.mm_errstr(7)
    assert lib.mm_errstr(0) == b"ok"


def _ran2_bigint(seed, n):
    """ran_2.f:1-42 with exact modular arithmetic instead of Schrage's
    decomposition, and the binary32-folded am of the reference."""
    im1, im2, ia1, ia2, ntab = 2147483563, 2147483399, 40014, 40692, 32
    imm1 = im1 - 1
    ndiv = 1 + imm1 // ntab
    am = float(np.float32(np.float32(1.0) / np.float32(im1)))
    rnmx = 1.0 - 1.2e-7
    idum = max(-seed, 1)
    idum2 = idum
    iv = [0] * ntab
    for j in range(ntab + 8, 0, -1):
        idum = (ia1 * idum) % im1
        if j <= ntab:
            iv[j - 1] = idum
    iy = iv[0]
    out = []
    for _ in range(n):
        idum = (ia1 * idum) % im1
        idum2 = (ia2 * idum2) % im2
        j = iy // ndiv
        iy = iv[j] - idum2
        iv[j] = idum
        if iy < 1:
            iy += imm1
        out.append(min(am * iy, rnmx))
    return out


def test_ran2_matches_exact_lecuyer_recurrences():
    lib = _lib.load()
    st = (ctypes.c_int32 * 35)()
    idum = ctypes.c_int32(-7)
    got = [lib.mm_ran2(ctypes.byref(idum), st) for _ in range(2000)]
    assert got == _ran2_bigint(-7, 2000)
    assert all(0.0 < g < 1.0 for g in got)


def test_ic_uniform_layout_and_range():
    lib = _lib.load()
    n, V, E = 4, 3, 1
    phi = np.empty(V * n ** 3)
    cons = np.empty(E * n ** 3)
    cav = np.array([0.4])
    D = ctypes.POINTER(ctypes.c_double)
    rc = lib.mm_ic_uniform(n, n, n, V, E, -3, 0.01, cav.ctypes.data_as(D), phi.ctypes.data_as(D),
                           cons.ctypes.data_as(D))
    assert rc == 0
    r = _ran2_bigint(-3, (V + E) * n ** 3)
    # per cell (x fastest): c_1..c_E then phi_1..phi_V
    assert cons[0] == 0.4 + 0.01 * (2 * r[0] - 1)
    assert phi[0] == 0.01 * (2 * r[1] - 1)
    assert phi[n ** 3] == 0.01 * (2 * r[2] - 1)
    assert phi[1] == 0.01 * (2 * r[5] - 1)
    assert np.abs(phi).max() <= 0.01 and np.abs(cons - 0.4).max() <= 0.01


def test_create_rejects_bad_parameters_before_touching_a_device():
    lib = _lib.load()
    p = MMParams()
    p.nx = p.ny = p.nz = 16
    p.vartot, p.numel, p.numph = 3, 1, 2
    p.nvar[0], p.nvar[1] = 1, 2  # sum(Nvar) != vartot + 1 (initerr 1)
    ctx = ctypes.c_void_p()
    assert lib.mm_create(ctypes.byref(p), 0, ctypes.byref(ctx)) == 20001
    assert lib.mm_create(ctypes.byref(p), 7, ctypes.byref(ctx)) == 20001


FLANG = "/opt/rocm/lib/llvm/bin/flang"


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not in this image")
def test_fortran_binding_compiles_and_types_match(tmp_path):
    """The ISO_C_BINDING module compiles and its bind(c) types have the byte
    sizes of the C structs."""
    import subprocess

    from meumapps_ss_amd.params import MMAdjGrads

    src = tmp_path / "sizes.f90"
    src.write_text(
        "program sizes\n  use mm_abi\n  type(mm_params) :: p\n  type(mm_diag) :: d\n"
        "  type(mm_adj_grads) :: g\n  print '(3I8)', c_sizeof(p), c_sizeof(d), c_sizeof(g)\nend program sizes\n")
    mod = os.path.join(ROOT, "meumapps_ss_amd", "fortran", "mm_abi.f90")
    exe = tmp_path / "sizes"
    libdir = os.path.join(ROOT, "meumapps_ss_amd")
    _lib.load()  # the library is built
    subprocess.run([FLANG, "-J", str(tmp_path), mod, str(src), "-o", str(exe), "-L", libdir, "-lmeumapps_ss",
                    "-Wl,-rpath," + libdir, "-Wl,--allow-shlib-undefined"], check=True, cwd=tmp_path,
                   capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert [int(x) for x in out] == [ctypes.sizeof(MMParams), ctypes.sizeof(MMDiag), ctypes.sizeof(MMAdjGrads)]
