"""The data_save records this project writes, read by the reference's own
reader: data_ext_modular.f compiled with flang from the reference tree into
oracle/_ref (oracle/Makefile, target `ref`; skipped where it is absent).
The reader takes data_extract_mult.in and data_save.0001, reads record 1
(ist, ien) and record 2 (Cons, phi, el_en) with its own array bounds, and
writes every point as text (shape_con: i, j, k, Cons(1:nel), el_en;
shape_phi: i, j, k, the last variant above 0.85, sum phi^2, sum phi; phi_var:
the variant fractions), data_ext_modular.f:11-98."""
import os
import subprocess

import numpy as np
import pytest

from meumapps_ss_amd.records import dump_step, read_data_save
from oracle import pyoracle

EXE = pyoracle.REF_DATA_EXT
pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="oracle/_ref not built (reference tree absent)")


def _rows(path):
    return np.loadtxt(path, skiprows=1)


@pytest.mark.parametrize("shape,V,E", [((6, 5, 4), 3, 2), ((8, 8, 8), 1, 1), ((4, 6, 10), 12, 1)])
def test_reference_reader_reads_our_dump(tmp_path, shape, V, E):
    nx, ny, nz = shape
    n3 = nx * ny * nz
    rng = np.random.default_rng(nx * 100 + V)
    st = dict(cons=rng.random(E * n3), phi=rng.random(V * n3), el_en=rng.random(n3) - 0.5,
              u=rng.random((3, n3)), gradu=rng.random((9, n3)))
    d = dump_step(str(tmp_path), 11, shape, V, E, st)
    r = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    # Fortran order: point (i, j, k) 1-based, x fastest; components fastest in the record
    cons = st["cons"].reshape(n3, E)
    phi = st["phi"].reshape(n3, V)
    con = _rows(os.path.join(d, "shape_con"))
    assert con.shape == (n3, 3 + E + 1)
    i, j, k = (con[:, c].astype(int) for c in range(3))
    idx = (i - 1) + nx * ((j - 1) + ny * (k - 1))
    assert np.array_equal(np.sort(idx), np.arange(n3))
    # e15.7 keeps 7 significant digits (0.dddddddE+xx): relative error <= 5e-7
    np.testing.assert_allclose(con[:, 3:3 + E], cons[idx], rtol=5.1e-7, atol=1e-30)
    np.testing.assert_allclose(con[:, 3 + E], st["el_en"][idx], rtol=5.1e-7, atol=1e-30)
    ph = _rows(os.path.join(d, "shape_phi"))
    above = phi[idx] > float(np.float32(0.85))
    last = np.where(above.any(axis=1), V - np.argmax(above[:, ::-1], axis=1), 0)
    assert np.array_equal(ph[:, 3].astype(int), last)
    np.testing.assert_allclose(ph[:, 4], (phi[idx] ** 2).sum(axis=1), rtol=5.1e-7)
    np.testing.assert_allclose(ph[:, 5], phi[idx].sum(axis=1), rtol=5.1e-7)
    frac = np.loadtxt(os.path.join(d, "phi_var")).ravel()
    np.testing.assert_allclose(frac, phi.sum(axis=0) / phi.sum(), atol=6e-5)
    # and the same file read back by this project's reader
    back = read_data_save(os.path.join(d, "data_save.0001"), V, E)
    for key in ("cons", "phi", "el_en"):
        assert np.array_equal(back[key].ravel(), np.ravel(st[key]))
