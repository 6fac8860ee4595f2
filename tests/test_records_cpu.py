"""The reference's restart / dump records (data_save.NNNN) and run-file rows:
round trips, the subrecord framing, and byte-for-byte agreement with the
unformatted I/O and the E15.7 editing of flang (the Fortran compiler of this
image) running the reference's own read / write statements."""
import os
import subprocess

import numpy as np
import pytest

from meumapps_ss_amd.records import (fort98_row, read_data_save, write_data_save, write_extract_in)

FLANG = "/opt/rocm/lib/llvm/bin/flang"


def _arrays(shape, V, E, seed=0):
    rng = np.random.default_rng(seed)
    n3 = int(np.prod(shape))
    return dict(cons=rng.standard_normal(E * n3), phi=rng.standard_normal(V * n3), el_en=rng.standard_normal(n3),
                u=rng.standard_normal((3, n3)), gradu=rng.standard_normal((9, n3)))


@pytest.mark.parametrize("max_sub", [2 ** 31 - 9, 100])
def test_round_trip(tmp_path, max_sub):
    shape, V, E = (4, 6, 8), 3, 2
    a = _arrays(shape, V, E)
    p = tmp_path / "data_save.0001"
    write_data_save(str(p), shape, a["cons"], a["phi"], a["el_en"], a["u"], a["gradu"], max_sub=max_sub)
    b = read_data_save(str(p), V, E)
    assert b["shape"] == shape
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_decomposed_file_is_refused(tmp_path):
    import struct
    p = tmp_path / "d"
    with open(p, "wb") as f:
        f.write(struct.pack("<i6ii", 24, 1, 1, 5, 4, 4, 8, 24))
    with pytest.raises(ValueError):
        read_data_save(str(p), 1, 1)


FPROG = """
program rt
  implicit none
  integer, parameter :: V = 3, E = 2, nx = 4, ny = 6, nz = 8
  integer :: ist(3), ien(3)
  real*8 :: cons(E, nx, ny, nz), phi(V, nx, ny, nz), el_en(nx, ny, nz)
  real*8 :: u1(nx, ny, nz), u2(nx, ny, nz), u3(nx, ny, nz)
  real*8 :: g11(nx, ny, nz), g12(nx, ny, nz), g13(nx, ny, nz), g21(nx, ny, nz), g22(nx, ny, nz)
  real*8 :: g23(nx, ny, nz), g31(nx, ny, nz), g32(nx, ny, nz), g33(nx, ny, nz)
  open(5, file='data_save.0001', status='old', form='unformatted')
  read(5) ist, ien
  read(5) cons, phi, el_en
  read(5) u1, u2, u3
  read(5) g11, g12, g13
  read(5) g21, g22, g23
  read(5) g31, g32, g33
  close(5)
  open(4, file='data_save.copy', form='unformatted')
  rewind(4)
  write(4) ist, ien
  write(4) cons, phi, el_en
  write(4) u1, u2, u3
  write(4) g11, g12, g13
  write(4) g21, g22, g23
  write(4) g31, g32, g33
  close(4)
  print '(6I6)', ist, ien
  print '(3ES26.17E3)', cons(2, 3, 4, 5), phi(3, 1, 6, 8), g32(2, 2, 2)
  write(*, '(i8, 10e15.7)') 123, 1.0d0, -2.5d-3, 6.02214076d23, 0.0d0, -1.0d0, 9.99999999d0, 1.0d-30, &
       0.31415926535d0, -123456.789d0
end program rt
"""


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not in this image")
def test_flang_reads_and_writes_the_same_bytes(tmp_path):
    shape, V, E = (4, 6, 8), 3, 2
    nx, ny, nz = shape
    a = _arrays(shape, V, E, seed=3)
    p = tmp_path / "data_save.0001"
    write_data_save(str(p), shape, a["cons"], a["phi"], a["el_en"], a["u"], a["gradu"])
    (tmp_path / "rt.f90").write_text(FPROG)
    subprocess.run([FLANG, "rt.f90", "-o", "rt"], cwd=tmp_path, check=True, capture_output=True)
    out = subprocess.run(["./rt"], cwd=tmp_path, check=True, capture_output=True, text=True).stdout.splitlines()
    assert (tmp_path / "data_save.copy").read_bytes() == p.read_bytes()
    assert [int(x) for x in out[0].split()] == [1, 1, 1, nx, ny, nz]
    vals = [float(x) for x in out[1].split()]

    def aos(arr, ncomp, c, x, y, z):  # 1-based Fortran indices
        return arr[(c - 1) + ncomp * ((x - 1) + nx * ((y - 1) + ny * (z - 1)))]
    assert vals[0] == aos(a["cons"], E, 2, 3, 4, 5)
    assert vals[1] == aos(a["phi"], V, 3, 1, 6, 8)
    # g32 = grad_3_u2: record 6, second array
    assert vals[2] == a["gradu"][7][(2 - 1) + nx * ((2 - 1) + ny * (2 - 1))]
    row = fort98_row(dict(step=123, fint=1.0, fw=-2.5e-3, fch=6.02214076e23, elen=0.0, etot=-1.0,
                          glob_min=9.99999999, glob_max=1.0e-30, phi_max=0.31415926535, solid_tot=-123456.789))
    assert row == out[2]


def test_extract_in(tmp_path):
    p = tmp_path / "data_extract_mult.in"
    write_extract_in(str(p), (64, 32, 16), 1, 3, 1, (1, 1))
    lines = p.read_text().split("\n")
    assert lines[0].split() == ["64", "32", "16", "1", "3", "1"] and lines[1].split() == ["1", "1"]
