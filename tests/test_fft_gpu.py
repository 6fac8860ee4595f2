"""GPU FFT (f_trans / inv_trans replacement) against numpy's pocketfft.

f_trans (ModFunc.f:15-26) = r2c 3D DFT with the forward sign, times 1/N^3.
inv_trans (ModFunc.f:30-40) = unnormalised c2r, imaginary parts of the kx = 0
and kx = Nx/2 entries ignored after the y and z passes.
Tolerances: binary64 rel. 1e-13 of the spectrum norm; binary32 rel. 2e-6.
"""
import ctypes

import numpy as np
import pytest

from meumapps_ss_amd import _lib

pytestmark = pytest.mark.gpu

SIZES = [(8, 8, 8), (16, 16, 16), (12, 12, 12), (24, 16, 8), (32, 32, 32), (48, 48, 48),
         (64, 64, 64), (96, 96, 96), (128, 128, 128), (16, 192, 24), (256, 64, 32),
         (8, 384, 16), (16, 16, 512), (512, 16, 8)]


def _dev(nbytes):
    import torch

    return torch.empty(nbytes, dtype=torch.uint8, device="cuda")


def _spec_to_numpy(buf, nx, ny, nz):
    H = nx // 2 + 1
    return buf.reshape(nz, H, ny).transpose(0, 2, 1)  # [kz, ky, kx]


def _numpy_to_spec(a):
    return np.ascontiguousarray(a.transpose(0, 2, 1))  # [kz][kx][ky]


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("variant", [None, "0", "1", "2"])
@pytest.mark.parametrize("shape", SIZES)
def test_forward_inverse(gpu_lib, shape, prec, variant):
    """variant None: the autotuner's plans; 0/1/2: every pass forced to the
    default, lean (radix 8) or lean-4 (radix 4) plan of its length."""
    import os

    import torch

    nx, ny, nz = shape
    rng = np.random.default_rng(nx * 1000 + ny * 10 + nz)
    x = rng.standard_normal((nz, ny, nx))
    h = ctypes.c_void_p()
    if variant is not None:
        os.environ["MM_FFT_VARIANT"] = variant
    try:
        _lib.check(gpu_lib.mm_fft_create(nx, ny, nz, prec, ctypes.byref(h)), "mm_fft_create")
    finally:
        os.environ.pop("MM_FFT_VARIANT", None)
    try:
        H = nx // 2 + 1
        nc = gpu_lib.mm_fft_spectrum_elems(h)
        assert nc == H * ny * nz
        tdt = torch.float64 if prec == 0 else torch.float32
        xr = torch.from_numpy(x).to(device="cuda", dtype=tdt).contiguous()
        spec = torch.zeros(2 * nc, dtype=tdt, device="cuda")
        _lib.check(gpu_lib.mm_fft_forward(h, xr.data_ptr(), spec.data_ptr(), None), "forward")
        torch.cuda.synchronize()
        s = spec.cpu().numpy().astype(np.float64).view(np.complex128)
        got = _spec_to_numpy(s, nx, ny, nz)
        ref = np.fft.rfftn(x) / (nx * ny * nz)
        tol = 1e-13 if prec == 0 else 2e-6
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert err < tol, f"forward rel err {err}"
        # inverse of a random (non-Hermitian) half spectrum: FFTW c2r semantics
        c = rng.standard_normal((nz, ny, H)) + 1j * rng.standard_normal((nz, ny, H))
        cs = torch.from_numpy(_numpy_to_spec(c).view(np.float64).ravel()).to(device="cuda", dtype=tdt)
        out = torch.zeros(nz * ny * nx, dtype=tdt, device="cuda")
        _lib.check(gpu_lib.mm_fft_inverse(h, cs.data_ptr(), out.data_ptr(), None), "inverse")
        torch.cuda.synchronize()
        got_r = out.cpu().numpy().astype(np.float64).reshape(nz, ny, nx)
        # y,z inverse c2c, drop Im of kx = 0 and Nyquist, then 1D c2r along x
        t = np.fft.ifft(np.fft.ifft(c, axis=0), axis=1) * (ny * nz)
        t[..., 0] = t[..., 0].real
        t[..., nx // 2] = t[..., nx // 2].real
        ref_r = np.fft.irfft(t, n=nx, axis=2) * nx
        err = np.abs(got_r - ref_r).max() / np.abs(ref_r).max()
        assert err < tol * 10, f"inverse rel err {err}"
    finally:
        gpu_lib.mm_fft_destroy(h)


def test_roundtrip_identity(gpu_lib):
    import torch

    nx = ny = nz = 64
    rng = np.random.default_rng(7)
    x = rng.standard_normal((nz, ny, nx))
    h = ctypes.c_void_p()
    _lib.check(gpu_lib.mm_fft_create(nx, ny, nz, 0, ctypes.byref(h)), "mm_fft_create")
    try:
        nc = gpu_lib.mm_fft_spectrum_elems(h)
        xr = torch.from_numpy(x).cuda()
        spec = torch.zeros(2 * nc, dtype=torch.float64, device="cuda")
        out = torch.zeros_like(xr)
        _lib.check(gpu_lib.mm_fft_forward(h, xr.data_ptr(), spec.data_ptr(), None), "fwd")
        _lib.check(gpu_lib.mm_fft_inverse(h, spec.data_ptr(), out.data_ptr(), None), "inv")
        torch.cuda.synchronize()
        assert np.abs(out.cpu().numpy() - x).max() < 1e-13
    finally:
        gpu_lib.mm_fft_destroy(h)
