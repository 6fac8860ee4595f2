"""mm_fft_forward_noalloc (the forward transform that allocates nothing)
against mm_fft_forward on the default plans: the same kernels with the same
twiddle values, so the spectra are equal bit for bit.  Two calls in a row,
the second on a smaller grid, in each precision; numpy's rfftn / N^3 as a
sanity check; sizes beyond the tables are refused."""
import ctypes
import os

import numpy as np
import pytest

from meumapps_ss_amd import _lib

pytestmark = pytest.mark.gpu


def _noalloc(lib):
    f = lib.mm_fft_forward_noalloc
    f.restype = ctypes.c_int32
    f.argtypes = [ctypes.c_int32] * 4 + [ctypes.c_void_p] * 3
    return f


def _planned_forward(lib, x_dev, shape, prec, nc, dt):
    import torch

    nx, ny, nz = shape
    h = ctypes.c_void_p()
    os.environ["MM_FFT_VARIANT"] = "0"  # the default plans, as the no-alloc path runs them
    try:
        _lib.check(lib.mm_fft_create(nx, ny, nz, prec, ctypes.byref(h)), "mm_fft_create")
    finally:
        os.environ.pop("MM_FFT_VARIANT", None)
    try:
        spec = torch.zeros(2 * nc, dtype=dt, device="cuda")
        _lib.check(lib.mm_fft_forward(h, x_dev.data_ptr(), spec.data_ptr(), None), "mm_fft_forward")
        torch.cuda.synchronize()
        return spec.cpu().numpy()
    finally:
        lib.mm_fft_destroy(h)


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_forward_equals_planned_forward(gpu_lib, prec):
    import torch

    fwd = _noalloc(gpu_lib)
    dt = torch.float64 if prec == 0 else torch.float32
    stream = torch.cuda.current_stream().cuda_stream
    for shape in [(96, 64, 128), (16, 24, 32)]:  # the second call on a smaller grid
        nx, ny, nz = shape
        print(f"noalloc forward {shape} precision {prec}", flush=True)
        rng = np.random.default_rng(nx + 7 * ny + 31 * nz)
        x = rng.standard_normal((nz, ny, nx))
        H = nx // 2 + 1
        nc = H * ny * nz
        xd = torch.from_numpy(x).to(device="cuda", dtype=dt).contiguous()
        spec = torch.full((2 * nc,), float("nan"), dtype=dt, device="cuda")
        rc = fwd(nx, ny, nz, prec, xd.data_ptr(), spec.data_ptr(), stream)
        assert rc == 0, gpu_lib.mm_errstr(rc)
        torch.cuda.synchronize()
        got = spec.cpu().numpy()
        ref = _planned_forward(gpu_lib, xd, shape, prec, nc, dt)
        assert np.array_equal(got, ref), f"{shape}: max diff {np.abs(got - ref).max():.3e}"
        c = got.astype(np.float64).view(np.complex128).reshape(nz, H, ny).transpose(0, 2, 1)
        want = np.fft.rfftn(x) / (nx * ny * nz)
        err = np.abs(c - want).max() / np.abs(want).max()
        assert err < (1e-13 if prec == 0 else 2e-6), err


def test_noalloc_rejects_sizes_beyond_the_tables(gpu_lib):
    import torch

    fwd = _noalloc(gpu_lib)
    buf = torch.zeros(16, dtype=torch.float64, device="cuda")
    for shape in [(1024, 16, 16), (16, 1024, 16), (16, 16, 1024), (10, 16, 16)]:
        assert fwd(*shape, 0, buf.data_ptr(), buf.data_ptr(), None) == 20001  # MM_ERR_ARG
    assert fwd(16, 16, 16, 2, buf.data_ptr(), buf.data_ptr(), None) == 20001
    assert fwd(16, 16, 16, 0, None, buf.data_ptr(), None) == 20001
