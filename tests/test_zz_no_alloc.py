"""The allocation-free forward x-transform (ekm_fft_x_to_fourier_dev_noalloc)
against the allocating one (ekm_fft_x_to_fourier_dev), through the C ABI, on
two calls in a row, the second with a smaller transform; results must agree
bit for bit (same twiddles, same kernels).  Sizes above the scratch's
maximum are refused on the host."""
import ctypes

import numpy as np
import pytest

from ekman_bl_amd import solver as S

pytestmark = pytest.mark.gpu
_vp = ctypes.c_void_p


def _lib():
    L = S.lib()
    L.ekm_fft_x_to_fourier_dev_noalloc.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp]
    L.ekm_fft_x_to_fourier_dev.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp]
    L.ekm_dev_alloc.argtypes = [ctypes.POINTER(_vp), ctypes.c_size_t]
    L.ekm_dev_free.argtypes = [_vp]
    L.ekm_memcpy_h2d.argtypes = [_vp, _vp, ctypes.c_size_t]
    L.ekm_memcpy_d2h.argtypes = [_vp, _vp, ctypes.c_size_t]
    return L


def _run(L, fn, host):
    d = _vp()
    assert L.ekm_dev_alloc(ctypes.byref(d), host.nbytes) == 0
    assert L.ekm_memcpy_h2d(d, host.ctypes.data_as(_vp), host.nbytes) == 0
    nx = host.shape[-1] - 2
    st = fn(d, nx, host.size // (nx + 2), None)
    out = np.empty_like(host)
    assert L.ekm_memcpy_d2h(out.ctypes.data_as(_vp), d, out.nbytes) == 0
    L.ekm_dev_free(d)
    return st, out


def test_noalloc_transform_matches_the_allocating_one():
    L = _lib()
    rng = np.random.default_rng(21)
    for nx, npen in ((512, 3000), (64, 700)):
        x = rng.standard_normal((npen, nx + 2))
        s1, a = _run(L, L.ekm_fft_x_to_fourier_dev, x.copy())
        s2, b = _run(L, L.ekm_fft_x_to_fourier_dev_noalloc, x.copy())
        print(f"[noalloc] nx {nx}, {npen} pencils: status {s1}/{s2}, max |diff| {np.abs(a - b).max():.1e}", flush=True)
        assert s1 == 0 and s2 == 0
        assert np.array_equal(a, b)
        # and it is the transform: mode 1 of cos(2 pi i / nx) is 1/2
        c = np.cos(2 * np.pi * np.arange(nx) / nx)
        y = np.zeros((1, nx + 2))
        y[0, :nx] = c
        s3, z = _run(L, L.ekm_fft_x_to_fourier_dev_noalloc, y)
        assert s3 == 0 and abs(z[0, 2] - 0.5) < 1e-14 and abs(z[0, 3]) < 1e-14


def test_noalloc_refuses_sizes_beyond_its_scratch():
    L = _lib()
    x = np.zeros((2, 2048 + 2))
    st, _ = _run(L, L.ekm_fft_x_to_fourier_dev_noalloc, x)
    assert st == -1
    assert b"1024" in L.ekm_last_error()
