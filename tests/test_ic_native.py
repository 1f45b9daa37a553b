"""The library's initial-condition generator (gkm_dhit_velocity, host code:
radix-2 and Bluestein DFTs, SplitMix64 random numbers) against the
properties the init_BS.m recipe fixes and against the numpy restatement
gkm_amd.ic.dhit.  The random field depends on the generator, the rescaled
spectrum and the derived scales do not:

* every shell |k| ~ m holds exactly E(m) = sqrt(2/pi)(16/3) Mt^2 m^4
  exp(-2 m^2/k0^2) N^2/k0^5 of energy (numpy's FFT of the output, so the
  library's inverse transform is checked by an independent one);
* k . u_hat = 0 (solenoidal), zero mean, Nyquist planes empty;
* q2, u_rms, c0, p0, mu0, t0 equal the numpy restatement's;
* a seed always gives the same field.
Extents cover powers of two (radix-2), odd and even non-powers of two
(Bluestein) and uneven boxes.
"""
import math

import numpy as np
import pytest

import gkm_amd
from gkm_amd.ic import dhit, dhit_native


def shell_spectrum(u, v, w):
    ny, nx, nz = u.shape
    ky = np.fft.fftfreq(ny, 1.0 / ny)[:, None, None]
    kx = np.fft.fftfreq(nx, 1.0 / nx)[None, :, None]
    kz = np.fft.fftfreq(nz, 1.0 / nz)[None, None, :]
    uk, vk, wk = np.fft.fftn(u), np.fft.fftn(v), np.fft.fftn(w)
    kmag = np.sqrt(kx * kx + ky * ky + kz * kz) * np.ones(u.shape)
    shell = np.rint(kmag).astype(np.int64)
    e = np.abs(uk) ** 2 + np.abs(vk) ** 2 + np.abs(wk) ** 2
    div = np.abs(kx * uk + ky * vk + kz * wk)
    return np.bincount(shell.ravel(), weights=e.ravel()), div, (uk, vk, wk), (kx, ky, kz)


def e_exact(m, Mt, k0, N):
    return math.sqrt(2 / math.pi) * (16 / 3) * Mt ** 2 * m ** 4 * math.exp(-2 * m * m / k0 ** 2) * N * N / k0 ** 5


SHAPES = [(16, 16, 16), (15, 15, 15), (12, 10, 20), (9, 24, 7), (32, 32, 32)]


@pytest.mark.parametrize("shape", SHAPES)
def test_spectrum_divergence_and_mean(shape):
    nx, ny, nz = shape
    Mt, k0 = 0.4, 3.0
    ic = dhit_native(nx, k0=k0, Mt=Mt, Re=40, seed=11, ny=ny, nz=nz)
    N = nx * ny * nz
    espec, div, (uk, vk, wk), _ = shell_spectrum(ic.u, ic.v, ic.w)
    scale = max(e_exact(m, Mt, k0, N) for m in range(len(espec)))
    for m, e in enumerate(espec):
        if e > 1e-12 * scale:
            assert abs(e - e_exact(m, Mt, k0, N)) <= 1e-10 * e_exact(m, Mt, k0, N), (m, e)
    amp = np.sqrt(np.abs(uk) ** 2 + np.abs(vk) ** 2 + np.abs(wk) ** 2).max()
    assert div.max() <= 1e-11 * amp * max(shape)
    for f in (ic.u, ic.v, ic.w):
        assert abs(f.mean()) <= 1e-14 * np.abs(f).max()
    # Nyquist planes of even extents carry nothing
    for ax, n in enumerate((ny, nx, nz)):
        if n % 2 == 0:
            sl = [slice(None)] * 3
            sl[ax] = n // 2
            assert np.abs(uk[tuple(sl)]).max() <= 1e-9 * amp


@pytest.mark.parametrize("shape", [(16, 16, 16), (12, 10, 20)])
def test_derived_scales_match_numpy_recipe(shape):
    nx, ny, nz = shape
    a = dhit_native(nx, k0=4, Mt=0.5, Re=70, seed=3, ny=ny, nz=nz, gamma=1.4)
    b = dhit(nx, k0=4, Mt=0.5, Re=70, seed=3, ny=ny, nz=nz, gamma=1.4)
    assert not np.allclose(a.u, b.u)  # different random fields
    for key in ("urms", "c0", "p0", "mu0", "t0"):
        assert abs(getattr(a, key) - getattr(b, key)) <= 1e-12 * abs(getattr(b, key)), key
    # u_rms / c0 = Mt / sqrt(3)
    assert abs(a.urms / a.c0 - 0.5 / math.sqrt(3)) <= 1e-14


def test_seed_determinism():
    a = dhit_native(12, seed=5)
    b = dhit_native(12, seed=5)
    c = dhit_native(12, seed=6)
    assert np.array_equal(a.u, b.u) and np.array_equal(a.w, b.w)
    assert not np.array_equal(a.u, c.u)


def test_argument_checks():
    import ctypes as C

    import gkm_amd._abi as A
    L = A.load()
    icp = A.GkmIc(1, 4.0, 0.0, 30.0, 1.0)  # Mt = 0
    x = np.empty(8)
    dp = C.POINTER(C.c_double)
    rc = L.gkm_dhit_velocity(2, 2, 2, C.byref(icp), C.c_double(1.4), x.ctypes.data_as(dp), x.ctypes.data_as(dp),
                             x.ctypes.data_as(dp), None)
    assert rc == A.GKM_E_ARG
    assert b"Mt" in L.gkm_last_error()
    with pytest.raises(gkm_amd.GkmError):
        dhit_native(8, k0=-1.0)
