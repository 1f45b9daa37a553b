"""Decaying isotropic turbulence initial condition.

Restates GKM/CUDA GKM/InitialConditions/init_BS.m (Erlebacher's method) with a
seeded generator in place of MATLAB's ``rand``:

* u, v, w ~ U[0, 1) on the N^3 grid, transformed to Fourier space (:24-26);
* projected onto the solenoidal subspace, u_hat -= (k . u_hat) k / |k|^2
  (:29-37);
* every shell |k| in [m - 1/2, m + 1/2) rescaled to
  E(k) = sqrt(2/pi) (16/3) Mt^2 k^4 exp(-2 k^2 / k0^2) N^6 / k0^5 (:40-83);
* back-transformed, real part (:89-91);
* u_rms = sqrt(<u_i u_i>/3), c0 = sqrt(3) u_rms / Mt, p0 = rho0 c0^2 / gamma,
  mu0 = u_rms (2/k0) rho0 / Re, eddy time t0 = (2/k0) / u_rms (:116-157).

The script assumes odd N (krange = -(N-1)/2 .. (N-1)/2).  For even N the
Nyquist planes (any |k_i| = N/2) are set to zero, so the field stays real.

Axis convention: arrays are indexed [j, i, k] = (y, x, z), the library's
layout; u (x-velocity) is paired with the wavenumber along axis 1, v with
axis 0, w with axis 2, which is how init_BS.m's meshgrid pairs them once the
file is read back by weno/src/FlowField.cpp:87-102.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class DhitIC:
    u: np.ndarray  # (ny, nx, nz) x-velocity
    v: np.ndarray
    w: np.ndarray
    rho0: float
    p0: float
    mu0: float
    t0: float       # eddy turnover time (2/k0)/u_rms
    urms: float
    c0: float
    Mt: float
    Re: float
    k0: float
    gamma: float


def _wavenumbers(n: int) -> np.ndarray:
    return np.fft.fftfreq(n, d=1.0 / n)


def dhit_velocity(n: int, k0: float = 4.0, Mt: float = 0.3, seed: int = 1234, ny: int | None = None,
                  nz: int | None = None):
    """Random solenoidal field with the init_BS.m spectrum on an (ny, n, nz) grid."""
    nx = n
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    rng = np.random.default_rng(seed)
    shape = (ny, nx, nz)
    u = rng.random(shape)
    v = rng.random(shape)
    w = rng.random(shape)
    uk, vk, wk = np.fft.fftn(u), np.fft.fftn(v), np.fft.fftn(w)
    ky = _wavenumbers(ny)[:, None, None]
    kx = _wavenumbers(nx)[None, :, None]
    kz = _wavenumbers(nz)[None, None, :]
    ksq = np.maximum(kx * kx + ky * ky + kz * kz, 1e-14)
    kdotv = kx * uk + ky * vk + kz * wk
    uk = uk - kdotv * kx / ksq
    vk = vk - kdotv * ky / ksq
    wk = wk - kdotv * kz / ksq
    # Nyquist planes of even extents cannot stay Hermitian: drop them
    nyq = np.zeros(shape, dtype=bool)
    if ny % 2 == 0:
        nyq |= np.broadcast_to(np.abs(ky) == ny // 2, shape)
    if nx % 2 == 0:
        nyq |= np.broadcast_to(np.abs(kx) == nx // 2, shape)
    if nz % 2 == 0:
        nyq |= np.broadcast_to(np.abs(kz) == nz // 2, shape)
    kmag = np.sqrt(kx * kx + ky * ky + kz * kz) * np.ones(shape)
    shell = np.rint(kmag).astype(np.int64)  # |k| = sqrt(int) is never a half-integer
    e = np.abs(uk) ** 2 + np.abs(vk) ** 2 + np.abs(wk) ** 2
    e[nyq] = 0.0
    nshell = int(shell.max()) + 1
    espec = np.bincount(shell.ravel(), weights=e.ravel(), minlength=nshell)
    espec = np.maximum(espec, 1e-14)
    npts = float(nx * ny * nz)
    kk = np.arange(nshell, dtype=np.float64)
    e_ex = np.sqrt(2 / np.pi) * (16.0 / 3.0) * Mt ** 2 * kk ** 4 * np.exp(-2 * kk ** 2 / k0 ** 2) * npts ** 2 / k0 ** 5
    fac = np.sqrt(e_ex / espec)[shell]
    fac[nyq] = 0.0
    u = np.real(np.fft.ifftn(uk * fac))
    v = np.real(np.fft.ifftn(vk * fac))
    w = np.real(np.fft.ifftn(wk * fac))
    return u, v, w


def dhit(n: int, k0: float = 4.0, Mt: float = 0.3, Re: float = 30.0, seed: int = 1234, rho0: float = 1.0,
         gamma: float = 1.4, ny: int | None = None, nz: int | None = None) -> DhitIC:
    u, v, w = dhit_velocity(n, k0, Mt, seed, ny, nz)
    q2 = float(np.mean(u * u + v * v + w * w))
    urms = np.sqrt(q2 / 3.0)
    c0 = np.sqrt(3.0) * urms / Mt
    p0 = rho0 * c0 * c0 / gamma
    lam = 2.0 / k0
    mu0 = urms * lam * rho0 / Re
    t0 = lam / urms
    return DhitIC(u=u, v=v, w=w, rho0=rho0, p0=p0, mu0=mu0, t0=t0, urms=urms, c0=c0, Mt=Mt, Re=Re, k0=k0,
                  gamma=gamma)


def dhit_native(n: int, k0: float = 4.0, Mt: float = 0.3, Re: float = 30.0, seed: int = 1234, rho0: float = 1.0,
                gamma: float = 1.4, ny: int | None = None, nz: int | None = None) -> DhitIC:
    """The same recipe generated by the library (gkm_dhit_velocity, host code,
    SplitMix64 random numbers): the field differs from dhit()'s for a given
    seed, the spectrum and the derived scales do not."""
    import ctypes as C

    from . import _abi

    nx = n
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    u = np.empty((ny, nx, nz))
    v = np.empty((ny, nx, nz))
    w = np.empty((ny, nx, nz))
    icp = _abi.GkmIc(int(seed) & 0xFFFFFFFFFFFFFFFF, float(k0), float(Mt), float(Re), float(rho0))
    info = _abi.GkmIcInfo()
    dp = C.POINTER(C.c_double)
    _abi.check(_abi.load().gkm_dhit_velocity(nx, ny, nz, C.byref(icp), C.c_double(gamma), u.ctypes.data_as(dp),
                                             v.ctypes.data_as(dp), w.ctypes.data_as(dp), C.byref(info)))
    return DhitIC(u=u, v=v, w=w, rho0=rho0, p0=info.p0, mu0=info.mu0, t0=info.t0, urms=info.urms, c0=info.c0, Mt=Mt,
                  Re=Re, k0=k0, gamma=gamma)


def conservative_padded(ic: DhitIC, ghost: int = 3):
    """Padded conservative planes W0..W4 (ghosts zero) from uniform rho0, p0 and
    the IC velocity, as weno/src/FlowField.cpp:97-102 forms them."""
    ny, nx, nz = ic.u.shape
    P = (ny + 2 * ghost, nx + 2 * ghost, nz + 2 * ghost)
    W = [np.zeros(P) for _ in range(5)]
    s = (slice(ghost, ghost + ny), slice(ghost, ghost + nx), slice(ghost, ghost + nz))
    d0 = np.full((ny, nx, nz), ic.rho0)
    pr0 = np.full((ny, nx, nz), ic.p0)
    vsqr = ic.u * ic.u + ic.v * ic.v + ic.w * ic.w
    W[0][s] = d0
    W[1][s] = ic.u * d0
    W[2][s] = ic.v * d0
    W[3][s] = ic.w * d0
    W[4][s] = d0 * (pr0 / (d0 * (ic.gamma - 1.0)) + 0.5 * vsqr)
    return [np.ascontiguousarray(x.reshape(-1)) for x in W]
