"""Closed-form pin of the oracle's interior statistics (`orc_stats`, CPU only).

The reference defines none of these quantities for the weno path (its PrintStats
uses spline gradients, weno/src/update.cpp:638-784, pinned in
test_print_stats.py); the GKM tree prints TKE and rho_rms/Mt^2
(GKM/CUDA GKM/src/GPUupdate.cu:425-483). `orc_stats` states its own
definitions, and these fields fix them by hand:

  * a Fourier mode f = A sin(kappa x) has the exact second-order central
    difference (f(x+h) - f(x-h)) / 2h = A sin(kappa h)/h cos(kappa x), and the
    grid means of sin^2, cos^2 over whole periods are 1/2, of cos^3 zero and of
    cos^4 3/8;
  * E_k, q^2 and Mt of such a mode follow by hand;
  * a velocity field whose every component is independent of its own axis has
    zero discrete dilatation; its enstrophy and S_ij S_ij are sums of the modes;
  * 1/(1 + e cos t) averages to 1/sqrt(1 - e^2) over a period (to e^n aliasing).

A wrong stride (d/dx taken along y), a swapped component, a missing factor
(the 4/3 of the dilatational dissipation, the 1/2 of S_ij) or a wrong
normalisation fails one of these.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

GAM, RGAS = 1.4, 287.0
NY, NX, NZ = 12, 16, 20       # padded index order (y, x, z), as in the library
DX = 2 * math.pi / 16


def _coords():
    j = np.arange(NY)[:, None, None] * DX
    i = np.arange(NX)[None, :, None] * DX
    k = np.arange(NZ)[None, None, :] * DX
    y = np.broadcast_to(j, (NY, NX, NZ))
    x = np.broadcast_to(i, (NY, NX, NZ))
    z = np.broadcast_to(k, (NY, NX, NZ))
    return x, y, z


def _kappa(m, n):
    """Wavenumber of m periods over n cells of width DX."""
    return 2 * math.pi * m / (n * DX)


def _state(rho, u, v, w, p):
    W = [np.zeros((NY + 6, NX + 6, NZ + 6)) for _ in range(5)]
    s = (slice(3, 3 + NY), slice(3, 3 + NX), slice(3, 3 + NZ))
    W[0][s] = rho
    W[1][s] = rho * u
    W[2][s] = rho * v
    W[3][s] = rho * w
    W[4][s] = p / (GAM - 1) + 0.5 * rho * (u * u + v * v + w * w)
    return [np.ascontiguousarray(a.reshape(-1)) for a in W]


def _stats(orc, W, T0, visc_law=1):
    P = orc.make_params(gamma=GAM, R=RGAS, mu0=2e-3, T0=T0, visc_law=visc_law)
    return orc.stats(orc.make_grid(NX, NY, NZ, DX), P, W)


def _zero(v, scale):
    assert abs(v) <= 1e-12 * scale, (v, scale)


@pytest.mark.parametrize("axis", ["x", "y", "z"])
@pytest.mark.parametrize("visc_law", [0, 1])
def test_single_longitudinal_mode(orc, axis, visc_law):
    """u_a = A sin(kappa a) along one axis a, uniform rho and p."""
    x, y, z = _coords()
    n = {"x": NX, "y": NY, "z": NZ}[axis]
    coord = {"x": x, "y": y, "z": z}[axis]
    # 8, 12 and 20 points per period: the grid reaches the crest (points per period
    # divisible by 4) and cos^4 does not alias (more than 4 points per period)
    m = {"x": 2, "y": 1, "z": 1}[axis]
    A, rho0, p0 = 0.3, 1.1, 0.9
    kap = _kappa(m, n)
    vel = A * np.sin(kap * coord)
    zero = np.zeros_like(vel)
    u, v, w = {"x": (vel, zero, zero), "y": (zero, vel, zero), "z": (zero, zero, vel)}[axis]
    T = p0 / (RGAS * rho0)
    s = _stats(orc, _state(np.full_like(vel, rho0), u, v, w, np.full_like(vel, p0)), T0=T, visc_law=visc_law)
    N = NX * NY * NZ
    c = math.sqrt(GAM * p0 / rho0)
    D = A * math.sin(kap * DX) / DX          # amplitude of the central difference
    # E_k, q^2, Mt by hand
    assert s["q2"] == pytest.approx(A * A / 2, rel=1e-13)
    assert s["tke"] == pytest.approx(rho0 * A * A / 2, rel=1e-13)
    assert s["ek"] == pytest.approx(rho0 * A * A / 4, rel=1e-13)
    assert s["c_mean"] == pytest.approx(c, rel=1e-13)
    assert s["mt"] == pytest.approx(math.sqrt(A * A / 2) / c, rel=1e-13)
    assert s["umax_plus_c"] == pytest.approx(A + c, rel=1e-13)
    assert s["rho_mean"] == pytest.approx(rho0, rel=1e-14)
    assert s["p_mean"] == pytest.approx(p0, rel=1e-13)
    assert s["T_mean"] == pytest.approx(T, rel=1e-13)
    _zero(s["rho_rms"], rho0)
    _zero(s["p_rms"], p0)
    # gradient moments: only the longitudinal derivative of component a is non-zero
    g2, g3, g4 = {"x": ("dudx2", "dudx3", "dudx4"), "y": ("dvdy2", "dvdy3", "dvdy4"),
                  "z": ("dwdz2", "dwdz3", "dwdz4")}[axis]
    sk, ku = {"x": ("skew_u", "kurt_u"), "y": ("skew_v", "kurt_v"), "z": ("skew_w", "kurt_w")}[axis]
    assert s[g2] == pytest.approx(D * D / 2, rel=1e-12)
    _zero(s[g3], D ** 3)
    assert s[g4] == pytest.approx(3 * D ** 4 / 8, rel=1e-12)
    _zero(s[sk], 1.0)                         # a symmetric mode has no skewness
    assert s[ku] == pytest.approx(1.5, rel=1e-12)
    for other in {"dudx2", "dvdy2", "dwdz2"} - {g2}:
        assert s[other] == 0.0
    mu = 2e-3                                 # T = T0, so mu = mu0 for both laws
    assert s["dil2"] == pytest.approx(D * D / 2, rel=1e-12)
    assert s["diss_dil"] == pytest.approx(4 / 3 * mu * D * D / 2, rel=1e-12)
    assert s["diss_ref"] == pytest.approx(mu * D * D / 2, rel=1e-12)
    assert s["sijsij"] == pytest.approx(D * D / 2, rel=1e-12)
    assert s["enstrophy"] == 0.0 and s["diss_sol"] == 0.0
    # raw sums
    assert s["mass"] == pytest.approx(N * rho0, rel=1e-13)
    _zero(s["mom_x"] + s["mom_y"] + s["mom_z"], N * A * rho0)
    assert s["energy"] == pytest.approx(N * (p0 / (GAM - 1) + rho0 * A * A / 4), rel=1e-13)


def test_solenoidal_shear_field(orc):
    """u = A sin(kappa y), v = B sin(kappa z), w = C sin(kappa x): no component
    varies along its own axis, so the discrete dilatation is zero; the
    vorticity and strain are the three shear modes."""
    x, y, z = _coords()
    A, B, C = 0.2, -0.15, 0.1
    ky, kz, kx = _kappa(1, NY), _kappa(2, NZ), _kappa(3, NX)
    u, v, w = A * np.sin(ky * y), B * np.sin(kz * z), C * np.sin(kx * x)
    rho0, p0 = 1.0, 1 / GAM
    T = p0 / (RGAS * rho0)
    s = _stats(orc, _state(np.full_like(u, rho0), u, v, w, np.full_like(u, p0)), T0=T)
    Dy, Dz, Dx = A * math.sin(ky * DX) / DX, B * math.sin(kz * DX) / DX, C * math.sin(kx * DX) / DX
    mu = 2e-3
    G = (Dx * Dx + Dy * Dy + Dz * Dz) / 2     # mean of the three squared shear derivatives
    assert s["dil2"] == 0.0 and s["diss_dil"] == 0.0
    assert s["dudx2"] == 0.0 and s["dvdy2"] == 0.0 and s["dwdz2"] == 0.0
    # omega = (dw/dy - dv/dz, du/dz - dw/dx, dv/dx - du/dy) = (-dv/dz, -dw/dx, -du/dy)
    assert s["enstrophy"] == pytest.approx(G, rel=1e-12)
    assert s["diss_sol"] == pytest.approx(mu * G, rel=1e-12)
    assert s["diss_ref"] == pytest.approx(mu * G, rel=1e-12)
    # S_ij S_ij = 2 * sum_{i<j} (1/2 (d_i u_j + d_j u_i))^2 = G / 2 here
    assert s["sijsij"] == pytest.approx(G / 2, rel=1e-12)
    q2 = (A * A + B * B + C * C) / 2
    assert s["q2"] == pytest.approx(q2, rel=1e-13)
    assert s["mt"] == pytest.approx(math.sqrt(q2) / math.sqrt(GAM * p0 / rho0), rel=1e-13)


def test_density_wave_means_and_rms(orc):
    """rho = rho0 (1 + e cos(kappa x)) at uniform p, at rest: rho_rms = rho0 e/sqrt(2),
    <1/rho> = 1/(rho0 sqrt(1 - e^2)), T = p/(R rho)."""
    x, _, _ = _coords()
    rho0, e, p0 = 1.3, 0.2, 0.8
    rho = rho0 * (1 + e * np.cos(_kappa(1, NX) * x))
    zero = np.zeros_like(rho)
    s = _stats(orc, _state(rho, zero, zero, zero, np.full_like(rho, p0)), T0=1.0)
    vmean = 1 / (rho0 * math.sqrt(1 - e * e))
    assert s["rho_mean"] == pytest.approx(rho0, rel=1e-14)
    assert s["rho_rms"] == pytest.approx(rho0 * e / math.sqrt(2), rel=1e-12)
    assert s["v_mean"] == pytest.approx(vmean, rel=1e-12)
    # <1/rho^2> = 1/(rho0^2 (1 - e^2)^(3/2)), so rms(1/rho)^2 = <1/rho^2> - <1/rho>^2
    v2 = 1 / (rho0 ** 2 * (1 - e * e) ** 1.5)
    assert s["v_rms"] == pytest.approx(math.sqrt(v2 - vmean ** 2), rel=1e-10)
    assert s["T_mean"] == pytest.approx(p0 / RGAS * vmean, rel=1e-12)
    assert s["T_rms"] == pytest.approx(p0 / RGAS * math.sqrt(v2 - vmean ** 2), rel=1e-10)
    assert s["p_mean"] == pytest.approx(p0, rel=1e-13)
    _zero(s["p_rms"], p0)
    assert s["q2"] == 0.0 and s["tke"] == 0.0 and s["mt"] == 0.0
    assert s["c_mean"] == pytest.approx(np.mean(np.sqrt(GAM * p0 / rho)), rel=1e-13)


def test_taylor_green_cells_in_three_planes(orc):
    """Three Taylor-Green cells, one per coordinate plane, with one wavenumber on
    every axis: the discrete dilatation cancels exactly, each vorticity component
    is -2 a s cos cos of its plane (so a sign slip in any of them changes the
    enstrophy), and sum-by-parts gives <S_ij S_ij> = <omega^2>/2 and
    <d_j u_i d_j u_i> = <omega^2> for a field without dilatation."""
    x, y, z = _coords()
    kap = _kappa(4, NX)                        # 4, 3 and 5 whole periods: 4 points per period
    assert kap == pytest.approx(_kappa(3, NY)) and kap == pytest.approx(_kappa(5, NZ))
    A, B, C = 0.2, 0.15, -0.1
    cx, sx, cy, sy, cz, sz = (np.cos(kap * x), np.sin(kap * x), np.cos(kap * y), np.sin(kap * y),
                              np.cos(kap * z), np.sin(kap * z))
    u = A * cx * sy - C * sz * cx
    v = -A * sx * cy + B * cy * sz
    w = -B * sy * cz + C * cz * sx
    rho0, p0 = 1.0, 1 / GAM
    T = p0 / (RGAS * rho0)
    s = _stats(orc, _state(np.full_like(u, rho0), u, v, w, np.full_like(u, p0)), T0=T)
    D = math.sin(kap * DX) / DX
    ens = (A * A + B * B + C * C) * D * D     # <(2 a D cos cos)^2> = a^2 D^2
    mu = 2e-3
    assert s["dil2"] <= 1e-28 and abs(s["diss_dil"]) <= 1e-30
    assert s["enstrophy"] == pytest.approx(ens, rel=1e-12)
    assert s["diss_sol"] == pytest.approx(mu * ens, rel=1e-12)
    assert s["sijsij"] == pytest.approx(ens / 2, rel=1e-12)
    assert s["diss_ref"] == pytest.approx(mu * ens, rel=1e-12)


@pytest.mark.parametrize("phi", [0.0, 0.7, 2.5])
def test_two_mode_derivative_skewness(orc, phi):
    """u = A sin(kx) + B sin(2kx + phi): the central difference is
    D1 cos(kx) + D2 cos(2kx + phi), whose grid moments are <d^2> = (D1^2 + D2^2)/2
    and <d^3> = 3/4 D1^2 D2 cos(phi) (16 points per period, no aliasing)."""
    x, _, _ = _coords()
    kap = _kappa(1, NX)
    A, B = 0.3, 0.12
    u = A * np.sin(kap * x) + B * np.sin(2 * kap * x + phi)
    zero = np.zeros_like(u)
    rho0, p0 = 1.0, 1 / GAM
    s = _stats(orc, _state(np.full_like(u, rho0), u, zero, zero, np.full_like(u, p0)), T0=p0 / (RGAS * rho0))
    D1 = A * math.sin(kap * DX) / DX
    D2 = B * math.sin(2 * kap * DX) / DX
    m2 = (D1 * D1 + D2 * D2) / 2
    m3 = 0.75 * D1 * D1 * D2 * math.cos(phi)
    assert s["dudx2"] == pytest.approx(m2, rel=1e-12)
    assert s["dudx3"] == pytest.approx(m3, rel=1e-11, abs=1e-13 * m2 ** 1.5)
    assert s["skew_u"] == pytest.approx(m3 / m2 ** 1.5, rel=1e-11, abs=1e-13)
