"""The GKM tree's smooth-flow scheme in the oracle (CPU checks).

GKM/CUDA GKM/src/GKMflux.cpp does not compile as shipped (its param.h
declares `typedef float double`), so the oracle's restatement (orc_gks_flux,
sweep_gks) is pinned by an independent evaluation of the same definition:

* the flux integral F = rho0 ∫ u psi [p0 + p1 (u ax.psi + v ay.psi + w az.psi)
  + (p1 + p2) A.psi] g0, with ax/ay/az/A the Maxwellian slope solves and A
  from the compatibility condition, evaluated by 3-D Gauss–Hermite
  quadrature in velocity plus the exact chi^2_K moments of the internal
  energy, the 5x5 slope systems solved with numpy (Pr = 1: the slope solve is
  the exact inverse of the moment matrix; 1/Pr on a[4] is pinned separately
  in test_oracle.py);
* closed forms: a uniform state with zero slopes gives dt x the Euler flux; a
  pure shear gives the Navier–Stokes stress -dt mu dv/dx;
* the sweep restated from the reference text's explicit gathers
  (CPUupdate.cpp:29-192) on single interfaces equals orc_fluxes bit for bit.
"""
import math

import numpy as np
import pytest
from numpy.polynomial.hermite_e import hermegauss

from _util import cons, smooth_box


def _params(orc, **kw):
    base = dict(gamma=1.4, R=1.0, Pr=1.0, mu0=1e-2, T0=1.0, visc_exp=0.76, K=2, visc_law=1, recon=0,
                energy_recon=0, zbug_compat=1, precision=0, flux_scheme=1)
    base.update(kw)
    return orc.make_params(**base)


def _quadrature_flux(W0, DWx, DWy, DWz, mu, dt, gam=1.4, K=2, nq=10):
    rho = W0[0]
    U = np.array(W0[1:4]) / rho
    P = (rho * (gam - 1)) * (W0[4] / rho - 0.5 * U @ U)
    lam = 0.5 * rho / P
    sig = math.sqrt(1.0 / (2 * lam))
    x, w = hermegauss(nq)
    w = w / w.sum()
    # velocity nodes (u, v, w) and weights of the normalised Maxwellian
    uu = U[0] + sig * x[:, None, None] + 0 * x[None, :, None] + 0 * x[None, None, :]
    vv = U[1] + sig * x[None, :, None] + 0 * uu
    ww = U[2] + sig * x[None, None, :] + 0 * uu
    wt = w[:, None, None] * w[None, :, None] * w[None, None, :]
    # internal energy S = xi^2: two-point rule with E[S] = K s^2, E[S^2] = (K^2 + 2K) s^4
    m1, m2 = K * sig ** 2, (K * K + 2 * K) * sig ** 4
    d = math.sqrt(m2 - m1 * m1)
    nodes = []
    for s_val in (m1 - d, m1 + d):
        psi = np.stack([np.ones_like(uu), uu, vv, ww, 0.5 * (uu * uu + vv * vv + ww * ww + s_val)])
        nodes.append(psi)

    def E(f):  # f(psi) -> array over the grid; expectation over velocity and S
        return 0.5 * sum(np.sum(f(psi) * wt, axis=(-3, -2, -1)) for psi in nodes)

    Mmat = np.array([[E(lambda p, i=i, j=j: p[i] * p[j]) for j in range(5)] for i in range(5)])
    solve = lambda b: np.linalg.solve(Mmat, b)  # noqa: E731
    ax, ay, az = solve(np.array(DWx) / rho), solve(np.array(DWy) / rho), solve(np.array(DWz) / rho)

    def mom(mult, a):  # <mult psi (a . psi)>
        return np.array([E(lambda p, q=q: mult(p) * p[q] * np.tensordot(a, p, axes=1)) for q in range(5)])

    B = -(mom(lambda p: p[1], ax) + mom(lambda p: p[2], ay) + mom(lambda p: p[3], az))
    A = solve(B)
    tau = mu / P
    p0, p1, p2 = dt, -tau * dt, dt * dt / 2
    Mu = np.array([E(lambda p, q=q: p[1] * p[q]) for q in range(5)])
    F = rho * ((p1 + p2) * mom(lambda p: p[1], A) + p1 * mom(lambda p: p[1] * p[1], ax)
               + p1 * mom(lambda p: p[1] * p[2], ay) + p1 * mom(lambda p: p[1] * p[3], az) + p0 * Mu)
    return F


def _rec(orc, W0, DWx, DWy, DWz, mu, dt):
    return np.concatenate([W0, DWx, DWy, DWz, [mu, dt]])[None, :]


@pytest.mark.parametrize("seed", range(6))
def test_gks_flux_matches_quadrature(orc, seed):
    rng = np.random.default_rng(seed)
    W0 = cons(rng.uniform(0.6, 1.6), *rng.normal(0, 0.6, 3), rng.uniform(0.5, 2.0))
    D = [rng.normal(size=5) * 0.4 for _ in range(3)]
    mu, dt = 10 ** rng.uniform(-3, -1), 10 ** rng.uniform(-3, -1.5)
    F = orc.gks_flux_points(_params(orc), _rec(orc, W0, *D, mu, dt))[0]
    Fq = _quadrature_flux(W0, *D, mu, dt)
    assert np.max(np.abs(F - Fq)) <= 1e-12 * np.max(np.abs(Fq)), (F, Fq)


def test_gks_uniform_state_is_dt_times_euler_flux(orc):
    gam = 1.4
    W0 = cons(1.3, 0.4, -0.2, 0.1, 0.9)
    rho, u, v, w = W0[0], W0[1] / W0[0], W0[2] / W0[0], W0[3] / W0[0]
    p = (gam - 1) * (W0[4] - 0.5 * rho * (u * u + v * v + w * w))
    euler = np.array([rho * u, rho * u * u + p, rho * u * v, rho * u * w, u * (W0[4] + p)])
    z = np.zeros(5)
    F = orc.gks_flux_points(_params(orc), _rec(orc, W0, z, z, z, 0.01, 1e-3))[0]
    assert np.max(np.abs(F - 1e-3 * euler)) <= 1e-15 * np.max(np.abs(euler))


def test_gks_shear_stress_is_navier_stokes(orc):
    """At rest with dv/dx = s (v = 0 at the interface, uniform rho and p), the
    y-momentum flux is -dt mu s and nothing else moves."""
    rho, p, s, mu, dt = 1.0, 1.0 / 1.4, 0.3, 0.02, 1e-3
    W0 = cons(rho, 0.0, 0.0, 0.0, p)
    DWx = np.array([0.0, 0.0, rho * s, 0.0, 0.0])
    z = np.zeros(5)
    F = orc.gks_flux_points(_params(orc), _rec(orc, W0, DWx, z, z, mu, dt))[0]
    assert abs(F[2] - (-dt * mu * s)) <= 1e-14
    assert abs(F[1] - dt * p) <= 1e-14
    assert abs(F[0]) <= 1e-16 and abs(F[3]) <= 1e-16 and abs(F[4]) <= 1e-16


def _explicit_gks_fluxes(W, shape, gp, dt, dx):
    """derivsX/Y/Z of GKM/CUDA GKM/src/CPUupdate.cpp:29-192 with its own
    neighbour naming (N/S/F/B), the component swaps of the Y (1<->2) and Z
    (1<->3) sweeps, mu from the interface average (:62-65), on a few
    interfaces of each direction, mapped to the 3-ghost layout."""
    ny, nx, nz = shape
    Wp = [w.reshape(ny + 6, nx + 6, nz + 6) for w in W]
    out = {}

    def at(j, i, k):
        return np.array([w[j, i, k] for w in Wp])

    def mu_of(W0):
        d = W0[0]
        u = W0[1:4] / d
        P = (d * (gp["gamma"] - 1)) * (W0[4] / d - .5 * (u @ u))
        T = P / (gp["R"] * d)
        return gp["mu0"] * (T / gp["T0"]) ** gp["visc_exp"]

    for (j, i, k) in [(3, 2, 3), (4, 5, 6), (ny + 2, nx + 2, nz + 2)]:
        # X: L = (j,i,k), R = (j,i+1,k); N/S = j+-1; F/B = k+-1
        L, R = at(j, i, k), at(j, i + 1, k)
        W0 = .5 * (R + L)
        W0N, W0S = .5 * (at(j + 1, i + 1, k) + at(j + 1, i, k)), .5 * (at(j - 1, i + 1, k) + at(j - 1, i, k))
        W0F, W0B = .5 * (at(j, i + 1, k + 1) + at(j, i, k + 1)), .5 * (at(j, i + 1, k - 1) + at(j, i, k - 1))
        out[("x", j, i, k)] = (W0, (R - L) / dx, 0.5 * (W0N - W0S) / dx, 0.5 * (W0F - W0B) / dx, mu_of(W0), None)
    for (j, i, k) in [(2, 3, 3), (5, 4, 6), (ny + 2, nx + 2, nz + 2)]:
        # Y: L = (j,i,k), R = (j+1,i,k); N = i-1, S = i+1, DWy = (W0S - W0N)/(2dx); F/B = k+-1
        L, R = at(j, i, k), at(j + 1, i, k)
        W0 = .5 * (R + L)
        W0N, W0S = .5 * (at(j + 1, i - 1, k) + at(j, i - 1, k)), .5 * (at(j + 1, i + 1, k) + at(j, i + 1, k))
        W0F, W0B = .5 * (at(j + 1, i, k + 1) + at(j, i, k + 1)), .5 * (at(j + 1, i, k - 1) + at(j, i, k - 1))
        out[("y", j, i, k)] = (W0, (R - L) / dx, 0.5 * (W0S - W0N) / dx, 0.5 * (W0F - W0B) / dx, mu_of(W0), (1, 2))
    for (j, i, k) in [(3, 3, 2), (6, 4, 5), (ny + 2, nx + 2, nz + 2)]:
        # Z: L = (j,i,k), R = (j,i,k+1); N/S = i+-1; F/B = j+-1
        L, R = at(j, i, k), at(j, i, k + 1)
        W0 = .5 * (R + L)
        W0N, W0S = .5 * (at(j, i + 1, k + 1) + at(j, i + 1, k)), .5 * (at(j, i - 1, k + 1) + at(j, i - 1, k))
        W0F, W0B = .5 * (at(j + 1, i, k + 1) + at(j + 1, i, k)), .5 * (at(j - 1, i, k + 1) + at(j - 1, i, k))
        out[("z", j, i, k)] = (W0, (R - L) / dx, 0.5 * (W0N - W0S) / dx, 0.5 * (W0F - W0B) / dx, mu_of(W0), (1, 3))
    return out


def test_gks_sweeps_match_explicit_reference_gathers(orc):
    nx, ny, nz = 9, 8, 10
    dx = 2 * math.pi / nx
    gp = dict(gamma=1.4, R=1.0, mu0=5e-3, T0=0.7, visc_exp=0.76)
    W = smooth_box(nx, ny, nz, seed=12, discontinuous=True)
    grid = orc.make_grid(nx, ny, nz, dx)
    orc.fill_ghosts(grid, W)
    P = _params(orc, mu0=gp["mu0"], T0=gp["T0"])
    dt = 2e-3
    F = orc.fluxes(grid, P, [w.copy() for w in W], dt)
    Fp = [f.reshape(ny + 6, nx + 6, nz + 6) for f in F]
    for (d, j, i, k), (W0, Dx, Dy, Dz, mu, swap) in _explicit_gks_fluxes(W, (ny, nx, nz), gp, dt, dx).items():
        W0, Dx, Dy, Dz = W0.copy(), Dx.copy(), Dy.copy(), Dz.copy()
        if swap:
            a, b = swap
            for v in (W0, Dx, Dy, Dz):
                v[a], v[b] = v[b], v[a]
        Fs = orc.gks_flux_points(P, _rec(orc, W0, Dx, Dy, Dz, mu, dt))[0]
        if swap:
            Fs[a], Fs[b] = Fs[b], Fs[a]
        base = {"x": 0, "y": 5, "z": 10}[d]
        got = np.array([Fp[base + q][j, i, k] for q in range(5)])
        assert np.array_equal(got, Fs), (d, j, i, k, got, Fs)


def test_gks_step_conserves_and_keeps_uniform_state(orc):
    nx, ny, nz = 8, 6, 7
    grid = orc.make_grid(nx, ny, nz, 2 * math.pi / nx)
    P = _params(orc, mu0=3e-3, T0=1.0 / 1.4)
    W = smooth_box(nx, ny, nz, seed=2)
    orc.fill_ghosts(grid, W)
    inner = (slice(3, -3),) * 3
    before = [w.reshape(ny + 6, nx + 6, nz + 6)[inner].sum() for w in W]
    orc.step(grid, P, W, 2e-3, 3, 1)
    after = [w.reshape(ny + 6, nx + 6, nz + 6)[inner].sum() for w in W]
    for a, b in zip(before, after):
        assert abs(a - b) <= 1e-12 * max(abs(a), 1.0)
    U = [np.zeros((ny + 6) * (nx + 6) * (nz + 6)) for _ in range(5)]
    for q, val in enumerate(cons(1.1, 0.3, -0.2, 0.4, 0.8)):
        U[q][:] = val
    U0 = [u.copy() for u in U]
    orc.step(grid, P, U, 2e-3, 2, 1)
    assert all(np.array_equal(a, b) for a, b in zip(U, U0))
