"""Shared helpers of the test suite: seeded states and interface records."""
from __future__ import annotations

import math

import numpy as np


def cons(rho, u, v, w, p, gam=1.4):
    return np.array([rho, rho * u, rho * v, rho * w, p / (gam - 1) + 0.5 * rho * (u * u + v * v + w * w)])


def flux_records(n: int, seed: int = 0, jumps: bool = True, mach: float = 0.6,
                 dt_over_tau: tuple | None = None) -> np.ndarray:
    """Random interface records (n, 52): WL WR Wl Wr DWxl DWxr DWyl DWyr DWzl DWzr tau dt.
    Smooth cases have Wl ~ WL ~ WR; jump cases have independent left/right states.
    dt_over_tau = (lo, hi) draws log-uniform dt/tau in that range (default: tau and dt
    independent, log-uniform in [1e-4, 1e-1] and [1e-4, 1e-2])."""
    rng = np.random.default_rng(seed)
    recs = []
    for r in range(n):
        def state():
            return cons(rng.uniform(0.5, 2.0), *rng.normal(0, mach, 3), rng.uniform(0.5, 3.0))
        WL = state()
        WR = state() if (jumps and r % 2) else WL * (1 + 0.02 * rng.normal(size=5))
        Wl = WL + 0.03 * rng.normal(size=5) * np.abs(WL)
        Wr = WR + 0.03 * rng.normal(size=5) * np.abs(WR)
        Wl[4] = max(Wl[4], 0.5 * (Wl[1] ** 2 + Wl[2] ** 2 + Wl[3] ** 2) / Wl[0] + 0.3)
        Wr[4] = max(Wr[4], 0.5 * (Wr[1] ** 2 + Wr[2] ** 2 + Wr[3] ** 2) / Wr[0] + 0.3)
        DW = [rng.normal(size=5) * 0.3 for _ in range(6)]
        tau = 10 ** rng.uniform(-4, -1)
        dt = 10 ** rng.uniform(-4, -2)
        if dt_over_tau is not None:
            dt = tau * 10 ** rng.uniform(np.log10(dt_over_tau[0]), np.log10(dt_over_tau[1]))
        recs.append(np.concatenate([WL, WR, Wl, Wr] + DW + [[tau, dt]]))
    return np.array(recs)


def smooth_box(nx, ny, nz, seed=0, amp=0.05, mach=0.3, gam=1.4, discontinuous=False):
    """Padded conservative planes of a periodic smooth (or jumpy) state; ghosts filled
    by the caller.  Interior (ny, nx, nz)."""
    rng = np.random.default_rng(seed)
    P = (ny + 6, nx + 6, nz + 6)
    j = np.arange(ny)[:, None, None]
    i = np.arange(nx)[None, :, None]
    k = np.arange(nz)[None, None, :]
    th = [2 * math.pi * i / nx, 2 * math.pi * j / ny, 2 * math.pi * k / nz]

    def field(a):
        f = np.zeros((ny, nx, nz))
        for _ in range(3):
            ph = rng.uniform(0, 2 * math.pi, 3)
            m = rng.integers(1, 3, 3)
            f += a * np.sin(m[0] * th[0] + ph[0]) * np.cos(m[1] * th[1] + ph[1]) * np.sin(m[2] * th[2] + ph[2])
        return f

    rho = 1.0 + field(amp)
    u, v, w = field(mach), field(mach), field(mach)
    p = 1.0 / gam + field(amp / gam)
    if discontinuous:
        rho = rho + 0.3 * (((i + j + k) % 7) < 3)
        p = p + 0.2 * (((2 * i + k) % 5) < 2)
    W = [np.zeros(P) for _ in range(5)]
    s = (slice(3, 3 + ny), slice(3, 3 + nx), slice(3, 3 + nz))
    W[0][s] = rho
    W[1][s] = rho * u
    W[2][s] = rho * v
    W[3][s] = rho * w
    W[4][s] = p / (gam - 1) + 0.5 * rho * (u * u + v * v + w * w)
    return [np.ascontiguousarray(x.reshape(-1)) for x in W]


def interior(W, shape):
    ny, nx, nz = shape
    return [w.reshape(ny + 6, nx + 6, nz + 6)[3:-3, 3:-3, 3:-3] for w in W]


def rel_l2(a, b):
    a = np.concatenate([np.ravel(x) for x in a])
    b = np.concatenate([np.ravel(x) for x in b])
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def rel_l2_per_field(a, b):
    return [float(np.linalg.norm(np.ravel(x) - np.ravel(y)) / np.linalg.norm(np.ravel(y))) for x, y in zip(a, b)]
