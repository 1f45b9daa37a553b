"""The GKM tree's smooth-flow scheme (flux_scheme = "gks") on the device
against the oracle (sweep_gks / orc_gks_flux, pinned in tests/test_gks.py).

Bars as for the weno scheme: FP64 relative L2 <= 1e-12 on whole states after
1 and 10 steps, points within 1e-12 (FP64) and 1e-5 (FP32); mixed vs FP64
E_k and Mt within 1e-3; batch and graph bitwise.
"""
import math
import os

import numpy as np
import pytest

from _util import cons, rel_l2, smooth_box

pytestmark = pytest.mark.gpu

NTHREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def gk():
    import gkm_amd
    gkm_amd.load()
    return gkm_amd


def _gas(gk, **kw):
    base = dict(gamma=1.4, R=1.0, Pr=1.0, mu0=4e-3, T0=1.0 / 1.4, visc_law="power", recon="symoo",
                energy_recon="T", zbug_compat=True, precision="fp64", flux_scheme="gks")
    base.update(kw)
    return gk.GasParams(**base)


def _orc_params(orc, gas):
    import gkm_amd._abi as A
    return orc.make_params(gamma=gas.gamma, R=gas.R, Pr=gas.Pr, mu0=gas.mu0, T0=gas.T0, visc_exp=gas.visc_exp,
                           K=gas.dof(), visc_law=A.VISC[gas.visc_law], recon=A.RECON[gas.recon],
                           energy_recon=A.ENERGY[gas.energy_recon], zbug_compat=int(gas.zbug_compat), precision=0,
                           flux_scheme=A.SCHEME[gas.flux_scheme])


def _records(n, seed):
    rng = np.random.default_rng(seed)
    rec = []
    for _ in range(n):
        W0 = cons(rng.uniform(0.5, 2.0), *rng.normal(0, 0.7, 3), rng.uniform(0.4, 2.5))
        D = [rng.normal(size=5) * 0.3 for _ in range(3)]
        rec.append(np.concatenate([W0, *D, [10 ** rng.uniform(-4, -1.5), 10 ** rng.uniform(-4, -2)]]))
    return np.array(rec)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("mixed", 1e-5)])
def test_gks_points(gk, orc, prec, tol):
    rec = _records(400, 21)
    gas = _gas(gk, precision=prec, Pr=0.7)
    dev = gk.gks_points(gas, rec)
    ref = orc.gks_flux_points(_orc_params(orc, gas), rec)
    err = np.abs(dev - ref) / np.abs(ref).max(axis=1, keepdims=True)
    print("gks points", prec, "max rel err", err.max())
    assert err.max() < tol


CASES = [
    ((12, 12, 12), dict()),
    ((10, 14, 9), dict(zbug_compat=False, visc_law="sutherland", Pr=0.7)),
]


@pytest.mark.parametrize("shape,opts", CASES)
@pytest.mark.parametrize("nsteps", [1, 10])
def test_gks_step_parity_fp64(gk, orc, shape, opts, nsteps):
    ny, nx, nz = shape
    print("gks step", shape, opts, nsteps, flush=True)
    gas = _gas(gk, **opts)
    dx = 2 * math.pi / nx
    grid = orc.make_grid(nx, ny, nz, dx)
    W0 = smooth_box(nx, ny, nz, seed=31, discontinuous=True)
    orc.fill_ghosts(grid, W0)
    dt = 0.2 * dx / 3.0
    with gk.Solver(nx, ny, nz, gas=gas) as s:
        s.set_state(W0)
        s.evolve_eulerian(nsteps, dt=dt)
        Wd = s.get_state()
    Wo = [w.copy() for w in W0]
    orc.step(grid, _orc_params(orc, gas), Wo, dt, nsteps, NTHREADS)
    err = rel_l2(Wd, Wo)
    print("rel L2", err)
    assert err <= 1e-12


def test_gks_uniform_state_fixed_and_conservation(gk):
    n = 10
    gas = _gas(gk, precision="mixed")
    U = [np.full((n + 6) ** 3, v) for v in cons(1.1, 0.3, -0.2, 0.4, 0.8)]
    with gk.Solver(n, gas=gas) as s:
        s.set_state(U)
        U0 = s.get_state()  # the hi/lo pair holds ~48 bits of each value
        s.evolve_eulerian(3, dt=1e-3)
        V = s.get_state()
    assert all(np.array_equal(a, b) for a, b in zip(U0, V))
    with gk.Solver.from_dhit(16, Mt=0.4, Re=40, seed=3, flux_scheme="gks", Pr=1.0) as s:
        a = s.stats()
        s.evolve_eulerian(5, dt=s.timestep(0.4))
        b = s.stats()
    for key in ("mass", "mom_x", "mom_y", "mom_z", "energy"):
        assert abs(b[key] - a[key]) <= 1e-12 * (abs(a["mass"]) + abs(a["energy"])), key


def test_gks_mixed_close_to_fp64(gk):
    kw = dict(k0=4, Mt=0.5, Re=70, seed=22, visc_law="power", flux_scheme="gks", Pr=1.0)
    a = gk.Solver.from_dhit(32, precision="fp64", **kw)
    b = gk.Solver.from_dhit(32, precision="mixed", **kw)
    dt = a.timestep(0.3)
    for _ in range(4):
        a.evolve_eulerian(10, dt=dt)
        b.evolve_eulerian(10, dt=dt)
        sa, sb = a.stats(), b.stats()
        for key in ("ek", "mt"):
            assert abs(sb[key] - sa[key]) <= 1e-3 * abs(sa[key]), (key, sa[key], sb[key])
    a.close()
    b.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_gks_batch_and_graph_match_single(gk, prec):
    ics = [gk.dhit(12, Mt=0.3 + 0.1 * i, Re=40, seed=60 + i) for i in range(3)]
    kw = dict(precision=prec, visc_law="power", flux_scheme="gks", Pr=1.0)
    b = gk.Batch.from_ics(ics, **kw)
    b.set_graph(True)
    dts = b.timestep(0.4)
    b.step(2)
    b.stats(0)
    b.step(2)
    for i, ic in enumerate(ics):
        with gk.Solver.from_ic(ic, **kw) as s:
            s.evolve_eulerian(4, dt=dts[i])
            for x, y in zip(s.get_state(), b.get_state(i)):
                assert np.array_equal(x, y), f"box {i}"
    b.close()
