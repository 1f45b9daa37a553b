"""Parity of the HIP path (libgkm.so on the MI355X) against the CPU oracle.

FP64 bar: relative L2 <= 1e-12 on whole states after 1 and 10 steps (the
device reorders sums, contracts to FMA and uses the ROCm device libm, so bits
differ at the last place); flux and WENO points within 1e-12 relative.
Mixed bar: statistics within 1e-3 relative of the FP64 run.
"""
import math
import os

import numpy as np
import pytest

from _util import flux_records, interior, rel_l2, rel_l2_per_field, smooth_box

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
NTHREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def gk():
    import gkm_amd
    gkm_amd.load()
    return gkm_amd


def _gas(gk, **kw):
    base = dict(gamma=1.4, R=287.0, Pr=0.7, mu0=2e-3, T0=1.0 / (1.4 * 287.0), visc_law="sutherland",
                recon="symoo", energy_recon="T", zbug_compat=True, precision="fp64")
    base.update(kw)
    return gk.GasParams(**base)


def _orc_params(orc, gas):
    import gkm_amd._abi as A
    return orc.make_params(gamma=gas.gamma, R=gas.R, Pr=gas.Pr, mu0=gas.mu0, T0=gas.T0, visc_exp=gas.visc_exp,
                           K=gas.dof(), visc_law=A.VISC[gas.visc_law], recon=A.RECON[gas.recon],
                           energy_recon=A.ENERGY[gas.energy_recon], zbug_compat=int(gas.zbug_compat),
                           precision=0)


def test_flux_points_fp64(gk, orc):
    rec = flux_records(400, seed=11)
    gas = _gas(gk)
    dev = gk.flux_points(gas, rec, 0.05)
    ref = orc.flux_points(_orc_params(orc, gas), rec, 0.05)
    scale = np.abs(ref).max(axis=1, keepdims=True)
    err = np.abs(dev - ref) / scale
    print("flux fp64 max rel err", err.max())
    assert err.max() < 1e-12


@pytest.mark.parametrize("dt_over_tau", [(0.5, 100.0), (0.01, 0.5), (1e-3, 0.01)])
def test_flux_points_fp32(gk, orc, dt_over_tau):
    """FP32 flux against the FP64 oracle over the collision regimes: the
    continuum one the 256^3 Mt = 0.5 benchmark runs in (dt/tau ~ 4 at CFL 0.5),
    the reference's CFL 0.02 (dt/tau ~ 0.16 at 256^3) and below.  As dt/tau -> 0
    the reference's time-integral coefficients (GKMflux.cpp:155-157, 186-190)
    cancel; the FP32 form takes them from the exponential-integrator functions
    there (gkm_device.h, time_coef), so the bar holds down to dt/tau = 1e-3."""
    rec = flux_records(400, seed=12, dt_over_tau=dt_over_tau)
    gas = _gas(gk, precision="mixed")
    dev = gk.flux_points(gas, rec, 0.05)
    ref = orc.flux_points(_orc_params(orc, gas), rec, 0.05)
    scale = np.abs(ref).max(axis=1, keepdims=True)
    err = np.abs(dev - ref) / scale
    print("flux fp32 dt/tau", dt_over_tau, "max rel err", err.max(), "median", np.median(err))
    assert np.median(err) < 1e-6
    assert err.max() < 1e-4


@pytest.mark.parametrize("recon", ["symoo", "js"])
@pytest.mark.parametrize("energy", ["T", "E"])
def test_weno_points(gk, orc, recon, energy):
    rng = np.random.default_rng(5)
    n = 300
    cells = np.empty((n, 6, 5))
    for r in range(n):
        rho = rng.uniform(0.5, 2, 6)
        u, v, w = rng.normal(0, 0.5, (3, 6))
        p = rng.uniform(0.5, 3, 6)
        if r % 3 == 0:  # smooth stencil
            x = np.arange(6)
            rho = 1 + 0.1 * np.sin(0.3 * x + r)
            p = 1 + 0.1 * np.cos(0.3 * x + r)
        cells[r] = np.stack([rho, rho * u, rho * v, rho * w, p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)], axis=1)
    gas = _gas(gk, recon=recon, energy_recon=energy)
    dev = gk.weno_points(gas, cells)
    P = _orc_params(orc, gas)
    ref = np.array([np.concatenate(orc.weno(P, cells[r])) for r in range(n)])
    err = np.abs(dev - ref) / (np.abs(ref) + 1e-300)
    print("weno max rel err", err.max())
    assert err.max() < 1e-12


def _run_pair(gk, orc, shape, gas, W0, dt, nsteps):
    ny, nx, nz = shape
    dx = 2 * math.pi / nx
    s = gk.Solver(nx, ny, nz, dx=dx, gas=gas)
    s.set_state(W0)
    s.evolve_eulerian(nsteps, dt=dt)
    Wd = s.get_state()
    s.close()
    Wo = [w.copy() for w in W0]
    orc.step(orc.make_grid(nx, ny, nz, dx), _orc_params(orc, gas), Wo, dt, nsteps, NTHREADS)
    return Wd, Wo


CASES = [
    # shape (ny, nx, nz), options, discontinuous
    ((16, 16, 16), dict(), False),
    ((12, 8, 16), dict(recon="js", energy_recon="E", visc_law="power", zbug_compat=False), True),
    ((10, 14, 12), dict(Pr=1.0), True),
    # the smallest boxes the handle accepts: every ghost is an image of the
    # interior three cells away, and a z line holds 4 interfaces, so the fused
    # Z kernel's wavefronts span 15 lines
    ((3, 3, 3), dict(recon="js", visc_law="power"), True),
    ((5, 3, 4), dict(), True),
]


@pytest.mark.parametrize("shape,opts,disc", CASES)
@pytest.mark.parametrize("nsteps", [1, 10])
def test_step_parity_fp64(gk, orc, shape, opts, disc, nsteps):
    ny, nx, nz = shape
    gas = _gas(gk, **opts)
    W0 = smooth_box(nx, ny, nz, seed=3, discontinuous=disc)
    orc.fill_ghosts(orc.make_grid(nx, ny, nz, 2 * math.pi / nx), W0)
    dt = 0.2 * (2 * math.pi / nx) / 3.0
    Wd, Wo = _run_pair(gk, orc, shape, gas, W0, dt, nsteps)
    err = rel_l2(Wd, Wo)
    print(shape, opts, nsteps, "rel L2", err, rel_l2_per_field(Wd, Wo))
    assert err <= FP64_TOL


def test_step_parity_dhit32(gk, orc):
    """32^3 decaying-turbulence initial condition (Mt = 0.3, k0 = 4), 10 steps."""
    s = gk.Solver.from_dhit(32, k0=4, Mt=0.3, Re=30, seed=7, visc_law="power")
    dt = s.timestep(0.5)
    W0 = s.get_state()
    s.evolve_eulerian(10)
    Wd = s.get_state()
    gas = s.gas
    s.close()
    Wo = [w.copy() for w in W0]
    grid = orc.make_grid(32, 32, 32, 2 * math.pi / 32)
    assert abs(orc.timestep(grid, _orc_params(orc, gas), Wo, 0.5) - dt) <= 1e-14 * dt
    orc.step(grid, _orc_params(orc, gas), Wo, dt, 10, NTHREADS)
    err = rel_l2(Wd, Wo)
    print("dhit32 rel L2", err)
    assert err <= FP64_TOL


def test_step_parity_dhit128(gk, orc):
    """The 128^3 drop-in case: decaying-turbulence initial condition (k0 = 4,
    Mt = 0.3, Re = 50, gamma = 1.4, Pr = 0.7, power-law viscosity, WENO-JS),
    10 FP64 steps at CFL 0.5 against the oracle, relative L2 <= 1e-12."""
    s = gk.Solver.from_dhit(128, k0=4, Mt=0.3, Re=50, seed=1234, visc_law="power", recon="js", Pr=0.7)
    dt = s.timestep(0.5)
    W0 = s.get_state()
    s.evolve_eulerian(10)
    Wd = s.get_state()
    gas = s.gas
    s.close()
    print("dhit128: device done, oracle on", NTHREADS, "threads", flush=True)
    Wo = [w.copy() for w in W0]
    grid = orc.make_grid(128, 128, 128, 2 * math.pi / 128)
    orc.step(grid, _orc_params(orc, gas), Wo, dt, 10, NTHREADS)
    err = rel_l2(Wd, Wo)
    print("dhit128 rel L2", err)
    assert err <= FP64_TOL


def test_cpu_case_32_statistics_every_step(gk, orc):
    """The reference's CPU-runnable case: 32^3 (k0 = 4, Mt = 0.3, Re = 30,
    gamma = 1.4, Pr = 0.7, power-law viscosity, WENO-JS, CFL 0.5), 200 FP64
    steps with E_k, Mt and the dissipation taken every step, on the device and
    on the oracle side by side.  Differences of 1e-16 grow through the
    nonlinear dynamics; the bar is 1e-10 relative at every step."""
    s = gk.Solver.from_dhit(32, k0=4, Mt=0.3, Re=30, seed=1234, visc_law="power", recon="js", Pr=0.7)
    dt = s.timestep(0.5)
    gas = s.gas
    P = _orc_params(orc, gas)
    grid = orc.make_grid(32, 32, 32, 2 * math.pi / 32)
    Wo = s.get_state()
    worst = {"ek": 0.0, "mt": 0.0, "diss_ref": 0.0}
    for step in range(201):
        d = s.stats()
        o = orc.stats(grid, P, Wo)
        for k in worst:
            worst[k] = max(worst[k], abs(d[k] - o[k]) / abs(o[k]))
        if step < 200:
            s.evolve_eulerian(1, dt=dt)
            orc.step(grid, P, Wo, dt, 1, NTHREADS)
    s.close()
    print("32^3 x 200 steps, worst rel diff", worst)
    for k, v in worst.items():
        assert v <= 1e-10, (k, v)


def test_ghost_fill_bitwise(gk, orc):
    nx, ny, nz = 9, 7, 11
    W = [np.random.default_rng(q).random((ny + 6) * (nx + 6) * (nz + 6)) for q in range(5)]
    s = gk.Solver(nx, ny, nz, gas=_gas(gk))
    s.set_state(W)
    s.periodic_bc()
    Wd = s.get_state()
    s.close()
    Wo = [w.copy() for w in W]
    orc.fill_ghosts(orc.make_grid(nx, ny, nz, 1.0), Wo)
    for a, b in zip(Wd, Wo):
        assert np.array_equal(a, b)


def test_state_roundtrip(gk):
    nx, ny, nz = 8, 9, 10
    W = [np.random.default_rng(q).normal(size=(ny + 6) * (nx + 6) * (nz + 6)) for q in range(5)]
    for prec in ("fp64", "mixed"):
        s = gk.Solver(nx, ny, nz, gas=_gas(gk, precision=prec))
        s.set_state(W)
        back = s.get_state()
        s.close()
        for a, b in zip(W, back):
            if prec == "fp64":
                assert np.array_equal(a, b)
            else:  # hi/lo float pair keeps ~48 bits
                assert np.max(np.abs(a - b) / (np.abs(a) + 1e-30)) < 1e-13


def test_uniform_state_fixed(gk):
    nx = ny = nz = 8
    W = smooth_box(nx, ny, nz, amp=0.0, mach=0.0)
    W = [w.reshape(ny + 6, nx + 6, nz + 6) for w in W]
    for q, val in enumerate([1.3, 0.2, -0.1, 0.05, 3.0]):
        W[q][...] = val
    W = [np.ascontiguousarray(w.reshape(-1)) for w in W]
    for prec in ("fp64", "mixed"):
        s = gk.Solver(nx, ny, nz, gas=_gas(gk, precision=prec))
        s.set_state(W)
        s.evolve_eulerian(3, dt=1e-3)
        out = s.get_state()
        s.close()
        for a, b in zip(out, W):
            if prec == "fp64":
                assert np.array_equal(a, b)
            else:
                assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_conservation(gk, prec):
    nx, ny, nz = 16, 12, 20
    W = smooth_box(nx, ny, nz, seed=9, discontinuous=True)
    s = gk.Solver(nx, ny, nz, gas=_gas(gk, precision=prec))
    s.set_state(W)
    s.periodic_bc()
    a = s.stats()
    s.evolve_eulerian(20, dt=0.02)
    b = s.stats()
    s.close()
    for key in ("mass", "mom_x", "mom_y", "mom_z", "energy"):
        scale = abs(a["mass"]) + abs(a["energy"])
        tol = 1e-13 if prec == "fp64" else 1e-9
        assert abs(b[key] - a[key]) <= tol * scale, (key, a[key], b[key])


def test_stats_and_dt(gk, orc):
    nx, ny, nz = 12, 10, 14
    gas = _gas(gk, visc_law="power")
    W = smooth_box(nx, ny, nz, seed=4, discontinuous=True)
    grid = orc.make_grid(nx, ny, nz, 2 * math.pi / nx)
    orc.fill_ghosts(grid, W)
    s = gk.Solver(nx, ny, nz, gas=gas)
    s.set_state(W)
    d = s.stats()
    dt = s.timestep(0.5)
    s.close()
    o = orc.stats(grid, _orc_params(orc, gas), W)
    for k, v in o.items():
        assert abs(d[k] - v) <= 1e-11 * max(abs(v), 1e-300) + 1e-14, (k, d[k], v)
    assert abs(dt - orc.timestep(grid, _orc_params(orc, gas), W, 0.5)) <= 1e-15 * dt


def test_mixed_statistics_close_to_fp64(gk):
    """Mixed form vs FP64 on a 32^3 Mt = 0.5 box over 40 steps: E_k and Mt within 1e-3."""
    kw = dict(k0=4, Mt=0.5, Re=70, seed=21, visc_law="power", recon="js")
    a = gk.Solver.from_dhit(32, precision="fp64", **kw)
    b = gk.Solver.from_dhit(32, precision="mixed", **kw)
    dt = a.timestep(0.5)
    for _ in range(4):
        a.evolve_eulerian(10, dt=dt)
        b.evolve_eulerian(10, dt=dt)
        sa, sb = a.stats(), b.stats()
        for key in ("ek", "mt"):
            assert abs(sb[key] - sa[key]) <= 1e-3 * abs(sa[key]), (key, sa[key], sb[key])
    a.close()
    b.close()


@pytest.mark.parametrize("dt_over_tau", [0.16, 0.02])
def test_mixed_close_to_fp64_at_small_dt_over_tau(gk, dt_over_tau):
    """Mixed form vs FP64 where the collision time is long against the step:
    dt/tau = 0.16 is the reference's CFL 0.02 on the 256^3 Mt = 0.5 box; here a
    64^3 box of the same recipe with dt set from tau0 = mu0/p0.  E_k and Mt
    within 1e-3 relative over 200 steps, sampled every 40."""
    ic = gk.dhit(64, k0=4, Mt=0.5, Re=70, seed=78)
    dt = dt_over_tau * ic.mu0 / ic.p0
    kw = dict(visc_law="power", recon="js")
    print("mixed vs fp64 at dt/tau", dt_over_tau, "dt", dt, flush=True)
    a = gk.Solver.from_ic(ic, precision="fp64", **kw)
    b = gk.Solver.from_ic(ic, precision="mixed", **kw)
    worst = {"ek": 0.0, "mt": 0.0}
    for _ in range(5):
        a.evolve_eulerian(40, dt=dt)
        b.evolve_eulerian(40, dt=dt)
        sa, sb = a.stats(), b.stats()
        for key in worst:
            worst[key] = max(worst[key], abs(sb[key] - sa[key]) / abs(sa[key]))
    print("max rel err", worst)
    a.close()
    b.close()
    assert worst["ek"] <= 1e-3 and worst["mt"] <= 1e-3


def test_mixed_close_to_fp64_over_four_eddy_times(gk):
    """Mixed form vs FP64 from one 64^3 Mt = 0.5 initial condition to t/t0 = 4
    (the decay the mixed 256^3 workload is compared over), sampled every 20
    steps: E_k and Mt within 1e-3 relative at every sample; the dilatational
    dissipation and the velocity-derivative skewness within 1e-2 (a sanity bar:
    no tolerance is set for them, their errors are printed)."""
    ic = gk.dhit(64, k0=4, Mt=0.5, Re=70, seed=77)
    r = gk.mixed_vs_fp64(ic, t_end=4.0, cfl=0.5, every=20, visc_law="power", recon="js")
    print("steps", r["steps"], "max rel err", r["max_rel_err"], "q2 decay", r["q2_decay"])
    assert r["samples"][-1]["t_over_t0"] >= 4.0
    assert r["max_rel_err"]["ek"] <= 1e-3
    assert r["max_rel_err"]["mt"] <= 1e-3
    assert r["max_rel_err"]["diss_dil"] <= 1e-2
    assert r["max_rel_err"]["skew"] <= 1e-2
    # the FP64 run decays monotonically
    q = [s["fp64"]["q2"] for s in r["samples"]]
    assert all(b < a for a, b in zip(q, q[1:]))
