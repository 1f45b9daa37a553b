"""PrintStats on the device (gkm_print_stats: natural-spline gradients over
the padded box + FP64 reductions) against the CPU oracle's orc_print_stats,
which tests/test_print_stats.py pins against scipy's natural cubic spline.

Bar: 1e-11 relative per quantity for the FP64 state; the mixed form is
compared on its own state (read back as hi + lo), to 1e-10.
"""
import math

import numpy as np
import pytest

from _util import smooth_box

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gk():
    import gkm_amd
    gkm_amd.load()
    return gkm_amd


def _gas(gk, **kw):
    base = dict(gamma=1.4, R=287.0, Pr=0.7, mu0=3e-3, T0=1.0 / (1.4 * 287.0), visc_law="sutherland",
                recon="symoo", energy_recon="T", zbug_compat=True, precision="fp64")
    base.update(kw)
    return gk.GasParams(**base)


def _orc_params(orc, gas):
    import gkm_amd._abi as A
    return orc.make_params(gamma=gas.gamma, R=gas.R, Pr=gas.Pr, mu0=gas.mu0, T0=gas.T0, visc_exp=gas.visc_exp,
                           K=gas.dof(), visc_law=A.VISC[gas.visc_law], recon=A.RECON[gas.recon],
                           energy_recon=A.ENERGY[gas.energy_recon], zbug_compat=int(gas.zbug_compat), precision=0)


def _rough_state(orc, nx, ny, nz, seed):
    """Smooth field plus cell noise, so skewness and kurtosis are far from
    their symmetric values; ghosts filled."""
    rng = np.random.default_rng(seed)
    grid = orc.make_grid(nx, ny, nz, 2 * math.pi / nx)
    W = smooth_box(nx, ny, nz, seed=seed, mach=0.4)
    orc.fill_ghosts(grid, W)
    W[0] = W[0] * (1 + 0.2 * rng.random(W[0].size))
    for q in (1, 2, 3):
        W[q] = W[q] + 0.3 * rng.normal(size=W[q].size)
    W[4] = W[4] + 0.5 * (W[1] ** 2 + W[2] ** 2 + W[3] ** 2) / W[0]
    orc.fill_ghosts(grid, W)
    return grid, W


def _check(dev, ref, tol):
    for k, v in ref.items():
        assert abs(dev[k] - v) <= tol * abs(v) + 1e-300, (k, dev[k], v)


@pytest.mark.parametrize("shape", [(12, 10, 14), (7, 9, 5), (40, 33, 36)])
@pytest.mark.parametrize("law", ["sutherland", "power"])
def test_print_stats_fp64(gk, orc, shape, law):
    nx, ny, nz = shape
    print("print_stats fp64", shape, law, flush=True)
    grid, W = _rough_state(orc, nx, ny, nz, seed=nx + 3 * ny)
    gas = _gas(gk, visc_law=law)
    with gk.Solver(nx, ny, nz, gas=gas) as s:
        s.set_state(W)
        dev = s.print_stats()
    ref = orc.print_stats(grid, _orc_params(orc, gas), W)
    _check(dev, ref, 1e-11)


@pytest.mark.parametrize("shape", [(335, 6, 7), (336, 6, 7), (5, 340, 4), (4, 5, 340)])
def test_print_stats_long_lines(gk, orc, shape):
    """Lines of up to 341 padded nodes (335 cells) are solved whole in LDS,
    longer ones in 16-node chunks through the gradient planes: one axis at a
    time on either side of that limit, against the oracle."""
    nx, ny, nz = shape
    print("print_stats long lines", shape, flush=True)
    grid, W = _rough_state(orc, nx, ny, nz, seed=nx + 7 * nz)
    gas = _gas(gk)
    with gk.Solver(nx, ny, nz, gas=gas) as s:
        s.set_state(W)
        dev = s.print_stats()
    ref = orc.print_stats(grid, _orc_params(orc, gas), W)
    _check(dev, ref, 1e-11)


@pytest.mark.parametrize("shape", [(6, 5, 3), (5, 6, 4), (4, 5, 335), (5, 4, 334)])
def test_print_stats_z_line_lengths(gk, orc, shape):
    """The z lines are solved from both ends toward row m/2 (m = nz + 6 padded
    nodes), the upper half one row longer when m is even: the shortest lines
    the grid allows (m = 9, 10) and the longest solved whole in LDS (m = 341,
    340), against the oracle."""
    nx, ny, nz = shape
    print("print_stats z line lengths", shape, flush=True)
    grid, W = _rough_state(orc, nx, ny, nz, seed=nz + 11)
    gas = _gas(gk)
    with gk.Solver(nx, ny, nz, gas=gas) as s:
        s.set_state(W)
        dev = s.print_stats()
    ref = orc.print_stats(grid, _orc_params(orc, gas), W)
    _check(dev, ref, 1e-11)


def test_print_stats_mixed(gk, orc):
    nx, ny, nz = 16, 12, 20
    print("print_stats mixed", (nx, ny, nz), flush=True)
    grid, W = _rough_state(orc, nx, ny, nz, seed=5)
    gas = _gas(gk, precision="mixed")
    with gk.Solver(nx, ny, nz, gas=gas) as s:
        s.set_state(W)
        Wm = s.get_state()
        dev = s.print_stats()
    ref = orc.print_stats(grid, _orc_params(orc, gas), Wm)
    _check(dev, ref, 1e-10)


def test_print_stats_after_steps_and_line(gk, orc):
    """DHIT 24^3 after 5 FP64 steps: device vs oracle on the evolved state, and
    the formatted line carries the same numbers."""
    print("print_stats dhit 24^3", flush=True)
    s = gk.Solver.from_dhit(24, k0=4, Mt=0.4, Re=50, seed=8, visc_law="sutherland", recon="js")
    dt = s.timestep(0.5)
    s.evolve_eulerian(5, dt=dt)
    d = s.print_stats()
    W = s.get_state()
    grid = orc.make_grid(24, 24, 24, 2 * math.pi / 24)
    ref = orc.print_stats(grid, _orc_params(orc, s.gas), W)
    _check({k: d[k] for k in ref}, ref, 1e-11)
    cols = [float(x) for x in d["line"].split()]
    ic = s.ic
    g = s.gas.gamma
    want = [s.info.t / s.info.t0, ref["tke"], ref["diss"], ref["skew_u"], ref["skew_v"], ref["skew_w"],
            ref["kurt_u"], ref["kurt_v"], ref["kurt_w"], ref["p_rms"] / (g * ic.p0 * ic.Mt ** 2),
            ref["T_rms"] / ((g - 1) * s.gas.T0 * ic.Mt ** 2), ref["v_rms"] / (ic.Mt ** 2 / ic.rho0)]
    assert len(cols) == 12
    for a, b in zip(cols, want):
        assert abs(a - b) <= 1e-8 * abs(b), (a, b)
    s.close()


def test_print_stats_batch_equals_single(gk, orc):
    """Box b of a batch gives bit-for-bit the single-box result."""
    nx, ny, nz = 10, 12, 9
    print("print_stats batch", flush=True)
    states = [_rough_state(orc, nx, ny, nz, seed=40 + b)[1] for b in range(3)]
    gases = [_gas(gk, mu0=1e-3 * (b + 1)) for b in range(3)]
    b = gk.Batch(nx, gases, ny=ny, nz=nz)
    for i, W in enumerate(states):
        b.set_state(i, W)
    got = [b.print_stats(i) for i in range(3)]
    b.close()
    for i, W in enumerate(states):
        with gk.Solver(nx, ny, nz, gas=gases[i]) as s:
            s.set_state(W)
            d = s.print_stats()
            assert {k: d[k] for k in got[i]} == got[i]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("scheme", ["weno_bgk", "gks"])
def test_step_print_stats_matches_calls_in_turn(gk, prec, scheme):
    """gkm_step_print_stats (statistics of each state on a second stream beside
    the step that reads it) gives, bitwise, the statistics and states of
    gkm_print_stats and gkm_step called in turn: eager and graph steps, and
    the slab-exchange mode, which runs them in turn."""
    kw = dict(precision=prec, visc_law="power")
    if scheme == "gks":
        kw.update(flux_scheme="gks", Pr=1.0)
    else:
        kw.update(recon="js")
    ic = gk.dhit(14, ny=12, nz=10, Mt=0.3, Re=40, seed=123)
    for graph, halo in [(False, "gather"), (True, "gather"), (False, "pack")]:
        print("step_print_stats", prec, scheme, graph, halo, flush=True)
        a = gk.Solver.from_ic(ic, **kw)
        b = gk.Solver.from_ic(ic, **kw)
        a.set_graph(graph)
        b.set_graph(graph)
        a.set_halo_mode(halo)
        b.set_halo_mode(halo)
        dt = a.timestep(0.4)
        got = a.step_print_stats(5, dt)
        want = []
        for _ in range(5):
            want.append(b.print_stats())
            b.evolve_eulerian(1, dt=dt)
        for g_, w_ in zip(got, want):
            assert g_ == w_
        for x, y in zip(a.get_state(), b.get_state()):
            assert np.array_equal(x, y)
        # and again from where they are (second call reuses the stream and buffers)
        got = a.step_print_stats(2, dt)
        for g_ in got:
            w_ = b.print_stats()
            b.evolve_eulerian(1, dt=dt)
            assert g_ == w_
        a.close()
        b.close()
