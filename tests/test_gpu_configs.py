"""The two largest workloads at their full sizes (MI355X).

1. The 96-realization ensemble of uneven boxes (32 x 32^3, 32 x 48^3,
   24 x 64^3, 8 x 96^3; Mt ~ U[0.2, 0.5], k0 = 4, Re 70), mixed form, advanced
   by one captured graph launch per step over every box (gkm_group_step):
     * the first and last box of every size class equal a single-box run of
       the same initial condition and dt, bit for bit;
     * every box's E_k and Mt stay within 1e-3 relative of the same ensemble
       in FP64 over 10 steps.
2. One rank's 1024 x 128 x 1024 slab of a 1024^3 box (x, y, z; y is the
   slab-normal axis whose ghost planes a distributed run exchanges), mixed
   form, with the y faces packed, exchanged and unpacked on the device
   (GKM_HALO_PACK; on one GPU the exchange is the periodic self-copy that
   stands in for the neighbour ranks, weno/src/update.cpp:894-1054):
     * pack mode equals the per-cell periodic gather bit for bit;
     * mass, momentum and energy are conserved over 3 steps (1e-9 of the mass
       + energy scale: the FP32 flux differences do not telescope exactly);
     * the state shifted periodically by (37, 301, 517) cells steps to the
       shifted result, bit for bit: a wrong face offset, a 32-bit offset that
       wraps or a stride slip at this size would move values between cells;
     * one mixed step's increment is within 1e-4 relative L2 of the FP64 step's.
   The slab starts from a smooth analytic state (Fourier modes in every axis,
   Mt ~ 0.6) so that no 10^8-cell spectral initial condition is generated on
   the host inside the test.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gk():
    import gkm_amd
    gkm_amd.load()
    return gkm_amd


# --------------------------------------------------------------------------- ensemble
def test_baseline_ensemble_full_size(gk):
    reals = gk.baseline_ensemble()
    assert [r.n for r in reals].count(32) == 32 and [r.n for r in reals].count(96) == 8
    kw = dict(recon="js", energy_recon="T", visc_law="power")
    classes = {}
    for n in sorted({r.n for r in reals}):
        members = [r for r in reals if r.n == n]
        ics = [gk.dhit(n, k0=r.k0, Mt=r.Mt, Re=r.Re, seed=r.seed) for r in members]
        classes[n] = ics
    print("ensemble ICs made", {n: len(v) for n, v in classes.items()}, flush=True)
    mixed = {n: gk.Batch.from_ics(ics, precision="mixed", **kw) for n, ics in classes.items()}
    fp64 = {n: gk.Batch.from_ics(ics, precision="fp64", **kw) for n, ics in classes.items()}
    dts = {n: b.timestep(0.5) for n, b in fp64.items()}
    for n in classes:
        mixed[n].dt[:] = dts[n]
    group = gk.Group(list(mixed.values()))
    worst = {"ek": 0.0, "mt": 0.0}
    for _ in range(2):
        group.step(5)
        for n, b in fp64.items():
            b.step(5, dts[n])
        group.sync()
        for n in classes:
            for k in range(len(classes[n])):
                sm, sd = mixed[n].stats(k), fp64[n].stats(k)
                for key in worst:
                    worst[key] = max(worst[key], abs(sm[key] - sd[key]) / abs(sd[key]))
    print("ensemble mixed vs fp64 max rel err", worst, flush=True)
    assert worst["ek"] <= 1e-3 and worst["mt"] <= 1e-3
    for n, ics in classes.items():
        for k in (0, len(ics) - 1):
            s = gk.Solver.from_ic(ics[k], precision="mixed", **kw)
            s.evolve_eulerian(10, dt=float(dts[n][k]))
            want = s.get_state()
            s.close()
            got = mixed[n].get_state(k)
            assert all(np.array_equal(x, y) for x, y in zip(got, want)), (n, k)
    group.close()
    for b in list(mixed.values()) + list(fp64.values()):
        b.close()


# --------------------------------------------------------------------------- slab
NX, NY, NZ = 1024, 128, 1024
SHIFT = (37, 301, 517)  # cells along (y, x, z), the array axes of the padded planes
SLAB_STEPS = 3


def _slab_gas(gk, precision):
    gam, R = 1.4, 287.0
    p0 = 1.0 / gam
    # tau0 = mu0/p0 = 4.9e-4 against dt ~ 2e-3 at CFL 0.5: the continuum regime
    return gk.GasParams(gamma=gam, R=R, mu0=3.5e-4, T0=p0 / R, visc_law="power", recon="js",
                        energy_recon="T", precision=precision)


@pytest.fixture(scope="module")
def slab_state():
    """Padded conservative planes of a smooth periodic state on the slab: each
    field a sum of separable Fourier modes with whole periods along every axis
    (the slab's y extent is an eighth of the x and z extents)."""
    dx = 2 * math.pi / NX
    x = np.arange(NX) * dx
    y = np.arange(NY) * dx * 8.0      # one period of sin(y') over the slab's 128 cells
    z = np.arange(NZ) * dx

    def sep(a, fx, fy, fz):
        return a * fy[:, None, None] * fx[None, :, None] * fz[None, None, :]

    gam = 1.4
    rho = 1.0 + sep(0.1, np.cos(x), np.sin(y), np.cos(2 * z))
    # rms velocity ~0.6 against a sound speed of 1 (Mt ~ 0.6; locally supersonic)
    u = sep(1.15, np.sin(x), np.cos(y), np.cos(z)) + sep(0.23, np.cos(3 * x), np.ones(NY), np.sin(z))
    v = sep(0.69, np.cos(2 * x), np.sin(y), np.sin(z + 0.3))
    w = sep(-0.92, np.cos(x), np.cos(y), np.sin(3 * z)) + sep(0.23, np.ones(NX), np.sin(2 * y), np.cos(z))
    p = (1.0 / gam) * (1.0 + sep(0.1, np.sin(2 * x), np.cos(y), np.cos(z)))
    W = [np.zeros((NY + 6, NX + 6, NZ + 6)) for _ in range(5)]
    s = (slice(3, 3 + NY), slice(3, 3 + NX), slice(3, 3 + NZ))
    W[0][s] = rho
    W[1][s] = rho * u
    W[2][s] = rho * v
    W[3][s] = rho * w
    W[4][s] = p / (gam - 1) + 0.5 * rho * (u * u + v * v + w * w)
    del rho, u, v, w, p
    return [a.reshape(-1) for a in W]


def _interior(W):
    return [w.reshape(NY + 6, NX + 6, NZ + 6)[3:-3, 3:-3, 3:-3] for w in W]


def _run(gk, W, precision, mode, dt=None, steps=SLAB_STEPS, stats=False):
    s = gk.Solver(NX, NY, NZ, dx=2 * math.pi / NX, gas=_slab_gas(gk, precision))
    s.set_halo_mode(mode)
    s.set_state(W)
    s.periodic_bc()
    if dt is None:
        dt = s.timestep(0.5)
    s0 = s.stats() if stats else None
    s.evolve_eulerian(steps, dt=dt)
    s1 = s.stats() if stats else None
    out = s.get_state()
    s.close()
    return out, dt, s0, s1


def test_slab_pack_gather_conservation_translation(gk, slab_state):
    print("slab pack run", flush=True)
    Wa, dt, s0, s1 = _run(gk, slab_state, "mixed", "pack", stats=True)
    scale = abs(s0["mass"]) + abs(s0["energy"])
    for key in ("mass", "mom_x", "mom_y", "mom_z", "energy"):
        assert abs(s1[key] - s0[key]) <= 1e-9 * scale, (key, s0[key], s1[key])
    assert 0.5 < s0["mt"] < 0.7
    print("slab gather run", flush=True)
    Wb, _, _, _ = _run(gk, slab_state, "mixed", "gather", dt=dt)
    for q, (x, y) in enumerate(zip(Wa, Wb)):
        assert np.array_equal(x, y), ("pack vs gather", q)
    del Wb
    print("slab shifted run", flush=True)
    roll = dict(shift=SHIFT, axis=(0, 1, 2))
    Ws = [np.zeros_like(w) for w in slab_state]
    for src, dst in zip(_interior(slab_state), _interior(Ws)):
        dst[...] = np.roll(src, **roll)
    Wc, _, _, _ = _run(gk, Ws, "mixed", "pack", dt=dt)
    del Ws
    for q, (x, y) in enumerate(zip(_interior(Wa), _interior(Wc))):
        assert np.array_equal(np.roll(x, **roll), y), ("shift", q)


def test_slab_mixed_step_matches_fp64_step(gk, slab_state):
    print("slab fp64 step", flush=True)
    Wd, dt, _, _ = _run(gk, slab_state, "fp64", "pack", steps=1)
    print("slab mixed step", flush=True)
    Wm, _, _, _ = _run(gk, slab_state, "mixed", "pack", dt=dt, steps=1)
    num = den = 0.0
    for a, b, w0 in zip(_interior(Wm), _interior(Wd), _interior(slab_state)):
        inc = b - w0
        num += float(np.sum((a - b) ** 2))
        den += float(np.sum(inc * inc))
    err = math.sqrt(num / den)
    print("slab one-step increment rel L2 mixed vs fp64", err, flush=True)
    assert err <= 1e-4
