"""The 256^3, Mt = 0.5 decaying-turbulence box (the benchmark's workload) checked
at its full size through properties that need no oracle run at that size:

- conservation: the periodic step moves mass, momentum and energy between
  cells only, so their box sums stay put over 10 steps (FP64 to 1e-13 of the
  mass + energy scale, the mixed form to 1e-9: its FP32 flux differences do
  not telescope exactly);
- translation: the scheme has no preferred cell, so the initial condition
  shifted periodically by (37, 101, 5) cells steps to the shifted state, bit
  for bit. Every interface's arithmetic sees the same stencil values in both
  runs. A wrong stride, a 32-bit offset that wraps, a missed ghost or a tile
  edge that reads the wrong neighbour would move values between cells, and
  the shifted run would then disagree with the unshifted one.
The oracle's own parity at these sizes is covered at 128^3 (test_gpu_parity).
"""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 256
SHIFT = (37, 101, 5)  # cells along (y, x, z), the array axes of the IC fields
STEPS = 10


@pytest.fixture(scope="module")
def gk():
    import gkm_amd
    gkm_amd.load()
    return gkm_amd


@pytest.fixture(scope="module")
def ic(gk):
    return gk.dhit(N, k0=4, Mt=0.5, Re=70, seed=1234)


def _interior(W):
    P = N + 6
    return [w.reshape(P, P, P)[3:-3, 3:-3, 3:-3] for w in W]


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_full_size_conservation_and_translation(gk, ic, prec):
    kw = dict(precision=prec, visc_law="power", recon="js")
    print("fullsize", prec, N, flush=True)
    a = gk.Solver.from_ic(ic, **kw)
    dt = a.timestep(0.5)
    s0 = a.stats()
    a.evolve_eulerian(STEPS, dt=dt)
    s1 = a.stats()
    Wa = _interior(a.get_state())
    a.close()
    scale = abs(s0["mass"]) + abs(s0["energy"])
    tol = 1e-13 if prec == "fp64" else 1e-9
    for key in ("mass", "mom_x", "mom_y", "mom_z", "energy"):
        assert abs(s1[key] - s0[key]) <= tol * scale, (key, s0[key], s1[key])
    assert all(np.all(np.isfinite(w)) for w in Wa)

    roll = dict(shift=SHIFT, axis=(0, 1, 2))
    shifted = dataclasses.replace(ic, u=np.roll(ic.u, **roll), v=np.roll(ic.v, **roll), w=np.roll(ic.w, **roll))
    b = gk.Solver.from_ic(shifted, **kw)
    b.evolve_eulerian(STEPS, dt=dt)
    Wb = _interior(b.get_state())
    b.close()
    for q, (x, y) in enumerate(zip(Wa, Wb)):
        assert np.array_equal(np.roll(x, **roll), y), (prec, q)
