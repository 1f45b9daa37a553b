"""gkm_init_dhit on the device: the state it writes is the library's host
generator's field in the conservative layout (bitwise), the ghosts are
filled, the box's mu0 and T0 are the recipe's, and a run from it equals a run
from the same field set through gkm_set_state (bitwise)."""
import numpy as np
import pytest

from _util import rel_l2  # noqa: F401

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gk():
    import gkm_amd
    gkm_amd.load()
    return gkm_amd


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_init_dhit_state_and_params(gk, orc, prec):
    import ctypes as C
    import math

    import gkm_amd._abi as A
    nx, ny, nz = 20, 16, 24
    gas = gk.GasParams(precision=prec, visc_law="power", recon="js")
    s = gk.Solver(nx, ny, nz, gas=gas)
    info = s.init_dhit(k0=4, Mt=0.4, Re=60, seed=42)
    W = s.get_state()
    ic = gk.dhit_native(nx, k0=4, Mt=0.4, Re=60, seed=42, ny=ny, nz=nz, gamma=gas.gamma)
    ref = gk.conservative_padded(ic)
    grid = orc.make_grid(nx, ny, nz, 2 * math.pi / nx)
    orc.fill_ghosts(grid, ref)
    if prec == "fp64":
        assert all(np.array_equal(a, b) for a, b in zip(W, ref))
    else:
        assert rel_l2(W, ref) < 1e-13
    p = A.GkmParams()
    A.check(s._L.gkm_get_params_box(s._h, 0, C.byref(p)))
    assert p.mu0 == info["mu0"] == ic.mu0
    assert p.T0 == ic.p0 / (gas.R * ic.rho0)
    # a run from it equals a run from the same state and constants set by hand
    dt = s.timestep(0.4)
    s.evolve_eulerian(3, dt=dt)
    a = s.get_state()
    s.close()
    with gk.Solver.from_ic(ic, precision=prec, visc_law="power", recon="js") as t:
        t.evolve_eulerian(3, dt=dt)
        b = t.get_state()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))


def test_set_params_box_changes_the_viscosity(gk):
    """gkm_set_params_box on box 1 of a batch changes that box only; a captured
    graph picks the new constants up without a new capture."""
    import ctypes as C

    import gkm_amd._abi as A
    ics = [gk.dhit(12, Mt=0.3, Re=40, seed=70 + i) for i in range(2)]
    b = gk.Batch.from_ics(ics, precision="fp64", visc_law="power")
    b.set_graph(True)
    dts = b.timestep(0.4)
    b.step(1)
    p = A.GkmParams()
    A.check(b._L.gkm_get_params_box(b._h, 1, C.byref(p)))
    p.mu0 *= 3.0
    A.check(b._L.gkm_set_params_box(b._h, 1, C.byref(p)))
    bad = A.GkmParams.from_buffer_copy(p)
    bad.recon = 1 - bad.recon
    assert b._L.gkm_set_params_box(b._h, 1, C.byref(bad)) == A.GKM_E_ARG
    b.step(2)
    got = [b.get_state(i) for i in range(2)]
    b.close()
    for i, ic in enumerate(ics):
        with gk.Solver.from_ic(ic, precision="fp64", visc_law="power") as s:
            s.evolve_eulerian(1, dt=dts[i])
            if i == 1:
                s.gas.mu0 *= 3.0
                A.check(s._L.gkm_set_params_box(s._h, 0, C.byref(s.gas.to_c())))
            s.evolve_eulerian(2, dt=dts[i])
            assert all(np.array_equal(x, y) for x, y in zip(s.get_state(), got[i])), i
