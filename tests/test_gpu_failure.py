"""Failure detection in the step (MI355X).

The reference checks nothing (weno/src/GPUupdate.cu:11-20 is commented out): a
negative density or pressure runs on as NaNs.  Here the fused Z kernel raises
a per-box flag on any new cell whose rho or T = p/(R rho) is not positive and
finite, and the host calls report GKM_E_NONPHYSICAL (include/gkm.h, gkm_step).

The non-physical states are planted from the host through gkm_set_state: the
kernels then compute with them and produce NaNs, which is ordinary arithmetic
(no out-of-range index or wait depends on the values).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gk():
    import gkm_amd
    gkm_amd.load()
    return gkm_amd


def _plant(W, shape, case):
    ny, nx, nz = shape
    W = [w.copy() for w in W]
    c = np.ravel_multi_index((3 + ny // 2, 3 + nx // 3, 3 + nz // 4), (ny + 6, nx + 6, nz + 6))
    if case == "neg_rho":
        W[0][c] = -abs(W[0][c])
    elif case == "neg_p":
        ke = 0.5 * (W[1][c] ** 2 + W[2][c] ** 2 + W[3][c] ** 2) / W[0][c]
        W[4][c] = ke - 0.5 * abs(W[4][c] - ke)     # internal energy < 0
    else:
        W[2][c] = np.nan
    return W


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["neg_rho", "neg_p", "nan"])
def test_nonphysical_state_is_reported(gk, precision, case):
    import gkm_amd._abi as A
    n = 10
    print(f"nonphysical {precision} {case} {n}^3", flush=True)
    s = gk.Solver.from_dhit(n, Mt=0.3, Re=30, seed=5, precision=precision, visc_law="power")
    dt = s.timestep(0.5)
    good = s.get_state()
    # a healthy run raises nothing
    s.evolve_eulerian(3, dt)
    s.sync()
    s.set_state(_plant(good, (n, n, n), case))
    s.evolve_eulerian(1, dt)  # queued: the flag is raised on the device
    with pytest.raises(A.GkmError) as e:
        s.sync()
    assert e.value.code == A.GKM_E_NONPHYSICAL and "non-physical" in str(e.value)
    # the flag copy has landed: the next step is refused without waiting and without stepping
    before = s.get_state()
    with pytest.raises(A.GkmError) as e:
        s.evolve_eulerian(1, dt)
    assert e.value.code == A.GKM_E_NONPHYSICAL
    after = s.get_state()
    assert all(np.array_equal(a, b, equal_nan=True) for a, b in zip(before, after))
    for call in (s.stats, lambda: s.timestep(0.5), s.print_stats):
        with pytest.raises(A.GkmError) as e:
            call()
        assert e.value.code == A.GKM_E_NONPHYSICAL
    # a new state clears the flag
    s.set_state(good)
    s.evolve_eulerian(2, dt)
    s.sync()
    assert np.isfinite(s.stats()["ek"])
    s.close()


def test_nonphysical_flag_with_graph_and_lagged_report(gk):
    """With captured steps the flag is raised inside the graph; a second stepping
    call reports it without a host wait once the first call's flag copy landed."""
    import gkm_amd._abi as A
    n = 12
    print(f"nonphysical graph {n}^3", flush=True)
    s = gk.Solver.from_dhit(n, Mt=0.3, Re=30, seed=6, precision="mixed", visc_law="power")
    s.set_graph(True)
    dt = s.timestep(0.5)
    good = s.get_state()
    s.evolve_eulerian(4, dt)
    s.sync()
    s.set_state(_plant(good, (n, n, n), "neg_rho"))
    s.evolve_eulerian(2, dt)
    assert s._L.gkm_sync(s._h) == A.GKM_E_NONPHYSICAL
    with pytest.raises(A.GkmError):
        s.evolve_eulerian(1, dt)
    s.set_state(good)
    s.evolve_eulerian(3, dt)
    s.sync()
    s.close()


def test_batch_names_the_box(gk):
    """In a batch only the planted box's flag is raised, and the message names it."""
    import gkm_amd._abi as A
    from gkm_amd.ic import dhit
    n = 10
    print(f"nonphysical batch 3 x {n}^3", flush=True)
    ics = [dhit(n, k0=4, Mt=0.3, Re=30, seed=40 + b) for b in range(3)]
    b = gk.Batch.from_ics(ics, precision="fp64", visc_law="power")
    dts = b.timestep(0.5)
    W1 = b.get_state(1)
    b.set_state(1, _plant(W1, (n, n, n), "neg_p"))
    b.step(1, dts)
    with pytest.raises(A.GkmError) as e:
        b.sync()
    assert e.value.code == A.GKM_E_NONPHYSICAL and "box 1" in str(e.value)
    b.set_state(1, W1)
    b.step(1, dts)
    b.sync()
    b.close()
