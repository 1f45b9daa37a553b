"""Batched boxes, captured-graph replay and launch-shape tuning on the GPU.

Every box of a batch runs the same per-interface arithmetic as a single-box
handle, and a replayed graph runs the same kernels as eager launches, so these
comparisons are bitwise.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gk():
    import gkm_amd
    gkm_amd.load()
    return gkm_amd


def _ics(gk, n=12, k=3):
    return [gk.dhit(n, Mt=0.25 + 0.1 * i, Re=40 + 10 * i, seed=50 + i) for i in range(k)]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batch_matches_single_boxes(gk, prec):
    ics = _ics(gk)
    b = gk.Batch.from_ics(ics, precision=prec, recon="symoo", visc_law="power")
    dts = b.timestep(0.5)
    assert len(set(dts.tolist())) == 3  # each box has its own dt
    b.step(5)
    for i, ic in enumerate(ics):
        s = gk.Solver.from_ic(ic, precision=prec, recon="symoo", visc_law="power")
        s.evolve_eulerian(5, dt=dts[i])
        a, c = s.get_state(), b.get_state(i)
        s.close()
        for x, y in zip(a, c):
            assert np.array_equal(x, y), f"box {i} differs"
        assert b.stats(i)["ek"] > 0
    b.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_graph_replay_matches_eager(gk, prec):
    ic = gk.dhit(16, Mt=0.4, Re=50, seed=8)
    eager = gk.Solver.from_ic(ic, precision=prec)
    dt = eager.timestep(0.5)
    ref = []
    for _ in range(4):
        eager.evolve_eulerian(2, dt=dt)
        ref.append(eager.get_state())
    eager.close()
    g = gk.Solver.from_ic(ic, precision=prec)
    g.set_graph(True)
    for r in range(4):
        g.evolve_eulerian(2, dt=dt)           # replays of the captured step
        got = g.get_state()                   # eager copies + sync between replays
        g.stats()                             # eager kernels between replays
        for x, y in zip(got, ref[r]):
            assert np.array_equal(x, y), f"replay round {r} differs"
    g.close()


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_tuning_does_not_change_results(gk, prec):
    """Every launch shape the autotuner may pick gives the same bits: X/Y flux
    block and WENO chunk (gkm_set_tuning), the fused Z kernel's block and strip
    width (gkm_set_tuning_z; 20 columns of i make the last strip partial for
    widths 8 and 16, and 32 is wider than the box)."""
    ic = gk.dhit(20, Mt=0.5, Re=70, seed=9)
    outs = []
    for shape in ((256, 8), (64, 4), (128, 16), (256, 8, 64, 8), (64, 16, 256, 32), (128, 4, 128, 4)):
        s = gk.Solver.from_ic(ic, precision=prec, recon="js")
        s.set_tuning(*shape)
        s.evolve_eulerian(3, dt=0.01)
        outs.append(s.get_state())
        s.close()
    for o in outs[1:]:
        for x, y in zip(o, outs[0]):
            assert np.array_equal(x, y)


def test_small_ensemble_mixed_vs_fp64(gk):
    reals = [gk.Realization(n=n, Mt=m, seed=s) for n, m, s in ((12, 0.3, 1), (12, 0.45, 2), (16, 0.25, 3))]
    em = gk.Ensemble(reals, precision="mixed")
    ed = gk.Ensemble(reals, precision="fp64")
    dts = ed.timestep(0.5)
    em.set_dt(dts)
    for _ in range(3):
        em.step(5)
        ed.step(5)
    sm, sd = em.stats(), ed.stats()
    for a, b in zip(sm, sd):
        assert abs(a["ek"] - b["ek"]) <= 1e-3 * abs(b["ek"])
        assert abs(a["mt"] - b["mt"]) <= 1e-3 * abs(b["mt"])
    em.close()
    ed.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_pack_halo_matches_gather(gk, prec):
    """y-face pack / exchange / unpack + x/z fill gives the same ghosts, hence
    the same steps, as the per-cell periodic gather (uneven batch of 2)."""
    ics = [gk.dhit(10, ny=7, nz=12, Mt=0.3 + 0.1 * i, seed=70 + i) for i in range(2)]
    a = gk.Batch.from_ics(ics, precision=prec)
    b = gk.Batch.from_ics(ics, precision=prec)
    b.set_halo_mode("pack")
    dts = a.timestep(0.5)
    a.step(4, dts)
    b.step(4, dts)
    for i in range(2):
        for x, y in zip(a.get_state(i), b.get_state(i)):
            assert np.array_equal(x, y)
    a.close()
    b.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_group_step_matches_separate_steps(gk, prec):
    """A group of a 12^3 batch, a 16x12x10 box and a GKS box, stepped by one
    captured graph per step (gkm_group_step), equals the same handles stepped
    one by one, bitwise, over both state-buffer parities and with eager work
    (statistics, a lone step of one member) between the group calls."""
    import gkm_amd
    kw = dict(precision=prec, visc_law="power", recon="js")
    ics = [gk.dhit(12, Mt=0.3 + 0.1 * i, Re=40, seed=80 + i) for i in range(2)]
    ic1 = gk.dhit(16, Mt=0.4, Re=50, seed=90, ny=12, nz=10)
    ic2 = gk.dhit(10, Mt=0.3, Re=40, seed=91)

    def make():
        b = gk.Batch.from_ics(ics, **kw)
        s1 = gk.Solver.from_ic(ic1, **kw)
        s2 = gk.Solver.from_ic(ic2, flux_scheme="gks", Pr=1.0, precision=prec, visc_law="power")
        b.timestep(0.4)
        s1.timestep(0.4)
        s2.timestep(0.4)
        return b, s1, s2

    b, s1, s2 = make()
    g = gkm_amd.Group([b, s1, s2])
    g.step(3)
    b.stats(0)
    s1.evolve_eulerian(1)  # a lone step flips s1 only: the group recaptures
    g.step(2)
    g.sync()
    got = [b.get_state(0), b.get_state(1), s1.get_state(), s2.get_state()]
    g.close()
    for x in (b, s1, s2):
        x.close()

    b, s1, s2 = make()
    b.step(3)
    s1.evolve_eulerian(4)
    s2.evolve_eulerian(3)
    b.step(2)
    s1.evolve_eulerian(2)
    s2.evolve_eulerian(2)
    want = [b.get_state(0), b.get_state(1), s1.get_state(), s2.get_state()]
    for x in (b, s1, s2):
        x.close()
    for k, (x, y) in enumerate(zip(got, want)):
        assert all(np.array_equal(p, q) for p, q in zip(x, y)), k


def test_group_needs_dt(gk):
    import ctypes as C

    import gkm_amd._abi as A
    s = gk.Solver.from_dhit(8, Mt=0.3, Re=30, seed=1)
    L = A.load()
    g = C.c_void_p()
    arr = (C.c_void_p * 1)(s._h.value)
    A.check(L.gkm_group_create(arr, 1, C.byref(g)))
    assert L.gkm_group_step(g, 1) == A.GKM_E_ARG
    assert b"no dt" in L.gkm_last_error()
    arr2 = (C.c_void_p * 2)(s._h.value, s._h.value)
    g2 = C.c_void_p()
    assert L.gkm_group_create(arr2, 2, C.byref(g2)) == A.GKM_E_ARG
    L.gkm_group_destroy(g)
    s.close()


@pytest.mark.parametrize("scheme", ["weno_bgk", "gks"])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_fused_ghosts_match_ghost_fill(gk, prec, scheme):
    """With the periodic fill, a step's Z sweep also writes the next state's
    ghost cells and its Tw / tau planes; with the slab exchange the ghost
    fill and the primitives run at the start of every step instead.  Both
    give the same states, ghosts included, bitwise: boxes down to 3 cells a
    side (every cell then has images on both sides of every axis), graph and
    eager steps, a set_state in between (which makes the next step fill the
    ghosts again) and statistics between steps."""
    kw = dict(precision=prec, visc_law="power")
    if scheme == "gks":
        kw.update(flux_scheme="gks", Pr=1.0)
    else:
        kw.update(recon="js")
    for shape in [(3, 3, 3), (5, 3, 4), (7, 6, 9)]:
        nx, ny, nz = shape
        print("fused ghosts", prec, scheme, shape, flush=True)
        ic = gk.dhit(nx, ny=ny, nz=nz, Mt=0.3, Re=40, seed=nx + 10 * ny + 100 * nz)
        a = gk.Batch.from_ics([ic], **kw)
        b = gk.Batch.from_ics([ic], **kw)
        b.set_halo_mode("pack")
        a.set_graph(True)
        dts = a.timestep(0.4)
        for _ in range(2):
            a.step(3, dts)
            b.step(3, dts)
            a.stats(0)
            b.stats(0)
            for x, y in zip(a.get_state(0), b.get_state(0)):
                assert np.array_equal(x, y), shape
        W = a.get_state(0)
        a.set_state(0, W)
        b.set_state(0, W)
        a.step(2, dts)
        b.step(2, dts)
        for x, y in zip(a.get_state(0), b.get_state(0)):
            assert np.array_equal(x, y), shape
        a.close()
        b.close()
