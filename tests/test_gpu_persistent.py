"""Persistent group step (gkm_group_step_persistent) on the GPU.

One cooperative launch runs every phase of every step for every box of every
member handle, with grid barriers between the phases; each phase runs the same
block functions as the per-stage kernels, so every comparison with the
captured-graph group step (gkm_group_step) is bitwise.  The cases cover both
precisions and both reconstructions, the temperature and energy modes, uneven
classes with several boxes each (every box its own dt), odd and even step
counts (the state buffers alternate inside the launch), eager statistics
between calls, WENO chunk choices per class, the non-physical flag raised
inside the launch, and the combinations the persistent step refuses.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gk():
    import gkm_amd
    gkm_amd.load()
    return gkm_amd


def _classes(gk):
    """Three size classes: 2 x 12^3, 3 x 10x14x12 (x, y, z), 1 x 16^3."""
    return [
        [gk.dhit(12, Mt=0.3 + 0.05 * i, Re=40, seed=110 + i) for i in range(2)],
        [gk.dhit(10, ny=14, nz=12, Mt=0.35 + 0.05 * i, Re=50, seed=120 + i) for i in range(3)],
        [gk.dhit(16, Mt=0.45, Re=60, seed=130)],
    ]


def _batches(gk, classes, **kw):
    out = []
    for ics in classes:
        b = gk.Batch.from_ics(ics, **kw)
        b.timestep(0.5)  # each box its own dt
        out.append(b)
    return out


def _same(ba, bb, what):
    for b1, b2 in zip(ba, bb):
        for i in range(b1.nbox):
            for q, (x, y) in enumerate(zip(b1.get_state(i), b2.get_state(i))):
                assert np.array_equal(x, y), f"{what}: box {i} field {q} differs"


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
@pytest.mark.parametrize("recon,energy", [("js", "T"), ("symoo", "E")])
def test_persistent_matches_group_graph(gk, prec, recon, energy):
    classes = _classes(gk)
    kw = dict(precision=prec, recon=recon, energy_recon=energy, visc_law="power")
    print("persistent", prec, recon, energy, flush=True)
    ba, bb = _batches(gk, classes, **kw), _batches(gk, classes, **kw)
    ga, gb = gk.Group(ba), gk.Group(bb)
    for r, n in enumerate((1, 3, 2, 1)):
        ga.step(n)
        gb.step(n, persistent=True)
        ga.sync()
        gb.sync()
        _same(ba, bb, f"call {r} ({n} steps)")
        for b in bb:                      # eager work between calls
            b.stats(0)
    ga.close()
    gb.close()
    for b in ba + bb:
        b.close()


def test_persistent_weno_chunks_per_class(gk):
    """Each class keeps its own WENO chunk (4, 16, 8) inside the launch."""
    classes = _classes(gk)
    ba, bb = _batches(gk, classes, precision="mixed"), _batches(gk, classes, precision="mixed")
    for b1, b2, rc in zip(ba, bb, (4, 16, 8)):
        b1.set_tuning(128, rc)
        b2.set_tuning(256, rc)
    ga, gb = gk.Group(ba), gk.Group(bb)
    ga.step(3)
    gb.step(3, persistent=True)
    ga.sync()
    gb.sync()
    _same(ba, bb, "chunks")
    ga.close()
    gb.close()
    for b in ba + bb:
        b.close()


def test_persistent_reports_nonphysical_box(gk):
    import gkm_amd._abi as A
    classes = _classes(gk)
    bs = _batches(gk, classes, precision="mixed")
    g = gk.Group(bs)
    g.step(2, persistent=True)
    g.sync()
    W = bs[1].get_state(2)
    c = np.ravel_multi_index((3 + 7, 3 + 5, 3 + 6), (14 + 6, 10 + 6, 12 + 6))
    W[0][c] = -abs(W[0][c])
    bs[1].set_state(2, W)
    g.step(1, persistent=True)
    with pytest.raises(A.GkmError) as e:
        g.sync()
    assert e.value.code == A.GKM_E_NONPHYSICAL and "box 2" in str(e.value)
    g.close()
    for b in bs:
        b.close()


def test_persistent_refuses_what_it_does_not_run(gk):
    import gkm_amd._abi as A
    classes = _classes(gk)[:2]
    cases = {
        "gks": [dict(precision="mixed", flux_scheme="gks", Pr=1.0), dict(precision="mixed", flux_scheme="gks", Pr=1.0)],
        "mixed_precisions": [dict(precision="mixed"), dict(precision="fp64")],
        "mixed_recon": [dict(precision="mixed", recon="js"), dict(precision="mixed", recon="symoo")],
    }
    for name, kws in cases.items():
        bs = []
        for ics, kw in zip(classes, kws):
            b = gk.Batch.from_ics(ics, **kw)
            b.timestep(0.5)
            bs.append(b)
        g = gk.Group(bs)
        with pytest.raises(A.GkmError) as e:
            g.step(1, persistent=True)
        assert e.value.code == A.GKM_E_ARG, name
        g.step(1)  # the per-stage group step runs them
        g.sync()
        g.close()
        for b in bs:
            b.close()
    bs = _batches(gk, classes, precision="mixed")
    bs[0].set_halo_mode("pack")
    g = gk.Group(bs)
    with pytest.raises(A.GkmError) as e:
        g.step(1, persistent=True)
    assert e.value.code == A.GKM_E_ARG
    g.close()
    for b in bs:
        b.close()
