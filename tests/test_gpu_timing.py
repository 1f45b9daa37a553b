"""Device timing of the step (gkm_set_timing / gkm_last_timing / gkm_last_sweep_timing).

GKM_TIMING_ALL records an event around every launch of a timed step, so every
entry of the breakdown is set; GKM_TIMING_FLUX records only the four events
around the X and Y flux kernels, so only those entries are set and the rest
read -1.  Timing never changes results: a timed run is bitwise equal to an
untimed one.
"""
import numpy as np
import pytest

from _util import smooth_box

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gk():
    import gkm_amd
    gkm_amd.load()
    return gkm_amd


def _solver(gk, prec):
    W = smooth_box(16, 12, 14, seed=3)
    s = gk.Solver(16, 12, 14, dx=2 * np.pi / 16, gas=gk.GasParams(precision=prec))
    s.set_state(W)
    return s


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_timing_modes(gk, prec):
    dt = 1e-3
    ref = _solver(gk, prec)
    ref.evolve_eulerian(3, dt=dt)
    W_ref = ref.get_state()
    ref.close()
    for mode in ("all", "flux"):
        print("timing mode", mode, prec, flush=True)
        s = _solver(gk, prec)
        s.set_timing(mode)
        s.evolve_eulerian(3, dt=dt)
        t = s.last_timing()
        W = s.get_state()
        s.set_timing(False)
        s.close()
        assert t["steps"] == 3
        assert t["flux_ms"] > 0 and t["flux_x_ms"] > 0 and t["flux_y_ms"] > 0
        if mode == "all":
            assert t["recon_ms"] > 0 and t["flux_z_ms"] > 0
            assert t["flux_ms"] >= t["flux_x_ms"] + t["flux_y_ms"] + t["flux_z_ms"] - 1e-6
        else:
            assert t["recon_ms"] == -1.0 and t["ghost_ms"] == -1.0 and t["update_ms"] == -1.0
            assert t["flux_z_ms"] == -1.0
            assert abs(t["flux_ms"] - (t["flux_x_ms"] + t["flux_y_ms"])) < 1e-6
        for a, b in zip(W, W_ref):
            assert np.array_equal(a, b)
