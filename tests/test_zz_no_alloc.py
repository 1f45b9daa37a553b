"""gkm_flux_points_noalloc: the flux-points entry point with a static device
scratch array instead of a per-call allocation.  Checked against
gkm_flux_points (the same kernel) bit for bit, on two calls in a row, the
second with fewer records, in both precisions; sizes beyond the array's
maximum are rejected before any GPU work."""
import ctypes as C

import numpy as np
import pytest

from _util import flux_records

pytestmark = pytest.mark.gpu

MAX_RECORDS = 8192


@pytest.fixture(scope="module")
def lib():
    import gkm_amd._abi as A
    return A.load()


def _call(L, fn, gas, rec, dx):
    import gkm_amd._abi as A
    rec = np.ascontiguousarray(rec, dtype=np.float64)
    F = np.full((rec.shape[0], 5), np.nan)
    p = gas.to_c()
    A.check(getattr(L, fn)(C.byref(p), rec.shape[0], rec.ctypes.data_as(A.DP), float(dx), F.ctypes.data_as(A.DP)))
    return F


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_matches_flux_points(lib, precision):
    import gkm_amd
    gas = gkm_amd.GasParams(precision=precision, Pr=0.7)
    for n, seed in ((MAX_RECORDS, 31), (777, 32)):
        print(f"noalloc {precision} n={n}", flush=True)
        rec = flux_records(n, seed=seed, dt_over_tau=(0.01, 50.0))
        want = _call(lib, "gkm_flux_points", gas, rec, 0.05)
        got = _call(lib, "gkm_flux_points_noalloc", gas, rec, 0.05)
        assert np.all(np.isfinite(want))
        assert np.array_equal(got, want)


def test_noalloc_rejects_oversize(lib):
    import gkm_amd
    import gkm_amd._abi as A
    p = gkm_amd.GasParams().to_c()
    rec = np.zeros(52)
    F = np.zeros(5)
    rc = lib.gkm_flux_points_noalloc(C.byref(p), MAX_RECORDS + 1, rec.ctypes.data_as(A.DP), 0.05,
                                     F.ctypes.data_as(A.DP))
    assert rc == A.GKM_E_ARG
    assert b"static scratch" in lib.gkm_last_error()
