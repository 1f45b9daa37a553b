"""CPU checks of the C ABI library and of the host-side pieces.

* libgkm.so loads and exports every entry point include/gkm.h declares (no
  compute calls: there is no GPU here);
* the ctypes structs match the C structs field for field;
* the DHIT initial condition has the init_BS.m properties: solenoidal, zero
  mean, the prescribed shell spectrum, u_rms / c0 = Mt/sqrt(3).
"""
import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "gkm.h"


def declared_functions():
    """Every function declared by the headers under include/."""
    names = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(gkm_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import gkm_amd._abi as A
    L = A.load()
    names = declared_functions()
    assert len(names) >= 35
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(A.EXPORTS)
    assert L.gkm_version().decode().startswith("gkm-mi355x")


def test_error_path_without_device_is_reported():
    """Argument errors are caught on the host, before any HIP call."""
    import gkm_amd._abi as A
    from gkm_amd.solver import GasParams
    L = A.load()
    grid = A.GkmGrid(8, 8, 8, 2, 0.1)  # ghost must be 3
    p = GasParams().to_c()
    h = C.c_void_p()
    assert L.gkm_create(C.byref(grid), C.byref(p), C.byref(h)) == A.GKM_E_ARG
    assert b"ghost" in L.gkm_last_error()
    assert not h.value


def test_struct_layouts_match_header():
    import gkm_amd._abi as A
    from oracle import oracle as O
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    body = re.search(r"typedef struct gkm_params \{(.*?)\} gkm_params;", text, flags=re.S).group(1)
    fields = re.findall(r"(double|int)\s+([\w\s,]+);", body)
    names = [n.strip() for _, group in fields for n in group.split(",")]
    assert [f for f, _ in A.GkmParams._fields_] == names
    assert [f for f, _ in O.Params._fields_] == names
    body = re.search(r"typedef struct gkm_stats_s \{(.*?)\} gkm_stats_t;", text, flags=re.S).group(1)
    snames = [n.strip() for _, group in re.findall(r"(double)\s+([\w\s,]+);", body) for n in group.split(",")]
    assert [f for f, _ in A.GkmStats._fields_] == snames


def test_gas_dof_matches_reference_formula():
    from gkm_amd.solver import GasParams
    assert GasParams(gamma=1.4).dof() == 2
    assert GasParams(gamma=5.0 / 3.0).dof() == 0


@pytest.mark.parametrize("n", [17, 24])
def test_dhit_initial_condition(n):
    from gkm_amd.ic import conservative_padded, dhit
    Mt, k0 = 0.3, 4.0
    ic = dhit(n, k0=k0, Mt=Mt, Re=30, seed=5)
    u, v, w = ic.u, ic.v, ic.w
    # solenoidal in the spectral sense
    k = np.fft.fftfreq(n, 1.0 / n)
    ky, kx, kz = k[:, None, None], k[None, :, None], k[None, None, :]
    div = kx * np.fft.fftn(u) + ky * np.fft.fftn(v) + kz * np.fft.fftn(w)
    assert np.max(np.abs(div)) < 1e-9 * n ** 3
    # zero mean (the k = 0 shell has E = 0)
    for f in (u, v, w):
        assert abs(f.mean()) < 1e-14
    # shell spectrum follows E(k) = sqrt(2/pi) (16/3) Mt^2 k^4 exp(-2k^2/k0^2)/k0^5 (init_BS.m:65)
    e = (np.abs(np.fft.fftn(u)) ** 2 + np.abs(np.fft.fftn(v)) ** 2 + np.abs(np.fft.fftn(w)) ** 2) / n ** 6
    kk = np.rint(np.sqrt(kx ** 2 + ky ** 2 + kz ** 2) * np.ones_like(e)).astype(int)
    spec = np.bincount(kk.ravel(), weights=e.ravel())
    for m in range(1, 6):
        target = math.sqrt(2 / math.pi) * 16 / 3 * Mt ** 2 * m ** 4 * math.exp(-2 * m * m / k0 ** 2) / k0 ** 5
        assert spec[m] == pytest.approx(target, rel=1e-10)
    # c0 = sqrt(3) u_rms / Mt, p0 = rho0 c0^2 / gamma, t0 = (2/k0)/u_rms (init_BS.m:146-157)
    q2 = np.mean(u * u + v * v + w * w)
    assert ic.urms == pytest.approx(math.sqrt(q2 / 3), rel=1e-14)
    assert math.sqrt(q2) / ic.c0 == pytest.approx(Mt, rel=1e-14)
    assert ic.p0 == pytest.approx(ic.c0 ** 2 / 1.4, rel=1e-14)
    assert ic.t0 == pytest.approx(0.5 / ic.urms, rel=1e-14)
    W = conservative_padded(ic)
    W3 = [x.reshape(n + 6, n + 6, n + 6)[3:-3, 3:-3, 3:-3] for x in W]
    p = 0.4 * (W3[4] - 0.5 * (W3[1] ** 2 + W3[2] ** 2 + W3[3] ** 2) / W3[0])
    assert np.allclose(p, ic.p0, rtol=1e-12)


def test_dhit_uneven_box():
    from gkm_amd.ic import dhit
    ic = dhit(12, ny=8, nz=16, seed=2)
    assert ic.u.shape == (8, 12, 16)
    assert np.all(np.isfinite(ic.u))


def _create(**changes):
    import gkm_amd._abi as A
    from gkm_amd.solver import GasParams
    L = A.load()
    grid = A.GkmGrid(8, 8, 8, 3, 0.1)
    p = GasParams().to_c()
    for k, v in changes.items():
        if hasattr(grid, k):
            setattr(grid, k, v)
        else:
            setattr(p, k, v)
    h = C.c_void_p()
    rc = L.gkm_create(C.byref(grid), C.byref(p), C.byref(h))
    return rc, L.gkm_last_error().decode(), h


@pytest.mark.parametrize("changes,word", [
    (dict(nx=2), "too small"), (dict(nz=0), "too small"), (dict(dx=0.0), "dx"), (dict(dx=float("nan")), "dx"),
    (dict(gamma=1.0), "gamma"), (dict(Pr=0.0), "Pr"), (dict(mu0=-1.0), "mu0"), (dict(T0=0.0), "T0"),
    (dict(K=-1), "K"), (dict(visc_law=7), "visc_law"), (dict(recon=5), "recon"),
    (dict(energy_recon=-1), "energy_recon"), (dict(precision=3), "precision"), (dict(flux_scheme=9), "flux_scheme"),
])
def test_parameter_checks_run_before_any_device_call(changes, word):
    """gkm_create validates the grid and the gas parameters on the host
    (GKM_E_ARG with a message naming the field) and leaves *out null."""
    import gkm_amd._abi as A
    rc, msg, h = _create(**changes)
    assert rc == A.GKM_E_ARG, (changes, rc, msg)
    assert word in msg, msg
    assert not h.value


def test_batch_checks_run_before_any_device_call():
    import gkm_amd._abi as A
    from gkm_amd.solver import GasParams
    L = A.load()
    grid = A.GkmGrid(8, 8, 8, 3, 0.1)
    ps = (A.GkmParams * 2)(GasParams().to_c(), GasParams(recon="js").to_c())
    h = C.c_void_p()
    assert L.gkm_create_batch(C.byref(grid), ps, 2, C.byref(h)) == A.GKM_E_ARG
    assert b"must match box 0" in L.gkm_last_error()
    assert L.gkm_create_batch(C.byref(grid), ps, 0, C.byref(h)) == A.GKM_E_ARG
    big = A.GkmGrid(1024, 1024, 1024, 3, 0.1)  # 1030^3 padded cells > 2^28
    assert L.gkm_create_batch(C.byref(big), ps, 1, C.byref(h)) == A.GKM_E_ARG
    assert b"2^28" in L.gkm_last_error()
    assert not h.value


def test_null_handles_are_rejected():
    """Every entry point taking a handle returns GKM_E_ARG for a null one
    (gkm_destroy(NULL) is a no-op, as free(NULL))."""
    import gkm_amd._abi as A
    L = A.load()
    d = C.c_double()
    n = C.c_int()
    st = A.GkmStats()
    assert L.gkm_destroy(None) == A.GKM_OK
    assert L.gkm_step(None, C.c_double(1e-3), 1) == A.GKM_E_ARG
    assert L.gkm_compute_dt(None, C.c_double(0.5), C.byref(d)) == A.GKM_E_ARG
    assert L.gkm_stats(None, C.byref(st)) == A.GKM_E_ARG
    assert L.gkm_fill_ghosts(None) == A.GKM_E_ARG
    assert L.gkm_sync(None) == A.GKM_E_ARG
    assert L.gkm_batch_size(None, C.byref(n)) == A.GKM_E_ARG
    assert L.gkm_set_graph(None, 1) == A.GKM_E_ARG
    assert L.gkm_set_timing(None, 1) == A.GKM_E_ARG
    assert L.gkm_set_tuning(None, 256, 8) == A.GKM_E_ARG
    assert L.gkm_set_tuning_z(None, 128, 16) == A.GKM_E_ARG


def test_tuning_cache_key_names_the_built_library():
    """The autotuner's cache key carries the library version and a hash of
    libgkm.so, so launch shapes tuned for one build are not reused by another."""
    import hashlib

    import gkm_amd._abi as A
    from gkm_amd import tune
    lid = tune.library_id()
    assert lid.startswith(A.load().gkm_version().decode() + "@")
    assert lid.endswith(hashlib.sha1(A.LIB_PATH.read_bytes()).hexdigest()[:16])
    assert tune._key(64, "mixed", "js", 4) == f"{lid}/64/mixed/js/4"
