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
