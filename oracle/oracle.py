"""ctypes wrapper of the CPU oracle (liboracle.so) and of the reference flux
compiled from weno/src/GKMflux.cpp (libref_flux.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
the cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline, never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "_build" / "liboracle.so"
REF_SO = HERE / "_ref" / "libref_flux.so"
OPCOUNT_SO = HERE / "_build" / "libopcount.so"

DP = C.POINTER(C.c_double)


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("ghost", C.c_int), ("dx", C.c_double)]


class Params(C.Structure):
    _fields_ = [
        ("gamma", C.c_double), ("R", C.c_double), ("Pr", C.c_double), ("mu0", C.c_double),
        ("T0", C.c_double), ("visc_exp", C.c_double), ("K", C.c_int), ("visc_law", C.c_int),
        ("recon", C.c_int), ("energy_recon", C.c_int), ("zbug_compat", C.c_int), ("precision", C.c_int),
        ("flux_scheme", C.c_int),
    ]


STAT_FIELDS = [
    "tke", "ek", "q2", "mt", "rho_mean", "rho_rms", "p_mean", "p_rms", "T_mean", "T_rms",
    "v_mean", "v_rms", "c_mean", "umax_plus_c",
    "dudx2", "dudx3", "dudx4", "dvdy2", "dvdy3", "dvdy4", "dwdz2", "dwdz3", "dwdz4",
    "skew_u", "skew_v", "skew_w", "kurt_u", "kurt_v", "kurt_w",
    "enstrophy", "dil2", "diss_sol", "diss_dil", "diss_ref", "sijsij",
    "mass", "mom_x", "mom_y", "mom_z", "energy",
]


class Stats(C.Structure):
    _fields_ = [(f, C.c_double) for f in STAT_FIELDS]

    def as_dict(self):
        return {f: getattr(self, f) for f in STAT_FIELDS}


PRINT_STAT_FIELDS = [
    "tke", "diss", "sijsij", "skew_u", "skew_v", "skew_w", "kurt_u", "kurt_v", "kurt_w",
    "p_mean", "T_mean", "v_mean", "p_rms", "T_rms", "v_rms", "dudx2", "dvdy2", "dwdz2",
]


class PrintStats(C.Structure):
    _fields_ = [(f, C.c_double) for f in PRINT_STAT_FIELDS]

    def as_dict(self):
        return {f: getattr(self, f) for f in PRINT_STAT_FIELDS}


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists():
            raise RuntimeError(f"oracle not built: {ORACLE_SO} (run `make -C oracle`)")
        L = C.CDLL(str(ORACLE_SO))
        L.orc_flux.argtypes = [C.POINTER(Params)] + [DP] * 10 + [C.c_double] * 3 + [DP]
        L.orc_gks_flux.argtypes = [C.POINTER(Params), DP, DP, DP, DP, C.c_double, C.c_double, DP]
        L.orc_weno.argtypes = [C.POINTER(Params), DP, DP, DP]
        L.orc_weno_scalar.argtypes = [C.c_int, DP, DP, DP]
        L.orc_moment_tables.argtypes = [C.c_double, C.c_double, C.c_int, DP, DP, DP, DP, DP]
        L.orc_slopesolver.argtypes = [C.POINTER(Params), DP, DP, C.c_double, DP]
        L.orc_fill_ghosts.argtypes = [C.POINTER(Grid), C.POINTER(DP)]
        L.orc_step.argtypes = [C.POINTER(Grid), C.POINTER(Params), C.POINTER(DP), C.c_double, C.c_int, C.c_int]
        L.orc_fluxes.argtypes = [C.POINTER(Grid), C.POINTER(Params), C.POINTER(DP), C.c_double, C.POINTER(DP),
                                 C.c_int]
        L.orc_timestep.argtypes = [C.POINTER(Grid), C.POINTER(Params), C.POINTER(DP), C.c_double]
        L.orc_timestep.restype = C.c_double
        L.orc_stats.argtypes = [C.POINTER(Grid), C.POINTER(Params), C.POINTER(DP), C.POINTER(Stats)]
        L.orc_spline_slopes.argtypes = [DP, C.c_long, C.c_int, C.c_double, DP]
        L.orc_print_stats.argtypes = [C.POINTER(Grid), C.POINTER(Params), C.POINTER(DP), C.POINTER(PrintStats)]
        _lib = L
    return _lib


def ref_available() -> bool:
    return REF_SO.exists()


def ref_lib():
    global _ref
    if _ref is None:
        if not REF_SO.exists():
            raise RuntimeError(f"reference flux not built: {REF_SO} (run `make -C oracle ref`)")
        L = C.CDLL(str(REF_SO))
        L.ref_flux_points.argtypes = [C.c_int, C.c_double, C.c_int, DP, C.c_double, DP]
        _ref = L
    return _ref


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(DP)


def _ptrs(W):
    arr = (DP * len(W))()
    for q, w in enumerate(W):
        assert w.dtype == np.float64 and w.flags.c_contiguous
        arr[q] = w.ctypes.data_as(DP)
    return arr


def make_grid(nx, ny, nz, dx):
    return Grid(nx, ny, nz, 3, dx)


def make_params(**kw) -> Params:
    from_dict = dict(gamma=1.4, R=287.0, Pr=0.7, mu0=1e-3, T0=1.0, visc_exp=0.76, K=2, visc_law=0,
                     recon=0, energy_recon=0, zbug_compat=1, precision=0, flux_scheme=0)
    from_dict.update(kw)
    return Params(**from_dict)


def flux_points(params: Params, rec: np.ndarray, dx: float) -> np.ndarray:
    """rec: (n, 52) records WL WR Wl Wr DWxl DWxr DWyl DWyr DWzl DWzr tau dt."""
    rec = np.ascontiguousarray(rec, dtype=np.float64)
    n = rec.shape[0]
    out = np.zeros((n, 5))
    L = lib()
    for r in range(n):
        x = rec[r]
        v = [np.ascontiguousarray(x[5 * a:5 * a + 5]) for a in range(10)]
        F = np.zeros(5)
        L.orc_flux(C.byref(params), *[_p(t) for t in v], float(x[50]), float(x[51]), float(dx), _p(F))
        out[r] = F
    return out


def gks_flux_points(params: Params, rec: np.ndarray) -> np.ndarray:
    """rec: (n, 22) records W0 DWx DWy DWz mu dt (rotated slot order)."""
    rec = np.ascontiguousarray(rec, dtype=np.float64)
    out = np.zeros((rec.shape[0], 5))
    L = lib()
    for r in range(rec.shape[0]):
        x = rec[r]
        v = [np.ascontiguousarray(x[5 * a:5 * a + 5]) for a in range(4)]
        F = np.zeros(5)
        L.orc_gks_flux(C.byref(params), *[_p(t) for t in v], float(x[20]), float(x[21]), _p(F))
        out[r] = F
    return out


def ref_flux_points(params: Params, rec: np.ndarray, dx: float) -> np.ndarray:
    rec = np.ascontiguousarray(rec, dtype=np.float64)
    n = rec.shape[0]
    out = np.zeros((n, 5))
    ref_lib().ref_flux_points(int(params.K), float(params.gamma), n, _p(rec), float(dx), _p(out))
    return out


def weno(params: Params, cells: np.ndarray):
    """cells: (6, 5) -> (Wl[5], Wr[5])."""
    cells = np.ascontiguousarray(cells, dtype=np.float64)
    wl = np.zeros(5)
    wr = np.zeros(5)
    lib().orc_weno(C.byref(params), _p(cells), _p(wl), _p(wr))
    return wl, wr


def weno_scalar(recon: int, c) -> tuple[float, float]:
    c = np.ascontiguousarray(c, dtype=np.float64)
    a = C.c_double()
    b = C.c_double()
    lib().orc_weno_scalar(int(recon), _p(c), C.byref(a), C.byref(b))
    return a.value, b.value


def moment_tables(U, lam, K=2):
    hp, hn, fu = np.zeros(7), np.zeros(7), np.zeros(7)
    e2, e4 = C.c_double(), C.c_double()
    lib().orc_moment_tables(float(U), float(lam), int(K), _p(hp), _p(hn), _p(fu), C.byref(e2), C.byref(e4))
    return hp, hn, fu, e2.value, e4.value


def slopesolver(params: Params, b, U, lam):
    b = np.ascontiguousarray(b, dtype=np.float64)
    U = np.ascontiguousarray(U, dtype=np.float64)
    a = np.zeros(5)
    lib().orc_slopesolver(C.byref(params), _p(b), _p(U), float(lam), _p(a))
    return a


def fill_ghosts(grid: Grid, W):
    lib().orc_fill_ghosts(C.byref(grid), _ptrs(W))


def step(grid: Grid, params: Params, W, dt: float, nsteps: int, nthreads: int = 1):
    lib().orc_step(C.byref(grid), C.byref(params), _ptrs(W), float(dt), int(nsteps), int(nthreads))


def fluxes(grid: Grid, params: Params, W, dt: float, nthreads: int = 1):
    n = W[0].size
    F = [np.zeros(n) for _ in range(15)]
    lib().orc_fluxes(C.byref(grid), C.byref(params), _ptrs(W), float(dt), _ptrs(F), int(nthreads))
    return F


def timestep(grid: Grid, params: Params, W, cfl: float) -> float:
    return lib().orc_timestep(C.byref(grid), C.byref(params), _ptrs(W), float(cfl))


def stats(grid: Grid, params: Params, W) -> dict:
    s = Stats()
    lib().orc_stats(C.byref(grid), C.byref(params), _ptrs(W), C.byref(s))
    return s.as_dict()


def spline_slopes(f, h: float) -> np.ndarray:
    """Natural cubic spline slopes at the nodes of a uniformly spaced line."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    D = np.zeros(f.size)
    lib().orc_spline_slopes(_p(f), 1, int(f.size), float(h), _p(D))
    return D


def print_stats(grid: Grid, params: Params, W) -> dict:
    s = PrintStats()
    lib().orc_print_stats(C.byref(grid), C.byref(params), _ptrs(W), C.byref(s))
    return s.as_dict()


OPCOUNT_FIELDS = ("add", "mul", "div", "sqrt", "exp", "erfc", "pow")


def opcount(params: Params) -> dict:
    """Operation counts of the oracle (oracle/opcount.cpp) for one interface's
    WENO, one interface's flux and one cell's update:
    {"weno": {...}, "flux": {...}, "update": {...}} with add (incl. sub), mul,
    div, sqrt (the arithmetic flops) and exp, erfc, pow (transcendentals)."""
    if not OPCOUNT_SO.exists():
        raise RuntimeError(f"op counter not built: {OPCOUNT_SO} (run `make -C oracle`)")
    L = C.CDLL(str(OPCOUNT_SO))
    L.orc_opcount.argtypes = [C.POINTER(Params), C.POINTER(C.c_ulonglong)]
    out = (C.c_ulonglong * 21)()
    L.orc_opcount(C.byref(params), out)
    return {part: dict(zip(OPCOUNT_FIELDS, out[7 * k:7 * k + 7]))
            for k, part in enumerate(("weno", "flux", "update"))}


def flops_per_cell_update(params: Params, n: int | None = None) -> dict:
    """Algorithmic work per cell-update of the scheme as the oracle states it:
    three sweeps of (WENO per interface + flux per interface) plus the update.
    With n, the interface counts of an n^3 box (WENO on the extended set
    (n+1)(n+2)^2, flux on the core set (n+1) n^2, per sweep); without, the
    large-box limit of one of each per cell and sweep."""
    c = opcount(params)
    arith = {k: c[k]["add"] + c[k]["mul"] + c[k]["div"] + c[k]["sqrt"] for k in c}
    trans = {k: c[k]["exp"] + c[k]["erfc"] + c[k]["pow"] for k in c}
    if n is None:
        we = fl = 1.0
    else:
        we = (n + 1) * (n + 2) ** 2 / n ** 3
        fl = (n + 1) / n
    return {"flops": 3 * (we * arith["weno"] + fl * arith["flux"]) + arith["update"],
            "transcendentals": 3 * (we * trans["weno"] + fl * trans["flux"]) + trans["update"],
            "per_part": c}


def build(ref: bool = True) -> None:
    """Compile liboracle.so (and the reference flux when its sources exist)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    if ref and Path("/root/reference/weno/src/GKMflux.cpp").exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)


if __name__ == "__main__":  # pragma: no cover
    build()
    print("oracle:", ORACLE_SO, "ref:", REF_SO if REF_SO.exists() else None, os.getpid())
