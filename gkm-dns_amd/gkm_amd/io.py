"""The reference's initial-condition and restart files, read and written by
libgkm.so's host code (include/gkm_io.h), plus an exact checkpoint of a
solver's state.

weno case folder (the reference's InitialConditions/):
    params                     GetParams, weno/src/FlowField.cpp:9-57
    Init%i%i%i                 one per rank: Initialize_DHIT (FlowField.cpp:59-112);
                               WriteReInit writes the same (update.cpp:1352-1391)
GKM:
    Init.dat                   init_BS.m:160-175 / Macro::Initialize_DHIT
                               (GKM/CUDA GKM/src/FlowField.cpp:178-235)

Arrays of the interior are indexed [j, i, k] = (y, x, z) like the library's
padded planes; the files hold them with j fastest (MATLAB order).
"""
from __future__ import annotations

import ctypes as C
import json
import math
from dataclasses import asdict, dataclass
from pathlib import Path

import numpy as np

from . import _abi

R_WENO = 287.0  # weno/src/param.h:18


@dataclass
class WenoParams:
    nc: int
    gam: float
    den0: float
    Mt: float
    Re: float
    mu0: float
    p0: float
    t0: float

    # globals GetParams derives (FlowField.cpp:40-49)
    def K(self) -> int:
        return int((5.0 - 3.0 * self.gam) / (self.gam - 1.0))

    def dx(self) -> float:
        return 2.0 * math.pi / self.nc

    def T0(self, R: float = R_WENO) -> float:
        return self.p0 / (R * self.den0)

    def to_c(self) -> _abi.GkmWenoParams:
        return _abi.GkmWenoParams(**asdict(self))


def _path(p) -> bytes:
    return str(p).encode()


def _pd(proc_dim):
    pd = (C.c_int * 3)(*[int(x) for x in proc_dim])
    return pd


def read_weno_params(path) -> WenoParams:
    c = _abi.GkmWenoParams()
    _abi.check(_abi.load().gkm_read_weno_params(_path(path), C.byref(c)))
    return WenoParams(nc=c.nc, gam=c.gam, den0=c.den0, Mt=c.Mt, Re=c.Re, mu0=c.mu0, p0=c.p0, t0=c.t0)


def write_weno_params(path, p: WenoParams) -> None:
    c = p.to_c()
    _abi.check(_abi.load().gkm_write_weno_params(_path(path), C.byref(c)))


def read_weno_init(case_dir, nc: int, gam: float, proc_dim=(1, 1, 1)):
    """Padded conservative planes (ghosts zero) from the rank files."""
    P = (nc + 6) ** 3
    W = [np.zeros(P) for _ in range(5)]
    _abi.check(_abi.load().gkm_read_weno_init(_path(case_dir), int(nc), _pd(proc_dim), float(gam),
                                              _abi.ptr_array(W)))
    return W


def write_weno_init(case_dir, W, nc: int, gam: float, proc_dim=(1, 1, 1)) -> None:
    W = [np.ascontiguousarray(np.asarray(w, dtype=np.float64).reshape(-1)) for w in W]
    Path(case_dir).mkdir(parents=True, exist_ok=True)
    _abi.check(_abi.load().gkm_write_weno_init(_path(case_dir), int(nc), _pd(proc_dim), float(gam),
                                               _abi.ptr_array(W)))


def write_weno_case(case_dir, ic, proc_dim=(1, 1, 1)) -> WenoParams:
    """A weno case folder (params + Init rank files) holding a DHIT initial
    condition (gkm_amd.ic.DhitIC on a cubic grid)."""
    from .ic import conservative_padded
    ny, nx, nz = ic.u.shape
    if not nx == ny == nz:
        raise ValueError("the weno params file describes a cubic box")
    p = WenoParams(nc=nx, gam=ic.gamma, den0=ic.rho0, Mt=ic.Mt, Re=ic.Re, mu0=ic.mu0, p0=ic.p0, t0=ic.t0)
    Path(case_dir).mkdir(parents=True, exist_ok=True)
    write_weno_params(Path(case_dir) / "params", p)
    write_weno_init(case_dir, conservative_padded(ic), nx, ic.gamma, proc_dim)
    return p


@dataclass
class InitDat:
    n: int
    rho0: float
    Mt: float
    Re: float
    k0: float
    urms: float
    mu: float
    crms: float
    p0: float
    eddy_time: float


def _to_file_order(a: np.ndarray) -> np.ndarray:
    # [j, i, k] -> flat with j fastest
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).transpose(2, 1, 0).reshape(-1))


def _from_file_order(a: np.ndarray, n: int) -> np.ndarray:
    return np.ascontiguousarray(a.reshape(n, n, n).transpose(2, 1, 0))


def read_init_dat(path):
    """(header, u, v, w) with u, v, w float32 arrays indexed [j, i, k]."""
    h = _abi.GkmInitDat()
    L = _abi.load()
    _abi.check(L.gkm_read_init_dat(_path(path), C.byref(h), None, None, None))
    n = h.n
    fp = C.POINTER(C.c_float)
    f = [np.empty(n ** 3, dtype=np.float32) for _ in range(3)]
    _abi.check(L.gkm_read_init_dat(_path(path), C.byref(h), *[x.ctypes.data_as(fp) for x in f]))
    hdr = InitDat(n=h.n, **{k: getattr(h, k) for k in ("rho0", "Mt", "Re", "k0", "urms", "mu", "crms", "p0",
                                                          "eddy_time")})
    return hdr, *[_from_file_order(x, n) for x in f]


def write_init_dat(path, hdr: InitDat, u, v, w) -> None:
    h = _abi.GkmInitDat(**asdict(hdr))
    fp = C.POINTER(C.c_float)
    f = [_to_file_order(x) for x in (u, v, w)]
    if any(x.size != hdr.n ** 3 for x in f):
        raise ValueError(f"velocity fields must hold {hdr.n}^3 values")
    _abi.check(_abi.load().gkm_write_init_dat(_path(path), C.byref(h), *[x.ctypes.data_as(fp) for x in f]))


def init_dat_from_ic(ic) -> tuple:
    """Init.dat header and FP32 fields of a DHIT IC, as init_BS.m writes them."""
    n = ic.u.shape[0]
    hdr = InitDat(n=n, rho0=ic.rho0, Mt=ic.Mt, Re=ic.Re, k0=ic.k0, urms=ic.urms, mu=ic.mu0, crms=ic.c0, p0=ic.p0,
                  eddy_time=ic.t0)
    return hdr, ic.u.astype(np.float32), ic.v.astype(np.float32), ic.w.astype(np.float32)


def init_dat_state(hdr: InitDat, u, v, w, gam: float = 1.4):
    """Padded FP64 conservative planes of Macro::Initialize_DHIT's conversion."""
    h = _abi.GkmInitDat(**asdict(hdr))
    fp = C.POINTER(C.c_float)
    f = [_to_file_order(x) for x in (u, v, w)]
    W = [np.zeros((hdr.n + 6) ** 3) for _ in range(5)]
    _abi.check(_abi.load().gkm_init_dat_state(C.byref(h), *[x.ctypes.data_as(fp) for x in f], float(gam),
                                              _abi.ptr_array(W)))
    return W


# -- exact checkpoint of a solver ------------------------------------------------------
def save_checkpoint(path, solver) -> None:
    """The padded FP64 state, grid, gas parameters and run clock in one .npz
    (plain arrays plus a JSON header; no pickled objects).  Restores the FP64
    form bit for bit; the mixed form to within one FP64 rounding of its
    hi + lo value."""
    W = solver.get_state()
    meta = {"nx": solver.nx, "ny": solver.ny, "nz": solver.nz, "dx": solver.dx, "gas": asdict(solver.gas),
            "t": solver.info.t, "t0": solver.info.t0, "dt": solver.info.dt, "steps": solver.info.steps,
            "norm": getattr(solver, "norm", None)}
    np.savez(path, W0=W[0], W1=W[1], W2=W[2], W3=W[3], W4=W[4], meta=np.frombuffer(json.dumps(meta).encode(),
                                                                                      dtype=np.uint8))


def load_checkpoint(path):
    """A new Solver holding the checkpointed state and clock."""
    from .solver import GasParams, Solver
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(bytes(z["meta"]).decode())
        W = [np.array(z[f"W{q}"]) for q in range(5)]
    s = Solver(meta["nx"], meta["ny"], meta["nz"], dx=meta["dx"], gas=GasParams(**meta["gas"]))
    s.set_state(W)
    s.info.t, s.info.t0, s.info.dt, s.info.steps = meta["t"], meta["t0"], meta["dt"], meta["steps"]
    if meta.get("norm"):
        s.norm = meta["norm"]
    return s
