"""Host-side mirror of the reference solver interface over libgkm.so.

The reference drives one rank's segment through free functions and globals
(weno/src/Main.cpp:10-49, FlowField.cpp, update.cpp):

    GetParams()            -> Solver(...) / Solver.from_dhit(...)
    Initialize_DHIT(W)     -> Solver.initialize_dhit(ic) / Solver.set_state(W)
    GetParams + Initialize_DHIT from a case folder -> Solver.from_weno_case(dir)
    WriteReInit(W)         -> Solver.write_reinit(dir)
    PeriodicBC(W)          -> Solver.periodic_bc()
    timestep(W)            -> Solver.timestep(cfl)            (sets Solver.dt)
    evolveEulerian(W, dev) -> Solver.evolve_eulerian(nsteps)  (one step each)
    PrintStats(W, splines) -> Solver.print_stats()            (stats dict + line)
    Evolve(W, comm3d)      -> Solver.evolve(t_end)

The state lives on the GPU between calls; ``get_state`` copies it back in the
reference layout (five padded FP64 planes, Is(j, i, k)).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from .ic import DhitIC, conservative_padded, dhit


@dataclass
class GasParams:
    gamma: float = 1.4
    R: float = 287.0          # weno/src/param.h:18
    Pr: float = 0.7           # applied in the slope solver as GKMflux.cpp:289 does
    mu0: float = 1e-3
    T0: float = 1.0
    visc_law: str = "power"   # "power" (mu ~ T^visc_exp) or "sutherland"
    visc_exp: float = 0.76
    recon: str = "symoo"      # "symoo" (reference) or "js"
    energy_recon: str = "T"   # "T" (CPU path) or "E" (GPU path)
    zbug_compat: bool = True  # reference Z-sweep tangential pairing
    precision: str = "fp64"   # "fp64" or "mixed"
    K: int | None = None      # internal DOF; default int((5 - 3 gamma)/(gamma - 1))
    flux_scheme: str = "weno_bgk"  # "weno_bgk" (weno tree) or "gks" (GKM tree smooth-flow scheme)

    def dof(self) -> int:
        if self.K is not None:
            return int(self.K)
        return int((5.0 - 3.0 * self.gamma) / (self.gamma - 1.0))  # FlowField.cpp:41

    def to_c(self) -> _abi.GkmParams:
        return _abi.GkmParams(
            gamma=self.gamma, R=self.R, Pr=self.Pr, mu0=self.mu0, T0=self.T0, visc_exp=self.visc_exp,
            K=self.dof(), visc_law=_abi.VISC[self.visc_law], recon=_abi.RECON[self.recon],
            energy_recon=_abi.ENERGY[self.energy_recon], zbug_compat=int(bool(self.zbug_compat)),
            precision=_abi.PRECISION[self.precision], flux_scheme=_abi.SCHEME[self.flux_scheme])


@dataclass
class RunInfo:
    t: float = 0.0
    t0: float = 1.0      # eddy turnover time used to print t/t0
    dt: float = 0.0
    steps: int = 0
    history: list = field(default_factory=list)


class Solver:
    """One periodic box (nx, ny, nz interior cells, 3 ghost layers) on one GPU."""

    def __init__(self, nx: int, ny: int | None = None, nz: int | None = None, dx: float | None = None,
                 gas: GasParams | None = None):
        self.nx = int(nx)
        self.ny = int(nx if ny is None else ny)
        self.nz = int(nx if nz is None else nz)
        # dx = 2 pi / nc (FlowField.cpp:45); for uneven boxes the x extent sets it
        self.dx = float(dx) if dx is not None else 2.0 * math.pi / self.nx
        self.gas = gas or GasParams()
        self.shape = (self.ny + 6, self.nx + 6, self.nz + 6)
        self.ncell = int(np.prod(self.shape))
        self.info = RunInfo()
        self._L = _abi.load()
        self._h = C.c_void_p()
        grid = _abi.GkmGrid(self.nx, self.ny, self.nz, 3, self.dx)
        self._params = self.gas.to_c()
        _abi.check(self._L.gkm_create(C.byref(grid), C.byref(self._params), C.byref(self._h)))

    # -- construction -----------------------------------------------------------------
    @classmethod
    def from_dhit(cls, n: int, k0: float = 4.0, Mt: float = 0.3, Re: float = 30.0, seed: int = 1234,
                  ny: int | None = None, nz: int | None = None, gamma: float = 1.4, **gas_kw) -> "Solver":
        """Box of decaying isotropic turbulence with the init_BS.m recipe."""
        ic = dhit(n, k0=k0, Mt=Mt, Re=Re, seed=seed, gamma=gamma, ny=ny, nz=nz)
        return cls.from_ic(ic, **gas_kw)

    @classmethod
    def from_ic(cls, ic: DhitIC, precision: str = "fp64", visc_law: str = "power", recon: str = "symoo",
                energy_recon: str = "T", zbug_compat: bool = True, Pr: float = 0.7, R: float = 287.0,
                visc_exp: float = 0.76, flux_scheme: str = "weno_bgk") -> "Solver":
        """Solver on the IC's grid with mu0 from the IC and T0 = p0/(R rho0)
        (weno/src/FlowField.cpp:46)."""
        ny, nx, nz = ic.u.shape
        gas = GasParams(gamma=ic.gamma, R=R, Pr=Pr, mu0=ic.mu0, T0=ic.p0 / (R * ic.rho0), visc_law=visc_law,
                        visc_exp=visc_exp, recon=recon, energy_recon=energy_recon, zbug_compat=zbug_compat,
                        precision=precision, flux_scheme=flux_scheme)
        s = cls(nx, ny, nz, gas=gas)
        s.initialize_dhit(ic)
        return s

    def init_dhit(self, k0: float = 4.0, Mt: float = 0.3, Re: float = 30.0, seed: int = 1234,
                  rho0: float = 1.0) -> dict:
        """GetParams + Initialize_DHIT in the library (gkm_init_dhit): the
        init_BS.m recipe generated by libgkm on the host, written into the box,
        ghosts filled, and the box's mu0 and T0 set from it.  Returns the
        derived scales (q2, urms, c0, p0, mu0, t0)."""
        icp = _abi.GkmIc(int(seed) & 0xFFFFFFFFFFFFFFFF, float(k0), float(Mt), float(Re), float(rho0))
        info = _abi.GkmIcInfo()
        _abi.check(self._L.gkm_init_dhit(self._h, C.byref(icp), C.byref(info)))
        d = info.as_dict()
        self.gas.mu0 = d["mu0"]
        self.gas.T0 = d["p0"] / (self.gas.R * rho0)
        self._params = self.gas.to_c()
        self.info.t0 = d["t0"]
        self.norm = {"Mt": Mt, "p0": d["p0"], "rho0": rho0}
        return d

    def initialize_dhit(self, ic: DhitIC) -> None:
        """Initialize_DHIT (FlowField.cpp:59-112) followed by the first halo fill."""
        self.ic = ic
        self.info.t0 = ic.t0
        self.norm = {"Mt": ic.Mt, "p0": ic.p0, "rho0": ic.rho0}
        self.set_state(conservative_padded(ic))
        self.periodic_bc()

    @classmethod
    def from_weno_case(cls, case_dir, proc_dim=(1, 1, 1), precision: str = "fp64", visc_law: str = "sutherland",
                       recon: str = "symoo", energy_recon: str = "T", zbug_compat: bool = True, Pr: float = 0.7,
                       R: float = 287.0, visc_exp: float = 0.76) -> "Solver":
        """GetParams + Initialize_DHIT from a reference case folder (params and
        the Init%i%i%i rank files of a proc_dim decomposition); defaults are
        the weno code's (Sutherland, SYMOO, T reconstruction, Pr 0.7)."""
        from pathlib import Path

        from . import io
        p = io.read_weno_params(Path(case_dir) / "params")
        gas = GasParams(gamma=p.gam, R=R, Pr=Pr, mu0=p.mu0, T0=p.T0(R), visc_law=visc_law, visc_exp=visc_exp,
                        recon=recon, energy_recon=energy_recon, zbug_compat=zbug_compat, precision=precision,
                        K=p.K())
        s = cls(p.nc, p.nc, p.nc, dx=p.dx(), gas=gas)
        s.set_state(io.read_weno_init(case_dir, p.nc, p.gam, proc_dim))
        s.periodic_bc()
        s.info.t0 = p.t0
        s.norm = {"Mt": p.Mt, "p0": p.p0, "rho0": p.den0}
        s.case = p
        return s

    def write_reinit(self, case_dir, proc_dim=(1, 1, 1)) -> None:
        """WriteReInit (update.cpp:1352-1391): Init%i%i%i rank files of the
        current state (a cubic box), readable by from_weno_case."""
        from . import io
        if not self.nx == self.ny == self.nz:
            raise ValueError("the rank-file format describes a cubic box")
        io.write_weno_init(case_dir, self.get_state(), self.nx, self.gas.gamma, proc_dim)

    def save_checkpoint(self, path) -> None:
        """Exact checkpoint (.npz): state, grid, gas parameters, clock."""
        from . import io
        io.save_checkpoint(path, self)

    @classmethod
    def load_checkpoint(cls, path) -> "Solver":
        from . import io
        return io.load_checkpoint(path)

    # -- state ------------------------------------------------------------------------
    def set_state(self, W) -> None:
        W = [np.ascontiguousarray(np.asarray(w, dtype=np.float64).reshape(-1)) for w in W]
        if len(W) != 5 or any(w.size != self.ncell for w in W):
            raise ValueError(f"expected 5 planes of {self.ncell} values")
        _abi.check(self._L.gkm_set_state(self._h, _abi.ptr_array(W)))

    def get_state(self):
        W = [np.empty(self.ncell) for _ in range(5)]
        _abi.check(self._L.gkm_get_state(self._h, _abi.ptr_array(W)))
        return W

    def periodic_bc(self) -> None:
        _abi.check(self._L.gkm_fill_ghosts(self._h))

    # -- stepping ---------------------------------------------------------------------
    def timestep(self, cfl: float = 0.02) -> float:
        """dt = cfl dx / max(|u| + c) (update.cpp:1057-1079); cfl default as param.h:17."""
        dt = C.c_double()
        _abi.check(self._L.gkm_compute_dt(self._h, float(cfl), C.byref(dt)))
        self.info.dt = dt.value
        return dt.value

    def evolve_eulerian(self, nsteps: int = 1, dt: float | None = None) -> None:
        dt = self.info.dt if dt is None else dt
        if not dt > 0:
            raise ValueError("dt is not set: call timestep() or pass dt")
        _abi.check(self._L.gkm_step(self._h, float(dt), int(nsteps)))
        self.info.t += nsteps * dt
        self.info.steps += nsteps

    def sync(self) -> None:
        _abi.check(self._L.gkm_sync(self._h))

    def stats(self) -> dict:
        s = _abi.GkmStats()
        _abi.check(self._L.gkm_stats(self._h, C.byref(s)))
        return s.as_dict()

    def _line(self, s: dict, t: float, Mt=None, p0=None, T0=None, rho0=None) -> str:
        norm = getattr(self, "norm", None) or {}
        Mt = Mt if Mt is not None else norm.get("Mt", 1.0)
        p0 = p0 if p0 is not None else norm.get("p0", s["p_mean"])
        rho0 = rho0 if rho0 is not None else norm.get("rho0", 1.0)
        T0 = T0 if T0 is not None else self.gas.T0
        return _abi.format_print_stats(s, t / self.info.t0, self.gas.gamma, p0, T0, Mt, rho0)

    def step_print_stats(self, nsteps: int, dt: float | None = None) -> list:
        """nsteps of the Evolve loop body, PrintStats then the step, with the
        statistics of each state overlapping the step that reads it
        (gkm_step_print_stats).  Returns the statistics dicts (with their
        "line") of the states before each step."""
        dt = self.info.dt if dt is None else dt
        if not dt > 0:
            raise ValueError("dt is not set: call timestep() or pass dt")
        out = (_abi.GkmPrintStats * max(1, int(nsteps)))()
        _abi.check(self._L.gkm_step_print_stats(self._h, float(dt), int(nsteps), out))
        res = []
        for k in range(int(nsteps)):
            s = out[k].as_dict()
            s["line"] = self._line(s, self.info.t)
            res.append(s)
            self.info.t += dt
            self.info.steps += 1
        return res

    def print_stats(self, Mt: float | None = None, p0: float | None = None, T0: float | None = None,
                    rho0: float | None = None, file=None) -> dict:
        """PrintStats (update.cpp:638-784): natural-spline velocity gradients
        over the padded box, the reference's sums, and its output line
        t/t0 TKE Diss SKEWu SKEWv SKEWw KURTu KURTv KURTw Prms Trms Vrms
        (gkm_print_stats + gkm_format_print_stats).  Mt, p0, den0 default to
        the initial condition's or case's (the reference's params globals)."""
        st = _abi.GkmPrintStats()
        _abi.check(self._L.gkm_print_stats(self._h, C.byref(st)))
        s = st.as_dict()
        s["line"] = self._line(s, self.info.t, Mt, p0, T0, rho0)
        if file is not None:
            print(s["line"], file=file)
        return s

    def evolve(self, t_end: float, cfl: float = 0.02, stats_every: int = 1, recompute_dt: bool = False,
               callback=None) -> list:
        """The Evolve loop (update.cpp:106-150): dt fixed from the initial
        state unless recompute_dt; statistics before each step; t += dt."""
        self.periodic_bc()
        if self.info.dt <= 0 or recompute_dt:
            self.timestep(cfl)
        if stats_every == 1 and not recompute_dt and callback is None:
            # statistics before every step at a fixed dt: the overlapped loop body
            n, t = 0, self.info.t
            while t <= t_end:
                n += 1
                t += self.info.dt
            t = self.info.t
            for s in self.step_print_stats(n):
                self.info.history.append((t, s))
                t += self.info.dt
            s = self.print_stats()
            self.info.history.append((self.info.t, s))
            return self.info.history
        while self.info.t <= t_end:
            if self.info.steps % stats_every == 0:
                s = self.print_stats()
                self.info.history.append((self.info.t, s))
                if callback:
                    callback(self, s)
            self.evolve_eulerian(1)
            if recompute_dt:
                self.timestep(cfl)
        s = self.print_stats()
        self.info.history.append((self.info.t, s))
        return self.info.history

    # -- launch options ---------------------------------------------------------------
    def set_graph(self, on: bool = True) -> None:
        """Capture one step in a hipGraph and replay it (gkm_set_graph)."""
        _abi.check(self._L.gkm_set_graph(self._h, int(on)))

    def set_tuning(self, flux_block: int = 256, recon_chunk: int = 8, zflux_block: int | None = None,
                   z_strip: int | None = None) -> None:
        """Launch shapes (gkm_set_tuning; with zflux_block and z_strip also
        gkm_set_tuning_z).  Results never depend on them."""
        _abi.check(self._L.gkm_set_tuning(self._h, int(flux_block), int(recon_chunk)))
        if zflux_block is not None:
            _abi.check(self._L.gkm_set_tuning_z(self._h, int(zflux_block), int(z_strip or 16)))

    def set_halo_mode(self, mode: str = "gather") -> None:
        """"gather" (periodic image per ghost cell) or "pack" (y-face pack /
        exchange / unpack, then x/z ghosts; the slab-exchange path)."""
        _abi.check(self._L.gkm_set_halo_mode(self._h, _abi.HALO[mode]))

    # -- timing -----------------------------------------------------------------------
    def set_timing(self, on: bool | str = True) -> None:
        """Device timing of the following steps: True or "all" (an event around
        every launch), "flux" (events around the X and Y flux kernels only, the
        least perturbing), False or "off"."""
        mode = {"off": 0, "all": 1, "flux": 2}[on] if isinstance(on, str) else int(bool(on))
        _abi.check(self._L.gkm_set_timing(self._h, mode))

    def last_timing(self):
        ms = (C.c_double * 5)()
        n = C.c_int()
        _abi.check(self._L.gkm_last_timing(self._h, ms, C.byref(n)))
        sw = (C.c_double * 3)()
        _abi.check(self._L.gkm_last_sweep_timing(self._h, sw))
        return {"flux_ms": ms[0], "recon_ms": ms[1], "ghost_ms": ms[2], "update_ms": ms[3], "prim_ms": ms[4],
                "flux_x_ms": sw[0], "flux_y_ms": sw[1], "flux_z_ms": sw[2], "steps": n.value}

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._L.gkm_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Batch:
    """nbox independent boxes of one grid advanced by the same launches
    (gkm_create_batch).  Each box has its own gas parameters and dt."""

    def __init__(self, nx: int, gases, ny: int | None = None, nz: int | None = None, dx: float | None = None):
        self.nx = int(nx)
        self.ny = int(nx if ny is None else ny)
        self.nz = int(nx if nz is None else nz)
        self.dx = float(dx) if dx is not None else 2.0 * math.pi / self.nx
        self.gases = list(gases)
        self.nbox = len(self.gases)
        self.shape = (self.ny + 6, self.nx + 6, self.nz + 6)
        self.ncell = int(np.prod(self.shape))
        self.dt = np.zeros(self.nbox)
        self.t = np.zeros(self.nbox)
        self.ics = [None] * self.nbox
        self._L = _abi.load()
        self._h = C.c_void_p()
        grid = _abi.GkmGrid(self.nx, self.ny, self.nz, 3, self.dx)
        arr = (_abi.GkmParams * self.nbox)(*[g.to_c() for g in self.gases])
        _abi.check(self._L.gkm_create_batch(C.byref(grid), arr, self.nbox, C.byref(self._h)))

    @classmethod
    def from_ics(cls, ics, **gas_kw) -> "Batch":
        """One box per DHIT initial condition (same grid), mu0 and T0 from each IC."""
        R = gas_kw.pop("R", 287.0)
        gases = [GasParams(gamma=ic.gamma, R=R, mu0=ic.mu0, T0=ic.p0 / (R * ic.rho0), **gas_kw) for ic in ics]
        ny, nx, nz = ics[0].u.shape
        b = cls(nx, gases, ny, nz)
        for i, ic in enumerate(ics):
            b.ics[i] = ic
            b.set_state(i, conservative_padded(ic))
        return b

    def set_state(self, box: int, W) -> None:
        W = [np.ascontiguousarray(np.asarray(w, dtype=np.float64).reshape(-1)) for w in W]
        if len(W) != 5 or any(w.size != self.ncell for w in W):
            raise ValueError(f"expected 5 planes of {self.ncell} values")
        _abi.check(self._L.gkm_set_state_box(self._h, int(box), _abi.ptr_array(W)))

    def get_state(self, box: int):
        W = [np.empty(self.ncell) for _ in range(5)]
        _abi.check(self._L.gkm_get_state_box(self._h, int(box), _abi.ptr_array(W)))
        return W

    def periodic_bc(self) -> None:
        _abi.check(self._L.gkm_fill_ghosts(self._h))

    def timestep(self, cfl: float) -> np.ndarray:
        self.periodic_bc()
        for b in range(self.nbox):
            d = C.c_double()
            _abi.check(self._L.gkm_compute_dt_box(self._h, b, float(cfl), C.byref(d)))
            self.dt[b] = d.value
        return self.dt.copy()

    def step(self, nsteps: int = 1, dt=None) -> None:
        dt = self.dt if dt is None else np.asarray(dt, dtype=np.float64)
        dt = np.ascontiguousarray(np.broadcast_to(dt, (self.nbox,)), dtype=np.float64)
        _abi.check(self._L.gkm_step_batch(self._h, dt.ctypes.data_as(_abi.DP), int(nsteps)))
        self.t += nsteps * dt

    def set_dt(self, dt=None) -> None:
        """dt per box for later steps, without stepping (gkm_set_dt_batch)."""
        if dt is not None:
            self.dt[:] = np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.nbox,))
        d = np.ascontiguousarray(self.dt, dtype=np.float64)
        _abi.check(self._L.gkm_set_dt_batch(self._h, d.ctypes.data_as(_abi.DP)))

    def stats(self, box: int) -> dict:
        s = _abi.GkmStats()
        _abi.check(self._L.gkm_stats_box(self._h, int(box), C.byref(s)))
        return s.as_dict()

    def print_stats(self, box: int) -> dict:
        """PrintStats quantities of one box (gkm_print_stats_box)."""
        s = _abi.GkmPrintStats()
        _abi.check(self._L.gkm_print_stats_box(self._h, int(box), C.byref(s)))
        return s.as_dict()

    def sync(self) -> None:
        _abi.check(self._L.gkm_sync(self._h))

    def set_graph(self, on: bool = True) -> None:
        _abi.check(self._L.gkm_set_graph(self._h, int(on)))

    def set_tuning(self, flux_block: int = 256, recon_chunk: int = 8, zflux_block: int | None = None,
                   z_strip: int | None = None) -> None:
        _abi.check(self._L.gkm_set_tuning(self._h, int(flux_block), int(recon_chunk)))
        if zflux_block is not None:
            _abi.check(self._L.gkm_set_tuning_z(self._h, int(zflux_block), int(z_strip or 16)))

    def set_halo_mode(self, mode: str = "gather") -> None:
        _abi.check(self._L.gkm_set_halo_mode(self._h, _abi.HALO[mode]))

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._L.gkm_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Group:
    """Handles stepped together by one captured graph per step (gkm_group_*):
    every box of every member (Solver or Batch) advances with one graph
    launch, each member with the dt it was last given."""

    def __init__(self, members):
        self.members = list(members)
        self._L = _abi.load()
        self._g = C.c_void_p()
        arr = (C.c_void_p * len(self.members))(*[m._h.value for m in self.members])
        _abi.check(self._L.gkm_group_create(arr, len(self.members), C.byref(self._g)))

    def step(self, nsteps: int = 1, persistent: bool = False) -> None:
        """nsteps steps of every member: one captured graph launch per step
        (gkm_group_step), or with persistent=True one persistent cooperative
        launch for all of them (gkm_group_step_persistent; same results)."""
        for m in self.members:  # each member's current dt (uploaded only when it changed)
            d = np.ascontiguousarray(m.dt if isinstance(m, Batch) else [m.info.dt], dtype=np.float64)
            _abi.check(self._L.gkm_set_dt_batch(m._h, d.ctypes.data_as(_abi.DP)))
        if persistent:
            _abi.check(self._L.gkm_group_step_persistent(self._g, int(nsteps)))
        else:
            _abi.check(self._L.gkm_group_step(self._g, int(nsteps)))
        for m in self.members:
            if isinstance(m, Batch):
                m.t += nsteps * m.dt
            else:
                m.info.t += nsteps * m.info.dt
                m.info.steps += nsteps

    def sync(self) -> None:
        _abi.check(self._L.gkm_group_sync(self._g))

    def close(self) -> None:
        if getattr(self, "_g", None) and self._g.value:
            self._L.gkm_group_destroy(self._g)
            self._g = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def flux_points(gas: GasParams, rec: np.ndarray, dx: float) -> np.ndarray:
    """Device BGK flux on n records (n, 52) -> (n, 5) (gkm_flux_points)."""
    L = _abi.load()
    rec = np.ascontiguousarray(rec, dtype=np.float64)
    out = np.zeros((rec.shape[0], 5))
    p = gas.to_c()
    _abi.check(L.gkm_flux_points(C.byref(p), rec.shape[0], rec.ctypes.data_as(_abi.DP), float(dx),
                                 out.ctypes.data_as(_abi.DP)))
    return out


def gks_points(gas: GasParams, rec: np.ndarray) -> np.ndarray:
    """Device GKS flux on (n, 22) records W0 DWx DWy DWz mu dt -> (n, 5)."""
    rec = np.ascontiguousarray(rec, dtype=np.float64)
    out = np.zeros((rec.shape[0], 5))
    p = gas.to_c()
    _abi.check(_abi.load().gkm_gks_points(C.byref(p), rec.shape[0], rec.ctypes.data_as(_abi.DP),
                                          out.ctypes.data_as(_abi.DP)))
    return out


def weno_points(gas: GasParams, rec: np.ndarray) -> np.ndarray:
    """Device WENO on n records (n, 6, 5) -> (n, 10) = Wl[5] Wr[5] (gkm_weno_points)."""
    L = _abi.load()
    rec = np.ascontiguousarray(rec, dtype=np.float64).reshape(-1, 30)
    out = np.zeros((rec.shape[0], 10))
    p = gas.to_c()
    _abi.check(L.gkm_weno_points(C.byref(p), rec.shape[0], rec.ctypes.data_as(_abi.DP),
                                 out.ctypes.data_as(_abi.DP)))
    return out
