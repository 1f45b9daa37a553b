"""Ensembles of independent decaying-turbulence realizations of uneven box
sizes on one GPU.

Realizations of the same size share one batch handle (gkm_create_batch): every
stage of a step (ghost fill + primitives, three WENO/flux sweeps, update) is a
single launch over all boxes of the batch, with each box's dt and gas
constants read from a device table.  Each batch replays its step as a
captured hipGraph on its own stream, so the size classes run concurrently.
With group=True one captured graph holds a step of every size class instead
(gkm_group_step), so one graph launch advances the whole ensemble; on MI355X
that measured ~3 % slower than the per-class streams at the 96-box baseline
ensemble, so it is not the default.  use_group(persistent=True) advances the
whole ensemble by one persistent cooperative launch per call to step(), all
steps included (gkm_group_step_persistent; same results bit for bit).  Launch
shapes per size come from the autotuner (tune.py).
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from .ic import dhit
from .solver import Batch, Group


@dataclass
class Realization:
    n: int
    Mt: float
    seed: int
    Re: float = 70.0
    k0: float = 4.0


def baseline_ensemble(seed: int = 2024) -> list[Realization]:
    """96 realizations: 32 x 32^3, 32 x 48^3, 24 x 64^3, 8 x 96^3, Mt ~ U[0.2, 0.5]."""
    rng = np.random.default_rng(seed)
    out = []
    for n, count in ((32, 32), (48, 32), (64, 24), (96, 8)):
        for _ in range(count):
            out.append(Realization(n=n, Mt=float(rng.uniform(0.2, 0.5)), seed=1234 + len(out)))
    return out


class Ensemble:
    def __init__(self, reals: list[Realization], precision: str = "mixed", recon: str = "js",
                 energy_recon: str = "T", visc_law: str = "power", tuning: dict | None = None,
                 graph: bool = True, group: bool = False):
        self.reals = list(reals)
        self.sizes = sorted({r.n for r in self.reals})
        self.batches: dict[int, Batch] = {}
        self.index: dict[int, list[int]] = {}
        for n in self.sizes:
            idx = [i for i, r in enumerate(self.reals) if r.n == n]
            ics = [dhit(n, k0=self.reals[i].k0, Mt=self.reals[i].Mt, Re=self.reals[i].Re, seed=self.reals[i].seed)
                   for i in idx]
            b = Batch.from_ics(ics, precision=precision, recon=recon, energy_recon=energy_recon, visc_law=visc_law)
            if tuning and n in tuning:
                b.set_tuning(*tuning[n])
            b.set_graph(graph)
            self.batches[n] = b
            self.index[n] = idx
        self.cells = sum(r.n ** 3 for r in self.reals)
        # group: one captured graph holds a step of every size class, so one
        # graph launch advances the whole ensemble (gkm_group_step)
        self.group = Group(list(self.batches.values())) if group else None
        self.persistent = False

    def use_group(self, on: bool = True, persistent: bool = False) -> None:
        """Step through one group graph per step (on), one persistent launch
        per step() call (on, persistent), or each batch on its own stream."""
        if on and self.group is None:
            self.group = Group(list(self.batches.values()))
        elif not on and self.group is not None:
            self.group.close()
            self.group = None
        self.persistent = bool(on and persistent)

    def timestep(self, cfl: float) -> dict[int, np.ndarray]:
        return {n: b.timestep(cfl) for n, b in self.batches.items()}

    def set_dt(self, dts: dict[int, np.ndarray]) -> None:
        for n, b in self.batches.items():
            b.dt[:] = dts[n]

    def step(self, nsteps: int = 1) -> None:
        if self.group is not None:
            self.group.step(nsteps, persistent=self.persistent)
            return
        for b in self.batches.values():
            b.step(nsteps)

    def sync(self) -> None:
        if self.group is not None:
            self.group.sync()
        for b in self.batches.values():
            b.sync()

    def run_timed(self, nsteps: int) -> float:
        """Seconds for nsteps steps of every realization (all batches queued, then synced)."""
        self.sync()
        t0 = time.perf_counter()
        self.step(nsteps)
        self.sync()
        return time.perf_counter() - t0

    def stats(self) -> list[dict]:
        out = [None] * len(self.reals)
        for n, b in self.batches.items():
            for k, i in enumerate(self.index[n]):
                out[i] = b.stats(k)
        return out

    def close(self) -> None:
        if self.group is not None:
            self.group.close()
        for b in self.batches.values():
            b.close()
