"""Autotuner for the launch shapes of the step kernels.

For a box size (and precision / reconstruction), the launch shapes are chosen
in two stages, each timing every candidate on a batch of boxes of that size
(wall time over a few steps after a warm-up step):
  1. X/Y flux threads per block in {64, 128, 256} x WENO interfaces per thread
     of the strided sweeps in {4, 8, 16}, the fused Z kernel at its default;
  2. with the best of stage 1, the fused Z kernel's threads per block in
     {64, 128, 256} x the width of the strips of columns its z lines go in,
     {8, 16, 32}.
The fastest of all is kept: (flux_block, recon_chunk, zflux_block, z_strip).
Results are cached as JSON, keyed by the built library's identity (its
version string and a hash of libgkm.so, so a rebuilt kernel is timed again)
and "n/precision/recon/nbox", so later runs read them instead of timing again.
The arithmetic of every candidate is the same: the choice changes speed, never
results.
"""
from __future__ import annotations

import functools
import hashlib
import itertools
import json
import os
import time
from pathlib import Path

import numpy as np

from . import _abi
from .ic import dhit
from .solver import Batch

FLUX_BLOCKS = (64, 128, 256)
RECON_CHUNKS = (4, 8, 16)
ZFLUX_BLOCKS = (64, 128, 256)
Z_STRIPS = (8, 16, 32)
# the library's defaults (gkm_create) per precision
DEFAULT = {"mixed": (256, 8, 128, 16), "fp64": (64, 8, 64, 16)}


def cache_path() -> Path:
    env = os.environ.get("GKM_TUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "gkm_amd" / "tuning.json"


@functools.lru_cache(maxsize=None)
def library_id() -> str:
    """gkm_version() and the first 16 hex digits of the SHA-1 of libgkm.so."""
    L = _abi.load()
    digest = hashlib.sha1(Path(_abi.LIB_PATH).read_bytes()).hexdigest()[:16]
    return f"{L.gkm_version().decode()}@{digest}"


def _key(n: int, precision: str, recon: str, nbox: int) -> str:
    return f"{library_id()}/{n}/{precision}/{recon}/{nbox}"


def _load(path: Path) -> dict:
    try:
        return json.loads(path.read_text())
    except (OSError, ValueError):
        return {}


def _store(path: Path, data: dict) -> None:
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(json.dumps(data, indent=1, sort_keys=True))
    except OSError:
        pass


def tune_size(n: int, precision: str = "mixed", recon: str = "js", nbox: int = 4, steps: int = 3,
              use_cache: bool = True, verbose: bool = False) -> tuple[int, int, int, int]:
    key = _key(n, precision, recon, nbox)
    path = cache_path()
    cache = _load(path) if use_cache else {}
    if key in cache:
        return tuple(cache[key]["best"])
    ics = [dhit(n, Mt=0.3 + 0.05 * i, Re=70, seed=900 + i) for i in range(nbox)]
    b = Batch.from_ics(ics, precision=precision, recon=recon, energy_recon="T", visc_law="power")
    dt = b.timestep(0.5)
    times = {}

    def timed(shape):
        b.set_tuning(*shape)
        b.step(1, dt)
        b.sync()
        t0 = time.perf_counter()
        b.step(steps, dt)
        b.sync()
        times[shape] = (time.perf_counter() - t0) / steps
        if verbose:
            print(f"tune n={n} flux_block={shape[0]} recon_chunk={shape[1]} zflux_block={shape[2]} "
                  f"z_strip={shape[3]}: {1e3 * times[shape]:.3f} ms/step")

    zdef = DEFAULT[precision][2:]
    for fb, rc in itertools.product(FLUX_BLOCKS, RECON_CHUNKS):
        timed((fb, rc) + zdef)
    first = min(times, key=times.get)
    for zb, zs in itertools.product(ZFLUX_BLOCKS, Z_STRIPS):
        timed(first[:2] + (zb, zs))
    b.close()
    best = min(times, key=times.get)
    cache[key] = {"best": list(best), "ms_per_step": {"x".join(map(str, k)): 1e3 * v for k, v in times.items()},
                  "default_ms": 1e3 * times[DEFAULT[precision]]}
    _store(path, cache)
    return best


def tune_sizes(sizes, **kw) -> dict[int, tuple[int, int, int, int]]:
    return {int(n): tune_size(int(n), **kw) for n in sizes}


def speedup_report(sizes, precision: str = "mixed", recon: str = "js", nbox: int = 4) -> dict:
    cache = _load(cache_path())
    rep = {}
    for n in sizes:
        e = cache.get(_key(n, precision, recon, nbox))
        if e:
            best = min(e["ms_per_step"].values())
            rep[int(n)] = {"best": e["best"], "best_ms": best, "default_ms": e["default_ms"],
                           "gain": e["default_ms"] / best if best > 0 else float(np.nan)}
    return rep
