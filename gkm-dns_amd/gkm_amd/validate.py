"""Mixed form against the FP64 form over a decay run.

Both forms start from the same initial condition and advance with the same
fixed dt (the reference computes dt once per run, weno/src/update.cpp:112) to
t = t_end * t0; every `every` steps the FP64 statistic sums (gkm_stats) of both
are compared.  The reported statistics are E_k, Mt, the dilatational
dissipation 4/3 <mu (div u)^2> and the velocity-derivative skewness (mean of
the three axes), plus q^2/q^2(0) = <u_i u_i>(t)/<u_i u_i>(0), the quantity the
GKM tree prints as TKE (GKM/CUDA GKM/src/GPUupdate.cu:425-483 divides
<u_i u_i> by (c0 Mt)^2, which is <u_i u_i> at t = 0 in the init_BS.m recipe).
"""
from __future__ import annotations

import math
import time

from .ic import DhitIC
from .solver import Solver

KEYS = ("ek", "mt", "diss_dil", "skew")

# The GKM tree's own decay run at Mt = 0.5, Re = 72 (its central-difference
# smooth-flow scheme; initial condition not shipped): q^2/q^2(0) at t/t0 = 1, 2,
# 3, 4, from GKM/CUDA GKM/output_Mach-0.5-Re72.dat lines 2450, 4899, 7349, 9798.
PUBLISHED_MT05_RE72 = {1.0: 0.7800612348, 2.0: 0.5465512549, 3.0: 0.367382937, 4.0: 0.2516527394}


def _pick(st: dict) -> dict:
    return {"ek": st["ek"], "mt": st["mt"], "q2": st["q2"], "diss_dil": st["diss_dil"],
            "skew": (st["skew_u"] + st["skew_v"] + st["skew_w"]) / 3.0}


def mixed_vs_fp64(ic: DhitIC, t_end: float = 4.0, cfl: float = 0.5, every: int = 25, **gas_kw) -> dict:
    """Run both forms to t_end eddy-turnover times; return the samples, the
    largest relative error of each statistic over the run (|mixed - fp64| /
    |fp64| at each sample), and q^2/q^2(0) of the FP64 run at whole t/t0."""
    if every < 1:
        raise ValueError("every must be >= 1")
    t_begin = time.perf_counter()
    a = Solver.from_ic(ic, precision="fp64", **gas_kw)
    b = Solver.from_ic(ic, precision="mixed", **gas_kw)
    try:
        dt = a.timestep(cfl)
        t0 = ic.t0
        nsteps = int(math.ceil(t_end * t0 / dt))
        a0 = _pick(a.stats())
        samples = [{"step": 0, "t_over_t0": 0.0, "fp64": a0, "mixed": _pick(b.stats())}]
        done = 0
        while done < nsteps:
            k = min(every, nsteps - done)
            a.evolve_eulerian(k, dt=dt)
            b.evolve_eulerian(k, dt=dt)
            done += k
            samples.append({"step": done, "t_over_t0": done * dt / t0, "fp64": _pick(a.stats()),
                            "mixed": _pick(b.stats())})
    finally:
        a.close()
        b.close()
    err = {key: 0.0 for key in KEYS}
    for s in samples[1:]:
        for key in KEYS:
            ref = s["fp64"][key]
            err[key] = max(err[key], abs(s["mixed"][key] - ref) / abs(ref))
    # q^2/q^2(0) of the FP64 run at whole t/t0, linear in t between samples
    decay = {}
    for whole in range(1, int(t_end) + 1):
        for s0, s1 in zip(samples, samples[1:]):
            if s0["t_over_t0"] <= whole <= s1["t_over_t0"]:
                f = (whole - s0["t_over_t0"]) / (s1["t_over_t0"] - s0["t_over_t0"])
                q = s0["fp64"]["q2"] + f * (s1["fp64"]["q2"] - s0["fp64"]["q2"])
                decay[float(whole)] = q / a0["q2"]
                break
    return {"t_end_over_t0": t_end, "dt": dt, "t0": t0, "steps": nsteps, "sample_every": every,
            "samples": samples, "max_rel_err": err, "q2_decay": decay,
            "seconds": time.perf_counter() - t_begin}
