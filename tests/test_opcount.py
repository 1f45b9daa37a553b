"""Operation counts of the oracle (oracle/opcount.cpp: gkm_oracle.cpp compiled
with a counting FP64 type), the algorithmic work the VALU roofline of the
benchmark is priced against (CPU only).

Pinned by hand from the reference's expressions:
  * update (update.cpp:455-469): per cell and field W -= (1/dx)(Fx - Fx' + Fy -
    Fy' + Fz - Fz'): 6 add/sub, 1 mul, 1 div;
  * flux transcendentals (GKMflux.cpp:54-57, 155-157, 186-190): erfc and
    2 sqrt in each of the 4 half-space tables (left u>0, right u<0,
    equilibrium both), exp in each of those 4 plus the 9 exp(-dt/tau) of the
    time coefficients as the reference writes them, and the power-law pow;
  * WENO temperature mode (update.cpp:560-621) costs the five lines of the
    energy mode plus 6 cell temperatures (13 operations each) and 2 energy
    rebuilds (12 each).
The benchmark's constant (bench.py ALG_FLOPS_PER_CELL) must equal the count.
"""
import importlib.util
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def counts(orc):
    return {(r, e, v): orc.opcount(orc.make_params(recon=r, energy_recon=e, visc_law=v))
            for r in (0, 1) for e in (0, 1) for v in (0, 1)}


def test_update_counts(counts):
    for c in counts.values():
        u = c["update"]
        assert (u["add"], u["mul"], u["div"]) == (30, 5, 5)
        assert u["sqrt"] == u["exp"] == u["erfc"] == u["pow"] == 0


def test_flux_transcendentals(counts):
    for (r, e, v), c in counts.items():
        f = c["flux"]
        assert f["erfc"] == 4 and f["sqrt"] == 8 and f["exp"] == 4 + 9
        assert f["pow"] == 1  # mu(T): one pow in either viscosity law
    # the flux does not depend on the reconstruction options
    base = counts[(0, 0, 1)]["flux"]
    assert all(c["flux"] == base for (r, e, v), c in counts.items() if v == 1)


def test_weno_temperature_mode_increment(counts):
    for r in (0, 1):
        t = counts[(r, 0, 1)]["weno"]
        en = counts[(r, 1, 1)]["weno"]
        d = {k: t[k] - en[k] for k in t}
        assert d == {"add": 6 * 4 + 2 * 4, "mul": 6 * 6 + 2 * 6, "div": 6 * 3 + 2 * 2,
                     "sqrt": 0, "exp": 0, "erfc": 0, "pow": 0}
        line = sum(en[k] for k in ("add", "mul", "div"))
        assert line % 5 == 0  # five identical lines in energy mode


def test_counts_do_not_depend_on_the_state(orc):
    P = orc.make_params(recon=1, energy_recon=0, visc_law=1)
    assert orc.opcount(P) == orc.opcount(P)


def test_bench_constant_is_the_count(orc):
    spec = importlib.util.spec_from_file_location("bench_mod", ROOT / "bench.py")
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    P = orc.make_params(recon=1, energy_recon=0, visc_law=1)  # the benchmark's WENO-JS, T mode, power law
    got = orc.flops_per_cell_update(P)
    assert bench.ALG_FLOPS_PER_CELL == got["flops"]
    assert bench.ALG_TRANS_PER_CELL == got["transcendentals"]
