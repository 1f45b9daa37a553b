"""gkm_amd — MI355X-native GKM-DNS time step (WENO + BGK flux + update,
periodic ghost fill, CFL and statistics reductions) behind the C ABI of
libgkm.so (include/gkm.h)."""
from ._abi import LIB_PATH, GkmError, load  # noqa: F401
from .ic import DhitIC, conservative_padded, dhit, dhit_native, dhit_velocity  # noqa: F401
from .solver import Batch, GasParams, Group, Solver, flux_points, gks_points, weno_points  # noqa: F401
from .ensemble import Ensemble, Realization, baseline_ensemble  # noqa: F401
from .validate import PUBLISHED_MT05_RE72, mixed_vs_fp64  # noqa: F401

__all__ = ["Solver", "Batch", "Group", "Ensemble", "Realization", "baseline_ensemble", "GasParams", "dhit", "dhit_velocity",
           "conservative_padded", "DhitIC", "dhit_native", "flux_points", "gks_points", "weno_points", "GkmError", "load", "LIB_PATH",
           "mixed_vs_fp64", "PUBLISHED_MT05_RE72"]
