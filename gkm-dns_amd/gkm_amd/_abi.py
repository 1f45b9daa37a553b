"""ctypes binding of libgkm.so (the C ABI declared in include/gkm.h).

The library is built in-tree (``make -C gkm-dns_amd/csrc``) and loaded from
this folder.  There is no fallback: if the library is missing or no HIP device
is usable, the calls raise.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libgkm.so"

GKM_OK = 0
GKM_E_ARG = -1
GKM_E_HIP = -2
GKM_E_NONPHYSICAL = -3
GKM_E_NOMEM = -4
GKM_E_UNSUPPORTED = -5
GKM_E_IO = -6

VISC = {"sutherland": 0, "power": 1}
RECON = {"symoo": 0, "js": 1}
ENERGY = {"T": 0, "E": 1}
PRECISION = {"fp64": 0, "mixed": 1}
SCHEME = {"weno_bgk": 0, "gks": 1}

DP = C.POINTER(C.c_double)

# every entry point include/gkm.h declares
EXPORTS = (
    "gkm_create", "gkm_destroy", "gkm_set_state", "gkm_get_state", "gkm_fill_ghosts", "gkm_step",
    "gkm_compute_dt", "gkm_stats", "gkm_sync", "gkm_set_timing", "gkm_last_timing", "gkm_flux_points",
    "gkm_weno_points", "gkm_last_error", "gkm_version",
    "gkm_create_batch", "gkm_batch_size", "gkm_set_state_box", "gkm_get_state_box", "gkm_step_batch",
    "gkm_compute_dt_box", "gkm_stats_box", "gkm_set_graph", "gkm_set_tuning", "gkm_set_tuning_z", "gkm_set_halo_mode",
    "gkm_print_stats", "gkm_print_stats_box", "gkm_step_print_stats", "gkm_format_print_stats", "gkm_gks_points",
    "gkm_last_sweep_timing", "gkm_get_grid", "gkm_get_params_box", "gkm_set_params_box",
    "gkm_set_dt_batch", "gkm_group_create", "gkm_group_step", "gkm_group_step_persistent", "gkm_group_sync",
    "gkm_group_destroy",
    "gkm_flux_points_noalloc",
    # include/gkm_ic.h
    "gkm_dhit_velocity", "gkm_init_dhit_box", "gkm_init_dhit",
    # include/gkm_io.h
    "gkm_read_weno_params", "gkm_write_weno_params", "gkm_read_weno_init", "gkm_write_weno_init",
    "gkm_read_init_dat", "gkm_write_init_dat", "gkm_init_dat_state",
)

HALO = {"gather": 0, "pack": 1}


class GkmGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("ghost", C.c_int), ("dx", C.c_double)]


class GkmParams(C.Structure):
    _fields_ = [
        ("gamma", C.c_double), ("R", C.c_double), ("Pr", C.c_double), ("mu0", C.c_double),
        ("T0", C.c_double), ("visc_exp", C.c_double), ("K", C.c_int), ("visc_law", C.c_int),
        ("recon", C.c_int), ("energy_recon", C.c_int), ("zbug_compat", C.c_int), ("precision", C.c_int),
        ("flux_scheme", C.c_int),
    ]


class GkmIc(C.Structure):
    _fields_ = [("seed", C.c_ulonglong), ("k0", C.c_double), ("Mt", C.c_double), ("Re", C.c_double),
                ("rho0", C.c_double)]


IC_INFO_FIELDS = ("q2", "urms", "c0", "p0", "mu0", "t0")


class GkmIcInfo(C.Structure):
    _fields_ = [(f, C.c_double) for f in IC_INFO_FIELDS]

    def as_dict(self):
        return {f: getattr(self, f) for f in IC_INFO_FIELDS}


STAT_FIELDS = (
    "tke", "ek", "q2", "mt", "rho_mean", "rho_rms", "p_mean", "p_rms", "T_mean", "T_rms",
    "v_mean", "v_rms", "c_mean", "umax_plus_c",
    "dudx2", "dudx3", "dudx4", "dvdy2", "dvdy3", "dvdy4", "dwdz2", "dwdz3", "dwdz4",
    "skew_u", "skew_v", "skew_w", "kurt_u", "kurt_v", "kurt_w",
    "enstrophy", "dil2", "diss_sol", "diss_dil", "diss_ref", "sijsij",
    "mass", "mom_x", "mom_y", "mom_z", "energy",
)


class GkmStats(C.Structure):
    _fields_ = [(f, C.c_double) for f in STAT_FIELDS]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f in STAT_FIELDS}


PRINT_STAT_FIELDS = (
    "tke", "diss", "sijsij", "skew_u", "skew_v", "skew_w", "kurt_u", "kurt_v", "kurt_w",
    "p_mean", "T_mean", "v_mean", "p_rms", "T_rms", "v_rms", "dudx2", "dvdy2", "dwdz2",
)


class GkmPrintStats(C.Structure):
    _fields_ = [(f, C.c_double) for f in PRINT_STAT_FIELDS]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f in PRINT_STAT_FIELDS}


class GkmWenoParams(C.Structure):
    _fields_ = [("nc", C.c_int), ("gam", C.c_double), ("den0", C.c_double), ("Mt", C.c_double), ("Re", C.c_double),
                ("mu0", C.c_double), ("p0", C.c_double), ("t0", C.c_double)]


class GkmInitDat(C.Structure):
    _fields_ = [("n", C.c_int)] + [(f, C.c_float) for f in
                                   ("rho0", "Mt", "Re", "k0", "urms", "mu", "crms", "p0", "eddy_time")]


class GkmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gkm error {code}: {msg}")
        self.code = code


_lib = None


def load() -> C.CDLL:
    """Load libgkm.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"libgkm.so is not built ({LIB_PATH}); run `make -C gkm-dns_amd/csrc`")
    L = C.CDLL(str(LIB_PATH))
    vp = C.c_void_p
    L.gkm_create.argtypes = [C.POINTER(GkmGrid), C.POINTER(GkmParams), C.POINTER(vp)]
    L.gkm_destroy.argtypes = [vp]
    L.gkm_set_state.argtypes = [vp, C.POINTER(DP)]
    L.gkm_get_state.argtypes = [vp, C.POINTER(DP)]
    L.gkm_fill_ghosts.argtypes = [vp]
    L.gkm_step.argtypes = [vp, C.c_double, C.c_int]
    L.gkm_compute_dt.argtypes = [vp, C.c_double, DP]
    L.gkm_stats.argtypes = [vp, C.POINTER(GkmStats)]
    L.gkm_sync.argtypes = [vp]
    L.gkm_set_timing.argtypes = [vp, C.c_int]
    L.gkm_last_timing.argtypes = [vp, DP, C.POINTER(C.c_int)]
    L.gkm_last_sweep_timing.argtypes = [vp, DP]
    L.gkm_flux_points.argtypes = [C.POINTER(GkmParams), C.c_int, DP, C.c_double, DP]
    L.gkm_flux_points_noalloc.argtypes = [C.POINTER(GkmParams), C.c_int, DP, C.c_double, DP]
    L.gkm_weno_points.argtypes = [C.POINTER(GkmParams), C.c_int, DP, DP]
    L.gkm_gks_points.argtypes = [C.POINTER(GkmParams), C.c_int, DP, DP]
    L.gkm_create_batch.argtypes = [C.POINTER(GkmGrid), C.POINTER(GkmParams), C.c_int, C.POINTER(vp)]
    L.gkm_batch_size.argtypes = [vp, C.POINTER(C.c_int)]
    L.gkm_set_state_box.argtypes = [vp, C.c_int, C.POINTER(DP)]
    L.gkm_get_state_box.argtypes = [vp, C.c_int, C.POINTER(DP)]
    L.gkm_step_batch.argtypes = [vp, DP, C.c_int]
    L.gkm_compute_dt_box.argtypes = [vp, C.c_int, C.c_double, DP]
    L.gkm_stats_box.argtypes = [vp, C.c_int, C.POINTER(GkmStats)]
    L.gkm_set_graph.argtypes = [vp, C.c_int]
    L.gkm_set_tuning.argtypes = [vp, C.c_int, C.c_int]
    L.gkm_set_tuning_z.argtypes = [vp, C.c_int, C.c_int]
    L.gkm_set_halo_mode.argtypes = [vp, C.c_int]
    L.gkm_print_stats.argtypes = [vp, C.POINTER(GkmPrintStats)]
    L.gkm_print_stats_box.argtypes = [vp, C.c_int, C.POINTER(GkmPrintStats)]
    L.gkm_step_print_stats.argtypes = [vp, C.c_double, C.c_int, C.POINTER(GkmPrintStats)]
    L.gkm_format_print_stats.argtypes = [C.POINTER(GkmPrintStats)] + [C.c_double] * 6 + [C.c_char_p, C.c_int]
    FP = C.POINTER(C.c_float)
    IP = C.POINTER(C.c_int)
    L.gkm_read_weno_params.argtypes = [C.c_char_p, C.POINTER(GkmWenoParams)]
    L.gkm_write_weno_params.argtypes = [C.c_char_p, C.POINTER(GkmWenoParams)]
    L.gkm_read_weno_init.argtypes = [C.c_char_p, C.c_int, IP, C.c_double, C.POINTER(DP)]
    L.gkm_write_weno_init.argtypes = [C.c_char_p, C.c_int, IP, C.c_double, C.POINTER(DP)]
    L.gkm_read_init_dat.argtypes = [C.c_char_p, C.POINTER(GkmInitDat), FP, FP, FP]
    L.gkm_write_init_dat.argtypes = [C.c_char_p, C.POINTER(GkmInitDat), FP, FP, FP]
    L.gkm_init_dat_state.argtypes = [C.POINTER(GkmInitDat), FP, FP, FP, C.c_double, C.POINTER(DP)]
    L.gkm_last_error.restype = C.c_char_p
    L.gkm_version.restype = C.c_char_p
    for name in EXPORTS:
        if name not in ("gkm_last_error", "gkm_version"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != GKM_OK:
        raise GkmError(rc, load().gkm_last_error().decode(errors="replace"))


def format_print_stats(s: dict, t_over_t0: float, gamma: float, p0: float, T0: float, Mt: float,
                       den0: float) -> str:
    """PrintStats' output line (weno/src/update.cpp:779-780) through
    gkm_format_print_stats."""
    st = GkmPrintStats(**{f: float(s[f]) for f in PRINT_STAT_FIELDS})
    buf = C.create_string_buffer(512)
    check(load().gkm_format_print_stats(C.byref(st), float(t_over_t0), float(gamma), float(p0), float(T0), float(Mt),
                                        float(den0), buf, len(buf)))
    return buf.value.decode()


def ptr_array(arrays) -> C.Array:
    arr = (DP * len(arrays))()
    for q, a in enumerate(arrays):
        if a.dtype.name != "float64" or not a.flags.c_contiguous:
            raise ValueError("state planes must be C-contiguous float64")
        arr[q] = a.ctypes.data_as(DP)
    return arr
