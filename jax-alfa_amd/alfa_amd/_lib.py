"""ctypes binding of libalfa.so (include/alfa.h).

The library is loaded from ``alfa_amd/lib/libalfa.so`` (built in-tree by
``jax-alfa_amd/build.py``).  There is no fallback: if the library is missing,
or a GPU call fails, the binding raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libalfa.so")
HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))),
                      "include", "alfa.h")

ALFA_FP64 = 0
ALFA_MIXED = 1

_CFG_INT = ["nx", "ny", "nz", "optSgs", "FGR", "dynamicSGS_call_time", "optSurfFlux",
            "optSurfBC", "optBuoyancy", "optDamping", "istep", "optMoisture",
            "optMoistureSurfBC"]
_CFG_DBL = ["l_x", "l_y", "l_z", "dt", "Ugal", "z0m", "z0T", "zTemperature",
            "SensibleHeatFlux", "Ug2", "Vg2", "f_coriolis", "inversion", "T_0", "Cs2",
            "Cs2PrRatio", "z_damping", "RelaxTime", "MoistureFlux", "zMoisture", "q_inversion"]


class AlfaConfig(C.Structure):
    _fields_ = ([(n, C.c_int32) for n in _CFG_INT] + [("reserved_i", C.c_int32 * 5)]
                + [(n, C.c_double) for n in _CFG_DBL] + [("reserved_d", C.c_double * 5)])


class AlfaDerived(C.Structure):
    _fields_ = ([(n, C.c_double) for n in (
        "z_scale", "dx", "dy", "dz", "L", "TFR", "dt_nondim", "g_nondim", "f_coriolis_nondim",
        "inversion_nondim", "z_damping_nondim", "RelaxTime_nondim")]
        + [(n, C.c_int32) for n in ("mx", "my", "nx_rfft", "ny_rfft", "optDealias", "cut0_x",
                                    "cut0_y", "cut1_x", "cut1_y", "cut2_x", "cut2_y",
                                    "precision")]
        + [("workspace_bytes", C.c_int64)])


class AlfaKernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 96), ("launches", C.c_int64), ("total_ms", C.c_double),
                ("bytes", C.c_double), ("ref_bytes", C.c_double)]


STATE_FIELDS = ["u", "v", "w", "TH", "RHS_u", "RHS_v", "RHS_w", "RHS_TH", "Cs2_3D",
                "Cs2PrRatio_3D", "psi_m", "psi_m0", "psi_h", "psi_h0", "fi_m", "fi_h"]
STATE_DIAG = ["Cs2_1D_avg1", "Cs2_1D_avg2", "beta1_1D", "Cs2PrRatio_1D", "beta2_1D", "scalars"]


STATE_MOIST = ["Q", "RHS_Q", "Q_sfc_val"]


class AlfaState(C.Structure):
    _fields_ = ([(n, C.c_void_p) for n in STATE_FIELDS] + [(n, C.c_void_p) for n in STATE_DIAG]
                + [(n, C.c_void_p) for n in STATE_MOIST])


def config_struct(cfg: dict) -> AlfaConfig:
    s = AlfaConfig()
    for n in _CFG_INT:
        setattr(s, n, int(cfg.get(n, 1 if n == "istep" else 0)))
    for n in _CFG_DBL:
        setattr(s, n, float(cfg.get(n, 0.0)))
    if int(cfg.get("optSgs", 0)) in (2, 4):
        # WL models start from Cwl / CwlPrRatio (src/config/DerivedVars.py:101-106)
        s.Cs2 = float(cfg.get("Cwl", 0.1 ** 2))
        s.Cs2PrRatio = float(cfg.get("CwlPrRatio", 0.1 ** 2))
    if "optBuoyancy" not in cfg:
        s.optBuoyancy = 1
    if "optDamping" not in cfg:
        s.optDamping = 1
    if "dynamicSGS_call_time" not in cfg:
        s.dynamicSGS_call_time = 1
    return s


_P = C.c_void_p
_SIGS = {
    "alfa_plan_create": [C.POINTER(AlfaConfig), C.c_int, C.POINTER(C.c_void_p)],
    "alfa_plan_destroy": [_P],
    "alfa_plan_derived": [_P, C.POINTER(AlfaDerived)],
    "alfa_plan_autotune": [_P, C.POINTER(AlfaState), C.c_int, C.c_double, _P, _P],
    "alfa_plan_set_tiles": [_P, _P],
    "alfa_plan_get_tiles": [_P, _P],
    "alfa_plan_set_fusion": [_P, C.c_int],
    "alfa_plan_get_fusion": [_P, _P],
    "alfa_state_init": [_P, C.POINTER(AlfaState), _P],
    "alfa_step": [_P, C.POINTER(AlfaState), C.c_int, C.c_double, _P],
    "alfa_step_sfc_dev": [_P, C.POINTER(AlfaState), C.c_int, _P, _P],
    "alfa_stats_accumulate": [_P, C.POINTER(AlfaState), _P, _P],
    "alfa_set_forcing": [_P, _P, _P, _P, _P, _P, _P, _P],
    "alfa_struct_sizes": [_P],
    "alfa_batch_step": [_P, _P, C.c_int, _P, _P, _P],
    "alfa_fft": [_P, _P, _P, _P],
    "alfa_fft_pad": [_P, _P, _P, _P],
    "alfa_filtering_explicit": [_P, _P, _P, _P, _P],
    "alfa_filtering_level1": [_P, _P, _P, _P],
    "alfa_filtering_level2": [_P, _P, _P, _P],
    "alfa_dealias1": [_P, _P, _P, _P],
    "alfa_dealias2": [_P, _P, _P, _P],
    "alfa_derivxy": [_P, _P, C.c_int, _P, _P],
    "alfa_derivz_m": [_P] * 11,
    "alfa_derivz_th": [_P] * 7,
    "alfa_derivz_generic_uvp": [_P] * 4,
    "alfa_derivz_generic_w": [_P] * 4,
    "alfa_planar_mean": [_P] * 4,
    "alfa_stag_grid_avg": [_P] * 4,
    "alfa_imfilter": [_P] * 4,
    "alfa_compute_beta": [_P, _P, C.c_int, _P, _P],
    "alfa_advection": [_P] * 14,
    "alfa_scalar_advection": [_P] * 9,
    "alfa_buoyancy": [_P] * 4,
    "alfa_div_stress": [_P] * 11,
    "alfa_div_flux": [_P] * 6,
    "alfa_rhs_momentum": [_P] * 15,
    "alfa_rhs_scalar": [_P] * 6,
    "alfa_pressure_rc": [_P] * 15,
    "alfa_thomas_pressure_solve": [_P] * 9,
    "alfa_ab2_uvw": [_P] * 11,
    "alfa_ab2_th": [_P] * 5,
    "alfa_strains_uvp": [_P, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), _P],
    "alfa_strains_w": [_P, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), _P],
    "alfa_lasdd_sm": [_P, _P, _P, _P, C.POINTER(C.c_void_p), _P, _P, _P, _P, _P],
    "alfa_lasdd_wl": [_P, _P, _P, _P, C.POINTER(C.c_void_p), _P, _P, _P, _P, _P],
    "alfa_scalar_lasdd": [_P, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                          C.POINTER(C.c_void_p), _P, C.POINTER(C.c_void_p), _P, _P, _P, _P, _P,
                          _P, _P],
    "alfa_sgs_stresses": [_P, C.POINTER(C.c_void_p), _P, _P, _P, _P, _P, _P,
                          C.POINTER(C.c_void_p), _P],
    "alfa_scalar_fluxes": [_P, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), _P, _P,
                           C.POINTER(C.c_void_p), _P],
    "alfa_div_stress_sgs": [_P, C.POINTER(C.c_void_p), _P, _P, _P, _P, _P, _P, C.c_int, _P, _P,
                            _P, _P, C.POINTER(C.c_void_p), _P],
    "alfa_div_flux_sgs": [_P, C.POINTER(C.c_void_p), _P, _P, _P, _P, C.POINTER(C.c_void_p), _P,
                          C.c_int, _P, _P, _P, _P, _P, _P],
    "alfa_wall": [_P] * 9,
    "alfa_surface_flux": [_P, _P, _P, _P, C.c_double] + [_P] * 12,
    "alfa_profile_begin": [_P, C.c_char_p, C.c_int],
    "alfa_profile_end": [_P, C.POINTER(AlfaKernelStat), C.c_int, C.POINTER(C.c_int)],
    "alfa_planar_mean_noalloc": [C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, _P],
    "alfa_group_create": [_P, _P, _P, C.c_int, _P],
    "alfa_group_step": [_P, C.c_int, _P],
    "alfa_group_launches": [_P, C.c_int, C.POINTER(C.c_int)],
    "alfa_group_destroy": [_P],
}

_lib = None


def header_symbols() -> list:
    """Every entry point declared in include/alfa.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(alfa_[a-z0-9_]+)\s*\(", txt)))


def load():
    """Load libalfa.so (raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libalfa.so not built: {LIB_PATH} (run jax-alfa_amd/build.py)")
    # torch first: torch ships its own HIP runtime under the same soname
    # (libamdhip64.so.7), and the library must bind the one torch uses.
    # Loaded the other way round, the library pulls in the system runtime
    # ahead of torch's and the process later finds no device at plan creation.
    import torch  # noqa: F401
    lib = C.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    lib.alfa_last_error.restype = C.c_char_p
    lib.alfa_last_error.argtypes = []
    _lib = lib
    return lib


def check(rc: int, what: str):
    if rc != 0:
        msg = load().alfa_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed (code {rc}): {msg}")
