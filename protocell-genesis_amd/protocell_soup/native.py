"""ctypes binding of libsoup_hip.so (the C-ABI in include/soup_abi.h).

The product path has exactly one implementation: the HIP library. Loading fails loudly when the
library is missing or when no GPU is visible; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

MAX_KINDS = 8
MAX_RULES = 4
CLASSES = 4
NONE_U32 = 0xFFFFFFFF

# soup_abi.h enums
SOUP_OK, SOUP_E_INVALID, SOUP_E_DEVICE, SOUP_E_NONFINITE, SOUP_E_OVERFLOW, SOUP_E_DRIFT, SOUP_E_NOMEM = 0, -1, -2, -3, -4, -5, -6
FIELD = {
    "POS": 0, "VEL": 1, "FORCE": 2, "BOND_SLOTS": 3, "CENTER_LINK": 4, "CENTER_HELD": 5, "FROZEN": 6, "CHARGE": 7,
    "BOND_RNG": 8, "THERMO_RNG": 9, "EVENTS": 10, "DESORB": 11, "VERLET_COUNT": 12, "VERLET_LIST": 13, "ES_COUNT": 14,
    "ES_LIST": 15, "CELL_OF": 16, "CELL_START": 17, "CELL_IDX": 18, "MAX_DRIFT_SQ": 19, "HEAD_IDX": 20,
    "VERLET_NEAR": 21, "LIST_READS": 22, "REBUILDS": 23,
}
FORCES_REBUILD, FORCES_BRUTE, FORCES_LIST = 0, 1, 2
REBUILD_CADENCE, REBUILD_DRIFT = 0, 1
PHASE = {"KICK_DRIFT": 1, "REBUILD": 2, "FORCE": 4, "DRIFT": 8, "BONDS": 16, "KICK_THERMO": 32}


class SoupConfig(C.Structure):
    _fields_ = [
        ("n", C.c_uint32), ("n_kinds", C.c_uint32), ("box", C.c_float * 3), ("dims", C.c_uint32 * 3),
        ("walk_radius", C.c_uint32), ("list_range", C.c_float), ("list_cap", C.c_uint32), ("rebuild_every", C.c_uint32),
        ("skin", C.c_float),
        ("sigma", C.c_float), ("epsilon", C.c_float), ("k_fene", C.c_float), ("r_inf", C.c_float), ("k_bend", C.c_float),
        ("r_bend", C.c_float), ("wc", C.c_float), ("kT", C.c_float), ("gamma", C.c_float), ("dt", C.c_float),
        ("b_tt", C.c_float), ("wca_factor", C.c_float),
        ("radius", C.c_float * MAX_KINDS), ("cls", C.c_uint32 * MAX_KINDS), ("attr", (C.c_float * CLASSES) * CLASSES),
        ("acid_soap", C.c_float),
        ("es_enabled", C.c_uint32), ("es_A", C.c_float), ("es_kappa", C.c_float), ("es_rc", C.c_float), ("es_Frc", C.c_float),
        ("es_split", C.c_float), ("es_list_range", C.c_float), ("es_cap", C.c_uint32), ("charged_kind", C.c_uint32),
        ("n_heads", C.c_uint32),
        ("ph_A", C.c_double), ("ph_kappa", C.c_double), ("ph_cutoff", C.c_double), ("ph_shiftU", C.c_double),
        ("ph_shiftF", C.c_double), ("ph_pKa", C.c_double), ("ph_pH", C.c_double), ("ph_kT", C.c_double), ("ph_q", C.c_double),
        ("as_scale", C.c_double), ("as_rc", C.c_double), ("as_wc", C.c_double), ("as_eps", C.c_double),
        ("sweep_every", C.c_uint32),
        ("n_rules", C.c_uint32), ("kindA", C.c_float * MAX_RULES), ("kindB", C.c_float * MAX_RULES),
        ("att_form", C.c_float * MAX_RULES), ("att_break", C.c_float * MAX_RULES), ("acc_form", C.c_float * MAX_RULES),
        ("acc_break", C.c_float * MAX_RULES), ("req_cat", C.c_float * MAX_RULES), ("roleA", C.c_float * MAX_RULES),
        ("roleB", C.c_float * MAX_RULES),
        ("catalyst_kind", C.c_float), ("head_terminal_only", C.c_float), ("head_chain_cap", C.c_float),
        ("max_hold_steps", C.c_float), ("bond_interval_f", C.c_float), ("bond_interval", C.c_uint32),
        ("capacity", C.c_uint32), ("rebuild_mode", C.c_uint32), ("drift_check_every", C.c_uint32),
        ("drift_trigger", C.c_float), ("list_stage", C.c_uint32), ("reserved", C.c_uint32 * 2),
    ]


class SoupInitState(C.Structure):
    _fields_ = [
        ("positions", C.c_void_p), ("velocities", C.c_void_p), ("bond_slots", C.c_void_p), ("center_link", C.c_void_p),
        ("center_held", C.c_void_p), ("frozen", C.c_void_p), ("bond_rng", C.c_void_p), ("thermo_rng", C.c_void_p),
        ("charges", C.c_void_p), ("events", C.c_void_p), ("desorb_events", C.c_void_p),
        ("global_step", C.c_uint64), ("next_sweep_at", C.c_uint64), ("protonation_rng", C.c_uint32), ("reserved", C.c_uint32),
    ]


class SoupSweepResult(C.Structure):
    _fields_ = [("attempts", C.c_uint32), ("accepted", C.c_uint32), ("heads", C.c_uint32), ("deprotonated", C.c_uint32),
                ("alpha", C.c_double), ("dEsMeanKT", C.c_double)]


class SoupListStats(C.Structure):
    _fields_ = [("verlet_mean", C.c_double), ("verlet_max", C.c_uint32), ("verlet_cap", C.c_uint32),
                ("es_mean", C.c_double), ("es_max", C.c_uint32), ("es_cap", C.c_uint32),
                ("verlet_pairs", C.c_uint64), ("es_pairs", C.c_uint64)]


class SoupPotentialBasis(C.Structure):
    """The doubles of makePotentialBasis (include/soup_abi.h)."""
    _fields_ = [("epsilon", C.c_double), ("fene_k", C.c_double), ("fene_r_inf", C.c_double), ("bend_k", C.c_double),
                ("bend_r0", C.c_double), ("wc", C.c_double), ("rc_attr", C.c_double), ("cutoff", C.c_double),
                ("wca_factor", C.c_double), ("acid_soap", C.c_double), ("b_radius", C.c_double * 8),
                ("attr", (C.c_double * 4) * 4), ("es_A", C.c_double), ("es_kappa", C.c_double), ("es_cutoff", C.c_double),
                ("es_shiftU", C.c_double), ("es_shiftF", C.c_double), ("cls", C.c_uint32 * 8), ("charged_kind", C.c_uint32),
                ("use_center_link", C.c_uint32), ("use_charges", C.c_uint32), ("reserved", C.c_uint32 * 5)]


class SoupSegmentSpec(C.Structure):
    _fields_ = [("steps", C.c_uint32), ("checkpoint_every", C.c_uint32), ("objective", C.c_uint32), ("reserved", C.c_uint32),
                ("strain", C.c_double * 3), ("scale", C.c_double)]


OBJ_END_ENERGY, OBJ_MEAN_STRAIN_VIRIAL = 0, 1  # soup_segment_adjoint objectives


class SoupWaterCounts(C.Structure):
    _fields_ = [("total_empty_cells", C.c_uint64), ("unreached_cells", C.c_uint64), ("reached_empty_cells", C.c_uint64),
                ("total_water", C.c_uint32), ("encapsulated", C.c_uint32), ("reached_empty_water", C.c_uint32),
                ("seed_cell", C.c_uint32), ("dims", C.c_uint32 * 3), ("reserved", C.c_uint32)]


PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libsoup_hip.so")

SET_BOX_FORCE = 1  # soup_set_box: run a force pass after the rebuild
NOALLOC_MAX_N, NOALLOC_MAX_CELLS = 65536, 262144  # soup_potential_at_noalloc limits

# every entry point include/soup_abi.h declares
EXPORTS = (
    "soup_create", "soup_destroy", "soup_last_error", "soup_step", "soup_step_batch", "soup_step_record", "soup_set_params", "soup_reset_particles", "soup_sample_insertion", "soup_force_candidate_stats", "soup_step_phases", "soup_global_step", "soup_read", "soup_write",
    "soup_field_bytes", "soup_forces", "soup_run_phases", "soup_protonation_sweep", "soup_protonation_rng",
    "soup_last_sweep", "soup_nonfinite", "soup_check_lists", "soup_list_stats", "soup_relax", "soup_potential", "soup_potential_at", "soup_potential_at_noalloc", "soup_energy_adjoint", "soup_energy_adjoint_strain", "soup_segment_adjoint", "soup_set_box", "soup_scale_molecules_rigid", "soup_proximity_pairs", "soup_analysis_create", "soup_analysis_destroy", "soup_analysis_last_error",
    "soup_analysis_load", "soup_analysis_load_system", "soup_analysis_amphiphiles", "soup_analysis_amphiphiles_read",
    "soup_analysis_clusters", "soup_analysis_water", "soup_analysis_cavity", "soup_profile", "soup_profile_read", "soup_use_graphs", "soup_version",
)
PROF_PHASES = ("kick_drift", "sort", "es_build", "verlet_build", "force", "es_far", "drift", "bonds", "kick_thermo",
               "proton", "sweep")

_lib: Optional[C.CDLL] = None


def load_library(path: Optional[str] = None) -> C.CDLL:
    """Loads libsoup_hip.so (SOUP_LIB_PATH names another build, such as an autotuner variant) and
    declares the signatures. Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("SOUP_LIB_PATH") or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"libsoup_hip.so is not built at {path}: run __graft_entry__.build() first")
    lib = C.CDLL(path)
    H = C.c_void_p
    sig = {
        "soup_create": (C.c_int, [C.POINTER(SoupConfig), C.POINTER(SoupInitState), C.POINTER(H)]),
        "soup_destroy": (C.c_int, [H]),
        "soup_last_error": (C.c_char_p, [H]),
        "soup_step": (C.c_int, [H, C.c_uint32]),
        "soup_step_batch": (C.c_int, [C.POINTER(H), C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32)]),
        "soup_step_record": (C.c_int, [H, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
        "soup_set_params": (C.c_int, [H, C.POINTER(SoupConfig)]),
        "soup_reset_particles": (C.c_int, [H, C.c_uint32, C.POINTER(SoupInitState)]),
        "soup_force_candidate_stats": (C.c_int, [H, C.POINTER(C.c_uint64)]),
        "soup_step_phases": (C.c_int, [H, C.c_uint32, C.POINTER(C.c_double)]),
        "soup_sample_insertion": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_double), C.c_uint32, C.c_double,
                                            C.POINTER(C.c_uint32), C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_double)]),
        "soup_global_step": (C.c_int, [H, C.POINTER(C.c_uint64)]),
        "soup_read": (C.c_int, [H, C.c_int, C.c_void_p, C.c_size_t]),
        "soup_write": (C.c_int, [H, C.c_int, C.c_void_p, C.c_size_t]),
        "soup_field_bytes": (C.c_size_t, [H, C.c_int]),
        "soup_forces": (C.c_int, [H, C.c_int]),
        "soup_run_phases": (C.c_int, [H, C.c_uint32, C.c_uint32]),
        "soup_protonation_sweep": (C.c_int, [H, C.POINTER(SoupSweepResult)]),
        "soup_protonation_rng": (C.c_int, [H, C.POINTER(C.c_uint32)]),
        "soup_last_sweep": (C.c_int, [H, C.POINTER(C.c_uint32), C.POINTER(SoupSweepResult)]),
        "soup_nonfinite": (C.c_int, [H, C.POINTER(C.c_uint32)]),
        "soup_check_lists": (C.c_int, [H]),
        "soup_list_stats": (C.c_int, [H, C.POINTER(SoupListStats)]),
        "soup_relax": (C.c_int, [H, C.c_uint32, C.c_double]),
        "soup_potential": (C.c_int, [H, C.POINTER(SoupPotentialBasis), C.POINTER(C.c_double)]),
        "soup_potential_at": (C.c_int, [H, C.POINTER(SoupPotentialBasis), C.c_void_p, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]),
        "soup_potential_at_noalloc": (C.c_int, [H, C.POINTER(SoupPotentialBasis), C.c_void_p, C.POINTER(C.c_double),
                                                C.POINTER(C.c_double)]),
        "soup_set_box": (C.c_int, [H, C.POINTER(SoupConfig), C.c_void_p, C.c_uint32]),
        "soup_energy_adjoint": (C.c_int, [H, C.POINTER(SoupPotentialBasis), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p]),
        "soup_energy_adjoint_strain": (C.c_int, [H, C.POINTER(SoupPotentialBasis), C.POINTER(C.c_double), C.c_void_p, C.c_void_p,
                                                 C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]),
        "soup_segment_adjoint": (C.c_int, [H, C.POINTER(SoupPotentialBasis), C.POINTER(SoupSegmentSpec), C.POINTER(C.c_double),
                                           C.c_void_p, C.c_void_p, C.c_void_p]),
        "soup_proximity_pairs": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_double), C.c_double, C.c_void_p, C.c_void_p,
                                           C.c_uint64, C.POINTER(C.c_uint64)]),
        "soup_scale_molecules_rigid": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.POINTER(C.c_double),
                                                 C.POINTER(C.c_double), C.c_void_p]),
        "soup_analysis_create": (C.c_int, [C.POINTER(H)]),
        "soup_analysis_destroy": (C.c_int, [H]),
        "soup_analysis_last_error": (C.c_char_p, [H]),
        "soup_analysis_load": (C.c_int, [H, C.c_uint32, C.c_void_p, C.c_void_p]),
        "soup_analysis_load_system": (C.c_int, [H, H]),
        "soup_analysis_amphiphiles": (C.c_int, [H, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "soup_analysis_amphiphiles_read": (C.c_int, [H, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
        "soup_analysis_clusters": (C.c_int, [H, C.c_void_p, C.c_uint32, C.POINTER(C.c_double), C.c_double, C.c_int, C.c_void_p]),
        "soup_analysis_water": (C.c_int, [H, C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                          C.c_double, C.c_double, C.POINTER(SoupWaterCounts)]),
        "soup_analysis_cavity": (C.c_int, [H, C.c_void_p, C.c_uint32, C.POINTER(C.c_double), C.c_double, C.c_double,
                                           C.POINTER(C.c_uint64)]),
        "soup_profile": (C.c_int, [H, C.c_int]),
        "soup_use_graphs": (C.c_int, [H, C.c_int]),
        "soup_profile_read": (C.c_int, [H, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
        "soup_version": (C.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class SoupError(RuntimeError):
    """Raised with the library's message; ``code`` is the SOUP_E_* status."""

    def __init__(self, msg: str, code: int):
        super().__init__(msg)
        self.code = code


def check(lib: C.CDLL, handle, code: int) -> None:
    if code != 0:
        msg = lib.soup_last_error(handle)
        raise SoupError(msg.decode() if msg else f"libsoup_hip error {code}", code)
