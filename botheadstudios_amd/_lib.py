"""ctypes binding of libihs_sph.so (include/ihs_sph.h). Fails loudly when the library is missing."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libihs_sph.so")

_lib = None


class IhsConfig(C.Structure):
    _fields_ = [
        ("capacity", C.c_int32), ("leaf_size", C.c_int32), ("precision", C.c_int32), ("h_mode", C.c_int32),
        ("theta", C.c_float), ("cfl", C.c_float), ("n_ngb", C.c_float), ("h_tol", C.c_float),
        ("h_max_iter", C.c_int32), ("h_margin", C.c_float), ("h_cap_factor", C.c_float),
        ("nbr_capacity", C.c_int32), ("use_graph", C.c_int32), ("h_margin_seeded", C.c_float),
        ("group_max", C.c_int32), ("quadrupole", C.c_int32), ("reuse_positions", C.c_int32),
        ("fused_walk", C.c_int32), ("theta_mono", C.c_float), ("quad_tol", C.c_float),
        ("theta_settled", C.c_float), ("reserved", C.c_int32 * 3),
    ]


class IhsStats(C.Structure):
    _fields_ = [
        ("ke", C.c_double), ("ie", C.c_double), ("pe", C.c_double), ("p", C.c_double * 3), ("L", C.c_double * 3),
        ("t", C.c_double), ("dt", C.c_double), ("steps", C.c_int64), ("force_evals", C.c_int64),
        ("mean_h_iters", C.c_double), ("mean_h_evals", C.c_double), ("mean_neighbours", C.c_double),
        ("nbr_overflow", C.c_uint32), ("walk_overflow", C.c_uint32), ("h_capped", C.c_uint32),
        ("h_unconverged", C.c_uint32), ("nonfinite", C.c_uint32), ("tie_runs", C.c_uint32),
        ("grav_cell_interactions", C.c_int64), ("grav_pp_interactions", C.c_int64),
        ("nbr_candidates", C.c_int64), ("walk_nodes", C.c_int64),
    ]


class IhsClump(C.Structure):
    _fields_ = [
        ("first", C.c_int64), ("n_members", C.c_int64), ("mass", C.c_double), ("com_pos", C.c_double * 3),
        ("com_vel", C.c_double * 3), ("rho", C.c_double), ("radius", C.c_double), ("internal_ke", C.c_double),
        ("self_pe", C.c_double), ("earth_mass", C.c_double), ("bound", C.c_int32), ("outside_roche", C.c_int32),
    ]


class IhsMeasure(C.Structure):
    _fields_ = [
        ("earth_frac", C.c_double), ("disk_kg", C.c_double), ("e_disk", C.c_double), ("t_disk", C.c_double),
        ("e_rem", C.c_double), ("t_rem", C.c_double), ("e_esc", C.c_double), ("t_esc", C.c_double),
        ("r_remnant", C.c_double), ("clump_kg", C.c_double), ("clump_earth_frac", C.c_double),
        ("clump_n", C.c_int64), ("n_clumps", C.c_int64), ("n_bound", C.c_int64),
        ("m_total", C.c_double), ("m_remnant", C.c_double), ("com", C.c_double * 3), ("v_com", C.c_double * 3),
        ("link", C.c_double), ("n_disk", C.c_int64),
    ]


class IhsDiskSummary(C.Structure):
    _fields_ = [
        ("disk", C.c_double), ("earth_pct", C.c_double), ("remnant_km", C.c_double), ("escaped", C.c_double),
        ("moon", C.c_double), ("e_disk", C.c_double), ("t_disk", C.c_double), ("esc", C.c_double),
        ("r_remnant", C.c_double), ("m_remnant", C.c_double), ("biggest", C.c_double), ("link", C.c_double),
        ("n_disk", C.c_int64), ("n_clumps", C.c_int64),
    ]


STAGES = ["sort", "gather", "tree", "nbr", "hsolve", "density", "forces", "gravity", "kick"]


# Every symbol include/ihs_sph.h declares (tests check the library exports all of them).
EXPORTS = [
    "ihs_default_config", "ihs_abi_version", "ihs_create", "ihs_destroy", "ihs_last_error", "ihs_upload",
    "ihs_set_dt", "ihs_set_av", "ihs_set_omega", "ihs_step_kdk", "ihs_relax", "ihs_force_eval", "ihs_readback",
    "ihs_count", "ihs_sync", "ihs_stats_get", "ihs_gravity", "ihs_debug_tree", "ihs_debug_field",
    "ihs_profile", "ihs_profile_read", "ihs_mlp_set_weights", "ihs_mlp_get_weights", "ihs_mlp_record",
    "ihs_mlp_samples", "ihs_mlp_train", "ihs_mlp_gradient", "ihs_mlp_predict", "ihs_direct_f64",
    "ihs_upload_batch", "ihs_segment_clock", "ihs_measure_impact", "ihs_disk_stats", "ihs_disk_clumps",
    "ihs_find_clumps", "ihs_body_bulk", "ihs_assemble", "ihs_checkpoint_bytes", "ihs_checkpoint_save",
    "ihs_checkpoint_load", "ihs_begin_readback", "ihs_take_readback", "ihs_segment_energy",
    "ihs_direct_f64_noalloc",
]

MLP_NPARAM = 1313
MLP_IN = 6
MLP_SAMPLE = 8


def load(path: str | None = None):
    """Load and declare the library (no device work happens here)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libihs_sph.so not built at {p}: run `python -m botheadstudios_amd.build` "
                           "(there is no CPU fallback)")
    L = C.CDLL(p)
    vp, i32, i64, f32, f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_float, C.c_double
    L.ihs_default_config.argtypes = [C.POINTER(IhsConfig)]
    L.ihs_default_config.restype = None
    L.ihs_abi_version.restype = C.c_int
    L.ihs_create.argtypes = [C.c_int, C.POINTER(IhsConfig), C.POINTER(vp)]
    L.ihs_destroy.argtypes = [vp]
    L.ihs_last_error.argtypes = [vp]
    L.ihs_last_error.restype = C.c_char_p
    L.ihs_upload.argtypes = [vp, vp, i64, vp, i32, f32]
    L.ihs_set_dt.argtypes = [vp, f64, f32]
    L.ihs_set_av.argtypes = [vp, f32, f32]
    L.ihs_set_omega.argtypes = [vp, f32]
    L.ihs_step_kdk.argtypes = [vp, i32]
    L.ihs_relax.argtypes = [vp, i32]
    L.ihs_force_eval.argtypes = [vp]
    L.ihs_readback.argtypes = [vp, vp, i64, i32]
    L.ihs_readback.restype = i64
    L.ihs_count.argtypes = [vp]
    L.ihs_count.restype = i64
    L.ihs_sync.argtypes = [vp]
    L.ihs_stats_get.argtypes = [vp, C.POINTER(IhsStats), i32]
    L.ihs_gravity.argtypes = [vp, vp, vp, i64, f64, f64, i32, vp, vp]
    L.ihs_debug_tree.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.ihs_debug_tree.restype = i64
    L.ihs_debug_field.argtypes = [vp, i32, vp]
    L.ihs_profile.argtypes = [vp, i32]
    L.ihs_profile_read.argtypes = [vp, vp, vp, i32]
    L.ihs_mlp_set_weights.argtypes = [vp, vp, i64]
    L.ihs_mlp_get_weights.argtypes = [vp, vp, i64]
    L.ihs_mlp_record.argtypes = [vp, i64]
    L.ihs_mlp_samples.argtypes = [vp, vp, i64]
    L.ihs_mlp_samples.restype = i64
    L.ihs_mlp_train.argtypes = [vp, vp, i64, i32, i32, f32, C.c_uint64, vp]
    L.ihs_mlp_gradient.argtypes = [vp, vp, vp, i64, vp]
    L.ihs_mlp_predict.argtypes = [vp, vp, i64, vp]
    L.ihs_direct_f64.argtypes = [vp, vp, i64, vp]
    L.ihs_upload_batch.argtypes = [vp, vp, vp, i32, vp, i32, vp]
    L.ihs_segment_clock.argtypes = [vp, i32, vp, vp, vp]
    L.ihs_measure_impact.argtypes = [vp, i32, i64, C.POINTER(IhsMeasure)]
    L.ihs_disk_stats.argtypes = [vp, i32, C.POINTER(IhsDiskSummary)]
    L.ihs_disk_clumps.argtypes = [vp, i32, i32, vp, vp, i64, vp]
    L.ihs_find_clumps.argtypes = [vp, i64, vp, vp, vp, vp, f64, f64, f64, vp, f64, f64, vp, i64, vp, vp]
    L.ihs_body_bulk.argtypes = [vp, i32, C.c_uint32, vp]
    L.ihs_assemble.argtypes = [vp, i32, vp, i32, vp, vp]
    L.ihs_checkpoint_bytes.argtypes = [vp]
    L.ihs_checkpoint_bytes.restype = i64
    L.ihs_checkpoint_save.argtypes = [vp, vp, i64]
    L.ihs_checkpoint_save.restype = i64
    L.ihs_checkpoint_load.argtypes = [vp, vp, i64]
    L.ihs_begin_readback.argtypes = [vp]
    L.ihs_take_readback.argtypes = [vp, vp, i64]
    L.ihs_take_readback.restype = i64
    L.ihs_segment_energy.argtypes = [vp, i32, vp]
    L.ihs_direct_f64_noalloc.argtypes = [vp, vp, i64, vp]
    for name in ("ihs_create", "ihs_destroy", "ihs_upload", "ihs_set_dt", "ihs_set_av", "ihs_set_omega",
                 "ihs_step_kdk", "ihs_relax", "ihs_force_eval", "ihs_sync", "ihs_stats_get", "ihs_gravity",
                 "ihs_debug_field", "ihs_profile", "ihs_profile_read", "ihs_mlp_set_weights", "ihs_mlp_get_weights",
                 "ihs_mlp_record", "ihs_mlp_train", "ihs_mlp_gradient", "ihs_mlp_predict", "ihs_direct_f64",
                 "ihs_upload_batch", "ihs_segment_clock", "ihs_measure_impact", "ihs_disk_stats", "ihs_disk_clumps",
                 "ihs_find_clumps", "ihs_body_bulk", "ihs_assemble", "ihs_checkpoint_load", "ihs_begin_readback",
                 "ihs_segment_energy", "ihs_direct_f64_noalloc"):
        getattr(L, name).restype = C.c_int
    if path is None:
        _lib = L
    return L
