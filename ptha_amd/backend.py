"""ctypes binding of libswals_hip.so (the C ABI in include/swals_hip.h).

This module only loads the library and declares its symbols. It raises if the
library has not been built: there is no fallback path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libswals_hip.so")

SWH_METHOD = {"rk2": 1, "rk2n": 2, "midpoint": 3, "euler": 4, "linear": 5, "leapfrog_linear_plus_nonlinear_friction": 6,
              "leapfrog_nonlinear": 8}
SWH_PREC = {"fp64": 0, "fp32k": 1}
MAXQ_BITS = {"max_stage": 1, "max_speed": 2, "max_flux": 4, "arrival_time": 8, "time_of_max_stage": 16,
             "min_stage": 32}

_D = ctypes.c_double
_I32 = ctypes.c_int32
_PD = ctypes.POINTER(_D)
_PI32 = ctypes.POINTER(_I32)
_PI64 = ctypes.POINTER(ctypes.c_int64)


class DomainDesc(ctypes.Structure):
    _fields_ = [
        ("nx", _I32), ("ny", _I32), ("timestepping_method", _I32), ("precision", _I32),
        ("linear_solver_is_truely_linear", _I32), ("include_coriolis", _I32), ("spherical", _I32),
        ("friction_type", _I32), ("exterior_cells_width", _I32), ("boundary_exterior", _I32 * 4),
        ("max_U_mask", _I32), ("max_U_update_frequency", _I32),
        ("msl_linear", _D), ("cfl", _D), ("theta", _D), ("linear_friction_coeff", _D), ("maximum_timestep", _D),
        ("arrival_stage_threshold_above_msl_linear", _D),
        ("distance_bottom_edge", _PD), ("distance_left_edge", _PD), ("area_cell_y", _PD), ("coriolis", _PD),
        ("coriolis_bottom_edge", _PD), ("tanlat_on_radius_earth", _PD),
    ]


class Scalars(ctypes.Structure):
    _fields_ = [
        ("time", _D), ("max_dt", _D), ("evolve_step_dt", _D), ("dt_last_update", _D),
        ("boundary_flux_store", _D * 4), ("boundary_flux_evolve_integral", _D),
        ("boundary_flux_time_integral", _D), ("negative_depth_fix_counter", _D), ("nsteps_advanced", ctypes.c_int64),
    ]


class DomainStats(ctypes.Structure):
    """swh_domain_stats (compute_domain_statistics_new's outputs)."""
    _fields_ = [("max_stage", _D), ("max_speed", _D), ("min_stage", _D), ("min_speed", _D),
                ("energy_potential_on_rho", _D), ("energy_kinetic_on_rho", _D), ("energy_total_on_rho", _D)]


class NestBox(ctypes.Structure):
    """swh_nest_box: one send or receive box of a multidomain (1-based indices)."""
    _fields_ = [
        ("inds", _I32 * 6), ("neighbour_domain_index", _I32), ("neighbour_comms_index", _I32),
        ("my_domain_is_finer", _I32), ("equal_cell_ratios", _I32), ("my_domain_staggered_grid", _I32),
        ("neighbour_domain_staggered_grid", _I32), ("use_wetdry_limiting", _I32), ("buffer_size", _I32),
        ("flux_start", _I32), ("time_index", _I32), ("flux_rows", _I32 * 4), ("flux_nvar", _I32),
        ("cell_ratios", _D * 2),
    ]


class BackendError(RuntimeError):
    pass


_lib = None

# (name, restype, argtypes) of every symbol declared in include/swals_hip.h
SYMBOLS = [
    ("swh_last_error", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    ("swh_domain_create", ctypes.c_int, [ctypes.POINTER(DomainDesc), ctypes.POINTER(ctypes.c_void_p)]),
    ("swh_domain_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("swh_upload_U", ctypes.c_int, [ctypes.c_void_p, _PD]),
    ("swh_upload_state", ctypes.c_int, [ctypes.c_void_p, _PD, _PD, _PD]),
    ("swh_download_U", ctypes.c_int, [ctypes.c_void_p, _PD]),
    ("swh_upload_manning_squared", ctypes.c_int, [ctypes.c_void_p, _PD]),
    ("swh_set_time", ctypes.c_int, [ctypes.c_void_p, _D]),
    ("swh_set_stage_boundary_sine", ctypes.c_int, [ctypes.c_void_p, _D, _D, _D, _PI32]),
    ("swh_set_flather_boundary", ctypes.c_int, [ctypes.c_void_p, _PI32, ctypes.c_int32, _D, _D, _D]),
    ("swh_set_boundary", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _PI32, ctypes.c_int32, _D, _D, _D]),
    ("swh_dispersion_enable", ctypes.c_int, [ctypes.c_void_p, _D, _D, _D, _D, ctypes.c_int32]),
    ("swh_dispersion_store_last_U", ctypes.c_int, [ctypes.c_void_p]),
    ("swh_dispersion_append_extrapolation", ctypes.c_int, [ctypes.c_void_p]),
    ("swh_dispersion_solve", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, _D]),
    ("swh_snapshot", ctypes.c_int, [ctypes.c_void_p]),
    ("swh_snapshot_download", ctypes.c_int, [ctypes.c_void_p, _PD, _PD]),
    ("swh_evolve_one_step", ctypes.c_int, [ctypes.c_void_p, _D]),
    ("swh_evolve_steps", ctypes.c_int, [ctypes.c_void_p, _D, _I32]),
    ("swh_get_scalars", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Scalars)]),
    ("swh_volume_interior", ctypes.c_int, [ctypes.c_void_p, _PD]),
    ("swh_stationary_timestep_max", ctypes.c_int, [ctypes.c_void_p, _PD]),
    ("swh_domain_statistics", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(DomainStats)]),
    ("swh_domain_statistics_noalloc", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(DomainStats)]),
    ("swh_domain_check_stability", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    ("swh_download_max_U", ctypes.c_int, [ctypes.c_void_p, _I32, _PD]),
    ("swh_gauges_set", ctypes.c_int, [ctypes.c_void_p, _I32, _PI32, _I32]),
    ("swh_gauges_sample", ctypes.c_int, [ctypes.c_void_p]),
    ("swh_gauges_download", ctypes.c_int, [ctypes.c_void_p, _PD, _PD, _PI32]),
    ("swh_synchronize", ctypes.c_int, [ctypes.c_void_p]),
    ("swh_batch_create", ctypes.c_int, [ctypes.POINTER(DomainDesc), _I32, ctypes.POINTER(ctypes.c_void_p)]),
    ("swh_batch_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("swh_batch_upload_elevation", ctypes.c_int, [ctypes.c_void_p, _PD]),
    ("swh_batch_upload_manning_squared", ctypes.c_int, [ctypes.c_void_p, _PD]),
    ("swh_batch_upload_source", ctypes.c_int, [ctypes.c_void_p, _I32, _PD]),
    ("swh_batch_download_source", ctypes.c_int, [ctypes.c_void_p, _I32, _PD]),
    ("swh_batch_set_nsteps", ctypes.c_int, [ctypes.c_void_p, _PI32]),
    ("swh_batch_gauges_set", ctypes.c_int, [ctypes.c_void_p, _I32, _PI32, _I32, _I32]),
    ("swh_batch_gauges_set_schedule", ctypes.c_int, [ctypes.c_void_p, _I32, _PI32, _I32, _PI64]),
    ("swh_batch_evolve", ctypes.c_int, [ctypes.c_void_p, _D, _I32]),
    ("swh_batch_gauges_download", ctypes.c_int, [ctypes.c_void_p, _I32, _PD, _PI32]),
    ("swh_batch_status", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), _PI32]),
    ("swh_batch_launch_counts", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                                               ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    ("swh_batch_set_launch", ctypes.c_int, [ctypes.c_void_p, _I32, _I32, _I32]),
    ("swh_batch_autotune", ctypes.c_int, [ctypes.c_void_p, _D, _I32, _PI32, _PD]),
    ("swh_domain_set_launch", ctypes.c_int, [ctypes.c_void_p, _I32, _I32, _I32]),
    ("swh_domain_set_launch2", ctypes.c_int, [ctypes.c_void_p, _I32, _I32]),
    ("swh_batch_synchronize", ctypes.c_int, [ctypes.c_void_p]),
    ("swh_batch_evolve_timed", ctypes.c_int, [ctypes.c_void_p, _D, _I32, _PD]),
    ("swh_set_forcing_patch", ctypes.c_int, [ctypes.c_void_p, _D, _D, _I32, _I32, _I32, _I32, _I32, _I32, _PD]),
    ("swh_clear_forcing_patch", ctypes.c_int, [ctypes.c_void_p]),
    ("swh_batch_superpose", ctypes.c_int, [ctypes.c_void_p, _I32, _PI32, _PI32, _PD, _PD, _D, _I32, _PD, _PI32,
                                            _PD]),
    ("swh_evolve_record_gauges", ctypes.c_int, [ctypes.c_void_p, _D, _I32, _I32]),
    ("swh_evolve_record_gauges_schedule", ctypes.c_int, [ctypes.c_void_p, _D, _I32, _PI64]),
    ("swh_checkpoint_save", ctypes.c_int, [ctypes.c_void_p, _I32]),
    ("swh_checkpoint_restore", ctypes.c_int, [ctypes.c_void_p, _I32]),
    ("swh_adjoint_run_schedule", ctypes.c_int, [ctypes.c_void_p, _D, _I32, _I32, _PI64, _PD]),
    ("swh_adjoint_reset", ctypes.c_int, [ctypes.c_void_p]),
    ("swh_adjoint_upload", ctypes.c_int, [ctypes.c_void_p, _PD]),
    ("swh_adjoint_download", ctypes.c_int, [ctypes.c_void_p, _PD]),
    ("swh_adjoint_step", ctypes.c_int, [ctypes.c_void_p, _D, _I32]),
    ("swh_adjoint_inject", ctypes.c_int, [ctypes.c_void_p, _PD]),
    ("swh_adjoint_run", ctypes.c_int, [ctypes.c_void_p, _D, _I32, _I32, _PD]),
    ("swh_probe_copy", ctypes.c_int, [ctypes.c_int64, _I32, _PD]),
    ("swh_probe_valu", ctypes.c_int, [_I32, _I32, _I32, _I32, _PD, _PD]),
    ("swh_probe_math", ctypes.c_int, [_I32, _PD, _PD, _PD, _PD, _PD, _PD, _PD]),
    ("swh_md_create", ctypes.c_int, [_I32, ctypes.POINTER(ctypes.c_void_p), _PD, _PI32, _PI32,
                                     ctypes.POINTER(ctypes.c_void_p)]),
    ("swh_md_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("swh_md_set_priority", ctypes.c_int, [ctypes.c_void_p, _I32, _PI32, _PI32]),
    ("swh_md_set_flux_correct_dry_cells", ctypes.c_int, [ctypes.c_void_p, _I32, _I32]),
    ("swh_md_add_box", ctypes.c_int, [ctypes.c_void_p, _I32, _I32, ctypes.POINTER(NestBox)]),
    ("swh_md_finalize", ctypes.c_int, [ctypes.c_void_p]),
    ("swh_md_make_initial_conditions_consistent", ctypes.c_int, [ctypes.c_void_p]),
    ("swh_md_evolve_one_step", ctypes.c_int, [ctypes.c_void_p, _D]),
    ("swh_md_evolve_steps", ctypes.c_int, [ctypes.c_void_p, _D, _I32]),
    ("swh_md_graph_capture", ctypes.c_int, [ctypes.c_void_p, _D]),
    ("swh_md_graph_launch", ctypes.c_int, [ctypes.c_void_p, _I32]),
    ("swh_md_graph_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("swh_md_volumes", ctypes.c_int, [ctypes.c_void_p, _PD]),
    ("swh_md_download_recv_buffer", ctypes.c_int, [ctypes.c_void_p, _I32, _I32, _PD]),
    ("swh_md_statistics", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(DomainStats), ctypes.POINTER(DomainStats),
                                         _PI32]),
    ("swh_md_check_stability", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    ("swh_md_synchronize", ctypes.c_int, [ctypes.c_void_p]),
]


def load():
    """Load libswals_hip.so (built in-tree by ptha_amd/csrc/Makefile)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BackendError("%s is missing: build it with `make -C ptha_amd/csrc` "
                           "(or __graft_entry__.build())" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in SYMBOLS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc, what):
    if rc != 0:
        buf = ctypes.create_string_buffer(1024)
        _lib.swh_last_error(buf, 1024)
        raise BackendError("%s failed (status %d): %s" % (what, rc, buf.value.decode(errors="replace")))
