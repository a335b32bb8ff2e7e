"""ctypes mirror of include/galic_types.h and include/galic_hip.h.

Loads galic_amd/libgalic_hip.so and fails loudly when it is missing: there is
no CPU path behind this package.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgalic_hip.so")

NT = 4  # types 1..3, index 0 unused
_d4 = ctypes.c_double * NT
_i4 = ctypes.c_int * NT


class GalicParams(ctypes.Structure):
    """Field-for-field copy of `galic_params` (include/galic_types.h)."""

    _fields_ = [
        ("G", ctypes.c_double), ("V200", ctypes.c_double), ("R200", ctypes.c_double), ("M200", ctypes.c_double),
        ("Halo_Mass", ctypes.c_double), ("Halo_A", ctypes.c_double),
        ("Disk_Mass", ctypes.c_double), ("Disk_H", ctypes.c_double), ("Disk_Z0", ctypes.c_double),
        ("Bulge_Mass", ctypes.c_double), ("Bulge_A", ctypes.c_double),
        ("Rmax", ctypes.c_double),
        ("MType", _d4), ("NType", _d4),
        ("SampleForceNhalo", ctypes.c_int), ("SampleForceNdisk", ctypes.c_int), ("SampleForceNbulge", ctypes.c_int),
        ("TypeOfVelocityStructure", _i4),
        ("TimeStepFactorOrbit", ctypes.c_double), ("TimeStepFactorCellCross", ctypes.c_double),
        ("MaxVelInUnitsVesc", ctypes.c_double),
        ("MaxStepsPerOrbit", ctypes.c_int),
        ("LowerDispLimit", ctypes.c_double),
        ("MinParticlesPerBinForDensityMeasurement", ctypes.c_double),
        ("MinParticlesPerBinForDispersionMeasurement", ctypes.c_double),
        ("fac_value_rs", _d4), ("fac_value_ts", _d4), ("fac_value_qs", _d4), ("fac_value_ps", _d4),
        ("FG_Nbin", ctypes.c_int), ("FG_Rmin", ctypes.c_double), ("FG_Fac", ctypes.c_double),
        ("DG_MaxLevel", ctypes.c_int), ("DG_Rmin", ctypes.c_double), ("DG_Fac", ctypes.c_double),
        ("EG_MaxLevel", ctypes.c_int), ("EG_Rmin", ctypes.c_double), ("EG_Fac", ctypes.c_double),
        ("ErrTolTheta", ctypes.c_double), ("ForceSoftening", ctypes.c_double),
    ]

    def copy(self):
        c = GalicParams()
        ctypes.memmove(ctypes.byref(c), ctypes.byref(self), ctypes.sizeof(GalicParams))
        return c


GALIC_OK = 0
ERRORS = {-1: "GALIC_ERR_ARG", -2: "GALIC_ERR_HIP", -3: "GALIC_ERR_STEPCAP", -4: "GALIC_ERR_STATE",
          -5: "GALIC_ERR_NOMEM"}

# every entry point declared in include/galic_hip.h
EXPORTS = [
    "galic_last_error", "galic_version", "galic_grid_create", "galic_density_geometry",
    "galic_create", "galic_destroy", "galic_set_force_tables", "galic_set_vstream", "galic_set_targets",
    "galic_set_disp_tables", "galic_sample_targets",
    "galic_write_fit_line", "galic_write_densfield", "galic_write_snapshot",
    "galic_set_particles", "galic_set_dispersions", "galic_get_particles", "galic_calc_all_response_fields", "galic_rep",
    "galic_get_stack", "galic_get_fit", "galic_orbit_responses", "galic_fp32_orbit_responses", "galic_field_at", "galic_field_at_static",
    "galic_fp32_calc_all_response_fields", "galic_fp32_rep", "galic_fp32_iterate", "galic_fp32_autotune", "galic_fp32_set_launch", "galic_fp32_set_fresh_trials", "galic_fp32_trial_merits", "galic_fp32_set_merit_form", "galic_fp32_set_schedule", "galic_fp32_set_merit_launch", "galic_fp32_autotune_schedule", "galic_fp32_merit_profile", "galic_tree_forces", "galic_tree_fill", "galic_direct_forces", "galic_direct_forces_mode", "galic_tree_last_timings",
    "galic_timings", "galic_reset_timings", "galic_copy_probe", "galic_bind",
    "produce_orbit_response_field_mod", "produce_orbit_response_field",
    "forcegrid_get_potential", "forcegrid_get_escape_speed", "forcegrid_get_acceleration", "smooth_stack",
    "forcegrid_get_cell", "densitygrid_get_cell", "energygrid_get_cell", "get_timestep",
    "calc_stack_difference", "calc_stack_difference_mod",
    "galic_set_tree_sources", "galic_sample_tree_sources", "galic_retabulate", "galic_get_force_tables",
    "forcedensitygrid_create", "forcedensitygrid_calculate", "galic_bind_tree_particles", "force_treebuild",
    "force_update_node_recursive", "force_treeevaluate", "gravity",
]

_P = ctypes.POINTER
_dp = _P(ctypes.c_double)
_ip = _P(ctypes.c_int)
_lp = _P(ctypes.c_longlong)
_vp = ctypes.c_void_p


class GalicError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib():
    """The loaded C-ABI library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(the HIP library is required; there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    L.galic_last_error.restype = ctypes.c_char_p
    L.galic_last_error.argtypes = []
    L.galic_version.restype = ctypes.c_int
    sig = {
        "galic_grid_create": [_P(GalicParams), ctypes.c_double, ctypes.c_double],
        "galic_density_geometry": [_P(GalicParams), _dp, _dp, _dp],
        "galic_create": [_P(GalicParams), ctypes.c_int, _P(_vp)],
        "galic_destroy": [_vp],
        "galic_set_force_tables": [_vp, _dp, _dp, _dp],
        "galic_set_vstream": [_vp, ctypes.c_int, _dp],
        "galic_set_targets": [_vp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp],
        "galic_set_disp_tables": [_vp, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, _dp, _dp],
        "galic_write_fit_line": [_vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_double, ctypes.c_longlong,
                                 ctypes.c_longlong],
        "galic_write_densfield": [_vp, ctypes.c_int, ctypes.c_char_p],
        "galic_write_snapshot": [_vp, ctypes.c_char_p, ctypes.c_int, _dp],
        "galic_sample_targets": [_vp, ctypes.c_int, ctypes.c_longlong, ctypes.c_uint, ctypes.c_uint, ctypes.c_double,
                                 _dp, _dp, _dp, _lp, _P(ctypes.c_float)],
        "galic_set_particles": [_vp, ctypes.c_int, _dp, _dp, _dp, _dp, _ip, _dp],
        "galic_get_particles": [_vp, _dp, _ip, _dp, _dp, _dp, _dp],
        "galic_set_dispersions": [_vp, ctypes.c_int, _dp],
        "galic_calc_all_response_fields": [_vp],
        "galic_rep": [_vp, ctypes.c_int, _ip, _dp, _ip, _dp],
        "galic_get_stack": [_vp, ctypes.c_int, ctypes.c_int, _dp],
        "galic_get_fit": [_vp, ctypes.c_int, _dp],
        "galic_orbit_responses": [_vp, ctypes.c_int, _dp, _dp, _dp, _dp, _ip, _dp, _dp, _ip, _lp, _dp, _dp],
        "galic_fp32_orbit_responses": [_vp, ctypes.c_int, _dp, _dp, _dp, _dp, _ip, _dp, _dp, _ip, _lp, _dp, _dp],
        "galic_field_at": [_vp, ctypes.c_int, _dp, _dp, _dp],
        "galic_field_at_static": [_vp, ctypes.c_int, _dp, _dp, _dp],
        "galic_fp32_calc_all_response_fields": [_vp],
        "galic_fp32_rep": [_vp, ctypes.c_int, _ip, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_ulonglong, _ip],
        "galic_fp32_iterate": [_vp, ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double,
                               ctypes.c_ulonglong, _ip],
        "galic_fp32_autotune": [_vp, ctypes.c_int, _ip, _ip],
        "galic_fp32_set_launch": [_vp, ctypes.c_int, ctypes.c_int],
        "galic_fp32_set_fresh_trials": [_vp, ctypes.c_int],
        "galic_fp32_trial_merits": [_vp, ctypes.c_int, ctypes.c_int, _dp],
        "galic_fp32_set_merit_form": [_vp, ctypes.c_int],
        "galic_fp32_set_schedule": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int],
        "galic_fp32_set_merit_launch": [_vp, ctypes.c_int],
        "galic_fp32_autotune_schedule": [_vp, ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_ulonglong, _ip, _ip, _ip],
        "galic_fp32_merit_profile": [_vp, ctypes.c_int, _P(ctypes.c_ulonglong)],
        "galic_tree_forces": [_P(GalicParams), ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp, _dp,
                              ctypes.c_int, _lp],
        "galic_tree_fill": [_P(GalicParams), ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, ctypes.c_int,
                            _lp],
        "galic_direct_forces": [_P(GalicParams), ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp, _dp],
        "galic_direct_forces_mode": [_P(GalicParams), ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int],
        "galic_tree_last_timings": [_dp],
        "galic_timings": [_vp, _dp],
        "galic_reset_timings": [_vp],
        "galic_copy_probe": [ctypes.c_int, ctypes.c_size_t, ctypes.c_int, _dp],
        "galic_bind": [_vp],
        "galic_set_tree_sources": [_vp, ctypes.c_int, _dp, _dp],
        "galic_sample_tree_sources": [_vp, ctypes.c_uint],
        "galic_retabulate": [_vp, ctypes.c_int, _lp],
        "galic_get_force_tables": [_vp, _dp, _dp, _dp],
        "galic_bind_tree_particles": [ctypes.c_int, _dp, _dp, _ip, _dp, _dp],
    }
    for name, args in sig.items():
        if not hasattr(L, name):
            continue
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    _lib = L
    return L


def check(rc):
    if rc != GALIC_OK:
        raise GalicError(rc, lib().galic_last_error().decode(errors="replace"))
    return rc


def dptr(a):
    """Pointer to a C-contiguous float64 array (or None)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "expected C-contiguous float64"
    return a.ctypes.data_as(_dp)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"], "expected C-contiguous int32"
    return a.ctypes.data_as(_ip)


def lptr(a):
    if a is None:
        return None
    assert a.dtype == np.int64 and a.flags["C_CONTIGUOUS"], "expected C-contiguous int64"
    return a.ctypes.data_as(_lp)
