"""ctypes binding of libsbslpm.so (declarations mirror include/sbslpm.h).

The library is the only compute path: if it is missing this module raises,
there is no fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsbslpm.so")

SBSLPM_OK, SBSLPM_E_ARG, SBSLPM_E_GEOMETRY, SBSLPM_E_HIP, SBSLPM_E_NOMEM = 0, 1, 2, 3, 4
ERROR_NAMES = {1: "SBSLPM_E_ARG", 2: "SBSLPM_E_GEOMETRY", 3: "SBSLPM_E_HIP", 4: "SBSLPM_E_NOMEM"}


class SbslpmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [
        ("levels", ctypes.c_int),
        ("interp_degree", ctypes.c_int),
        ("leaf_degree", ctypes.c_int),
        ("weight_rule", ctypes.c_int),
        ("rk4_combine", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("tracer_count", ctypes.c_int),
        ("flags", ctypes.c_int),
        ("theta", ctypes.c_double),
        ("kernel_eps", ctypes.c_double),
        ("omega", ctypes.c_double),
        ("delta_t", ctypes.c_double),
        ("topography", ctypes.c_int),
        ("swe_rule", ctypes.c_int),
        ("gravity", ctypes.c_double),
        ("mountain_height", ctypes.c_double),
        ("mountain_width", ctypes.c_double),
        ("mountain_lat", ctypes.c_double),
        ("mountain_lon", ctypes.c_double),
        ("swe_form", ctypes.c_int),
    ]


_V = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_D = ctypes.c_double
_PD = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int)
_PI64 = ctypes.POINTER(ctypes.c_int64)

# name -> (restype, argtypes); every function returns an int status unless noted
SIGNATURES = {
    "sbslpm_tree_build": (_I, [ctypes.POINTER(Params), ctypes.POINTER(_V)]),
    "sbslpm_tree_destroy": (None, [_V]),
    "sbslpm_tree_sizes": (_I, [_V, _PI64]),
    "sbslpm_tree_spec_points": (_I, [_V, _PD, _PD, _PD, _PD]),
    "sbslpm_tree_interaction_list": (_I, [_V, _PI]),
    "sbslpm_tree_panels": (_I, [_V, _PI, _PD]),
    "sbslpm_tree_locate": (_I, [_V, _D, _D, _D]),
    "sbslpm_tree_transfer_matrices": (_I, [_V, _PD, _PD, _PD]),
    "sbslpm_tree_dedupe": (_I, [_V, _PI, _PI, _PD, _PD, _PD, _PI64]),
    "sbslpm_cubed_sphere_midpoints": (_I, [_I, _PD, _PD, _PD, _PD]),
    "sbslpm_irreg_tree_build": (_I, [ctypes.c_int64, _PD, _PD, _PD, _I, _I, ctypes.c_double, _V]),
    "sbslpm_irreg_tree_destroy": (None, [_V]),
    "sbslpm_irreg_tree_sizes": (_I, [_V, _PI64]),
    "sbslpm_irreg_tree_panels": (_I, [_V, _PI, _PD]),
    "sbslpm_irreg_tree_points": (_I, [_V, _PI]),
    "sbslpm_irreg_tree_interaction_list": (_I, [_V, _PI]),
    "sbslpm_ctx_create": (_I, [ctypes.POINTER(Params), ctypes.POINTER(_V)]),
    "sbslpm_ctx_destroy": (None, [_V]),
    "sbslpm_ctx_set_stream": (_I, [_V, _V]),
    "sbslpm_ctx_set_precision": (_I, [_V, _I]),
    "sbslpm_last_error": (ctypes.c_char_p, [_V]),
    "sbslpm_synchronize": (_I, [_V]),
    "sbslpm_get_sizes": (_I, [_V, _PI64]),
    "sbslpm_spec_points": (_I, [_V, _PD, _PD, _PD, _PD]),
    "sbslpm_interaction_list": (_I, [_V, _PI]),
    "sbslpm_ctx_enable_timing": (_I, [_V, _I]),
    "sbslpm_kernel_time": (_I, [_V, _I, _PD, _PI64, _I]),
    "sbslpm_upward_pass": (_I, [_V, _V, _V, _V, _V]),
    "sbslpm_bve_vel_interactions": (_I, [_V] + [_V] * 14),
    "sbslpm_poisson_fmm_interactions": (_I, [_V, _V, _V, _V, _V]),
    "sbslpm_downward_pass": (_I, [_V, _V, _V, _V]),
    "sbslpm_downward_pass_3": (_I, [_V] + [_V] * 9),
    "sbslpm_bve_velocity": (_I, [_V] + [_V] * 11),
    "sbslpm_poisson_solve": (_I, [_V, _V, _V, _V]),
    "sbslpm_swe_vel_interactions": (_I, [_V] * 17),
    "sbslpm_swe_vel_interactions_2": (_I, [_V] * 17),
    "sbslpm_swe_velocity": (_I, [_V, _I] + [_V] * 12),
    "sbslpm_bve_back_rk4_step_ensemble": (_I, [_V, _I] + [_V] * 8 + [ctypes.c_double]),
    "sbslpm_bve_velocity_adjoint": (_I, [_V] * 6),
    "sbslpm_bve_back_rk4_step_adjoint": (_I, [_V] * 8),
    "sbslpm_bve_back_rk4_step_record": (_I, [_V] * 8),
    "sbslpm_bve_back_rk4_step_adjoint_stages": (_I, [_V] * 9),
    "sbslpm_bve_departure_to_target_adjoint": (_I, [_V] * 6),
    "sbslpm_swe_displacement": (_I, [_V, _V, _V, _V, ctypes.c_double] + [_V] * 6),
    "sbslpm_interp_points": (_I, [_V, ctypes.c_int64, _V, _V, _V, _I, _V, _V]),
    "sbslpm_interp_points_adjoint": (_I, [_V, ctypes.c_int64, _V, _V, _V, _I, _V, _V]),
    "sbslpm_autotune": (_I, [_V, _V, _V, _PD, _PI]),
    "sbslpm_autotune_ensemble": (_I, [_V, _I, _V, _V, _PD, _PI]),
    "sbslpm_unify_boundary_vals": (_I, [_V, _V, _V, ctypes.c_int64, _V]),
    "sbslpm_get_pair_shape": (_I, [_V, _PI, _PI]),
    "sbslpm_set_pair_shape": (_I, [_V, _I, _I]),
    "sbslpm_irreg_create": (_I, [_V, ctypes.c_int64, _PD, _PD, _PD, _I, _I, _V]),
    "sbslpm_irreg_destroy": (None, [_V]),
    "sbslpm_irreg_sizes": (_I, [_V, _PI64]),
    "sbslpm_irreg_upward_pass": (_I, [_V] * 4),
    "sbslpm_irreg_downward_pass_3": (_I, [_V] * 7),
    "sbslpm_irreg_swe_vel_interactions": (_I, [_V] * 12),
    "sbslpm_irreg_swe_velocity": (_I, [_V] * 7),
    "sbslpm_irreg_bve_velocity": (_I, [_V] * 6),
    "sbslpm_xyz_vel_to_uv_vel": (_I, [_V, ctypes.c_int64] + [_V] * 8),
    "sbslpm_ll_tendencies": (_I, [_V, _I, _I, _D, _D] + [_V] * 13),
    "sbslpm_ll_polar_filter": (_I, [_V, _I, _I, _D] + [_V] * 4),
    "sbslpm_ll_arrival_values": (_I, [_V, _I, _I, _D, _D] + [_V] * 14),
    "sbslpm_ll_field_averaging": (_I, [_V, _I, _I] + [_V] * 5 + [_D]),
    "sbslpm_ll_laplacian_diffusion": (_I, [_V, _I, _I, _D, _D] + [_V] * 5),
    "sbslpm_irreg_swe_vel_tend_ll": (_I, [_V, _I, _I, _D, _D] + [_V] * 14),
    "sbslpm_irreg_swe_back_rk4_step_ll": (_I, [_V, _I, _I, _D, _D, _D] + [_V] * 11 + [_D]),
    "sbslpm_swe_find_departure_points": (_I, [_V, _V, _V, _V, _V, ctypes.c_double, _V, _V, _V]),
    "sbslpm_swe_departure_to_target": (_I, [_V] * 8 + [_I] + [_V] * 4),
    "sbslpm_poisson_solve_targets": (_I, [_V, _V, _V, _I64, _V, _V, _V, _V]),
    "sbslpm_bve_velocity_targets": (_I, [_V, _V, _V, _I64, _V, _V, _V, _V, _V, _V]),
    "sbslpm_direct_sum_inv_lap": (_I, [_V, _I64, _V, _V, _V, _V, _V, _V, _I64, _I64]),
    "sbslpm_direct_sum_inv_lap_noalloc": (_I, [_V, _I64, _V, _V, _V, _V, _V, _V, _I64, _I64]),
    "sbslpm_direct_sum_inv_lap_targets": (_I, [_V, _I64, _V, _V, _V, _I64, _V, _V, _V, _V, _V, _V]),
    "sbslpm_direct_sum_bve_vel": (_I, [_V, _I64] + [_V] * 11 + [_I64, _I64]),
    "sbslpm_direct_sum_bve_vel_targets": (_I, [_V, _I64, _V, _V, _V, _I64] + [_V] * 8),
    "sbslpm_laplacian": (_I, [_V, _V, _V]),
    "sbslpm_laplacian_interactions": (_I, [_V] * 11),
    "sbslpm_apply_topography": (_I, [_V] * 12),
    "sbslpm_swe_tendency": (_I, [_V] * 14),
    "sbslpm_swe_height_correction": (_I, [_V] * 9 + [_D, _V]),
    "sbslpm_swe_back_rk4_step": (_I, [_V] * 13 + [_D]),
    "sbslpm_swe_back_rk4_step_record": (_I, [_V] * 12),
    "sbslpm_swe_back_rk4_step_adjoint_stages": (_I, [_V] * 15),
    "sbslpm_swe_back_rk4_step_adjoint": (_I, [_V] * 14),
    "sbslpm_bve_displacement": (_I, [_V] + [_V] * 7 + [_D] + [_V] * 7),
    "sbslpm_bve_find_departure_points": (_I, [_V, _V, _V, _V, _V, _D, _V, _V, _V]),
    "sbslpm_bve_departure_to_target": (_I, [_V, _V, _V, _V, _V, _V, _V, _I, _V, _V]),
    "sbslpm_bve_back_rk4_step": (_I, [_V] + [_V] * 10 + [_D]),
}

_lib = None


def load():
    """Loads libsbslpm.so (built in-tree by build.py); raises if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libsbslpm.so not built at {LIB_PATH}: run __graft_entry__.build() "
                               "(or python spectral-bslpm_amd/build.py)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def exported_symbols():
    return sorted(SIGNATURES)
