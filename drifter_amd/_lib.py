"""ctypes binding of libdrifter_hip.so (the C ABI declared in include/drifter_hip/drf.h).

The library is the only compute path: there is no CPU fallback.  Loading fails
loudly when the shared object has not been built.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdrifter_hip.so")

DRF_OK = 0
DRF_ERR_ARG = 1
DRF_ERR_HIP = 2
DRF_ERR_STATE = 3
DRF_ERR_UNSUPPORTED = 4
DRF_ERR_RANGE = 5

LAYOUT_PACKED = 0
LAYOUT_SOA = 1
LAYOUT_LINES = 2
LAYOUT_BLOCKS = 3
PREC_FP64 = 0
PREC_MIXED = 1

RK_EULER, RK_RK2, RK_SSPRK3, RK_RK4, RK_LSRK45, RK_DP45 = range(6)


class DrfPhysics(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("g", "nu_h", "nu_v", "kappa_v_T", "kappa_v_S", "rho_0", "alpha", "beta")]


# name -> (restype, argtypes); every symbol declared in drf.h
_P = C.c_void_p
_D = C.c_double
_I = C.c_int
_L = C.c_int64
_PD = C.POINTER(C.c_double)
SIGNATURES = {
    "drf_create": (_I, [_I, _I, _I, C.POINTER(_P)]),
    "drf_destroy": (None, [_P]),
    "drf_last_error": (C.c_char_p, []),
    "drf_default_physics": (None, [C.POINTER(DrfPhysics)]),
    "drf_set_physics": (_I, [_P, C.POINTER(DrfPhysics)]),
    "drf_basis": (_I, [_I, _PD, _PD, _PD]),
    "drf_initialize": (_I, [_P, _L, _PD, _PD, _PD, _D, _D, _D, _PD]),
    "drf_rhs": (_I, [_P, _D, _PD, _PD]),
    "drf_rhs_dev": (_I, [_P, _D, _P, _P, _I, _P]),
    "drf_layout_block_elements": (_I, [_I]),
    "drf_state_len": (C.c_int64, [_P, _I]),
    "drf_set_launch": (_I, [_P, _I, _I, _I]),
    "drf_get_launch": (_I, [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
    "drf_autotune": (_I, [_P, _P, _P, _I, _I, _P, C.POINTER(C.c_float)]),
    "drf_rk_step": (_I, [_P, _I, _D, _D, _PD]),
    "drf_rk_step_dev": (_I, [_P, _I, _D, _D, _P, _I, _I, _P]),
    "drf_max_wave_speed": (_I, [_P, _PD, _PD]),
    "drf_max_wave_speed_dev": (_I, [_P, _P, _I, _PD, _P]),
    "drf_max_wave_speed_derived_dev": (_I, [_P, _P, _P, _P, _L, _PD, _P]),
    "drf_max_wave_speed_noalloc_dev": (_I, [_P, _P, _I, _P, _P]),
    "drf_compute_dt": (_D, [_D, _D, _D, _D, _D]),
    "drf_mode_split_step": (_I, [_P, _D, _I, _PD]),
    "drf_mode_split_step_dev": (_I, [_P, _D, _I, _P, _P]),
    "drf_graph_rk": (_I, [_P, _I, _D, _P, _I, _I, _I, C.POINTER(_P)]),
    "drf_rk_step_adaptive_dev": (_I, [_P, _I, _D, _D, _D, _D, _P, _I, _P, _P, _P]),
    "drf_graph_rk_adaptive": (_I, [_P, _I, _D, _D, _D, _D, _P, _I, _I, _P, _P, C.POINTER(_P)]),
    "drf_graph_mode_split": (_I, [_P, _D, _I, _P, _I, C.POINTER(_P)]),
    "drf_graph_launch": (_I, [_P, _P]),
    "drf_graph_destroy": (None, [_P]),
    "drf_mesh_create": (_I, [_D, _D, _D, _D, _D, _D, C.POINTER(_P)]),
    "drf_mesh_destroy": (None, [_P]),
    "drf_mesh_build_uniform": (_I, [_P, _I, _I, _I]),
    "drf_mesh_refine": (_I, [_P, _L, _P, _P]),
    "drf_mesh_balance": (_I, [_P]),
    "drf_mesh_num_elements": (_L, [_P]),
    "drf_mesh_elements": (_I, [_P, _P, _P, _P]),
    "drf_mesh_element_bounds": (_I, [_P, _L, _PD]),
    "drf_mesh_face_connections": (_I, [_P, _P, _P, _P, _P, _P]),
    "drf_mesh_face_connections_2to1": (_I, [_P, _P, _P, _P, _P, _P]),
    "drf_face_plan_create": (_I, [_P, _L, _P, _P, _P, _P, _P, _P, _I, _PD, C.POINTER(_P)]),
    "drf_face_plan_destroy": (None, [_P]),
    "drf_face_plan_set_conforming_mode": (_I, [_P, _I]),
    "drf_face_plan_info": (_I, [_P, C.POINTER(_I), C.POINTER(_L), C.POINTER(_L)]),
    "drf_face_plan_groups": (_I, [_P, _P]),
    "drf_face_plan_lists": (_I, [_P, _P, _P]),
    "drf_face_term_dev": (_I, [_P, _P, _P, _I, _P]),
    "drf_face_group_fluxes_dev": (_I, [_P, _I, _P, _I, _P, _P, _P]),
    "drf_set_face_plan": (_I, [_P, _P]),
    "drf_flux_points": (_I, [_I, _L, _P, _P, _P, _PD, _P, _P]),
    "drf_face_map": (_I, [_I, _I, _P]),
    "drf_gather_face_traces": (_I, [_I, _L, _L, _L, _P, _P, _P, _P, _P]),
    "drf_conforming_flux_batch": (_I, [_I, _I, _PD, _L, _P, _P, _P, _P, _P, _P]),
    "drf_mortar_create": (_I, [_I, _I, _I, C.POINTER(_P)]),
    "drf_mortar_destroy": (None, [_P]),
    "drf_mortar_num_fine": (_I, [_P]),
    "drf_mortar_num_dofs": (_I, [_P]),
    "drf_mortar_operators": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "drf_mortar_flux_batch": (_I, [_P, _I, _PD, _L, _P, _P, _P, _P, _P, _P]),
    "drf_verify_mortar_conservation": (_I, [_P, _PD, _PD, _PD, _PD, _PD]),
    "drf_mortar_conservation_dev": (_I, [_P, _L, _P, _P, _P, _P, _P, _P]),
    "drf_estimator_create": (_I, [_I, _D, _D, _D, _I, _I, _I, C.POINTER(_P)]),
    "drf_estimator_destroy": (None, [_P]),
    "drf_element_char_length": (_I, [_L, _PD, _PD]),
    "drf_estimate_dev": (_I, [_P, _L, _P, _L, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "drf_project_to_children_dev": (_I, [_I, _I, _L, _P, _L, _P, _L, _P]),
    "drf_project_from_children_dev": (_I, [_I, _I, _L, _P, _L, _P, _L, _P]),
    "drf_transfer_solution_dev": (_I, [_I, _L, _P, _L, _L, _P, _P, _L, _P]),
    "drf_interpolate_boxes_dev": (_I, [_I, _L, _L, _P, _P, _P, _L, _P, _L, _P]),
    "drf_interpolate_boxes_fields_dev": (_I, [_I, _L, _L, _P, _P, _I, _P, _P, _P, _P, _P]),
    "drf_transfer_solution": (_I, [_I, _L, _PD, _L, _P, _PD, _I]),
    "drf_refine_transfer_map": (_I, [_L, _PD, _L, _PD, _P, _PD]),
    "drf_modal_matrices": (_I, [_I, _PD, _PD]),
    "drf_modal_coeffs_dev": (_I, [_I, _L, _P, _L, _P, _P]),
    "drf_modal_features_bf16_dev": (_I, [_I, _L, _P, _L, _D, _P, _P, _L, _I, _P]),
    "drf_mlp_create": (_I, [C.c_uint64, _I, C.POINTER(_P)]),
    "drf_mlp_destroy": (None, [_P]),
    "drf_mlp_set_params": (_I, [_P, _P]),
    "drf_mlp_get_params": (_I, [_P, _P]),
    "drf_mlp_forward_dev": (_I, [_P, _L, _P, _P, _P]),
    "drf_mlp_classify_dev": (_I, [_L, _P, _D, _D, _P, _P]),
    "drf_mlp_loss_grad_dev": (_I, [_P, _L, _P, _P, _P, _P, _P]),
    "drf_mlp_train_step_dev": (_I, [_P, _L, _P, _P, _D, _P, _P]),
    "drf_hydrostatic_pressure_dev": (_I, [_I, _L, _D, _P, _P, _L, _P, _P]),
    "drf_pressure_gradient_dev": (_I, [_I, _L, _I, _I, _D, _D] + [_P] * 8 + [_L, _P, _P, _P]),
    "drf_sigma_metrics_dev": (_I, [_L] + [_P] * 14),
    "drf_sigma_ops_metrics_dev": (_I, [_I, _L] + [_P] * 7 + [_L] + [_P] * 7),
    "drf_sigma_horizontal_gradient_dev": (_I, [_I, _L, _P, _P, _P, _L, _P, _P, _P]),
    "drf_sigma_vertical_gradient_dev": (_I, [_I, _L, _P, _P, _L, _P, _P]),
    "drf_sigma_divergence_dev": (_I, [_I, _L, _P, _P, _P, _L, _P, _P]),
    "drf_sigma_ale_correction_dev": (_I, [_I, _L, _P, _P, _P, _L, _P, _P]),
    "drf_sigma_material_derivative_dev": (_I, [_I, _L] + [_P] * 9 + [_L, _P, _P]),
    "drf_diagnose_omega_dev": (_I, [_I, _L, _P, _P, _L, _P, _P]),
    "drf_diagnose_omega_from_velocity_dev": (_I, [_I, _L, _P, _P, _P, _L, _P, _P]),
    "drf_omega_to_w_dev": (_I, [_L] + [_P] * 9),
    "drf_omega_to_w_simple_dev": (_I, [_L] + [_P] * 11),
    "drf_surface_omega": (_D, [_D] * 6),
    "drf_column_sigma_nodes": (_I, [_I, _PD]),
    "drf_diagnose_column_dev": (_I, [_I, _L, _P, _P, _P, _P, _P]),
    "drf_continuity_residual_dev": (_I, [_I, _L, _P, _P, _L, _P, _P]),
    "drf_vtk_create": (_I, [C.c_char_p, _I, _I, C.POINTER(_P)]),
    "drf_vtk_destroy": (None, [_P]),
    "drf_vtk_set_polynomial_order": (_I, [_P, _I]),
    "drf_vtk_set_mesh": (_I, [_P, _L, _PD]),
    "drf_vtk_add_field": (_I, [_P, C.c_char_p, _I, _I]),
    "drf_vtk_set_field": (_I, [_P, C.c_char_p, _I, _L, _PD]),
    "drf_vtk_set_point_field_dev": (_I, [_P, C.c_char_p, _L, _P, _L, _L, _P]),
    "drf_vtk_filename": (_I, [_P, _L, C.c_char_p, _L]),
    "drf_vtk_write": (_I, [_P, _L, _D]),
    "drf_vtk_write_timestep": (_I, [_P, _D]),
    "drf_vtk_finalize": (_I, [_P]),
    "drf_rhs_bytes": (_D, [_P]),
    "drf_num_elements": (_L, [_P]),
    "drf_order": (_I, [_P]),
}

_lib = None


class DrfError(RuntimeError):
    """Raised for a non-zero drf_status (the reference raises C++ exceptions)."""

    def __init__(self, status, msg):
        super().__init__(f"drf status {status}: {msg}")
        self.status = status


def load():
    """Load the HIP library (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m drifter_amd.build` "
            "(there is no CPU fallback)")
    # torch ships its own libamdhip64 (same soname, libamdhip64.so.7).  Loading
    # torch first makes the library bind to that already-loaded runtime, so a
    # process that also uses torch for device buffers has one HIP runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status):
    if status != DRF_OK:
        raise DrfError(status, load().drf_last_error().decode())
    return status


def dptr(a):
    """double* of a C-contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    assert a.dtype.name == "float64" and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_PD)
