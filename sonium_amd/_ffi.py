"""ctypes bindings of libsbh.so (the C ABI declared in include/sbh.h).

The library is loaded from sonium_amd/_lib/libsbh.so.  There is no fallback:
if the HIP library is missing, importing the bindings raises immediately.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libsbh.so")

MEM_HOST = 0
MEM_DEVICE = 1

KERNEL_CLASSES = {
    "assemble_far": 0,
    "assemble_near": 1,
    "assemble_self": 2,
    "rhs": 3,
    "matvec": 4,
    "krylov_vector": 5,
    "block": 6,
    "lowp": 7,
    "field": 8,
}


class c64(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


class physics_t(ctypes.Structure):
    _fields_ = [(name, ctypes.c_double) for name in (
        "frequency", "speed_of_sound", "density", "wave_number", "omega", "wave_length",
        "harmonic_factor", "pressure_factor", "tau")]


class nc_info_t(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int64) for name in (
        "num_nodes", "num_elements", "num_evaluation", "num_frequencies", "num_curve_points",
        "num_plane_waves", "num_point_sources", "num_boundary_specs", "num_node_files", "num_element_files")] + [
        ("mesh_loaded", ctypes.c_int32), ("main_params_i", ctypes.c_int32 * 9),
        ("main_params_ii", ctypes.c_double * 7), ("speed_of_sound", ctypes.c_double),
        ("density", ctypes.c_double), ("reference_pressure", ctypes.c_double),
        ("has_symmetry", ctypes.c_int32), ("symmetry_flags", ctypes.c_int32 * 3),
        ("symmetry_origin", ctypes.c_double * 3)]


class incident_source_t(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("pad", ctypes.c_int32), ("vec", ctypes.c_double * 3),
                ("re", ctypes.c_double), ("im", ctypes.c_double)]


class gmres_config_t(ctypes.Structure):
    _fields_ = [("max_iterations", ctypes.c_int64), ("restart", ctypes.c_int64),
                ("tolerance", ctypes.c_double), ("convergence_check_interval", ctypes.c_int64),
                ("exact_residual_check_interval", ctypes.c_int64), ("min_restart", ctypes.c_int64),
                ("max_restart", ctypes.c_int64)]


class gmres_solution_t(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int64), ("restarts", ctypes.c_int64),
                ("residual", ctypes.c_double), ("estimated_residual", ctypes.c_double),
                ("backward_error", ctypes.c_double), ("converged", ctypes.c_int32),
                ("status", ctypes.c_int32), ("spmv_count", ctypes.c_int64),
                ("reorthogonalizations", ctypes.c_int64), ("exact_residual_checks", ctypes.c_int64)]


class block_gmres_config_t(ctypes.Structure):
    _fields_ = [("max_iterations", ctypes.c_int64), ("restart", ctypes.c_int64), ("tolerance", ctypes.c_double),
                ("true_residual_check", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("exact_residual_check_interval", ctypes.c_int64)]


class block_gmres_info_t(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int64), ("converged", ctypes.c_int32), ("status", ctypes.c_int32),
                ("spmv_count", ctypes.c_int64), ("exact_residual_checks", ctypes.c_int64),
                ("restart_cycles", ctypes.c_int64), ("max_basis", ctypes.c_int64), ("deflated", ctypes.c_int64),
                ("cholesky_qr_blocks", ctypes.c_int64), ("chain_qr_blocks", ctypes.c_int64)]


class resource_budget_t(ctypes.Structure):
    _fields_ = [("max_host_bytes", ctypes.c_int64), ("max_device_bytes", ctypes.c_int64),
                ("max_active_work", ctypes.c_int64)]


class resource_estimate_t(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int64) for name in (
        "host_bytes", "device_bytes", "active_work", "operator_bytes", "rhs_bytes", "scratch_bytes")]


class gmres_ir_config_t(ctypes.Structure):
    _fields_ = [("max_outer_iterations", ctypes.c_int64), ("outer_tolerance", ctypes.c_double),
                ("max_inner_iterations", ctypes.c_int64), ("inner_tolerance", ctypes.c_double),
                ("restart_size", ctypes.c_int64)]


class gmres_ir_status_t(ctypes.Structure):
    _fields_ = [("outer_iterations", ctypes.c_int64), ("total_inner_iterations", ctypes.c_int64),
                ("estimated_residual", ctypes.c_double), ("final_outer_residual", ctypes.c_double),
                ("backward_error", ctypes.c_double), ("converged", ctypes.c_int32),
                ("status", ctypes.c_int32), ("restart_cycles", ctypes.c_int64),
                ("fp64_products", ctypes.c_int64), ("lowp_products", ctypes.c_int64)]


class batch_solve_config_t(ctypes.Structure):
    _fields_ = [("format", ctypes.c_int32), ("use_graph", ctypes.c_int32), ("restart", ctypes.c_int64),
                ("max_cycles", ctypes.c_int64), ("tolerance", ctypes.c_double),
                ("inner_tolerance", ctypes.c_double), ("use_guess", ctypes.c_int32), ("pad", ctypes.c_int32)]


class batch_column_status_t(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int64), ("cycles", ctypes.c_int64),
                ("reorthogonalizations", ctypes.c_int64), ("residual", ctypes.c_double),
                ("converged", ctypes.c_int32), ("pad", ctypes.c_int32), ("completion_ms", ctypes.c_double)]


class batch_tuning_t(ctypes.Structure):
    _fields_ = [("rows_per_item", ctypes.c_int32), ("blocks_per_cu", ctypes.c_int32),
                ("chunk_elems", ctypes.c_int32), ("pad", ctypes.c_int32)]


_vp = ctypes.c_void_p
_pp = ctypes.POINTER(ctypes.c_void_p)
_dp = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.c_int64
_int = ctypes.c_int

class gcro_dr_config_t(ctypes.Structure):
    _fields_ = [("max_iterations", ctypes.c_int64), ("restart", ctypes.c_int64), ("tolerance", ctypes.c_double)]


# sbh_apply_fn: int (*)(void* user, const sbh_c64* x, sbh_c64* y, int64_t n, int memory)
APPLY_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                            ctypes.c_int)

# name: (restype, argtypes)
SIGNATURES = {
    "sbh_last_error": (ctypes.c_char_p, []),
    "sbh_version": (ctypes.c_char_p, []),
    "sbh_physics_new": (_int, [ctypes.c_double, ctypes.c_double, ctypes.c_double, _int,
                               ctypes.POINTER(physics_t)]),
    "sbh_physics_from_wavenumber": (_int, [ctypes.c_double, ctypes.c_double, ctypes.c_double, _int,
                                           ctypes.POINTER(physics_t)]),
    "sbh_burton_miller_beta_scaled": (c64, [ctypes.POINTER(physics_t), ctypes.c_double]),
    "sbh_context_create": (_int, [_int, _pp]),
    "sbh_context_destroy": (None, [_vp]),
    "sbh_context_synchronize": (_int, [_vp]),
    "sbh_context_set_profiling": (_int, [_vp, _int]),
    "sbh_context_set_device_budget": (_int, [_vp, _i64]),
    "sbh_context_device_budget": (_int, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "sbh_admission_estimate": (_int, [_i64, _i64, _i64, _i64, ctypes.POINTER(resource_estimate_t)]),
    "sbh_admission_admit": (_int, [ctypes.POINTER(resource_budget_t), _i64, _i64, _i64, _i64,
                                   ctypes.POINTER(resource_estimate_t)]),
    "sbh_context_inject_failure": (_int, [_vp, _int]),
    "sbh_live_device_bytes": (_i64, []),
    "sbh_context_kernel_stats": (_int, [_vp, _int, ctypes.POINTER(_i64), _dp]),
    "sbh_context_reset_stats": (_int, [_vp]),
    "sbh_vector_create": (_int, [_vp, _i64, _pp]),
    "sbh_vector_destroy": (None, [_vp]),
    "sbh_vector_length": (_i64, [_vp]),
    "sbh_vector_data": (_vp, [_vp]),
    "sbh_vector_upload": (_int, [_vp, _vp, _i64, _i64]),
    "sbh_vector_download": (_int, [_vp, _vp, _i64, _i64]),
    "sbh_mesh_create": (_int, [_vp, _vp, _i64, _vp, _i64, _vp, _vp, _vp, _vp, _pp]),
    "sbh_mesh_destroy": (None, [_vp]),
    "sbh_mesh_num_elements": (_i64, [_vp]),
    "sbh_mesh_geometry": (_int, [_vp, _vp, _vp, _vp]),
    "sbh_workspace_create": (_int, [_vp, _vp, _int, _pp]),
    "sbh_workspace_destroy": (None, [_vp]),
    "sbh_workspace_stats": (_int, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64), _dp]),
    "sbh_system_create": (_int, [_vp, _i64, _pp]),
    "sbh_system_destroy": (None, [_vp]),
    "sbh_system_num_dofs": (_i64, [_vp]),
    "sbh_system_device_pointers": (_int, [_vp, _pp, ctypes.POINTER(_i64), _pp]),
    "sbh_system_download": (_int, [_vp, _i64, _i64, _vp, _vp]),
    "sbh_system_set_rhs": (_int, [_vp, _vp, _int]),
    "sbh_workspace_assemble": (_int, [_vp, ctypes.POINTER(physics_t), c64, _vp]),
    "sbh_build_tbem_system": (_int, [_vp, _vp, ctypes.POINTER(physics_t), _vp]),
    "sbh_incident_plane_wave_rhs": (_int, [_vp, _vp, ctypes.POINTER(physics_t), c64, ctypes.c_double,
                                           _vp, _vp, _i64, _vp, _int, _int]),
    "sbh_workspace_assemble_derivative": (_int, [_vp, ctypes.POINTER(physics_t), c64, c64, ctypes.c_double,
                                                 ctypes.c_double, ctypes.c_double, ctypes.c_double, _vp]),
    "sbh_mesh_create_oriented": (_int, [_vp, _vp, _i64, _vp, _i64, _vp, _vp, _vp, _vp, _int, _pp]),
    "sbh_nc_project_load": (_int, [ctypes.c_char_p, _pp]),
    "sbh_nc_parse": (_int, [ctypes.c_char_p, ctypes.c_char_p, _pp]),
    "sbh_nc_project_destroy": (None, [_vp]),
    "sbh_nc_project_info": (_int, [_vp, ctypes.POINTER(nc_info_t)]),
    "sbh_nc_project_string": (_int, [_vp, _int, _i64, ctypes.c_char_p, _i64]),
    "sbh_nc_project_values": (_int, [_vp, _int, _vp, _i64, ctypes.POINTER(ctypes.c_int64)]),
    "sbh_nc_project_mesh": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sbh_nc_project_evaluation_centers": (_int, [_vp, _vp]),
    "sbh_nc_project_surface_mesh": (_int, [_vp, _vp, _pp]),
    "sbh_incident_rhs": (_int, [_vp, _vp, ctypes.POINTER(physics_t), c64, ctypes.c_double, _vp, _i64, _vp, _int,
                                _int]),
    "sbh_field_pressure_sources": (_int, [_vp, _vp, ctypes.POINTER(physics_t), _vp, _vp, _vp, _i64, _vp, _i64, _vp,
                                          _vp, _vp, _int]),
    "sbh_field_pressure": (_int, [_vp, _vp, ctypes.POINTER(physics_t), _vp, _vp, _vp, _vp, _i64, _vp, _i64, _vp,
                                  _vp, _vp, _int]),
    "sbh_dense_operator_create": (_int, [_vp, _pp]),
    "sbh_exact_operator_create": (_int, [_vp, ctypes.POINTER(physics_t), c64, ctypes.c_double, _pp]),
    "sbh_operator_create_callback": (_int, [_vp, _i64, APPLY_FN, _vp, _int, _pp]),
    "sbh_operator_destroy": (None, [_vp]),
    "sbh_operator_num_rows": (_i64, [_vp]),
    "sbh_operator_num_cols": (_i64, [_vp]),
    "sbh_operator_transpose_capability": (_int, [_vp]),
    "sbh_operator_hermitian_capability": (_int, [_vp]),
    "sbh_operator_apply": (_int, [_vp, _vp, _vp, _int]),
    "sbh_operator_apply_transpose": (_int, [_vp, _vp, _vp, _int]),
    "sbh_operator_apply_hermitian": (_int, [_vp, _vp, _vp, _int]),
    "sbh_operator_apply_block": (_int, [_vp, _vp, _vp, _i64, _int]),
    "sbh_operator_apply_block_noalloc": (_int, [_vp, _vp, _vp, _i64]),
    "sbh_precond_identity": (_int, [_vp, _i64, _pp]),
    "sbh_precond_jacobi": (_int, [_vp, _pp]),
    "sbh_precond_create_callback": (_int, [_vp, _i64, APPLY_FN, _vp, _int, _pp]),
    "sbh_precond_destroy": (None, [_vp]),
    "sbh_precond_apply": (_int, [_vp, _vp, _vp, _int]),
    "sbh_gmres": (_int, [_vp, _vp, _vp, _vp, ctypes.POINTER(gmres_config_t), _vp, _int,
                         ctypes.POINTER(gmres_solution_t)]),
    "sbh_gmres_block": (_int, [_vp, _vp, _vp, _i64, ctypes.POINTER(gmres_config_t), _vp, _int,
                               ctypes.POINTER(gmres_solution_t)]),
    "sbh_lowp_matrix_create": (_int, [_vp, _int, _pp]),
    "sbh_lowp_matrix_update": (_int, [_vp]),
    "sbh_lowp_matrix_destroy": (None, [_vp]),
    "sbh_lowp_matrix_apply": (_int, [_vp, _vp, _vp, _i64, _int]),
    "sbh_lowp_matrix_download": (_int, [_vp, _vp, _vp, ctypes.POINTER(_i64)]),
    "sbh_block_gmres": (_int, [_vp, _vp, _vp, _vp, _i64, ctypes.POINTER(block_gmres_config_t), _vp, _int,
                               ctypes.POINTER(block_gmres_info_t), _vp, _vp, _vp]),
    "sbh_gcro_subspace_create": (_int, [_vp, _i64, _i64, _pp]),
    "sbh_gcro_subspace_destroy": (None, [_vp]),
    "sbh_gcro_subspace_size": (_i64, [_vp]),
    "sbh_gcro_subspace_reset": (_int, [_vp]),
    "sbh_gcro_subspace_download": (_int, [_vp, _vp]),
    "sbh_gcro_dr": (_int, [_vp, _vp, _vp, _vp, _vp, ctypes.POINTER(gcro_dr_config_t), _vp, _int,
                           ctypes.POINTER(gmres_solution_t)]),
    "sbh_gmres_ir": (_int, [_vp, _vp, _vp, _i64, ctypes.POINTER(gmres_ir_config_t), _vp, _int,
                            ctypes.POINTER(gmres_ir_status_t)]),
    "sbh_batch_create": (_int, [_vp, _i64, _vp, _int, _pp]),
    "sbh_batch_destroy": (None, [_vp]),
    "sbh_batch_set_problem": (_int, [_vp, _vp, _vp, _vp, _vp, ctypes.c_double]),
    "sbh_batch_assemble": (_int, [_vp]),
    "sbh_batch_convert": (_int, [_vp, _int]),
    "sbh_batch_solve": (_int, [_vp, ctypes.POINTER(batch_solve_config_t), _vp]),
    "sbh_batch_set_guess": (_int, [_vp, _i64, _vp]),
    "sbh_batch_apply": (_int, [_vp, _int, _vp, _vp]),
    "sbh_batch_set_guesses": (_int, [_vp, _vp]),
    "sbh_batch_set_guesses_device": (_int, [_vp, _vp]),
    "sbh_batch_store_solutions": (_int, [_vp, _vp]),
    "sbh_batch_guess_gram": (_int, [_vp, _int, _int, _vp, _vp, _vp, _vp]),
    "sbh_batch_guess_combine": (_int, [_vp, _int, _vp]),
    "sbh_batch_learned_guesses": (_int, [_vp, _vp, _vp]),
    "sbh_batch_download_lowp": (_int, [_vp, _i64, _vp, _vp, ctypes.POINTER(_i64)]),
    "sbh_batch_download_all": (_int, [_vp, _int, _vp]),
    "sbh_batch_download_solution": (_int, [_vp, _i64, _vp]),
    "sbh_batch_download_rhs": (_int, [_vp, _i64, _vp]),
    "sbh_batch_download_matrix": (_int, [_vp, _i64, _vp]),
    "sbh_batch_info": (_int, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                              ctypes.POINTER(_i64), _dp]),
    "sbh_batch_get_tuning": (_int, [_vp, ctypes.POINTER(batch_tuning_t)]),
    "sbh_batch_set_tuning": (_int, [_vp, ctypes.POINTER(batch_tuning_t)]),
    "sbh_batch_autotune": (_int, [_vp, _vp, _int, ctypes.POINTER(_int), ctypes.POINTER(batch_tuning_t)]),
    "sbh_mlp_create": (_int, [_vp, ctypes.c_int32, ctypes.c_uint64, _pp]),
    "sbh_mlp_destroy": (None, [_vp]),
    "sbh_mlp_num_params": (_i64, [_vp]),
    "sbh_mlp_get_params": (_int, [_vp, _vp]),
    "sbh_mlp_set_params": (_int, [_vp, _vp]),
    "sbh_mlp_loss_and_gradients": (_int, [_vp, _vp, _vp, _i64, _dp, _vp]),
    "sbh_mlp_set_dataset": (_int, [_vp, _vp, _vp, _i64]),
    "sbh_mlp_train": (_int, [_vp, _i64, _i64, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, _vp]),
    "sbh_mlp_predict": (_int, [_vp, _vp, _i64, _vp, _int]),
    "sbh_debug_subelements": (_int, [_vp, _vp, ctypes.c_double, _vp, ctypes.POINTER(_int)]),
    "sbh_debug_device_sincos": (_int, [_vp, _vp, _i64, _vp, _vp]),
    "sbh_debug_device_rsqrt": (_int, [_vp, _vp, _i64, _vp]),
    "sbh_debug_device_copy": (_int, [_vp, _i64, _int]),
    "sbh_debug_mlp_predict_layered": (_int, [_vp, _vp, _i64, _vp]),
    "sbh_debug_hbm_read_rate": (_int, [_vp, _i64, _int, ctypes.POINTER(ctypes.c_double)]),
    "sbh_debug_fp64_fma_rate": (_int, [_vp, ctypes.POINTER(ctypes.c_double)]),
}

_lib = None


def load():
    """Load libsbh.so; raises OSError (loudly) when the HIP build is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(
            f"sonium_amd: HIP library not built ({LIB_PATH} missing); run "
            "`python -c 'import __graft_entry__ as g; g.build()'` or `make -C sonium_amd/csrc`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
