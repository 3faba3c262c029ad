"""ctypes binding of include/ipt_hip.h (libipt_hip.so)."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(HERE, "libipt_hip.so")


def library_path():
    return _LIB


class IptError(RuntimeError):
    def __init__(self, status, where=""):
        self.status = status
        msg = "%s failed: status %d" % (where, status)
        try:
            msg += " (%s)" % lib().ipt_cuda_status_string(status).decode()
        except Exception:
            pass
        super().__init__(msg)


_c = ctypes
_HIST_FIELDS = [
    ("davidson_step", _c.c_int), ("active_pair_index", _c.c_int),
    ("residual_before", _c.c_double), ("residual_after", _c.c_double),
    ("accepted", _c.c_int), ("basis_cols", _c.c_int),
    ("max_relative_eigen_residual", _c.c_double), ("max_relative_eigen_residual_index", _c.c_int),
    ("pair_28_residual", _c.c_double), ("pair_29_residual", _c.c_double),
    ("orthogonality_max_abs_error", _c.c_double), ("restarted", _c.c_int),
    ("best_so_far_updated", _c.c_int), ("best_so_far_step", _c.c_int),
    ("best_so_far_max_residual", _c.c_double), ("pair28_best_residual", _c.c_double),
    ("pair29_best_residual", _c.c_double), ("accept_global_ok", _c.c_int),
    ("accept_active_ok", _c.c_int), ("accept_locked_safe", _c.c_int),
    ("accept_reason", _c.c_char * 64), ("accepted_steps", _c.c_int), ("rejected_steps", _c.c_int),
    ("early_jump_to_continuation", _c.c_int), ("gap28_29", _c.c_double),
    ("relative_gap28_29", _c.c_double), ("cluster_residual_fro_before", _c.c_double),
    ("cluster_residual_fro_after", _c.c_double), ("cluster_residual_max_before", _c.c_double),
    ("cluster_residual_max_after", _c.c_double), ("ritz_overlap_28_28", _c.c_double),
    ("ritz_overlap_28_29", _c.c_double), ("ritz_overlap_29_28", _c.c_double),
    ("ritz_overlap_29_29", _c.c_double), ("ritz_overlap_swap_detected", _c.c_int),
    ("pair28_lock_state", _c.c_char * 16), ("pair29_lock_state", _c.c_char * 16),
    ("cluster_hard_locked", _c.c_int), ("cluster_soft_locked_count", _c.c_int),
]


class IPTDavidsonHistoryEntry(ctypes.Structure):
    _fields_ = _HIST_FIELDS


class IPTDavidsonSelectionEntry(ctypes.Structure):
    _fields_ = [
        ("davidson_step", _c.c_int), ("pair_index", _c.c_int), ("residual", _c.c_double),
        ("selected_by_residual", _c.c_int), ("selected_forced", _c.c_int),
        ("selected_auto_cluster", _c.c_int), ("skipped_converged_in_cluster", _c.c_int),
        ("skipped_linear_dependent", _c.c_int),
        ("skipped_locked_old_logic_should_not_happen", _c.c_int),
        ("correction_norm_before_ortho", _c.c_double), ("correction_norm_after_ortho", _c.c_double),
    ]


class IPTDavidsonBlockHistoryEntry(ctypes.Structure):
    _fields_ = [
        ("davidson_step", _c.c_int), ("active_pairs", _c.c_char * 64),
        ("accepted_corrections", _c.c_int), ("rejected_corrections", _c.c_int),
        ("correction_norm_before_ortho", _c.c_char * 256),
        ("correction_norm_after_ortho", _c.c_char * 256),
        ("residual_before_global", _c.c_double), ("residual_after_global", _c.c_double),
        ("pair_28_before", _c.c_double), ("pair_28_after", _c.c_double),
        ("pair_29_before", _c.c_double), ("pair_29_after", _c.c_double),
        ("accepted", _c.c_int), ("reject_reason", _c.c_char * 64),
        ("basis_cols_before", _c.c_int), ("basis_cols_after", _c.c_int),
        ("orthogonality_error_before", _c.c_double), ("orthogonality_error_after", _c.c_double),
        ("best_so_far_updated", _c.c_int), ("best_so_far_step", _c.c_int),
        ("best_so_far_max_residual", _c.c_double), ("pair28_best_residual", _c.c_double),
        ("pair29_best_residual", _c.c_double), ("accept_global_ok", _c.c_int),
        ("accept_active_ok", _c.c_int), ("accept_locked_safe", _c.c_int),
        ("accept_reason", _c.c_char * 64), ("accepted_steps", _c.c_int), ("rejected_steps", _c.c_int),
        ("early_jump_to_continuation", _c.c_int), ("gap28_29", _c.c_double),
        ("relative_gap28_29", _c.c_double), ("cluster_residual_fro_before", _c.c_double),
        ("cluster_residual_fro_after", _c.c_double), ("cluster_residual_max_before", _c.c_double),
        ("cluster_residual_max_after", _c.c_double), ("ritz_overlap_28_28", _c.c_double),
        ("ritz_overlap_28_29", _c.c_double), ("ritz_overlap_29_28", _c.c_double),
        ("ritz_overlap_29_29", _c.c_double), ("ritz_overlap_swap_detected", _c.c_int),
        ("pair28_lock_state", _c.c_char * 16), ("pair29_lock_state", _c.c_char * 16),
        ("cluster_hard_locked", _c.c_int), ("cluster_soft_locked_count", _c.c_int),
    ]


class IPTCudaResult(ctypes.Structure):
    _fields_ = [
        ("n", _c.c_int), ("k", _c.c_int), ("iterations", _c.c_int),
        ("vectors", _c.POINTER(_c.c_double)), ("values", _c.POINTER(_c.c_double)),
        ("basis_cols", _c.c_int), ("matvecs", _c.c_longlong),
        ("preparation_time_sec", _c.c_double), ("transfer_setup_time_sec", _c.c_double),
        ("iteration_time_sec", _c.c_double), ("rayleigh_ritz_time_sec", _c.c_double),
        ("davidson_time_sec", _c.c_double), ("solve_time_sec", _c.c_double),
        ("fixed_point_residual", _c.c_double), ("oversample", _c.c_int),
        ("rayleigh_ritz_used_qr", _c.c_int),
        ("basis_orthogonality_frobenius_error", _c.c_double),
        ("basis_orthogonality_max_abs_error", _c.c_double),
        ("ritz_vectors_orthogonality_frobenius_error", _c.c_double),
        ("ritz_vectors_orthogonality_max_abs_error", _c.c_double),
        ("basis_has_nan_or_inf", _c.c_int), ("ritz_vectors_has_nan_or_inf", _c.c_int),
        ("davidson_attempted", _c.c_int), ("davidson_accepted", _c.c_int),
        ("davidson_target_index", _c.c_int), ("davidson_residual_before", _c.c_double),
        ("davidson_residual_after", _c.c_double), ("davidson_denom_clip", _c.c_double),
        ("adaptive_block_enabled", _c.c_int), ("adaptive_coupling_tau", _c.c_double),
        ("adaptive_limit_hit", _c.c_int), ("adaptive_target_block_start", _c.c_int),
        ("adaptive_target_block_end", _c.c_int), ("adaptive_added_indices", _c.c_char * 256),
        ("davidson_history_count", _c.c_int),
        ("davidson_history", _c.POINTER(IPTDavidsonHistoryEntry)),
        ("davidson_restart_count", _c.c_int), ("davidson_selection_history_count", _c.c_int),
        ("davidson_selection_history", _c.POINTER(IPTDavidsonSelectionEntry)),
        ("davidson_block_history_count", _c.c_int),
        ("davidson_block_history", _c.POINTER(IPTDavidsonBlockHistoryEntry)),
        ("best_so_far_enabled", _c.c_int), ("best_so_far_updated", _c.c_int),
        ("best_so_far_update_count", _c.c_int), ("best_so_far_step", _c.c_int),
        ("best_so_far_source", _c.c_char * 32), ("best_so_far_basis_cols", _c.c_int),
        ("best_so_far_max_residual", _c.c_double), ("best_so_far_max_residual_index", _c.c_int),
        ("final_returned_from_best_so_far", _c.c_int), ("pair28_best_residual", _c.c_double),
        ("pair29_best_residual", _c.c_double), ("accepted_steps", _c.c_int),
        ("rejected_steps", _c.c_int), ("early_jump_to_continuation", _c.c_int),
        ("gap28_29", _c.c_double), ("relative_gap28_29", _c.c_double),
        ("cluster_residual_fro", _c.c_double), ("cluster_residual_max", _c.c_double),
        ("pair28_lock_state", _c.c_char * 16), ("pair29_lock_state", _c.c_char * 16),
        ("cluster_hard_locked", _c.c_int), ("cluster_soft_locked_count", _c.c_int),
    ]


def _two_phase_struct(name, lo):
    """Result structs of include/ipt_hip_mixed.h (reference
    ipt_mixed_precision_cuda.cu:19-59): ints, doubles, device pointers."""
    return type(name, (_c.Structure,), {"_fields_": [
        ("n", _c.c_int), ("k", _c.c_int), ("%s_iterations" % lo, _c.c_int), ("fp64_iterations", _c.c_int),
        ("%s_time_sec" % lo, _c.c_double), ("fp64_time_sec", _c.c_double), ("total_time_sec", _c.c_double),
        ("%s_fixed_point_residual" % lo, _c.c_double), ("fp64_fixed_point_residual", _c.c_double),
        ("d_vectors", _c.c_void_p), ("d_values", _c.c_void_p)]})


IPTMixedCudaDeviceResult = _two_phase_struct("IPTMixedCudaDeviceResult", "fp16")
IPTMixedFP32CudaDeviceResult = _two_phase_struct("IPTMixedFP32CudaDeviceResult", "fp32")
IPTMixedTF32CudaDeviceResult = _two_phase_struct("IPTMixedTF32CudaDeviceResult", "tf32")


class IPTRayleighMixedCudaDeviceResult(_c.Structure):
    """Reference ipt_rayleigh_mixed_cuda.cu:19-34."""
    _fields_ = [
        ("n", _c.c_int), ("k", _c.c_int), ("tf32_iterations", _c.c_int), ("fp64_iterations", _c.c_int),
        ("tf32_time_sec", _c.c_double), ("orthogonalize_time_sec", _c.c_double),
        ("rayleigh_transform_time_sec", _c.c_double), ("fp64_time_sec", _c.c_double),
        ("backtransform_time_sec", _c.c_double), ("total_time_sec", _c.c_double),
        ("tf32_fixed_point_residual", _c.c_double), ("fp64_fixed_point_residual", _c.c_double),
        ("d_vectors", _c.c_void_p), ("d_values", _c.c_void_p)]


class ACXCudaResult(_c.Structure):
    """Reference IPT/src/ACX/acx_cuda.cu:20-33."""
    _fields_ = [
        ("n", _c.c_int), ("k", _c.c_int), ("iterations", _c.c_int), ("f_calls", _c.c_int),
        ("converged", _c.c_int), ("time_sec", _c.c_double), ("iteration_time_sec", _c.c_double),
        ("finalization_time_sec", _c.c_double), ("max_residual", _c.c_double),
        ("fixed_point_residual", _c.c_double), ("d_vectors", _c.c_void_p), ("d_values", _c.c_void_p)]


class ACXMixedTF32CudaResult(_c.Structure):
    """Reference IPT/src/ACX/acx_mixed_precision_cuda.cu:19-37."""
    _fields_ = [
        ("n", _c.c_int), ("k", _c.c_int), ("tf32_iterations", _c.c_int), ("fp64_iterations", _c.c_int),
        ("tf32_f_calls", _c.c_int), ("fp64_f_calls", _c.c_int), ("tf32_converged", _c.c_int),
        ("fp64_converged", _c.c_int), ("tf32_time_sec", _c.c_double), ("fp64_time_sec", _c.c_double),
        ("total_time_sec", _c.c_double), ("finalization_time_sec", _c.c_double),
        ("tf32_fixed_point_residual", _c.c_double), ("fp64_fixed_point_residual", _c.c_double),
        ("fp64_max_residual", _c.c_double), ("d_vectors", _c.c_void_p), ("d_values", _c.c_void_p)]


class ACXRayleighMixedTF32CudaResult(_c.Structure):
    """Reference IPT/src/ACX/acx_rayleigh_mixed_precision_cuda.cu:19-40."""
    _fields_ = [
        ("n", _c.c_int), ("k", _c.c_int), ("tf32_iterations", _c.c_int), ("fp64_iterations", _c.c_int),
        ("tf32_f_calls", _c.c_int), ("fp64_f_calls", _c.c_int), ("tf32_converged", _c.c_int),
        ("fp64_converged", _c.c_int), ("tf32_time_sec", _c.c_double), ("orthogonalize_time_sec", _c.c_double),
        ("rayleigh_transform_time_sec", _c.c_double), ("fp64_time_sec", _c.c_double),
        ("finalization_time_sec", _c.c_double), ("backtransform_time_sec", _c.c_double),
        ("total_time_sec", _c.c_double), ("tf32_fixed_point_residual", _c.c_double),
        ("fp64_fixed_point_residual", _c.c_double), ("fp64_max_residual", _c.c_double),
        ("d_vectors", _c.c_void_p), ("d_values", _c.c_void_p)]


class IPTFciCsc(_c.Structure):
    """include/ipt_hip_fci.h"""
    _fields_ = [
        ("n", _c.c_int), ("nnz", _c.c_int), ("col_ptr", _c.POINTER(_c.c_int)), ("row_ind", _c.POINTER(_c.c_int)),
        ("values", _c.POINTER(_c.c_double)), ("diagonal", _c.POINTER(_c.c_double)), ("perm", _c.POINTER(_c.c_int)),
        ("matrix_inf_norm", _c.c_double), ("norb", _c.c_int), ("nelec", _c.c_int), ("neleca", _c.c_int),
        ("nelecb", _c.c_int), ("na", _c.c_int), ("nb", _c.c_int)]


_dp = _c.POINTER(_c.c_double)
_ip = _c.POINTER(_c.c_int)
_dpp = _c.POINTER(_dp)
_vp = _c.c_void_p
_R = _c.POINTER(IPTCudaResult)

# name -> (restype, argtypes); every symbol declared in include/*.h
SIGNATURES = {
    "ipt_cuda_status_string": (_c.c_char_p, [_c.c_int]),
    "ipt_cuda_free_result": (None, [_R]),
    "ipt_cuda_free_device_result": (None, [_vp, _vp]),
    "ipt_cuda": (_c.c_int, [_dp, _c.c_int, _c.c_int, _c.c_int, _R]),
    "ipt_cuda_tol": (_c.c_int, [_dp, _c.c_int, _c.c_int, _c.c_double, _c.c_int, _R]),
    "ipt_cuda_with_initial": (_c.c_int, [_dp, _dp, _c.c_int, _c.c_int, _c.c_int, _R]),
    "ipt_cuda_with_initial_tol": (_c.c_int, [_dp, _dp, _c.c_int, _c.c_int, _c.c_double, _c.c_int, _R]),
    "ipt_cuda_device": (_c.c_int, [_vp, _c.c_int, _c.c_int, _c.c_int, _vp, _c.POINTER(_vp), _c.POINTER(_vp)]),
    "ipt_cuda_device_tol": (_c.c_int, [_vp, _c.c_int, _c.c_int, _c.c_double, _c.c_int, _vp,
                                       _c.POINTER(_vp), _c.POINTER(_vp), _ip, _dp]),
    "ipt_cuda_sparse_csc": (_c.c_int, [_ip, _ip, _dp, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _R]),
    "ipt_cuda_sparse_csc_tol": (_c.c_int, [_ip, _ip, _dp, _c.c_int, _c.c_int, _c.c_int, _c.c_double,
                                           _c.c_int, _R]),
    "ipt_cuda_sparse_csc_with_initial": (_c.c_int, [_ip, _ip, _dp, _dp, _c.c_int, _c.c_int, _c.c_int,
                                                    _c.c_int, _R]),
    "ipt_cuda_sparse_csc_with_initial_tol": (_c.c_int, [_ip, _ip, _dp, _dp, _c.c_int, _c.c_int, _c.c_int,
                                                        _c.c_double, _c.c_int, _R]),
    "ipt_cuda_sparse_csc_device": (_c.c_int, [_vp, _vp, _vp, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _vp,
                                              _c.POINTER(_vp), _c.POINTER(_vp)]),
    "ipt_cuda_sparse_csc_device_tol": (_c.c_int, [_vp, _vp, _vp, _c.c_int, _c.c_int, _c.c_int,
                                                  _c.c_double, _c.c_int, _vp, _c.POINTER(_vp),
                                                  _c.POINTER(_vp), _ip, _dp]),
    "ipt_hip_create_handle": (_c.c_int, [_c.POINTER(_vp), _vp]),
    "ipt_hip_destroy_handle": (_c.c_int, [_vp]),
    # include/ipt_hip_ext.h
    "ipt_hip_generate_dense": (_c.c_int, [_vp, _c.c_int, _c.c_ulonglong, _c.c_double, _c.c_double,
                                          _c.c_double, _vp]),
    "ipt_hip_generate_eps": (_c.c_int, [_vp, _c.c_int, _c.c_double, _c.c_ulonglong, _vp]),
    "ipt_hip_set_diagonal_gaps": (_c.c_int, [_vp, _c.c_int, _ip, _c.c_int, _c.c_double, _vp]),
    "ipt_hip_lowp_product": (_c.c_int, [_vp, _c.c_int, _c.c_int, _vp, _vp, _vp, _vp, _c.c_int, _vp, _vp]),
    "ipt_hip_dense_product": (_c.c_int, [_vp, _c.c_int, _vp, _c.c_int, _vp, _c.c_int, _vp]),
    "ipt_hip_profile_enable": (None, [_c.c_int]),
    "ipt_hip_profile_enable_only": (_c.c_int, [_c.c_char_p]),
    "ipt_hip_profile_reset": (None, []),
    "ipt_hip_profile_get": (_c.c_int, [_c.c_char_p, _dp, _c.POINTER(_c.c_longlong)]),
    "ipt_hip_profile_get_columns": (_c.c_int, [_c.c_char_p, _c.POINTER(_c.c_longlong)]),
    "ipt_hip_last_solve_stats": (_c.c_int, [_vp, _ip, _ip, _c.POINTER(_c.c_longlong),
                                            _c.POINTER(_c.c_longlong), _ip]),
    "ipt_hip_generate_orbital_csr": (_c.c_int, [_c.c_int, _c.c_int, _c.c_ulonglong, _c.c_int, _c.c_int, _c.c_double,
                                                _c.c_double, _c.c_double, _vp, _vp, _vp, _c.c_longlong,
                                                _c.POINTER(_c.c_longlong), _vp]),
    "ipt_hip_csr_spmm": (_c.c_int, [_vp, _vp, _vp, _c.c_int, _vp, _c.c_int, _vp, _c.c_int, _c.c_int, _vp]),
    "ipt_hip_csr_spmm_sliced": (_c.c_int, [_vp, _vp, _vp, _c.c_int, _vp, _c.c_int, _vp, _c.c_int, _c.c_int, _c.c_int, _vp]),
    "ipt_hip_csr_spmm_tiled": (_c.c_int, [_vp, _vp, _vp, _c.c_int, _vp, _c.c_int, _vp, _c.c_int, _c.c_int, _vp]),
    "ipt_hip_csr_row_stride": (_c.c_long, [_c.c_int, _c.c_int]),
    "ipt_hip_cholqr_mixed": (_c.c_int, [_vp, _c.c_int, _c.c_int, _vp]),
    "ipt_hip_last_davidson_stats": (_c.c_int, [_vp, _ip, _ip, _ip, _ip, _c.POINTER(_c.c_longlong), _dp]),
    "ipt_hip_last_cluster_ranges": (_c.c_int, [_vp, _ip, _c.c_int, _ip]),
    "ipt_hip_sym_eig": (_c.c_int, [_vp, _c.c_int, _c.c_int, _vp, _vp, _c.c_int, _vp]),
    "ipt_hip_last_cluster_states": (_c.c_int, [_vp, _ip, _c.c_int, _ip]),
    "ipt_hip_mlp_features": (_c.c_int, [_vp, _c.c_int, _vp, _vp, _vp, _vp, _c.c_int, _c.c_int, _vp, _vp, _vp]),
    "ipt_hip_mlp_infer": (_c.c_int, [_vp, _vp, _vp, _c.c_longlong, _vp, _vp]),
    "ipt_hip_mlp_gradient": (_c.c_int, [_vp, _vp, _vp, _vp, _vp, _c.c_longlong, _vp, _vp, _vp]),
    "ipt_hip_mlp_gradient_noalloc": (_c.c_int, [_vp, _vp, _vp, _vp, _vp, _c.c_longlong, _vp, _vp, _vp]),
    "ipt_hip_mlp_adam_step": (_c.c_int, [_vp, _vp, _vp, _vp, _c.c_float, _c.c_float, _c.c_int, _vp]),
    # include/ipt_hip_fci.h
    "ipt_hip_fci_build_csc": (_c.c_int, [_c.c_int, _c.c_int, _c.c_int, _dp, _dp, _c.c_double, _c.c_int,
                                         _c.POINTER(IPTFciCsc)]),
    "ipt_hip_fci_free_csc": (None, [_c.POINTER(IPTFciCsc)]),
    "ipt_hip_fcidump_read_header": (_c.c_int, [_c.c_char_p, _ip, _ip, _ip]),
    "ipt_hip_fcidump_read": (_c.c_int, [_c.c_char_p, _c.c_int, _dp, _dp, _dp]),
    "ipt_hip_fcidump_write": (_c.c_int, [_c.c_char_p, _c.c_int, _c.c_int, _c.c_int, _dp, _dp, _c.c_double]),
    "ipt_hip_csc_cache_write": (_c.c_int, [_c.c_char_p, _c.POINTER(IPTFciCsc)]),
    "ipt_hip_csc_cache_read": (_c.c_int, [_c.c_char_p, _c.POINTER(IPTFciCsc)]),
    # include/ipt_hip_mixed.h
    "ipt_mixed_precision_cuda_device_tol": (_c.c_int, [_vp, _c.c_int, _c.c_int, _c.c_double, _c.c_int,
                                                       _c.c_double, _c.c_int, _vp,
                                                       _c.POINTER(IPTMixedCudaDeviceResult)]),
    "ipt_mixed_precision_cuda_free_device_result": (None, [_c.POINTER(IPTMixedCudaDeviceResult)]),
    "ipt_mixed_precision_fp32_cuda_device_tol": (_c.c_int, [_vp, _c.c_int, _c.c_int, _c.c_double, _c.c_int,
                                                            _c.c_double, _c.c_int, _vp,
                                                            _c.POINTER(IPTMixedFP32CudaDeviceResult)]),
    "ipt_mixed_precision_fp32_cuda_free_device_result": (None, [_c.POINTER(IPTMixedFP32CudaDeviceResult)]),
    "ipt_mixed_precision_tf32_cuda_device_tol": (_c.c_int, [_vp, _c.c_int, _c.c_int, _c.c_double, _c.c_int,
                                                            _c.c_double, _c.c_int, _vp,
                                                            _c.POINTER(IPTMixedTF32CudaDeviceResult)]),
    "ipt_mixed_precision_tf32_cuda_free_device_result": (None, [_c.POINTER(IPTMixedTF32CudaDeviceResult)]),
    "ipt_rayleigh_mixed_tf32_cuda_device_tol": (_c.c_int, [_vp, _c.c_int, _c.c_int, _c.c_double, _c.c_int,
                                                           _c.c_double, _c.c_int, _vp,
                                                           _c.POINTER(IPTRayleighMixedCudaDeviceResult)]),
    "ipt_rayleigh_mixed_cuda_free_device_result": (None, [_c.POINTER(IPTRayleighMixedCudaDeviceResult)]),
    "ipt_hip_prepare_sparse_csc": (_c.c_int, [_ip, _ip, _dp, _c.c_int, _c.c_int, _c.c_int, _c.c_double, _dp, _ip,
                                              _ip]),
    "ipt_hip_prepared_vectors_to_original": (_c.c_int, [_ip, _ip, _dp, _c.c_int, _c.c_int, _c.c_int, _c.c_double,
                                                        _dp, _c.c_int]),
    # include/ipt_hip_acx.h
    "acx_cuda_status_string": (_c.c_char_p, [_c.c_int]),
    "acx_cuda_reset_result": (None, [_c.POINTER(ACXCudaResult)]),
    "acx_cuda_free_result": (None, [_c.POINTER(ACXCudaResult)]),
    "acx_cuda_device_tol": (_c.c_int, [_vp, _c.c_int, _c.c_int, _c.c_double, _c.c_int, _vp,
                                       _c.POINTER(ACXCudaResult)]),
    "acx_mixed_precision_tf32_cuda_reset_result": (None, [_c.POINTER(ACXMixedTF32CudaResult)]),
    "acx_mixed_precision_tf32_cuda_free_result": (None, [_c.POINTER(ACXMixedTF32CudaResult)]),
    "acx_mixed_precision_tf32_cuda_device_tol": (_c.c_int, [_vp, _c.c_int, _c.c_int, _c.c_double, _c.c_int,
                                                            _c.c_double, _c.c_int, _vp,
                                                            _c.POINTER(ACXMixedTF32CudaResult)]),
    "acx_rayleigh_mixed_tf32_cuda_reset_result": (None, [_c.POINTER(ACXRayleighMixedTF32CudaResult)]),
    "acx_rayleigh_mixed_tf32_cuda_free_result": (None, [_c.POINTER(ACXRayleighMixedTF32CudaResult)]),
    "acx_rayleigh_mixed_tf32_cuda_device_tol": (_c.c_int, [_vp, _c.c_int, _c.c_int, _c.c_double, _c.c_int,
                                                           _c.c_double, _c.c_int, _vp,
                                                           _c.POINTER(ACXRayleighMixedTF32CudaResult)]),
}

_lib = None


def lib():
    """Load libipt_hip.so; raise if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise ImportError("libipt_hip.so is not built: run `python -m block_ipt_amd.build` "
                              "(the HIP library is the only compute path)")
        l = ctypes.CDLL(_LIB)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib
