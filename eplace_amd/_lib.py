"""ctypes binding of the C ABI in include/eplace_amd.h (libeplace_amd.so).

The library is built in-tree (``python -c "import __graft_entry__ as g; g.build()"``
or ``make -C eplace_amd/csrc``).  There is no CPU fallback: if the library is
missing or cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libeplace_amd.so")

EP_OK = 0
EP_ERR_ARG, EP_ERR_HIP, EP_ERR_STATE, EP_ERR_NUMERIC, EP_ERR_NOMEM = 1, 2, 3, 4, 5
EP_STAGE_MGP, EP_STAGE_FILLERONLY, EP_STAGE_CGP = 0, 1, 2

EP_LOG_COLS = 24
EP_QP_NOALLOC_MAX_ROWS = 1 << 20
EP_LOG_ITER, EP_LOG_TAU, EP_LOG_LAMBDA, EP_LOG_HPWL, EP_LOG_STEP, EP_LOG_INV_GAMMA, EP_LOG_EVALS, \
    EP_LOG_BACKTRACKS, EP_LOG_FEATURES = range(9)
EP_LOG_STEP_PROP, EP_LOG_STEP_USED, EP_LOG_LAMBDA_BEFORE, EP_LOG_FALLBACKS = 20, 21, 22, 23
EP_CTRL_FEATURES, EP_CTRL_HIDDEN, EP_CTRL_OUTPUTS = 12, 64, 2
EP_CTRL_WEIGHTS = 12 * 64 + 64 + 64 * 64 + 64 + 2 * 64 + 2
EP_CTRL_PARAMS = EP_CTRL_WEIGHTS + 2 * 12 + 2 * 2

(EP_ARR_POSITIONS, EP_ARR_NODE_Q, EP_ARR_FILLER_Q, EP_ARR_RHO, EP_ARR_PHI, EP_ARR_EX, EP_ARR_EY,
 EP_ARR_GRAD_DENSITY, EP_ARR_GRAD_WL, EP_ARR_GRAD_TOTAL, EP_ARR_TERM_MAP, EP_ARR_BASE_MAP,
 EP_ARR_PIN_GRAD, EP_ARR_MAIN) = range(14)


class EPDesign(C.Structure):
    _fields_ = [
        ("num_nodes", C.c_int32), ("num_terminals", C.c_int32), ("num_nets", C.c_int32),
        ("num_rows", C.c_int32), ("num_pins", C.c_int64),
        ("node_w", C.c_void_p), ("node_h", C.c_void_p), ("node_x", C.c_void_p), ("node_y", C.c_void_p),
        ("term_w", C.c_void_p), ("term_h", C.c_void_p), ("term_x", C.c_void_p), ("term_y", C.c_void_p),
        ("term_ni", C.c_void_p), ("net_start", C.c_void_p), ("pin_obj", C.c_void_p),
        ("pin_dx", C.c_void_p), ("pin_dy", C.c_void_p),
        ("row_bottom", C.c_void_p), ("row_height", C.c_void_p), ("row_x0", C.c_void_p), ("row_x1", C.c_void_p),
        ("common_row_height", C.c_double),
    ]


class EPParams(C.Structure):
    _fields_ = [
        ("target_density", C.c_float), ("bins_x", C.c_int32), ("bins_y", C.c_int32),
        ("wirelength_model", C.c_int32), ("exp_mode", C.c_int32), ("filler_seed", C.c_uint32),
        ("filler_xy", C.c_void_p), ("filler_count", C.c_int32), ("poisson_mode", C.c_int32),
        ("device", C.c_int32),
    ]


class EPInfo(C.Structure):
    _fields_ = [
        ("num_nodes", C.c_int64), ("num_fillers", C.c_int64), ("num_objects", C.c_int64),
        ("num_terminals", C.c_int64), ("num_nets", C.c_int64), ("num_pins", C.c_int64),
        ("bins_x", C.c_int32), ("bins_y", C.c_int32), ("fixed_shift", C.c_int32),
        ("num_big_objects", C.c_int32), ("num_large_nets", C.c_int32), ("pad", C.c_int32),
        ("bin_step_x", C.c_double), ("bin_step_y", C.c_double), ("core_llx", C.c_double),
        ("core_lly", C.c_double), ("core_urx", C.c_double), ("core_ury", C.c_double),
        ("filler_w", C.c_double), ("filler_h", C.c_double), ("std_cell_area", C.c_double),
        ("macro_area", C.c_double), ("total_row_area", C.c_double),
    ]


class EPStats(C.Structure):
    _fields_ = [
        ("overflow", C.c_double), ("lambda_", C.c_double), ("last_hpwl", C.c_double), ("hpwl", C.c_double),
        ("inv_gamma_x", C.c_double), ("inv_gamma_y", C.c_double), ("wirelength", C.c_double),
        ("stage", C.c_int32), ("mgp_iterations", C.c_int32),
    ]


class EPOptParams(C.Structure):
    _fields_ = [
        ("method", C.c_int32), ("target_overflow", C.c_float), ("max_iterations", C.c_int32),
        ("max_backtracks", C.c_int32), ("sync_every", C.c_int32), ("log_capacity", C.c_int32),
        ("log", C.c_void_p), ("phase_seconds", C.c_void_p), ("controller", C.c_void_p),
    ]


class EPRunStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32), ("evaluations", C.c_int32), ("backtracks", C.c_int32),
        ("stopped", C.c_int32), ("final_overflow", C.c_double), ("final_hpwl", C.c_double),
        ("seconds", C.c_double), ("controller_fallbacks", C.c_int32), ("bb_short_steps", C.c_int32),
        ("gemm_seconds", C.c_double), ("gemm_launches", C.c_int32), ("numeric_fallbacks", C.c_int32),
    ]


class EPLegalParams(C.Structure):
    _fields_ = [("site_step", C.c_double), ("mode", C.c_int32), ("pad", C.c_int32)]


class EPLegalStats(C.Structure):
    _fields_ = [("cells", C.c_int64), ("subrows", C.c_int32), ("pad", C.c_int32),
                ("total_displacement", C.c_double), ("max_displacement", C.c_double)]


EP_DP_MAX_PASSES = 8


class EPDPParams(C.Structure):
    _fields_ = [("site_step", C.c_double), ("ism_time_limit", C.c_double), ("passes", C.c_int32),
                ("pad", C.c_int32)]


class EPDPStats(C.Structure):
    _fields_ = [("hpwl_before", C.c_double), ("hpwl_after", C.c_double), ("seconds", C.c_double),
                ("seconds_lr", C.c_double), ("seconds_ism", C.c_double), ("seconds_gs", C.c_double),
                ("hpwl_after_lr", C.c_double * EP_DP_MAX_PASSES), ("hpwl_after_ism", C.c_double * EP_DP_MAX_PASSES),
                ("hpwl_after_gs", C.c_double * EP_DP_MAX_PASSES), ("lr_windows", C.c_int64),
                ("ism_matchings", C.c_int64), ("gs_moves", C.c_int64)]


class EPLegalCheck(C.Structure):
    _fields_ = [("cells", C.c_int64), ("overlaps", C.c_int64), ("off_row", C.c_int64), ("off_site", C.c_int64),
                ("out_of_core", C.c_int64), ("in_blockage", C.c_int64), ("macro_overlaps", C.c_int64),
                ("overlap_area", C.c_double)]


class EPQpParams(C.Structure):
    _fields_ = [("max_iterations", C.c_int32), ("solver_iterations", C.c_int32), ("target_error", C.c_float),
                ("solver_tolerance", C.c_float), ("start_at_center", C.c_int32), ("device", C.c_int32)]


class EPQpStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("solver_iterations", C.c_int32), ("restarts", C.c_int32),
                ("pad", C.c_int32), ("initial_hpwl", C.c_double), ("final_hpwl", C.c_double),
                ("error_x", C.c_double), ("error_y", C.c_double), ("seconds", C.c_double)]


class EPQpMatrixOut(C.Structure):
    _fields_ = [("nnz", C.c_int64), ("capacity", C.c_int64), ("rowptr", C.c_void_p), ("col", C.c_void_p),
                ("val", C.c_void_p), ("diag", C.c_void_p), ("b", C.c_void_p), ("hpwl", C.c_double)]


class EPQpSystem(C.Structure):
    _fields_ = [("n", C.c_int32), ("ndim", C.c_int32), ("nnz", C.c_int64), ("rowptr", C.c_void_p),
                ("col", C.c_void_p), ("val", C.c_void_p), ("diag", C.c_void_p), ("b", C.c_void_p)]


class EPMlgParams(C.Structure):
    _fields_ = [("target_density", C.c_float), ("max_outer", C.c_int32), ("max_inner", C.c_int32),
                ("rand_seed", C.c_uint32), ("rand_skip", C.c_int64)]


class EPMlgStats(C.Structure):
    _fields_ = [("macros", C.c_int32), ("bins_x", C.c_int32), ("bins_y", C.c_int32), ("outer_iterations", C.c_int32),
                ("perturbations", C.c_int64), ("accepted", C.c_int64), ("overlap_free", C.c_int32),
                ("final_overlap", C.c_int32), ("initial_hpwl", C.c_double), ("final_hpwl", C.c_double),
                ("cell_area_covered", C.c_double), ("rand_draws", C.c_int64), ("seconds", C.c_double)]


SYMBOLS = {
    "ep_create": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "ep_destroy": ([C.c_void_p], C.c_int),
    "ep_last_error": ([C.c_void_p], C.c_char_p),
    "ep_get_info": ([C.c_void_p, C.c_void_p], C.c_int),
    "ep_total_gradient_update": ([C.c_void_p], C.c_int),
    "ep_get_position": ([C.c_void_p, C.c_void_p, C.c_int64], C.c_int),
    "ep_set_position": ([C.c_void_p, C.c_void_p, C.c_int64], C.c_int),
    "ep_get_gradient": ([C.c_void_p, C.c_void_p, C.c_int64], C.c_int),
    "ep_stage_size": ([C.c_void_p], C.c_int64),
    "ep_penalty_factor_init": ([C.c_void_p], C.c_int),
    "ep_update_penalty_factor": ([C.c_void_p], C.c_int),
    "ep_switch_to_filler_only": ([C.c_void_p], C.c_int),
    "ep_switch_to_cgp": ([C.c_void_p, C.c_int32], C.c_int),
    "ep_get_stats": ([C.c_void_p, C.c_void_p], C.c_int),
    "ep_hpwl": ([C.c_void_p, C.c_void_p], C.c_int),
    "ep_nesterov_run": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "ep_poisson": ([C.c_void_p] * 5, C.c_int),
    "ep_read": ([C.c_void_p, C.c_int32, C.c_void_p, C.c_int64], C.c_int),
    "ep_write": ([C.c_void_p, C.c_int32, C.c_void_p, C.c_int64], C.c_int),
    "ep_last_timings": ([C.c_void_p, C.c_void_p], C.c_int),
    "ep_host_last_error": ([], C.c_char_p),
    "ep_legalize_abacus": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "ep_check_legal": ([C.c_void_p, C.c_void_p, C.c_double, C.c_void_p], C.c_int),
    "ep_detailed_place": ([C.c_void_p] * 5, C.c_int),
    "ep_qp_last_error": ([], C.c_char_p),
    "ep_qplace": ([C.c_void_p] * 5, C.c_int),
    "ep_qp_matrix": ([C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p], C.c_int),
    "ep_qp_solve": ([C.c_void_p, C.c_void_p, C.c_int32, C.c_float, C.c_int32, C.c_void_p, C.c_void_p], C.c_int),
    "ep_qp_solve_noalloc": ([C.c_void_p, C.c_void_p, C.c_int32, C.c_float, C.c_int32, C.c_void_p, C.c_void_p,
                             C.c_void_p], C.c_int),
    "ep_mlg_last_error": ([], C.c_char_p),
    "ep_legalize_macros": ([C.c_void_p] * 5, C.c_int),
    "ep_macro_legalize": ([C.c_void_p] * 4, C.c_int),
    "ep_ctrl_create": ([C.c_int32, C.c_void_p, C.c_uint32, C.c_void_p], C.c_int),
    "ep_ctrl_destroy": ([C.c_void_p], C.c_int),
    "ep_ctrl_last_error": ([C.c_void_p], C.c_char_p),
    "ep_ctrl_get_params": ([C.c_void_p, C.c_void_p], C.c_int),
    "ep_ctrl_set_params": ([C.c_void_p, C.c_void_p], C.c_int),
    "ep_ctrl_train": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_float, C.c_void_p], C.c_int),
    "ep_ctrl_predict": ([C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p], C.c_int),
    "ep_ctrl_loss_grad": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p], C.c_int),
}

_lib = None


def load():
    """Load libeplace_amd.so; raises if it is missing (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C eplace_amd/csrc")
        lib = C.CDLL(LIB_PATH)
        for name, (args, res) in SYMBOLS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


class EPError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"eplace_amd error {code}: {msg}")
        self.code = code
