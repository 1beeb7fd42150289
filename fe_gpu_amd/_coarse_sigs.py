"""ctypes mirror of include/fehip.h (coarse-grained API)."""
import ctypes as C

I32 = C.c_int32
I64 = C.c_int64
U32 = C.c_uint32
VP = C.c_void_p
DP = C.POINTER(C.c_double)
MAX_FE = 8


class PanelInfo(C.Structure):
    _fields_ = [("n_obs", I64), ("n_fe", I32), ("primary_dim", I32), ("sorted_input", I32), ("max_segment", I32),
                ("group_sizes", I32 * MAX_FE), ("n_runs", I64 * MAX_FE), ("device_bytes", I64)]


class DemeanOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iter", I32), ("method", I32), ("fe_mask", U32), ("precision", I32),
                ("check_every", I32), ("reserved", I32)]


class DemeanStats(C.Structure):
    _fields_ = [("iterations", I32), ("converged", I32), ("max_update", C.c_double), ("seconds", C.c_double),
                ("bytes_moved", I64)]


class Opts(C.Structure):
    _fields_ = [("demean", DemeanOpts), ("n_cluster", I32), ("cluster_dims", I32 * 4), ("n_iv_reg", I32),
                ("iv_reg", C.POINTER(I32)), ("n_iv_inst", I32), ("iv_inst", C.POINTER(I32))]


class Result(C.Structure):
    _fields_ = [("converged", I32), ("fe_iterations", I32), ("solver_info", I32), ("rank", I32),
                ("dof_model", I32), ("dof_fes", I32), ("is_iv", I32), ("n_instruments", I32),
                ("dof_resid", I64), ("n_obs", I64), ("n_cluster_dims", I32), ("cluster_dims", I32 * 3),
                ("n_clusters", I64 * 7), ("min_single_cluster", I64), ("psd_fix_applied", I32), ("fe_refinements", I32),
                ("rss", C.c_double), ("tss_total", C.c_double), ("tss_within", C.c_double), ("r2", C.c_double),
                ("r2_adj", C.c_double), ("r2_within", C.c_double), ("f_stat", C.c_double),
                ("max_update", C.c_double), ("t_demean", C.c_double), ("t_regression", C.c_double),
                ("t_se", C.c_double)]


class Request(C.Structure):
    _fields_ = [("d_y", VP), ("d_W", VP), ("d_Z", VP), ("K", I32), ("L", I32), ("ld", I64), ("d_weights", VP),
                ("opts", Opts), ("result", C.POINTER(Result)), ("beta", DP), ("se", DP), ("vcov", DP)]


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget", I64), ("max_batch", I32), ("workers", I32), ("window_us", I32), ("reserved", I32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", I64), ("completed", I64), ("failed", I64), ("batches", I64), ("queued", I64),
                ("budget_bytes", I64), ("mean_batch", C.c_double), ("busy_s", C.c_double), ("span_s", C.c_double),
                ("p50_s", C.c_double), ("p99_s", C.c_double), ("max_s", C.c_double)]


class BinHeader(C.Structure):
    _fields_ = [("n_obs", I64), ("n_regressors", I32), ("n_instruments", I32), ("n_fe", I32), ("has_cluster", I32),
                ("has_weights", I32), ("precision", I32), ("new_format", I32), ("has_precision_flag", I32),
                ("data_offset", I64), ("metadata_bytes", I64)]


class DesignColumn(C.Structure):
    _fields_ = [("n_factors", I32), ("col", I32 * 4), ("level", I32 * 4), ("categorical", I32 * 4)]


COARSE = [
    ("fehip_panel_create", C.c_int, [C.POINTER(VP), C.c_int, I64, C.POINTER(VP)]),
    ("fehip_panel_destroy", C.c_int, [VP]),
    ("fehip_panel_get_info", C.c_int, [VP, C.POINTER(PanelInfo)]),
    ("fehip_panel_register_columns", C.c_int, [VP, VP, C.c_int, I64]),
    ("fehip_demean", C.c_int, [VP, VP, C.c_int, I64, VP, C.POINTER(DemeanOpts), C.POINTER(DemeanStats)]),
    ("fehip_estimate", C.c_int, [VP, VP, VP, C.c_int, VP, C.c_int, I64, VP, C.POINTER(Opts), C.POINTER(Result), DP,
                                 DP, DP]),
    ("fehip_estimate_batch", C.c_int, [VP, C.POINTER(Request), C.c_int]),
    ("fehip_server_create", C.c_int, [VP, C.POINTER(ServerOpts), C.POINTER(VP)]),
    ("fehip_server_submit", C.c_int, [VP, C.POINTER(Request), C.POINTER(I64)]),
    ("fehip_server_wait", C.c_int, [VP, I64, DP]),
    ("fehip_server_get_stats", C.c_int, [VP, C.POINTER(ServerStats)]),
    ("fehip_server_reset_stats", C.c_int, [VP]),
    ("fehip_server_destroy", C.c_int, [VP]),
    ("fehip_synth_worker_firm", C.c_int, [I64, I32, I32, C.c_double, C.c_uint64, VP, VP, VP]),
    ("fehip_synth_iv_wide", C.c_int, [I64, I32, I32, C.c_double, C.c_uint64, VP, VP, VP, VP]),
    ("fehip_synth_mask", C.c_int, [I64, C.c_double, C.c_uint64, VP]),
    ("fehip_kstats_enable", C.c_int, [C.c_int]),
    ("fehip_kstats_reset", C.c_int, []),
    ("fehip_kstats_get", C.c_int, [C.c_int, C.c_char_p, C.c_int, C.POINTER(I64), DP, DP, DP]),
    ("fehip_device_copy", C.c_int, [VP, VP, I64]),
    ("fehip_fe_values", C.c_int, [VP, VP, VP, C.c_int, I64, DP, C.POINTER(DemeanOpts), VP, VP,
                                  C.POINTER(DemeanStats)]),
    ("fehip_design_levels", C.c_int, [VP, I64, C.POINTER(I32), C.POINTER(I32), C.POINTER(I32)]),
    ("fehip_design_expand", C.c_int, [VP, I64, C.c_int, I64, C.POINTER(DesignColumn), C.c_int, VP, I64]),
    ("fehip_bin_read_header", C.c_int, [C.c_char_p, C.POINTER(BinHeader)]),
    ("fehip_bin_load", C.c_int, [C.c_char_p, C.POINTER(BinHeader), VP, VP, VP, VP, VP, VP]),
]


def declare(L):
    for name, res, args in COARSE:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
