"""ctypes binding of librfmf.so (include/rfmf.h).

Loading fails loudly: there is no Python fallback for any entry point."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(os.environ.get("RFMF_LIB", Path(__file__).resolve().parent / "lib" / "librfmf.so"))

i8p = C.POINTER(C.c_int8)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
f64p = C.POINTER(C.c_double)
f32p = C.POINTER(C.c_float)


class Csr(C.Structure):
    _fields_ = [("offsets", u32p), ("edge_to", i32p), ("edge_weight", i32p), ("n", C.c_int64),
                ("m2", C.c_int64)]


class ReadStats(C.Structure):
    _fields_ = [("num_vertices", C.c_int64), ("num_edges", C.c_int64), ("zero_dropped", C.c_int64),
                ("float_weights", C.c_int32), ("pad", C.c_int32), ("scale", C.c_double),
                ("max_abs_weighted_degree", C.c_int64)]


class BetaTables(C.Structure):
    _fields_ = [("run_end", i64p), ("run_val", i32p), ("chain_run_off", i64p), ("n_runs", C.c_int64)]


class LaunchCfg(C.Structure):
    _fields_ = [("corr", C.c_int32), ("device", C.c_int32), ("s_width", C.c_int32), ("reserved", C.c_int32)]


class Job(C.Structure):
    _fields_ = [("graph_id", C.c_int32), ("pad", C.c_int32), ("chains", C.c_int64),
                ("initial_spins", i8p), ("solver_seeds", u64p), ("beta", BetaTables),
                ("phases", u64p), ("iters", u64p), ("n_phases", C.c_int64),
                ("end_spins", i8p), ("best_spins", i8p), ("best_vals", i64p), ("final_vals", i64p)]


class Tuning(C.Structure):
    _fields_ = [("slice_sweeps", C.c_int64), ("grid_cap", C.c_int32), ("use_hip_graph", C.c_int32),
                ("threads_per_chain", C.c_int32), ("autotune", C.c_int32), ("score_table", C.c_int32),
                ("state_in_hbm", C.c_int32), ("rows_in_lds", C.c_int32)]


MAX_FLIPS_CANDIDATES = 16


class FlipsTune(C.Structure):
    _fields_ = [("sweeps", C.c_int64), ("pilot_chains", C.c_int64), ("seed_base", C.c_uint64),
                ("beta0_q20", C.c_int32), ("beta1_q20", C.c_int32), ("top_p", u64p),
                ("n_candidates", C.c_int32), ("reference", C.c_int32), ("quality_tol", C.c_double)]


class FlipsReport(C.Structure):
    _fields_ = [("chosen", C.c_int32), ("n_candidates", C.c_int32),
                ("median_best", C.c_double * MAX_FLIPS_CANDIDATES),
                ("chain_steps_per_s", C.c_double * MAX_FLIPS_CANDIDATES)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("rejected", C.c_int64),
                ("active_chains", C.c_int64), ("resident_bytes", C.c_int64), ("slices", C.c_int64),
                ("chain_steps", C.c_double), ("busy_kernel_ms", C.c_double)]


# name -> (restype, argtypes); every function rfmf.h declares.
SIGNATURES = {
    "rfmf_status_string": (C.c_char_p, [C.c_int]),
    "rfmf_last_error": (C.c_char_p, []),
    "rfmf_version": (C.c_char_p, []),
    "rfmf_graph_read_gset": (C.c_int, [C.c_char_p, C.c_double, C.POINTER(C.c_void_p), C.POINTER(ReadStats)]),
    "rfmf_graph_from_edges": (C.c_int, [i64p, i64p, i64p, C.c_int64, C.POINTER(C.c_void_p)]),
    "rfmf_graph_free": (None, [C.c_void_p]),
    "rfmf_graph_view": (C.c_int, [C.c_void_p, C.POINTER(Csr)]),
    "rfmf_graph_original_ids": (i64p, [C.c_void_p]),
    "rfmf_cut_value": (C.c_int64, [C.POINTER(Csr), i8p]),
    "rfmf_max_weighted_degree": (C.c_int64, [C.POINTER(Csr)]),
    "rfmf_score_safe_int32": (C.c_int32, [C.c_int64, C.c_double, C.c_int64]),
    "rfmf_beta_fixed_for_step": (C.c_int32, [C.c_uint64, C.c_double]),
    "rfmf_total_steps": (C.c_int, [u64p, C.c_int64, i64p]),
    "rfmf_initial_spins": (C.c_int, [C.c_uint64, C.c_int64, i8p]),
    "rfmf_beta_tables_log": (C.c_int, [f64p, C.c_int64, C.c_int64, i64p, i32p, C.c_int64, i64p, i64p, i64p]),
    "rfmf_beta_tables_linear": (C.c_int, [C.c_int32, C.c_int32, C.c_int64, i64p, i32p, C.c_int64, i64p]),
    "rfmf_anneal": (C.c_int, [C.POINTER(Csr), C.c_int64, i8p, C.c_int64, f64p, u64p, u64p, C.c_int64, u64p,
                              C.POINTER(LaunchCfg), i8p, i8p, f32p, i64p, i64p]),
    "rfmf_anneal_tables": (C.c_int, [C.POINTER(Csr), C.c_int64, i8p, C.c_int64, C.POINTER(BetaTables), u64p,
                                     u64p, C.c_int64, u64p, C.POINTER(LaunchCfg), i8p, i8p, f32p, i64p, i64p]),
    "rfmf_anneal_tables_noalloc": (C.c_int, [C.POINTER(Csr), C.c_int64, i8p, C.c_int64, C.POINTER(BetaTables),
                                             u64p, u64p, C.c_int64, u64p, C.POINTER(LaunchCfg), i8p, i8p, f32p,
                                             i64p, i64p]),
    "rfmf_ctx_create": (C.c_int, [C.c_int32, C.POINTER(C.c_void_p)]),
    "rfmf_ctx_destroy": (None, [C.c_void_p]),
    "rfmf_ctx_upload_graph": (C.c_int, [C.c_void_p, C.POINTER(Csr), i32p]),
    "rfmf_ctx_release_graph": (C.c_int, [C.c_void_p, C.c_int32]),
    "rfmf_ctx_device_info": (C.c_int, [C.c_void_p, i32p, i64p, i64p, C.c_char_p, C.c_int32]),
    "rfmf_ctx_copy_probe": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, f32p]),
    "rfmf_ctx_last_launch": (C.c_int, [C.c_void_p, i32p, i32p, i32p, i32p, f32p, i32p, i32p, i32p]),
    "rfmf_batch_run": (C.c_int, [C.c_void_p, C.POINTER(Job), C.c_int64, C.POINTER(Tuning), f32p]),
    "rfmf_batch_run_windows": (C.c_int, [C.c_void_p, C.POINTER(Job), C.c_int64, i64p, C.c_int64,
                                         C.POINTER(Tuning), f32p]),
    "rfmf_run_create": (C.c_int, [C.c_void_p, C.POINTER(Job), C.c_int64, C.POINTER(C.c_void_p)]),
    "rfmf_run_advance": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(Tuning), f32p, i32p]),
    "rfmf_run_collect": (C.c_int, [C.c_void_p]),
    "rfmf_run_destroy": (None, [C.c_void_p]),
    "rfmf_footprint_bytes": (C.c_int64, [C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int64]),
    "rfmf_flips_ladder": (C.c_int, [C.c_uint64, C.c_int64, u64p, u64p, C.c_int64, i64p]),
    "rfmf_tune_flips": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(FlipsTune), C.POINTER(FlipsReport)]),
    "rfmf_server_create": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.POINTER(C.c_void_p)]),
    "rfmf_server_submit": (C.c_int, [C.c_void_p, C.POINTER(Csr), C.POINTER(Job), i64p]),
    "rfmf_server_poll": (C.c_int, [C.c_void_p, C.c_int64, i32p, f64p, f64p, f64p]),
    "rfmf_server_wait": (C.c_int, [C.c_void_p, C.c_int64, C.c_double]),
    "rfmf_server_stats_get": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "rfmf_server_destroy": (None, [C.c_void_p]),
}

STATUS = {0: "RFMF_OK", 1: "RFMF_ERR_INVALID_ARG", 2: "RFMF_ERR_UNSUPPORTED", 3: "RFMF_ERR_HIP",
          4: "RFMF_ERR_NO_DEVICE", 5: "RFMF_ERR_OUT_OF_MEMORY", 6: "RFMF_ERR_BETA_CAP",
          7: "RFMF_ERR_STEP_CAP", 8: "RFMF_ERR_IO", 9: "RFMF_ERR_NOT_FOUND", 10: "RFMF_ERR_ADMISSION",
          11: "RFMF_ERR_STATE"}


class RfmfError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


_lib = None


def lib() -> C.CDLL:
    """The loaded library; raises if librfmf.so has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"librfmf.so not found at {LIB_PATH}; build it with "
                              "`python -m rfmf_amd.build` (no CPU fallback exists)")
        l = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def check(status: int) -> None:
    if status != 0:
        raise RfmfError(status, (lib().rfmf_last_error() or b"").decode())
