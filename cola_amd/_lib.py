"""ctypes binding of libcola_spmm.so (the C ABI declared in include/cola_spmm.h).

torch is imported before the library is opened: torch ships its own HIP
runtime under the same SONAME (libamdhip64.so.7), and loading torch first makes
the dynamic loader hand that one runtime to both, so device pointers and
streams from torch are valid in the library.  There is no CPU fallback: if the
library is missing, every call raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcola_spmm.so")

c_int = ctypes.c_int
c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_u64 = ctypes.c_uint64
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
c_char_p = ctypes.c_char_p
P_i32 = ctypes.POINTER(ctypes.c_int32)
P_i64 = ctypes.POINTER(ctypes.c_int64)
P_f32 = ctypes.POINTER(ctypes.c_float)

COLA_OK, COLA_ERR_INVALID, COLA_ERR_IO, COLA_ERR_HIP, COLA_ERR_UNSUPPORTED, COLA_ERR_NOMEM = range(6)
F32, BF16, F16, FP8_E4M3 = 0, 1, 2, 3
ACCUMULATE, OVERWRITE, KAHAN = 0, 1, 2

_ERR_NAMES = {1: "invalid argument", 2: "I/O error", 3: "HIP error", 4: "unsupported", 5: "out of memory"}


class ColaError(RuntimeError):
    """A non-zero cola_status; .code holds the status."""

    def __init__(self, code, msg):
        super().__init__(f"[{_ERR_NAMES.get(code, code)}] {msg}")
        self.code = code


class PartitionInfo(ctypes.Structure):
    _fields_ = [("num_nodes", c_i64), ("nGPU", c_i32), ("rank", c_i32), ("nodesPerPE", c_i32),
                ("num_rows", c_i32), ("L", c_i32), ("U_r", c_i32), ("nnz_l", c_i64),
                ("nnz_r", c_i64), ("U_l", c_i32), ("max_row_nnz", c_i32)]


class PartitionArrays(ctypes.Structure):
    _fields_ = [("csr_rowPtr_l", P_i32), ("csr_colIdx_l", P_i32), ("csr_rowIdx_l", P_i32),
                ("csr_val_l", P_f32), ("csc_colPtr_r", P_i32), ("csc_rowIdx_r", P_i32),
                ("csc_colIdx_r", P_i32), ("csc_val_r", P_f32)]


class PlanOpts(ctypes.Structure):
    _fields_ = [("hub_threshold", c_i32), ("chunk_nnz", c_i32), ("kahan_degree", c_i32),
                ("max_dim", c_i32), ("short_degree", c_i32), ("col_panels", c_i32),
                ("row_order", c_i32), ("hot_cols", c_i32)]


class PlanTuning(ctypes.Structure):
    _fields_ = [("col_tile", c_i32), ("rpw_short", c_i32), ("rpw_long", c_i32), ("blocks_per_cu", c_i32),
                ("xcd_remap", c_i32), ("split_remote", c_i32),
                ("unroll", c_i32), ("panels", c_i32), ("blocks_per_cu_short", c_i32), ("tiny", c_i32)]


class BatchProblem(ctypes.Structure):
    _fields_ = [("plan", c_vp), ("dim", c_i32), ("B", c_vp), ("ld_B", c_i64), ("scales", c_vp), ("packed", c_vp),
                ("ld_packed", c_i64), ("packed_scales", c_vp), ("C", c_vp), ("ld_C", c_i64)]


class BatchTuning(ctypes.Structure):
    _fields_ = [("col_tile", c_i32), ("rpw_short", c_i32), ("rpw_long", c_i32)]


class ServerOpts(ctypes.Structure):
    _fields_ = [("memory_budget", c_i64), ("max_batch", c_i32), ("max_inflight", c_i32), ("dtype", c_i32),
                ("flags", c_i32), ("plain", c_i32), ("blocks_per_cu", c_i32)]


class PlanInfo(ctypes.Structure):
    _fields_ = [("num_rows", c_i64), ("num_items", c_i64), ("num_hub_rows", c_i64),
                ("num_hub_chunks", c_i64), ("nnz", c_i64), ("U_r", c_i64), ("U_l", c_i64),
                ("device_bytes", c_i64), ("num_short_items", c_i64), ("num_long_items", c_i64),
                ("hot_nnz", c_i64), ("num_cold_items", c_i64), ("num_tiny_bundles", c_i64)]


# name -> argtypes (all return cola_status unless listed in _RESTYPES)
SIGNATURES = {
    "cola_last_error": [],
    "cola_version": [],
    "cola_rmat_graph": [c_int, c_dbl, c_dbl, c_dbl, c_dbl, c_u64, c_int, c_int, ctypes.POINTER(c_vp)],
    "cola_graph_from_csr": [c_i64, c_vp, c_vp, c_vp, ctypes.POINTER(c_vp)],
    "cola_graph_info": [c_vp, P_i64, P_i64, P_i64],
    "cola_graph_arrays": [c_vp, ctypes.POINTER(P_i64), ctypes.POINTER(P_i32), ctypes.POINTER(P_f32)],
    "cola_graph_destroy": [c_vp],
    "cola_graph_from_mtx": [c_char_p, c_int, ctypes.POINTER(c_vp)],
    "cola_graph_write_mtx": [c_vp, c_char_p, c_int],
    "cola_graph_reorder": [c_vp, c_int, c_vp, ctypes.POINTER(c_vp)],
    "cola_row_partition": [c_vp, c_int, c_int, ctypes.POINTER(c_vp)],
    "cola_partition_from_arrays": [c_i64, c_int, c_int, ctypes.POINTER(PartitionArrays), c_i64, c_i64,
                                   c_i64, c_i64, ctypes.POINTER(c_vp)],
    "cola_partition_get_info": [c_vp, ctypes.POINTER(PartitionInfo)],
    "cola_partition_get_arrays": [c_vp, ctypes.POINTER(PartitionArrays)],
    "cola_partition_destroy": [c_vp],
    "cola_serialize": [c_vp, c_char_p, c_char_p, c_int, c_int],
    "cola_serialize_ppl": [c_vp, c_char_p, c_char_p, c_int, c_int],
    "cola_partition_load": [c_char_p, c_char_p, c_int, c_int, ctypes.POINTER(c_vp)],
    "cola_nnz_split": [c_vp, c_i64, c_vp, c_i64, c_vp, c_vp, P_i64, P_i32],
    "cola_zip_create": [c_vp, c_i64, c_vp, c_i64, c_vp, c_vp, ctypes.POINTER(c_vp)],
    "cola_zip_device_arrays": [c_vp, ctypes.POINTER(c_vp), ctypes.POINTER(c_vp), ctypes.POINTER(c_vp),
                               ctypes.POINTER(c_vp), ctypes.POINTER(c_vp), P_i64, P_i64, P_i64],
    "cola_zip_destroy": [c_vp],
    "cola_zip_invalidate": [c_vp],
    "cola_pack_remote": [c_vp, c_i64, c_vp, c_vp, c_i64, c_int, c_int, c_vp, c_i64, c_vp, c_vp],
    "cola_ppl0_batch_spmm": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                             c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_vp],
    "cola_dist_spmm_ppl_batch": [c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_int, c_int, c_vp],
    "cola_ppl0_batch_spmm_global": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                    c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_vp],
    "cola_ppl0_plans_clear": [],
    "cola_warmup": [c_int, c_vp],
    "cola_plan_create": [c_vp, ctypes.POINTER(PlanOpts), ctypes.POINTER(c_vp)],
    "cola_plan_get_info": [c_vp, ctypes.POINTER(PlanInfo)],
    "cola_plan_remote_cols": [c_vp, ctypes.POINTER(c_vp)],
    "cola_plan_set_tuning": [c_vp, ctypes.POINTER(PlanTuning)],
    "cola_plan_get_tuning": [c_vp, ctypes.POINTER(PlanTuning)],
    "cola_plan_autotune": [c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, c_vp, c_i64, c_vp, c_int, c_int, c_int, c_int,
                           c_vp, ctypes.POINTER(PlanTuning), ctypes.POINTER(ctypes.c_float)],
    "cola_plan_destroy": [c_vp],
    "cola_spmm_partition": [c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, c_vp, c_i64, c_vp, c_int, c_int,
                            c_int, c_vp],
    "cola_spmm_partition_noalloc": [c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, c_vp, c_i64, c_vp, c_int, c_int,
                                    c_int, c_vp],
    "cola_batch_create": [ctypes.POINTER(BatchProblem), c_int, c_int, c_int, ctypes.POINTER(c_vp)],
    "cola_batch_set_tuning": [c_vp, c_int, ctypes.POINTER(BatchTuning), c_int],
    "cola_batch_get_tuning": [c_vp, c_int, ctypes.POINTER(BatchTuning), ctypes.POINTER(c_int)],
    "cola_batch_autotune": [c_vp, c_int, c_vp, ctypes.POINTER(ctypes.c_float)],
    "cola_batch_run": [c_vp, c_vp],
    "cola_batch_capture": [c_vp, c_vp],
    "cola_batch_replay": [c_vp, c_vp],
    "cola_batch_destroy": [c_vp],
    "cola_server_create": [ctypes.POINTER(ServerOpts), ctypes.POINTER(c_vp)],
    "cola_server_add_resident": [c_vp, c_vp, c_vp, c_i64, c_vp, c_int, ctypes.POINTER(c_int)],
    "cola_server_set_tuning": [c_vp, c_int, ctypes.POINTER(BatchTuning)],
    "cola_server_run_trace": [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, ctypes.POINTER(c_dbl), P_i64],
    "cola_server_destroy": [c_vp],
    "cola_probe_gather": [c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_int, c_int, c_int, c_i64, c_int, c_vp],
    "cola_probe_row_stream": [c_vp, c_i64, c_vp, c_int, c_int, c_int, c_vp, c_i64, c_vp],
    "cola_quantize_rows": [c_vp, c_i64, c_i64, c_int, c_int, c_vp, c_i64, c_vp, c_vp],
    "cola_error_norms": [c_vp, c_i64, c_vp, c_i64, c_i64, c_int, c_vp, c_vp],
}
_RESTYPES = {"cola_last_error": c_char_p, "cola_version": c_char_p}

_lib = None


def lib():
    """Opens libcola_spmm.so once (after torch) and declares every entry point."""
    global _lib
    if _lib is not None:
        return _lib
    import torch  # noqa: F401  (see module docstring)
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, c_int)
    _lib = L
    return L


def check(status):
    """Raises ColaError for a non-zero status, with cola_last_error()'s text."""
    if status != COLA_OK:
        msg = lib().cola_last_error()
        raise ColaError(status, msg.decode() if msg else "")
    return status
