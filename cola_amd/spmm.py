"""Device-side host API: device sparse structures, the reference-semantics
launcher, the degree-binned plan path, packing, precision conversion and
on-device error norms.  Device buffers are torch tensors on the GPU (torch is
used only for device memory and streams); every launch goes to the current
torch stream unless a stream is given.

Reference interfaces mirrored (paths relative to the CoLa tree):
  include/sparse_data.h:28-77            d_csc_zip / d_csr_zip
  include/dist_spmm_ppl_batch.cuh:202-245  ppl0_batch_SpMM
  include/dist_spmm_ppl_batch.cuh:851-918  distSpMM_ppl_batch_wrapper (single stage)
  include/dist_spmm_ppl_batch.cuh:58-73    batching_data_ppl (-> pack_remote)
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (ACCUMULATE, BF16, COLA_ERR_INVALID, F16, F32, FP8_E4M3, KAHAN, OVERWRITE, ColaError,  # noqa: F401
                   check, lib)

DTYPE_BYTES = {F32: 4, BF16: 2, F16: 2, FP8_E4M3: 1}
DTYPE_NAMES = {F32: "f32", BF16: "bf16", F16: "f16", FP8_E4M3: "fp8_e4m3"}


def _torch():
    import torch
    return torch


def _dptr(t):
    if t is None:
        return None
    if isinstance(t, int):
        return ctypes.c_void_p(t)
    if not t.is_cuda:
        raise ValueError("expected a GPU tensor")
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)


def _check_matrix(t, name, rows, dim, elem_bytes):
    """A row-major GPU matrix the kernels will index: at least `rows` rows
    of `dim` used columns, unit column stride, elements of `elem_bytes`
    bytes.  The C ABI takes raw pointers and cannot check this itself."""
    if t is None:
        if rows > 0:
            raise ValueError(f"{name} is required ({rows} rows)")
        return
    if not t.is_cuda:
        raise ValueError(f"{name} must be a GPU tensor")
    if t.element_size() != elem_bytes:
        raise ValueError(f"{name} has {t.element_size()}-byte elements; the storage type needs {elem_bytes}")
    if rows > 0:
        if t.dim() != 2 or t.stride(1) != 1:
            raise ValueError(f"{name} must be a 2-D tensor with contiguous rows")
        if t.shape[0] < rows or t.shape[1] < dim:
            raise ValueError(f"{name} is {tuple(t.shape)}; needs at least ({rows}, {dim})")


def _check_vector(t, name, n):
    if n > 0 and (t is None or not t.is_cuda or t.element_size() != 4 or t.numel() < n):
        raise ValueError(f"{name} must be a GPU float32 vector of at least {n} entries")


def _host(a, dtype):
    a = np.ascontiguousarray(np.asarray(a), dtype=dtype)
    return a


class _Zip:
    """Device copy of a zip CSR/CSC (hipMalloc + synchronous H2D, freed on delete)."""

    def __init__(self, ptr, idx, zip_ids, val=None):
        ptr, idx, zid = _host(ptr, np.int32), _host(idx, np.int32), _host(zip_ids, np.int32)
        va = None if val is None else _host(val, np.float32)
        # The C ABI reads nptr-1 ids and nnz values: check the lengths here.
        if len(zid) != max(len(ptr) - 1, 0) or (va is not None and len(va) != len(idx)):
            raise ColaError(COLA_ERR_INVALID, f"zip: {len(ptr)} pointers need {max(len(ptr) - 1, 0)} ids "
                                              f"(got {len(zid)}) and one value per index")
        h = ctypes.c_void_p()
        vp = lambda a: a.ctypes.data_as(ctypes.c_void_p) if a is not None and a.size else None  # noqa: E731
        check(lib().cola_zip_create(vp(ptr), len(ptr), vp(idx), len(idx), vp(zid), vp(va), ctypes.byref(h)))
        self._h = h
        # Largest ids, for the Python-side buffer checks of the entry points.
        self.max_id = int(zid.max()) if zid.size else -1
        self.max_idx = int(idx.max()) if idx.size else -1
        self._read()

    def _read(self):
        p = [ctypes.c_void_p() for _ in range(5)]
        n = [ctypes.c_int64() for _ in range(3)]
        check(lib().cola_zip_device_arrays(self._h, *[ctypes.byref(x) for x in p], *[ctypes.byref(x) for x in n]))
        self._ptr, self._idx, self._zip, self._val, self._pack = (x.value for x in p)
        self.nptr, self.nnz, self.n_unique = (x.value for x in n)

    @property
    def handle(self):
        return self._h

    def invalidate(self):
        """Drops the plan the single-stage wrapper cached from this (local)
        zip, e.g. after writing new values into its device arrays."""
        check(lib().cola_zip_invalidate(self._h))

    @property
    def num(self):
        """Number of (virtual) rows / columns."""
        return max(self.nptr - 1, 0)

    def __del__(self):
        if getattr(self, "_h", None) and _lib._lib is not None:
            _lib._lib.cola_zip_destroy(self._h)
            self._h = None


class d_csr_zip(_Zip):
    """sparse_data.h:53-77: d_csr_rowPtr, d_csr_colIdx, d_csr_rowIdx_zip, d_csr_val."""

    def __init__(self, rowPtr, colIdx, rowIdx_zip, val=None):
        super().__init__(rowPtr, colIdx, rowIdx_zip, val)

    d_csr_rowPtr = property(lambda s: s._ptr)
    d_csr_colIdx = property(lambda s: s._idx)
    d_csr_rowIdx_zip = property(lambda s: s._zip)
    d_csr_val = property(lambda s: s._val)


class d_csc_zip(_Zip):
    """sparse_data.h:28-52: d_csc_colPtr, d_csc_rowIdx, d_csc_colIdx_zip, d_csc_val,
    plus d_csc_packIdx (packed-buffer row of each virtual column)."""

    def __init__(self, colPtr, rowIdx, colIdx_zip, val=None):
        super().__init__(colPtr, rowIdx, colIdx_zip, val)

    d_csc_colPtr = property(lambda s: s._ptr)
    d_csc_rowIdx = property(lambda s: s._idx)
    d_csc_colIdx_zip = property(lambda s: s._zip)
    d_csc_val = property(lambda s: s._val)
    d_csc_packIdx = property(lambda s: s._pack)


def pack_remote(B_global, remote_cols, packed, dim=None, dtype=F32, scales_global=None,
                packed_scales=None, num_cols=None, stream=None):
    """packed[u, :dim] = B_global[remote_cols[u], :dim] (row strides from the tensors)."""
    dim = dim if dim is not None else B_global.shape[1]
    n = num_cols if num_cols is not None else (remote_cols.numel() if hasattr(remote_cols, "numel") else 0)
    if dtype not in DTYPE_BYTES:
        raise ValueError(f"unknown storage dtype {dtype}")
    _check_matrix(packed, "packed", n, dim, DTYPE_BYTES[dtype])
    _check_matrix(B_global, "B_global", 1 if n > 0 else 0, dim, DTYPE_BYTES[dtype])
    if hasattr(remote_cols, "numel") and remote_cols.numel() < n:
        raise ValueError(f"remote_cols holds {remote_cols.numel()} ids; {n} requested")
    cols_ptr = remote_cols if isinstance(remote_cols, int) else _dptr(remote_cols)
    check(lib().cola_pack_remote(_dptr(B_global), B_global.stride(0), _dptr(scales_global), cols_ptr, int(n),
                                 int(dim), int(dtype), _dptr(packed), packed.stride(0), _dptr(packed_scales),
                                 _stream(stream)))


def ppl0_batch_SpMM(output, input, packed, csr_zip_l, csc_zip_r, local_row_lb, remote_col_lb,
                    local_row_num, remote_col_num, dim, nodesPerPE, peid, stream=None):
    """Reference-semantics partition SpMM (FP32 atomics into output)."""
    if local_row_num + remote_col_num > 0:
        out_rows = max(csr_zip_l.max_id, csc_zip_r.max_idx) + 1
        _check_matrix(output, "output", out_rows, dim, 4)
        _check_matrix(input, "input", csr_zip_l.max_idx + 1 if local_row_num > 0 else 0, dim, 4)
        _check_matrix(packed, "packed", csc_zip_r.n_unique if remote_col_num > 0 else 0, dim, 4)
    check(lib().cola_ppl0_batch_spmm(
        _dptr(output), _dptr(input), _dptr(packed),
        csr_zip_l._ptr, csr_zip_l._idx, csr_zip_l._zip, csr_zip_l._val,
        csc_zip_r._ptr, csc_zip_r._idx, csc_zip_r._pack, csc_zip_r._val,
        int(local_row_lb), int(remote_col_lb), int(local_row_num), int(remote_col_num),
        int(dim), int(nodesPerPE), int(peid), _stream(stream)))


def ppl0_batch_SpMM_global(output, input, csr_zip_l, csc_zip_r, local_row_lb, remote_col_lb, local_row_num,
                           remote_col_num, dim, nodesPerPE, peid, literal=False, stream=None):
    """ppl0_batch_SpMM with the reference's parameter list (cuh:202-221):
    remote global row g is read at input + (g - peid*nodesPerPE)*dim, so
    `input` must be a row slice of one B_global holding every partition (a
    view B_global[peid*P:] is).  literal=False: cached plan (no atomics);
    literal=True: the virtual-row kernel with FP32 atomics."""
    if local_row_num + remote_col_num > 0:
        _check_matrix(output, "output", max(csr_zip_l.max_id, csc_zip_r.max_idx) + 1, dim, 4)
        _check_matrix(input, "input", csr_zip_l.max_idx + 1 if local_row_num > 0 else 0, dim, 4)
        if input.stride(0) != dim or output.stride(0) != dim:
            raise ValueError("input and output must have rows of exactly dim elements")
        base = input.data_ptr() - peid * nodesPerPE * dim * 4
        if remote_col_num > 0 and (base < input.untyped_storage().data_ptr() or
                                   base + (csc_zip_r.max_id + 1) * dim * 4 >
                                   input.untyped_storage().data_ptr() + input.untyped_storage().nbytes()):
            raise ValueError("input must be the partition's slice of a B_global holding every remote row")
    check(lib().cola_ppl0_batch_spmm_global(
        _dptr(output), _dptr(input), csr_zip_l._ptr, csr_zip_l._idx, csr_zip_l._zip, csr_zip_l._val,
        csc_zip_r._ptr, csc_zip_r._idx, csc_zip_r._zip, csc_zip_r._val,
        int(local_row_lb), int(remote_col_lb), int(local_row_num), int(remote_col_num),
        int(dim), int(nodesPerPE), int(peid), 1 if literal else 0, _stream(stream)))


def ppl0_plans_clear():
    """Drops the plans ppl0_batch_SpMM_global cached."""
    check(lib().cola_ppl0_plans_clear())


def dist_spmm_ppl_batch(output, B_global, packed, csr_zip_l, csc_zip_r, dim, nodesPerPE, peid, stream=None):
    """Single-stage distSpMM_ppl_batch_wrapper: pack the unique remote rows, then
    accumulate every virtual local row and remote column into output."""
    _check_matrix(output, "output", max(csr_zip_l.max_id, csc_zip_r.max_idx) + 1, dim, 4)
    b_rows = max(peid * nodesPerPE + csr_zip_l.max_idx + 1 if csr_zip_l.nnz else 0, csc_zip_r.max_id + 1)
    _check_matrix(B_global, "B_global", b_rows, dim, 4)
    _check_matrix(packed, "packed", csc_zip_r.n_unique, dim, 4)
    check(lib().cola_dist_spmm_ppl_batch(_dptr(output), _dptr(B_global), _dptr(packed), csr_zip_l.handle,
                                         csc_zip_r.handle, int(dim), int(nodesPerPE), int(peid),
                                         _stream(stream)))


class Plan:
    """Degree-binned, atomic-free plan of one partition (device resident)."""

    def __init__(self, partition, hub_threshold=0, chunk_nnz=0, kahan_degree=0, max_dim=0, short_degree=0,
                 col_panels=0, row_order=0, hot_cols=0):
        o = _lib.PlanOpts(int(hub_threshold), int(chunk_nnz), int(kahan_degree), int(max_dim), int(short_degree),
                          int(col_panels), int(row_order), int(hot_cols))
        h = ctypes.c_void_p()
        check(lib().cola_plan_create(partition.handle, ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        i = _lib.PlanInfo()
        check(lib().cola_plan_get_info(h, ctypes.byref(i)))
        self.info = {f: getattr(i, f) for f, _ in _lib.PlanInfo._fields_}
        rc = ctypes.c_void_p()
        check(lib().cola_plan_remote_cols(h, ctypes.byref(rc)))
        self.remote_cols_ptr = rc.value or 0
        self.U_r = self.info["U_r"]
        self.U_l = self.info["U_l"]
        self.nnz = self.info["nnz"]
        self.num_rows = self.info["num_rows"]
        # Largest remote (global) column, the last B row the pack reads: the
        # partition's remote columns are sorted.
        arrs = _lib.PartitionArrays()
        check(lib().cola_partition_get_arrays(partition.handle, ctypes.byref(arrs)))
        self.max_remote_col = int(arrs.csc_colIdx_r[self.U_r - 1]) if self.U_r > 0 else -1

    @property
    def handle(self):
        return self._h

    def __del__(self):
        if getattr(self, "_h", None) and _lib._lib is not None:
            _lib._lib.cola_plan_destroy(self._h)
            self._h = None

    def _check_call(self, C, B_local, packed, dim, dtype, scales_local, packed_scales):
        if dtype not in DTYPE_BYTES:
            raise ValueError(f"unknown storage dtype {dtype}")
        eb = DTYPE_BYTES[dtype]
        _check_matrix(C, "C", self.num_rows, dim, 4)
        _check_matrix(B_local, "B_local", self.num_rows if self.U_l > 0 else 0, dim, eb)
        _check_matrix(packed, "packed", self.U_r, dim, eb)
        if dtype == FP8_E4M3:
            _check_vector(scales_local, "scales_local", self.num_rows if self.U_l > 0 else 0)
            _check_vector(packed_scales, "packed_scales", self.U_r)

    def pack(self, B_global, packed, dim=None, dtype=F32, scales_global=None, packed_scales=None, stream=None):
        dim = dim if dim is not None else B_global.shape[1]
        if dtype not in DTYPE_BYTES:
            raise ValueError(f"unknown storage dtype {dtype}")
        _check_matrix(packed, "packed", self.U_r, dim, DTYPE_BYTES[dtype])
        _check_matrix(B_global, "B_global", self.max_remote_col + 1, dim, DTYPE_BYTES[dtype])
        if dtype == FP8_E4M3:
            _check_vector(packed_scales, "packed_scales", self.U_r)
        pack_remote(B_global, self.remote_cols_ptr, packed, dim=dim, dtype=dtype, scales_global=scales_global,
                    packed_scales=packed_scales, num_cols=self.U_r, stream=stream)

    def spmm(self, C, B_local, packed, dim=None, dtype=F32, flags=ACCUMULATE, scales_local=None,
             packed_scales=None, stream=None):
        dim = dim if dim is not None else C.shape[1]
        self._check_call(C, B_local, packed, dim, dtype, scales_local, packed_scales)
        ldp = packed.stride(0) if packed is not None and packed.dim() == 2 else dim
        ldb = B_local.stride(0) if B_local is not None else dim
        check(lib().cola_spmm_partition(self._h, _dptr(C), C.stride(0), _dptr(B_local), ldb,
                                        _dptr(scales_local), _dptr(packed), ldp, _dptr(packed_scales),
                                        int(dim), int(dtype), int(flags), _stream(stream)))

    TUNING_FIELDS = [f for f, _ in _lib.PlanTuning._fields_]

    @property
    def tuning(self):
        t = _lib.PlanTuning()
        check(lib().cola_plan_get_tuning(self._h, ctypes.byref(t)))
        return {f: getattr(t, f) for f in self.TUNING_FIELDS}

    @tuning.setter
    def tuning(self, d):
        cur = self.tuning
        cur.update(d)
        t = _lib.PlanTuning(*[int(cur[f]) for f in self.TUNING_FIELDS])
        check(lib().cola_plan_set_tuning(self._h, ctypes.byref(t)))

    def autotune(self, C, B_local, packed, dim=None, dtype=F32, flags=OVERWRITE, scales_local=None,
                 packed_scales=None, iters=3, stream=None):
        """Times candidate launch parameters on these buffers and keeps the
        fastest; C is overwritten with A_p·B.  Returns (tuning dict, ms)."""
        dim = dim if dim is not None else C.shape[1]
        self._check_call(C, B_local, packed, dim, dtype, scales_local, packed_scales)
        ldp = packed.stride(0) if packed is not None and packed.dim() == 2 else dim
        best = _lib.PlanTuning()
        ms = ctypes.c_float()
        ldb = B_local.stride(0) if B_local is not None else dim
        check(lib().cola_plan_autotune(self._h, _dptr(C), C.stride(0), _dptr(B_local), ldb,
                                       _dptr(scales_local), _dptr(packed), ldp, _dptr(packed_scales), int(dim),
                                       int(dtype), int(flags), int(iters), _stream(stream), ctypes.byref(best),
                                       ctypes.byref(ms)))
        return {f: getattr(best, f) for f in self.TUNING_FIELDS}, ms.value


class Batch:
    """Many independent SpMMs in one persistent launch.

    problems: list of dicts with keys plan, B, C and optionally dim, scales,
    packed, packed_scales (device tensors; kept alive by the batch)."""

    def __init__(self, problems, dtype=F32, flags=OVERWRITE):
        self._keep = []
        arr = (_lib.BatchProblem * max(1, len(problems)))()
        for k, pr in enumerate(problems):
            B, C = pr["B"], pr["C"]
            packed = pr.get("packed")
            pr["plan"]._check_call(C, B, packed, int(pr.get("dim", C.shape[1])), dtype, pr.get("scales"),
                                   pr.get("packed_scales"))
            arr[k].plan = pr["plan"].handle.value
            arr[k].dim = int(pr.get("dim", C.shape[1]))
            arr[k].B = B.data_ptr() if B is not None else None
            arr[k].ld_B = B.stride(0) if B is not None else arr[k].dim
            arr[k].scales = pr["scales"].data_ptr() if pr.get("scales") is not None else None
            arr[k].packed = packed.data_ptr() if packed is not None else None
            arr[k].ld_packed = packed.stride(0) if packed is not None else arr[k].dim
            arr[k].packed_scales = pr["packed_scales"].data_ptr() if pr.get("packed_scales") is not None else None
            arr[k].C = C.data_ptr()
            arr[k].ld_C = C.stride(0)
            self._keep.append(pr)
        h = ctypes.c_void_p()
        check(lib().cola_batch_create(arr, len(problems), int(dtype), int(flags), ctypes.byref(h)))
        self._h = h
        self.n = len(problems)

    def __del__(self):
        if getattr(self, "_h", None) and _lib._lib is not None:
            _lib._lib.cola_batch_destroy(self._h)
            self._h = None

    def run(self, stream=None):
        check(lib().cola_batch_run(self._h, _stream(stream)))

    def capture(self, stream):
        check(lib().cola_batch_capture(self._h, _stream(stream)))

    def replay(self, stream=None):
        check(lib().cola_batch_replay(self._h, _stream(stream)))

    def autotune(self, iters=3, stream=None):
        ms = ctypes.c_float()
        check(lib().cola_batch_autotune(self._h, int(iters), _stream(stream), ctypes.byref(ms)))
        return ms.value

    def set_tuning(self, problem=-1, col_tile=0, rpw_short=0, rpw_long=0, blocks_per_cu=0):
        t = _lib.BatchTuning(int(col_tile), int(rpw_short), int(rpw_long))
        check(lib().cola_batch_set_tuning(self._h, int(problem), ctypes.byref(t), int(blocks_per_cu)))

    def tuning(self, problem):
        t = _lib.BatchTuning()
        bpc = ctypes.c_int()
        check(lib().cola_batch_get_tuning(self._h, int(problem), ctypes.byref(t), ctypes.byref(bpc)))
        return dict(col_tile=t.col_tile, rpw_short=t.rpw_short, rpw_long=t.rpw_long, blocks_per_cu=bpc.value)


REQUEST_RECORD = np.dtype([("arrival_us", "f8"), ("dispatch_us", "f8"), ("done_us", "f8"), ("resident", "i4"),
                           ("batch_size", "i4")])


class Server:
    """Continuous-batching SpMM server over resident graphs (see
    cola_server_run_trace in include/cola_spmm.h)."""

    def __init__(self, memory_budget, dtype=BF16, max_batch=64, max_inflight=2, kahan=False, plain=False,
                 blocks_per_cu=0):
        o = _lib.ServerOpts(int(memory_budget), int(max_batch), int(max_inflight), int(dtype),
                            KAHAN if kahan else 0, 1 if plain else 0, int(blocks_per_cu))
        h = ctypes.c_void_p()
        check(lib().cola_server_create(ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        self._keep = []
        self._out_shape = {}
        self.dtype = dtype

    def __del__(self):
        if getattr(self, "_h", None) and _lib._lib is not None:
            _lib._lib.cola_server_destroy(self._h)
            self._h = None

    def add_resident(self, plan, B, dim=None, scales=None):
        rid = ctypes.c_int()
        dim = dim if dim is not None else B.shape[1]
        _check_matrix(B, "B", plan.num_rows if plan.U_l > 0 else 0, dim, DTYPE_BYTES[self.dtype])
        if self.dtype == FP8_E4M3:
            _check_vector(scales, "scales", plan.num_rows if plan.U_l > 0 else 0)
        check(lib().cola_server_add_resident(self._h, plan.handle, _dptr(B), B.stride(0), _dptr(scales), int(dim),
                                             ctypes.byref(rid)))
        self._keep.append((plan, B, scales))
        # Rows and dim of each resident's output, for the checks of run_trace.
        self._out_shape[rid.value] = (plan.num_rows, int(dim))
        return rid.value

    def set_tuning(self, dim, col_tile=0, rpw_short=0, rpw_long=0):
        t = _lib.BatchTuning(int(col_tile), int(rpw_short), int(rpw_long))
        check(lib().cola_server_set_tuning(self._h, int(dim), ctypes.byref(t)))

    def run_trace(self, arrival_us, resident, outputs=None):
        """Serves the trace in real time; returns (records, wall_us, peak_pool_bytes)."""
        arr = np.ascontiguousarray(arrival_us, np.float64)
        res = np.ascontiguousarray(resident, np.int32)
        n = len(arr)
        recs = np.zeros(n, REQUEST_RECORD)
        outp = None
        if outputs is not None:
            if len(outputs) != n:
                raise ValueError(f"outputs has {len(outputs)} entries; the trace has {n} requests")
            for i, o in enumerate(outputs):
                if o is None:
                    continue
                rid = int(res[i])
                if rid not in self._out_shape:
                    raise ValueError(f"request {i}: unknown resident id {rid}")
                rows, dim = self._out_shape[rid]
                # The server writes rows x dim floats with a row stride of dim.
                _check_matrix(o, f"outputs[{i}]", rows, dim, 4)
                if rows > 0 and o.stride(0) != dim:
                    raise ValueError(f"outputs[{i}] has row stride {o.stride(0)}; the server writes rows of {dim}")
            outp = (ctypes.c_void_p * n)(*[(o.data_ptr() if o is not None else None) for o in outputs])
        wall = ctypes.c_double()
        peak = ctypes.c_int64()
        check(lib().cola_server_run_trace(self._h, n, arr.ctypes.data_as(ctypes.c_void_p),
                                          res.ctypes.data_as(ctypes.c_void_p), outp,
                                          recs.ctypes.data_as(ctypes.c_void_p), ctypes.byref(wall),
                                          ctypes.byref(peak)))
        return recs, wall.value, peak.value


def quantize_rows(src, dtype, dst=None, scales=None, stream=None):
    """FP32 rows -> storage dtype (RNE; FP8 with per-row scale max|x|/448)."""
    torch = _torch()
    rows, dim = src.shape
    tdt = {F32: torch.float32, BF16: torch.uint16, F16: torch.uint16, FP8_E4M3: torch.uint8}[dtype]
    if dst is None:
        dst = torch.empty((rows, dim), dtype=tdt, device=src.device)
    if dtype == FP8_E4M3 and scales is None:
        scales = torch.empty(rows, dtype=torch.float32, device=src.device)
    check(lib().cola_quantize_rows(_dptr(src), src.stride(0), rows, dim, int(dtype), _dptr(dst), dst.stride(0),
                                   _dptr(scales), _stream(stream)))
    return dst, scales


def error_norms(C, R, out=None, stream=None):
    """(sum (C-R)^2, sum R^2, max_i |C_i-R_i|^2/|R_i|^2) computed on the device in FP64."""
    torch = _torch()
    rows, dim = C.shape
    if out is None:
        out = torch.empty(3, dtype=torch.float64, device=C.device)
    check(lib().cola_error_norms(_dptr(C), C.stride(0), _dptr(R), R.stride(0), rows, dim, _dptr(out),
                                 _stream(stream)))
    return out
