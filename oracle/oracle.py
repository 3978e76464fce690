"""Python side of the oracle — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the checker.  The product (cola_amd/) never imports it.

Contents:
  - ctypes binding of oracle/liboracle.so (oracle.c: FP64 SpMM, FP32
    restatement of the reference kernel, nnz_split, SAG + check_result,
    BF16 / FP8-e4m3 conversions, OpenMP CPU baseline);
  - row_partition_np: numpy restatement of prep_torch_par.py:24-79
    (row_partition) for one rank, written with numpy sorting instead of
    scipy's coo -> csr/csc conversion.  Pinned by the structural invariants the
    reference asserts (prep_torch_par.py:50,53,71,74) and by brute-force
    reconstruction of the row panel (tests/test_oracle.py).
  - par_files_from_mtx: prep_torch_par.py:14-101 for a MatrixMarket file,
    restated step by step with the same third-party routines the reference
    calls (scipy.io.mmread, coo_matrix(...).tocsr() / .tocsc(), np.unique):
    the global files and every rank's six partition files as int32 arrays.
    Pinned by the structural invariants of row_partition_np and by equality
    with row_partition_np on the deduplicated CSR (tests/test_mtx.py).
  - ppl_files_from_mtx: prep_torch_ppl.py:24-49 for a MatrixMarket file,
    restated with the routines it calls (scipy.io.mmread, np.unique): each
    stage's sorted unique columns.  Pinned by brute force: every stage's
    columns are exactly the columns of the stage's rows in the dense
    matrix, and the union over stages covers the rows the stages span
    (tests/test_mtx.py).
  - dense_partition_product: brute-force dense A_p·B for tiny graphs.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_spmm_f64.argtypes = [_vp, _vp, _vp, _i64, _i64, _vp, _i64, _int, _vp]
        L.oracle_spmm_f64.restype = None
        L.oracle_ppl0_f32.argtypes = [_vp, _vp, _int, _int, _int, _vp, _vp, _vp, _vp, _int, _vp, _vp, _vp,
                                      _vp, _int]
        L.oracle_ppl0_f32.restype = None
        L.oracle_nnz_split.argtypes = [_vp, _i64, _vp, _i64, _vp, _vp, ctypes.POINTER(ctypes.c_int32)]
        L.oracle_nnz_split.restype = _i64
        L.oracle_sag_unweighted.argtypes = [_i64, _i64, _int, _vp, _vp, _vp, _vp]
        L.oracle_sag_unweighted.restype = None
        L.oracle_check_result.argtypes = [_i64, _int, _vp, _vp]
        L.oracle_check_result.restype = _i64
        L.oracle_bf16_from_f32.argtypes = [ctypes.c_float]
        L.oracle_bf16_from_f32.restype = ctypes.c_uint16
        L.oracle_bf16_to_f32.argtypes = [ctypes.c_uint16]
        L.oracle_bf16_to_f32.restype = ctypes.c_float
        L.oracle_fp8e4m3_to_f32.argtypes = [ctypes.c_uint8]
        L.oracle_fp8e4m3_to_f32.restype = ctypes.c_float
        L.oracle_fp8e4m3_from_f32.argtypes = [ctypes.c_float]
        L.oracle_fp8e4m3_from_f32.restype = ctypes.c_uint8
        L.oracle_fp8_row_scale.argtypes = [_vp, _int]
        L.oracle_fp8_row_scale.restype = ctypes.c_float
        L.oracle_spmm_omp_f32.argtypes = [_vp, _vp, _vp, _i64, _i64, _vp, _i64, _int, _vp]
        L.oracle_spmm_omp_f32.restype = _int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_vp) if a is not None and a.size else None


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def spmm_f64(row_ptr, col_idx, val, lb, ub, B):
    """FP64 C = A[lb:ub, :]·B (weighted; val=None means 1.0)."""
    rp, ci = _c(row_ptr, np.int64), _c(col_idx, np.int32)
    va = None if val is None else _c(val, np.float32)
    B = _c(B, np.float32)
    N = B.shape[1]
    C = np.zeros((max(ub - lb, 0), N), dtype=np.float64)
    if ub > lb:
        lib().oracle_spmm_f64(_p(rp), _p(ci), _p(va), lb, ub, _p(B), B.shape[1], N, _p(C))
    return C


def ppl0_f32(C, B_global, dim, nodesPerPE, peid, rowPtr_l, colIdx_l, rowIdx_l, val_l, colPtr_r, rowIdx_r,
             colIdx_r, val_r):
    """FP32 restatement of the reference kernel (virtual rows / columns), accumulating into C."""
    assert C.dtype == np.float32 and C.flags.c_contiguous
    B = _c(B_global, np.float32)
    a = [_c(x, np.int32) for x in (rowPtr_l, colIdx_l, rowIdx_l)]
    vl = _c(val_l, np.float32)
    r = [_c(x, np.int32) for x in (colPtr_r, rowIdx_r, colIdx_r)]
    vr = _c(val_r, np.float32)
    L = max(len(a[0]) - 1, 0)
    R = max(len(r[0]) - 1, 0)
    lib().oracle_ppl0_f32(_p(C), _p(B), dim, nodesPerPE, peid, _p(a[0]), _p(a[1]), _p(a[2]), _p(vl), L,
                          _p(r[0]), _p(r[1]), _p(r[2]), _p(vr), R)
    return C


def nnz_split(rowIdx, rowPtr, nnz):
    ri, rp = _c(rowIdx, np.int32), _c(rowPtr, np.int32)
    gran = ctypes.c_int32()
    nv = lib().oracle_nnz_split(_p(ri), len(ri), _p(rp), nnz, None, None, ctypes.byref(gran))
    if len(ri) == 0:
        return np.zeros(0, np.int32), np.zeros(0, np.int32), 0
    t = np.zeros(max(nv, 1), np.int32)
    p = np.zeros(nv + 1, np.int32)
    lib().oracle_nnz_split(_p(ri), len(ri), _p(rp), nnz, _p(t), _p(p), ctypes.byref(gran))
    return t[:nv], p, gran.value


def sag_unweighted(lb, ub, row_ptr, col_idx, B):
    rp, ci, B = _c(row_ptr, np.int64), _c(col_idx, np.int32), _c(B, np.float32)
    N = B.shape[1]
    C = np.zeros((ub - lb, N), np.float32)
    lib().oracle_sag_unweighted(lb, ub, N, _p(rp), _p(ci), _p(B), _p(C))
    return C


def check_result(C, C_ref):
    C, C_ref = _c(C, np.float32), _c(C_ref, np.float32)
    return int(lib().oracle_check_result(C.shape[0], C.shape[1], _p(C), _p(C_ref)))


def bf16_from_f32(x):
    f = lib().oracle_bf16_from_f32
    return np.array([f(float(v)) for v in np.asarray(x, np.float32).ravel()], np.uint16).reshape(np.shape(x))


def fp8_to_f32(codes):
    f = lib().oracle_fp8e4m3_to_f32
    return np.array([f(int(c)) for c in np.asarray(codes).ravel()], np.float32).reshape(np.shape(codes))


def fp8_from_f32(x):
    f = lib().oracle_fp8e4m3_from_f32
    return np.array([f(float(v)) for v in np.asarray(x, np.float32).ravel()], np.uint8).reshape(np.shape(x))


def fp8_row_scale(row):
    r = _c(row, np.float32)
    return lib().oracle_fp8_row_scale(_p(r), len(r))


def spmm_omp_f32(row_ptr, col_idx, val, lb, ub, B, C=None):
    """OpenMP FP32 C = A[lb:ub,:]·B; returns (C, threads)."""
    rp, ci, va, B = _c(row_ptr, np.int64), _c(col_idx, np.int32), _c(val, np.float32), _c(B, np.float32)
    N = B.shape[1]
    if C is None:
        C = np.zeros((ub - lb, N), np.float32)
    th = lib().oracle_spmm_omp_f32(_p(rp), _p(ci), _p(va), lb, ub, _p(B), B.shape[1], N, _p(C))
    return C, th


def row_partition_np(row_ptr, col_idx, val, n, nGPU, rank):
    """prep_torch_par.py:24-79 for one rank, restated with numpy.

    Rows [rank*P, min(n,(rank+1)*P)) are split by column ownership.  Local part:
    CSR with empty rows collapsed (unique row pointer, unique row ids), columns
    taken mod P.  Remote part: CSC over the unique remote global columns, rows
    ascending within a column.  Values ride along (the reference writes none).
    """
    row_ptr = np.asarray(row_ptr, np.int64)
    col_idx = np.asarray(col_idx, np.int64)
    val = np.asarray(val, np.float32)
    P = (n + nGPU - 1) // nGPU
    lb, ub = rank * P, min(n, (rank + 1) * P)
    M = max(ub - lb, 0)
    k0, k1 = (row_ptr[lb], row_ptr[ub]) if M > 0 else (0, 0)
    rows = np.repeat(np.arange(M, dtype=np.int64), np.diff(row_ptr[lb:ub + 1]) if M > 0 else 0)
    cols = col_idx[k0:k1]
    vals = val[k0:k1]
    local = (cols >= rank * P) & (cols < (rank + 1) * P)
    # local CSR (stable order keeps the row-major, column-ascending input order)
    lr, lc, lv = rows[local], cols[local] % P, vals[local]
    o = np.lexsort((lc, lr))
    lr, lc, lv = lr[o], lc[o], lv[o]
    rowIdx_l = np.unique(lr)
    counts = np.bincount(lr, minlength=M)
    indptr = np.concatenate([[0], np.cumsum(counts)])
    rowPtr_l = np.unique(indptr)
    # remote CSC over unique columns
    rr, rc, rv = rows[~local], cols[~local], vals[~local]
    o = np.lexsort((rr, rc))
    rr, rc, rv = rr[o], rc[o], rv[o]
    colIdx_r = np.unique(rc)
    ccounts = np.bincount(np.searchsorted(colIdx_r, rc), minlength=len(colIdx_r)) if len(rc) else np.zeros(0, int)
    colPtr_r = np.concatenate([[0], np.cumsum(ccounts)])
    return dict(csr_rowPtr_l=rowPtr_l.astype(np.int32), csr_colIdx_l=lc.astype(np.int32),
                csr_rowIdx_l=rowIdx_l.astype(np.int32), csr_val_l=lv.astype(np.float32),
                csc_colPtr_r=colPtr_r.astype(np.int32), csc_rowIdx_r=rr.astype(np.int32),
                csc_colIdx_r=colIdx_r.astype(np.int32), csc_val_r=rv.astype(np.float32))


def par_files_from_mtx(path, nGPU):
    """prep_torch_par.py:14-101 for one MatrixMarket file and one GPU count.

    Returns {"global_rowPtr": .., "global_colIdx": .., (stem, rank): array}
    with the int32 contents the reference writes (file name stems as in
    prep_torch_par.py:56-58,77-79).  Values are not written by the reference.
    """
    import scipy.io as sio
    from scipy.sparse import coo_matrix
    A = sio.mmread(path)                                   # par.py:19
    A = coo_matrix(A)
    out = {}
    g = A.tocsr()                                          # par.py:97
    out["global_rowPtr"] = np.asarray(g.indptr).astype(np.int32)
    out["global_colIdx"] = np.asarray(g.indices).astype(np.int32)
    row, col = np.asarray(A.row), np.asarray(A.col)        # par.py:28-29
    num_nodes = A.shape[0]
    P = (num_nodes + nGPU - 1) // nGPU                     # par.py:31
    for rank in range(nGPU):
        lb, ub = rank * P, min((rank + 1) * P, num_nodes)
        m = (row >= lb) & (row < ub)
        r, c = row[m] - lb, col[m]
        lm = (rank * P <= c) & (c < (rank + 1) * P)        # par.py:40
        lr, lc = r[lm], c[lm]
        L = coo_matrix((np.ones(len(lr)), (lr, lc % P)), shape=(ub - lb, P)).tocsr()
        out[("csr_rowPtr_l", rank)] = np.unique(np.asarray(L.indptr, dtype=np.int32))
        out[("csr_colIdx_l", rank)] = np.asarray(L.indices, dtype=np.int32)
        out[("csr_rowIdx_l", rank)] = np.unique(lr).astype(np.int32)
        rr, rc = r[~lm], c[~lm]
        R = coo_matrix((np.ones(len(rr)), (rr, rc)), shape=(ub - lb, num_nodes)).tocsc()
        out[("csc_colIdx_r", rank)] = np.unique(np.asarray(rc, dtype=np.int32))
        out[("csc_colPtr_r", rank)] = np.unique(np.asarray(R.indptr, dtype=np.int32))
        out[("csc_rowIdx_r", rank)] = np.asarray(R.indices, dtype=np.int32)
    return out


def ppl_files_from_mtx(path, nGPU, gs):
    """prep_torch_ppl.py:24-49 row_partition for one MatrixMarket file, one
    GPU count and one group size: {stage: sorted unique columns of the
    stage's rows} with the dtype the reference writes (mmread's COO column
    dtype, not cast: int32 here)."""
    import scipy.io as sio
    A = sio.mmread(path)                                   # ppl.py:18
    row, col = np.asarray(A.row), np.asarray(A.col)        # ppl.py:30-31
    num_nodes = A.shape[0]                                 # ppl.py:32
    stages = nGPU // gs                                    # ppl.py:26
    P = (num_nodes + nGPU - 1) // nGPU                     # ppl.py:34
    out = {}
    for s in range(stages):                                # ppl.py:37-39
        lb, ub = s * gs * P, min((s + 1) * gs * P, num_nodes)
        m = (row >= lb) & (row < ub)                       # ppl.py:45
        u = np.unique(col[m])                              # ppl.py:47
        assert np.all(np.diff(u) > 0)                      # ppl.py:49
        out[s] = u
    return out


def dense_partition_product(row_ptr, col_idx, val, n, lb, ub, B):
    """Brute force: build the dense (ub-lb) x n panel and multiply in FP64."""
    A = np.zeros((ub - lb, n), np.float64)
    for i in range(lb, ub):
        for k in range(row_ptr[i], row_ptr[i + 1]):
            A[i - lb, col_idx[k]] += val[k] if val is not None else 1.0
    return A @ np.asarray(B, np.float64)
