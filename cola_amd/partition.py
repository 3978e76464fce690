"""Host-side graph and partition objects (mirrors of the reference's data prep
and load path, backed by the C++ library).

Reference behaviour mirrored (paths relative to the CoLa tree):
  prep_torch_par.py:14-20   load_mtx       -> Graph.from_mtx
  prep_torch_par.py:24-79   row_partition  -> Partition.row_partition / Graph.partition
  rabbit_order_write.py:46-77  reorder + mmwrite -> Graph.reorder / Graph.write_mtx
  prep_torch_par.py:93-110  serialize      -> serialize
  prep_torch_ppl.py:24-84   row_partition / serialize / process_graph (stage files)
                            -> serialize_ppl / process_graph_ppl, stage_cols
  src/distSpMM_ppl_batch.cu:49-99  partition load + 1.0 values -> Partition.load
  include/dist_spmm_ppl_batch.cuh:7-27   bin2vec   -> bin2vec
  include/dist_spmm_ppl_batch.cuh:29-56  nnz_split -> nnz_split
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import check, lib


def _np_copy(ptr, count, dtype):
    if count == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True)


def _i32(a):
    return np.ascontiguousarray(np.asarray(a), dtype=np.int32)


def _f32(a):
    return np.ascontiguousarray(np.asarray(a), dtype=np.float32)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None and a.size else None


class Graph:
    """Global CSR graph in host memory (int64 row pointer, int32 columns, FP32 values)."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle) if not isinstance(handle, ctypes.c_void_p) else handle

    @classmethod
    def rmat(cls, log_n, avg_degree, a=0.57, b=0.19, c=0.19, seed=1, reorder=False, gcn=True):
        """Seeded R-MAT graph, symmetrised, self-loops and duplicates dropped,
        GCN-normalised values 1/sqrt(d_i d_j) (or 1.0 when gcn=False);
        reorder=True (or "cm") applies the Cuthill-McKee locality order,
        "community" the community order (Rabbit Order's two steps)."""
        codes = {False: 0, None: 0, "none": 0, True: 1, "cm": 1, "community": 2}
        if reorder not in codes:
            raise ValueError(f"rmat: reorder must be one of {list(codes)}, got {reorder!r}")
        h = ctypes.c_void_p()
        check(lib().cola_rmat_graph(int(log_n), float(avg_degree), float(a), float(b), float(c),
                                    int(seed), codes[reorder], 1 if gcn else 0, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_csr(cls, row_ptr, col_idx, val=None):
        """Graph from CSR arrays; repeated (row, col) entries are summed."""
        rp = np.ascontiguousarray(np.asarray(row_ptr), dtype=np.int64)
        ci = _i32(col_idx)
        va = _f32(val) if val is not None else None
        h = ctypes.c_void_p()
        check(lib().cola_graph_from_csr(len(rp) - 1, _ptr(rp) if rp.size else None, _ptr(ci),
                                        _ptr(va), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_coo(cls, n, rows, cols, vals=None):
        rows = np.asarray(rows, dtype=np.int64)
        cols = np.asarray(cols, dtype=np.int64)
        order = np.argsort(rows, kind="stable")
        counts = np.bincount(rows, minlength=n)
        rp = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(counts, out=rp[1:])
        v = None if vals is None else np.asarray(vals, dtype=np.float32)[order]
        return cls.from_csr(rp, cols[order], v)

    @classmethod
    def from_mtx(cls, path, values="file"):
        """Graph from a MatrixMarket coordinate file (what scipy.io.mmread +
        tocsr() give: symmetric entries mirrored, duplicates summed).
        values: "file" (the file's values; pattern files 1.0), "one" (every
        entry 1.0, the reference driver's override) or "gcn"."""
        mode = {"file": 0, "one": 1, "gcn": 2}[values]
        h = ctypes.c_void_p()
        check(lib().cola_graph_from_mtx(os.fsencode(path), mode, ctypes.byref(h)))
        return cls(h)

    def write_mtx(self, path, with_values=True):
        """Writes a "coordinate real general" (or "pattern general") file."""
        check(lib().cola_graph_write_mtx(self._h, os.fsencode(path), 1 if with_values else 0))

    def reorder(self, method="cm"):
        """Relabelled copy and its new->old permutation.  method: "cm"
        (Cuthill-McKee), "rcm" (reversed), "degree" (degree descending) or
        "community" (Rabbit Order's aggregation and depth-first numbering)."""
        m = {"cm": 1, "rcm": 2, "degree": 3, "community": 4}[method]
        perm = np.empty(self.n, dtype=np.int32)
        h = ctypes.c_void_p()
        check(lib().cola_graph_reorder(self._h, m, _ptr(perm), ctypes.byref(h)))
        return Graph(h), perm

    def __del__(self):
        if getattr(self, "_h", None) and _lib._lib is not None:
            _lib._lib.cola_graph_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def info(self):
        n, nnz, md = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(lib().cola_graph_info(self._h, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(md)))
        return n.value, nnz.value, md.value

    @property
    def n(self):
        return self.info()[0]

    @property
    def nnz(self):
        return self.info()[1]

    def arrays(self):
        """(row_ptr int64, col_idx int32, val float32) copies."""
        n, nnz, _ = self.info()
        rp, ci, va = _lib.P_i64(), _lib.P_i32(), _lib.P_f32()
        check(lib().cola_graph_arrays(self._h, ctypes.byref(rp), ctypes.byref(ci), ctypes.byref(va)))
        return (_np_copy(rp, n + 1, np.int64), _np_copy(ci, nnz, np.int32), _np_copy(va, nnz, np.float32))

    def partition(self, nGPU, rank):
        return Partition.row_partition(self, nGPU, rank)


_FIELDS = ["csr_rowPtr_l", "csr_colIdx_l", "csr_rowIdx_l", "csr_val_l",
           "csc_colPtr_r", "csc_rowIdx_r", "csc_colIdx_r", "csc_val_r"]


class Partition:
    """One rank's row panel in the prep_torch_par.py layout (host memory)."""

    def __init__(self, handle):
        self._h = handle
        self._cache = None

    @classmethod
    def row_partition(cls, graph, nGPU, rank):
        h = ctypes.c_void_p()
        check(lib().cola_row_partition(graph.handle, int(nGPU), int(rank), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def load(cls, bin_path, graph_name, nGPU, rank):
        """Reads the six per-rank files (+ optional value files) with checks."""
        h = ctypes.c_void_p()
        check(lib().cola_partition_load(os.fsencode(bin_path), graph_name.encode(), int(nGPU), int(rank),
                                        ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_arrays(cls, num_nodes, nGPU, rank, csr_rowPtr_l, csr_colIdx_l, csr_rowIdx_l,
                    csc_colPtr_r, csc_rowIdx_r, csc_colIdx_r, csr_val_l=None, csc_val_r=None):
        arrs = {k: v for k, v in dict(csr_rowPtr_l=_i32(csr_rowPtr_l), csr_colIdx_l=_i32(csr_colIdx_l),
                                      csr_rowIdx_l=_i32(csr_rowIdx_l), csc_colPtr_r=_i32(csc_colPtr_r),
                                      csc_rowIdx_r=_i32(csc_rowIdx_r), csc_colIdx_r=_i32(csc_colIdx_r),
                                      csr_val_l=None if csr_val_l is None else _f32(csr_val_l),
                                      csc_val_r=None if csc_val_r is None else _f32(csc_val_r)).items()}
        # The C ABI takes the counts L, nnz_l, U_r, nnz_r and reads that many
        # elements from each array: check every length against them first.
        L, nl = len(arrs["csr_rowIdx_l"]), len(arrs["csr_colIdx_l"])
        U, nr = len(arrs["csc_colIdx_r"]), len(arrs["csc_rowIdx_r"])
        want = dict(csr_rowPtr_l=L + 1, csc_colPtr_r=U + 1, csr_val_l=nl, csc_val_r=nr)
        for k, n in want.items():
            if arrs[k] is not None and len(arrs[k]) != n:
                raise _lib.ColaError(_lib.COLA_ERR_INVALID, f"partition_from_arrays: len({k}) = {len(arrs[k])}, "
                                                            f"expected {n}")
        pa = _lib.PartitionArrays()
        keep = []
        for k in _FIELDS:
            a = arrs[k]
            if a is None:
                setattr(pa, k, None)
                continue
            if a.size == 0:  # non-NULL pointer for empty arrays
                a = np.zeros(1, dtype=a.dtype)
            keep.append(a)
            typ = _lib.P_f32 if "val" in k else _lib.P_i32
            setattr(pa, k, a.ctypes.data_as(typ))
        h = ctypes.c_void_p()
        check(lib().cola_partition_from_arrays(int(num_nodes), int(nGPU), int(rank), ctypes.byref(pa),
                                               len(arrs["csr_rowIdx_l"]), len(arrs["csr_colIdx_l"]),
                                               len(arrs["csc_colIdx_r"]), len(arrs["csc_rowIdx_r"]),
                                               ctypes.byref(h)))
        return cls(h)

    def __del__(self):
        if getattr(self, "_h", None) and _lib._lib is not None:
            _lib._lib.cola_partition_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def info(self):
        i = _lib.PartitionInfo()
        check(lib().cola_partition_get_info(self._h, ctypes.byref(i)))
        return {f: getattr(i, f) for f, _ in _lib.PartitionInfo._fields_}

    def arrays(self):
        """dict of numpy copies keyed like the reference's file stems."""
        if self._cache is None:
            info = self.info()
            pa = _lib.PartitionArrays()
            check(lib().cola_partition_get_arrays(self._h, ctypes.byref(pa)))
            L, U, nl, nr = info["L"], info["U_r"], info["nnz_l"], info["nnz_r"]
            sizes = dict(csr_rowPtr_l=L + 1, csr_colIdx_l=nl, csr_rowIdx_l=L, csr_val_l=nl,
                         csc_colPtr_r=U + 1, csc_rowIdx_r=nr, csc_colIdx_r=U, csc_val_r=nr)
            self._cache = {k: _np_copy(getattr(pa, k), sizes[k], np.float32 if "val" in k else np.int32)
                           for k in _FIELDS}
        return self._cache

    def __getattr__(self, name):
        if name in _FIELDS:
            return self.arrays()[name]
        raise AttributeError(name)


def serialize(graph, bin_path, graph_name, nGPU_list, write_values=True):
    """Writes global_rowPtr/colIdx and every rank's files for each nGPU."""
    for nGPU in nGPU_list:
        check(lib().cola_serialize(graph.handle, os.fsencode(bin_path), graph_name.encode(), int(nGPU),
                                   1 if write_values else 0))


def serialize_ppl(graph, bin_path, graph_name, nGPU_list, gs_list):
    """The stage files of the pipelined path (prep_torch_ppl.py:53-76
    serialize): for every nGPU and group size gs < nGPU, the sorted unique
    columns of each stage's rows under
    <bin_path>/<graph_name>/ppl<nGPU // gs>/colIdx_<nGPU>_<nGPU // gs>_<s>.bin."""
    for nGPU in nGPU_list:
        for gs in gs_list:
            if gs < nGPU:
                check(lib().cola_serialize_ppl(graph.handle, os.fsencode(bin_path), graph_name.encode(), int(nGPU),
                                               int(gs)))


def stage_cols(bin_path, graph_name, nGPU, gs, stage):
    """One stage file written by serialize_ppl (the array
    distSpMM_ppl_batch.cu:102-104 reads as colIdx_workgroup)."""
    S = nGPU // gs
    return bin2vec(os.path.join(bin_path, graph_name, f"ppl{S}", f"colIdx_{nGPU}_{S}_{stage}.bin"))


def process_graph_ppl(graph_name, mtx_path, bin_path, nGPU_list, gs_list):
    """load_mtx + serialize_ppl (prep_torch_ppl.py:79-84).  Returns the Graph."""
    g = load_mtx(graph_name, mtx_path)
    serialize_ppl(g, bin_path, graph_name, nGPU_list, gs_list)
    return g


def load_mtx(graph_name, mtx_path, values="one"):
    """<mtx_path>/<graph_name>.mtx as a Graph (prep_torch_par.py:14-20).
    values="one" keeps the reference's unweighted view."""
    path = os.path.join(mtx_path, f"{graph_name}.mtx")
    return Graph.from_mtx(path, values=values)


def process_graph(graph_name, mtx_path, bin_path, nGPU_list, write_values=False):
    """load_mtx + serialize for every GPU count (prep_torch_par.py:93-118);
    the partitions are written in the reference's layout under
    <bin_path>/<graph_name>/GPU<n>/.  Returns the Graph."""
    g = load_mtx(graph_name, mtx_path)
    serialize(g, bin_path, graph_name, nGPU_list, write_values=write_values)
    return g


def reorder_mtx(src_path, dst_path, method="cm", with_values=False):
    """Reads a MatrixMarket graph, relabels it for locality and writes the
    relabelled graph (the flow of rabbit_order_write.py:60-77 with this
    library's orders).  Returns the new->old permutation."""
    g = Graph.from_mtx(src_path, values="file")
    h, perm = g.reorder(method)
    h.write_mtx(dst_path, with_values=with_values)
    return perm


def bin2vec(path, dtype=np.int32):
    """Raw little-endian file -> numpy array; the size must be a whole number of elements."""
    size = os.path.getsize(path)
    item = np.dtype(dtype).itemsize
    if size % item:
        raise ValueError(f"{path}: size {size} is not a multiple of {item}")
    return np.fromfile(path, dtype=dtype)


def nnz_split(rowIdx, rowPtr, nnz):
    """Cuts each (zip) row into virtual rows of gran = ceil((nnz/nrows)/32)*32
    entries.  Returns (target, ptr, gran); an empty rowIdx gives empty
    target and ptr (the reference produces nothing in that case)."""
    ri, rp = _i32(rowIdx), _i32(rowPtr)
    if len(ri) and len(rp) != len(ri) + 1:
        raise _lib.ColaError(_lib.COLA_ERR_INVALID, f"nnz_split: len(rowPtr) = {len(rp)}, expected {len(ri) + 1}")
    nv, gran = ctypes.c_int64(), ctypes.c_int32()
    check(lib().cola_nnz_split(_ptr(ri), len(ri), _ptr(rp), int(nnz), None, None, ctypes.byref(nv),
                               ctypes.byref(gran)))
    if len(ri) == 0:
        return np.zeros(0, np.int32), np.zeros(0, np.int32), 0
    target = np.zeros(max(nv.value, 1), np.int32)
    ptr = np.zeros(nv.value + 1, np.int32)
    check(lib().cola_nnz_split(_ptr(ri), len(ri), _ptr(rp), int(nnz), _ptr(target), _ptr(ptr),
                               ctypes.byref(nv), ctypes.byref(gran)))
    return target[:nv.value], ptr, gran.value
