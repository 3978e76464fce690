"""Host-side product code (no GPU): graph generation, the partition layout
against the oracle's restatement of prep_torch_par.py:24-79, the file
contract (write / load / error paths), nnz_split, and plan bookkeeping."""
import hashlib
import os
import subprocess
import sys

import numpy as np
import pytest

import cola_amd as cola
from cola_amd import ColaError
from cola_amd import _lib
from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def check_graph_invariants(g):
    rp, ci, va = g.arrays()
    n = g.n
    assert rp[0] == 0 and np.all(np.diff(rp) >= 0) and rp[-1] == len(ci)
    rows = np.repeat(np.arange(n), np.diff(rp))
    assert not np.any(rows == ci), "self-loop"
    # sorted, unique within each row
    same_row = rows[1:] == rows[:-1]
    assert np.all(ci[1:][same_row] > ci[:-1][same_row])
    # symmetric: (i,j) present iff (j,i) present
    key = rows.astype(np.int64) * n + ci
    tkey = ci.astype(np.int64) * n + rows
    assert np.array_equal(np.sort(key), np.sort(tkey))
    return rp, ci, va


def test_rmat_invariants_and_gcn_values():
    g = cola.Graph.rmat(12, 16, seed=3)
    rp, ci, va = check_graph_invariants(g)
    deg = np.diff(rp).astype(np.float64)
    rows = np.repeat(np.arange(g.n), np.diff(rp))
    expect = (1.0 / np.sqrt(deg[rows] * deg[ci])).astype(np.float32)
    assert np.array_equal(va, expect)
    # R-MAT skew: low ids carry most of the edges (a = 0.57)
    assert rp[g.n // 4] > 0.4 * rp[-1]


def test_rmat_reorder_is_a_relabelling():
    g0 = cola.Graph.rmat(11, 12, seed=5)
    g1 = cola.Graph.rmat(11, 12, seed=5, reorder=True)
    rp0, _, _ = g0.arrays()
    rp1, _, _ = check_graph_invariants(g1)
    assert g0.nnz == g1.nnz
    assert np.array_equal(np.sort(np.diff(rp0)), np.sort(np.diff(rp1)))
    # isolated vertices are placed last
    d1 = np.diff(rp1)
    nz = np.nonzero(d1)[0]
    assert np.all(d1[nz.max() + 1:] == 0)


def _graph_digest(log_n, threads):
    code = ("import sys, hashlib; sys.path.insert(0, %r); import cola_amd as c; "
            "g = c.Graph.rmat(%d, 8, seed=9, reorder=True); rp, ci, va = g.arrays(); "
            "print(hashlib.sha256(rp.tobytes() + ci.tobytes() + va.tobytes()).hexdigest())") % (ROOT, log_n)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    return subprocess.check_output([sys.executable, "-c", code], env=env).decode().strip()


def test_rmat_is_deterministic_across_thread_counts():
    assert _graph_digest(14, 1) == _graph_digest(14, 4)


def test_rmat_seed_changes_graph():
    a = cola.Graph.rmat(10, 8, seed=1).arrays()[1]
    b = cola.Graph.rmat(10, 8, seed=2).arrays()[1]
    assert len(a) != len(b) or not np.array_equal(a, b)


@pytest.mark.parametrize("log_n,nGPU", [(10, 2), (12, 4), (12, 3), (11, 8)])
def test_partition_matches_oracle_restatement(log_n, nGPU):
    g = cola.Graph.rmat(log_n, 10, seed=log_n + nGPU, reorder=(nGPU % 2 == 0))
    rp, ci, va = g.arrays()
    for rank in range(nGPU):
        mine = g.partition(nGPU, rank).arrays()
        ref = orc.row_partition_np(rp, ci, va, g.n, nGPU, rank)
        for k in ref:
            assert mine[k].dtype == ref[k].dtype, k
            assert np.array_equal(mine[k], ref[k]), (k, rank)


def test_partition_info(tmp_path):
    g = cola.Graph.rmat(12, 8, seed=2)
    p = g.partition(4, 1)
    info = p.info()
    a = p.arrays()
    assert info["nodesPerPE"] == 1024 and info["num_rows"] == 1024
    assert info["L"] == len(a["csr_rowIdx_l"]) and info["U_r"] == len(a["csc_colIdx_r"])
    assert info["U_l"] == len(np.unique(a["csr_colIdx_l"]))


def test_serialize_load_round_trip(tmp_path):
    g = cola.Graph.rmat(11, 8, seed=4)
    rp, ci, _ = g.arrays()
    cola.serialize(g, str(tmp_path), "rmat11", [2, 4], write_values=True)
    gdir = tmp_path / "rmat11"
    assert np.array_equal(cola.bin2vec(gdir / "global_rowPtr.bin"), rp.astype(np.int32))
    assert np.array_equal(cola.bin2vec(gdir / "global_colIdx.bin"), ci)
    for nGPU in (2, 4):
        for rank in range(nGPU):
            mem = g.partition(nGPU, rank).arrays()
            d = gdir / f"GPU{nGPU}"
            for stem in ("csr_rowPtr_l", "csr_colIdx_l", "csr_rowIdx_l", "csc_colPtr_r", "csc_rowIdx_r",
                         "csc_colIdx_r"):
                f = d / f"{stem}_{nGPU}_{rank}.bin"
                assert f.read_bytes() == mem[stem].tobytes(), stem
            loaded = cola.Partition.load(str(tmp_path), "rmat11", nGPU, rank).arrays()
            for k in mem:
                assert np.array_equal(loaded[k], mem[k]), k


def test_load_without_value_files_uses_ones(tmp_path):
    g = cola.Graph.rmat(10, 8, seed=4)
    cola.serialize(g, str(tmp_path), "g", [2], write_values=False)
    p = cola.Partition.load(str(tmp_path), "g", 2, 0).arrays()
    assert np.all(p["csr_val_l"] == 1.0) and np.all(p["csc_val_r"] == 1.0)
    assert not (tmp_path / "g" / "GPU2" / "csr_val_l_2_0.bin").exists()


def test_load_error_paths(tmp_path):
    g = cola.Graph.rmat(10, 8, seed=4)
    cola.serialize(g, str(tmp_path), "g", [2], write_values=False)
    d = tmp_path / "g" / "GPU2"
    # missing file
    with pytest.raises(ColaError) as e:
        cola.Partition.load(str(tmp_path), "nope", 2, 0)
    assert e.value.code == _lib.COLA_ERR_IO
    # truncated file (not a whole number of int32)
    f = d / "csr_colIdx_l_2_0.bin"
    good = f.read_bytes()
    f.write_bytes(good[:-2])
    with pytest.raises(ColaError) as e:
        cola.Partition.load(str(tmp_path), "g", 2, 0)
    assert e.value.code == _lib.COLA_ERR_IO
    f.write_bytes(good)
    # non-monotone row pointer
    rpf = d / "csr_rowPtr_l_2_0.bin"
    rpb = np.fromfile(rpf, np.int32)
    bad = rpb.copy()
    bad[1], bad[2] = bad[2], bad[1]
    bad.tofile(rpf)
    with pytest.raises(ColaError) as e:
        cola.Partition.load(str(tmp_path), "g", 2, 0)
    assert e.value.code == _lib.COLA_ERR_INVALID
    rpb.tofile(rpf)
    # column out of range
    cf = d / "csr_colIdx_l_2_0.bin"
    cb = np.fromfile(cf, np.int32)
    bad = cb.copy()
    bad[0] = 10 ** 6
    bad.tofile(cf)
    with pytest.raises(ColaError):
        cola.Partition.load(str(tmp_path), "g", 2, 0)
    cb.tofile(cf)
    # a wrong-size value file
    (d / "csr_val_l_2_0.bin").write_bytes(b"\0" * 8)
    with pytest.raises(ColaError):
        cola.Partition.load(str(tmp_path), "g", 2, 0)
    os.remove(d / "csr_val_l_2_0.bin")
    cola.Partition.load(str(tmp_path), "g", 2, 0)
    # any index file written as int64 (numpy's default integer) instead of int32
    for stem in ("csr_rowPtr_l", "csr_colIdx_l", "csr_rowIdx_l", "csc_colPtr_r", "csc_rowIdx_r", "csc_colIdx_r"):
        f = d / f"{stem}_2_0.bin"
        good = np.fromfile(f, np.int32)
        good.astype(np.int64).tofile(f)
        with pytest.raises(ColaError):
            cola.Partition.load(str(tmp_path), "g", 2, 0)
        good.tofile(f)
    cola.Partition.load(str(tmp_path), "g", 2, 0)
    # rank out of range
    with pytest.raises(ColaError):
        cola.Partition.load(str(tmp_path), "g", 2, 2)


def test_partition_from_arrays_validates():
    g = cola.Graph.rmat(10, 8, seed=1)
    a = g.partition(2, 1).arrays()
    p = cola.Partition.from_arrays(g.n, 2, 1, a["csr_rowPtr_l"], a["csr_colIdx_l"], a["csr_rowIdx_l"],
                                   a["csc_colPtr_r"], a["csc_rowIdx_r"], a["csc_colIdx_r"], a["csr_val_l"],
                                   a["csc_val_r"])
    for k, v in p.arrays().items():
        assert np.array_equal(v, a[k])
    bad_cols = a["csc_colIdx_r"].copy()
    bad_cols[0] = 600  # inside the owned range [512, 1024)
    with pytest.raises(ColaError):
        cola.Partition.from_arrays(g.n, 2, 1, a["csr_rowPtr_l"], a["csr_colIdx_l"], a["csr_rowIdx_l"],
                                   a["csc_colPtr_r"], a["csc_rowIdx_r"], bad_cols)


@pytest.mark.parametrize("seed", [0, 1])
def test_nnz_split_matches_oracle(seed):
    g = cola.Graph.rmat(12, 12, seed=seed)
    a = g.partition(4, 0).arrays()
    for ri, rp, nnz in ((a["csr_rowIdx_l"], a["csr_rowPtr_l"], len(a["csr_colIdx_l"])),
                        (a["csc_colIdx_r"], a["csc_colPtr_r"], len(a["csc_rowIdx_r"]))):
        mine = cola.nnz_split(ri, rp, nnz)
        ref = orc.nnz_split(ri, rp, nnz)
        assert np.array_equal(mine[0], ref[0]) and np.array_equal(mine[1], ref[1]) and mine[2] == ref[2]
    t, p, g0 = cola.nnz_split(np.zeros(0, np.int32), np.zeros(1, np.int32), 0)
    assert len(t) == 0 and len(p) == 0 and g0 == 0


def test_graph_from_csr_sums_duplicates():
    g = cola.Graph.from_csr([0, 3, 4], [1, 1, 0, 0], [1.0, 2.0, 5.0, 4.0])
    rp, ci, va = g.arrays()
    assert list(rp) == [0, 2, 3] and list(ci) == [0, 1, 0] and list(va) == [5.0, 3.0, 4.0]
    with pytest.raises(ColaError):
        cola.Graph.from_csr([0, 1], [5])


def test_bin2vec_rejects_partial_elements(tmp_path):
    f = tmp_path / "x.bin"
    f.write_bytes(b"\1\0\0\0\2\0")
    with pytest.raises(ValueError):
        cola.bin2vec(f)


def test_graph_digest_is_stable():
    """The C1 graph (R-MAT 2^16, degree 16, seed 1) keeps its shape."""
    g = cola.Graph.rmat(16, 16, seed=1)
    n, nnz, md = g.info()
    assert n == 65536 and 0.85e6 < nnz < 1.05e6
    rp, ci, va = g.arrays()
    h = hashlib.sha256(ci[:1000].tobytes()).hexdigest()
    assert h == hashlib.sha256(cola.Graph.rmat(16, 16, seed=1).arrays()[1][:1000].tobytes()).hexdigest()


def test_partition_rejects_local_columns_past_the_last_node():
    """On the last rank (num_rows < nodesPerPE) a local column c is node
    rank*P + c, which must exist: columns in [M, P) are rejected."""
    g = cola.Graph.rmat(10, 8, seed=1)  # n = 1024, 3 ranks: P = 342, last M = 340
    a = g.partition(3, 2).arrays()
    info = g.partition(3, 2).info()
    assert info["num_rows"] < info["nodesPerPE"]
    bad = a["csr_colIdx_l"].copy()
    bad[0] = info["num_rows"]  # node 1024 does not exist
    with pytest.raises(ColaError):
        cola.Partition.from_arrays(g.n, 3, 2, a["csr_rowPtr_l"], bad, a["csr_rowIdx_l"],
                                   a["csc_colPtr_r"], a["csc_rowIdx_r"], a["csc_colIdx_r"])
