"""Property-based checks of the host side against the oracle's restatements,
on random graphs drawn by hypothesis (host only, a few seconds).

- nnz_split (cuh:29-56): library == oracle restatement, and its invariants;
- row_partition (prep_torch_par.py:24-79): library == numpy restatement for
  random graphs, GPU counts and ranks, and the two halves rebuild the row
  panel exactly;
- MatrixMarket (prep_torch_par.py:14-20): random matrices written by scipy
  read back equal to scipy.io.mmread + tocsr(); the library's writer and
  reader round-trip every value.
"""
import numpy as np
import pytest
import scipy.io as sio
import scipy.sparse as sp
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as st

import cola_amd as cola
from oracle import oracle as orc

SETTINGS = dict(max_examples=60, deadline=None, suppress_health_check=[HealthCheck.function_scoped_fixture])


@st.composite
def graphs(draw, max_n=300, max_deg=12):
    n = draw(st.integers(1, max_n))
    m = draw(st.integers(0, n * max_deg))
    seed = draw(st.integers(0, 2**31 - 1))
    rng = np.random.default_rng(seed)
    skew = draw(st.sampled_from([0.0, 0.5, 0.9]))
    # a mix of uniform and low-id-heavy endpoints (hub rows / columns)
    r = np.where(rng.random(m) < skew, rng.integers(0, max(1, n // 8), m), rng.integers(0, n, m))
    c = rng.integers(0, n, m)
    v = rng.uniform(-1, 1, m).astype(np.float32)
    return n, r.astype(np.int64), c.astype(np.int64), v


@settings(**SETTINGS)
@given(graphs(), st.integers(1, 9), st.data())
def test_row_partition_matches_restatement(g, nGPU, data):
    n, r, c, v = g
    G = cola.Graph.from_coo(n, r, c, v)
    rp, ci, va = G.arrays()
    rank = data.draw(st.integers(0, nGPU - 1))
    mine = G.partition(nGPU, rank).arrays()
    ref = orc.row_partition_np(rp, ci, va, n, nGPU, rank)
    for k in ref:
        assert np.array_equal(mine[k], ref[k]), k
    # the two halves rebuild the dense row panel
    P = -(-n // nGPU)
    lb, ub = rank * P, min(n, (rank + 1) * P)
    if ub <= lb:
        return
    D = np.zeros((ub - lb, n), np.float64)
    rows = np.repeat(mine["csr_rowIdx_l"], np.diff(mine["csr_rowPtr_l"]))
    np.add.at(D, (rows, mine["csr_colIdx_l"] + rank * P), mine["csr_val_l"])
    cols = np.repeat(mine["csc_colIdx_r"], np.diff(mine["csc_colPtr_r"]))
    np.add.at(D, (mine["csc_rowIdx_r"], cols), mine["csc_val_r"])
    E = np.zeros_like(D)
    for i in range(lb, ub):
        E[i - lb, ci[rp[i]:rp[i + 1]]] = va[rp[i]:rp[i + 1]]
    assert np.array_equal(D, E)


@settings(**SETTINGS)
@given(st.lists(st.integers(0, 400), min_size=0, max_size=60), st.data())
def test_nnz_split_matches_oracle(degrees, data):
    deg = np.array([d for d in degrees if d > 0], np.int64)
    rowIdx = np.sort(data.draw(st.lists(st.integers(0, 10**6), min_size=len(deg), max_size=len(deg),
                                        unique=True))).astype(np.int32) if len(deg) else np.zeros(0, np.int32)
    rowPtr = np.concatenate([[0], np.cumsum(deg)]).astype(np.int32)
    nnz = int(rowPtr[-1])
    t1, p1, g1 = cola.nnz_split(rowIdx, rowPtr, nnz)
    t2, p2, g2 = orc.nnz_split(rowIdx, rowPtr, nnz)
    assert np.array_equal(t1, t2) and np.array_equal(p1, p2) and g1 == g2
    if len(deg):
        assert g1 % 32 == 0 and g1 >= 32
        pieces = np.bincount(np.searchsorted(rowIdx, t1), minlength=len(deg))
        assert np.array_equal(pieces, -(-deg // g1))
        assert p1[0] == 0 and p1[-1] == nnz and np.all(np.diff(p1) > 0) and np.all(np.diff(p1) <= g1)


@settings(**dict(SETTINGS, max_examples=40))
@given(graphs(max_n=200), st.sampled_from(["general", "symmetric"]), st.booleans())
def test_mtx_reader_matches_scipy(tmp_path_factory, g, symmetry, pattern):
    n, r, c, v = g
    if symmetry == "symmetric":
        r, c = np.maximum(r, c), np.minimum(r, c)
    A = sp.coo_matrix((np.ones(len(r)) if pattern else v.astype(np.float64), (r, c)), shape=(n, n))
    path = str(tmp_path_factory.mktemp("mtx") / "g.mtx")
    sio.mmwrite(path, A, symmetry=symmetry, field="pattern" if pattern else "real")
    G = cola.Graph.from_mtx(path)
    rp, ci, va = G.arrays()
    S = sp.csr_matrix(sio.mmread(path))
    S.sum_duplicates()
    S.sort_indices()
    assert np.array_equal(rp, S.indptr) and np.array_equal(ci, S.indices)
    assert np.array_equal(va, S.data.astype(np.float32))


@settings(**dict(SETTINGS, max_examples=40))
@given(graphs(max_n=200))
def test_mtx_writer_round_trip(tmp_path_factory, g):
    n, r, c, v = g
    G = cola.Graph.from_coo(n, r, c, v)
    path = str(tmp_path_factory.mktemp("mtx") / "w.mtx")
    G.write_mtx(path)
    H = cola.Graph.from_mtx(path)
    for a, b in zip(G.arrays(), H.arrays()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("method", ["cm", "rcm", "degree"])
@settings(**dict(SETTINGS, max_examples=30))
@given(g=graphs(max_n=150))
def test_reorder_preserves_the_product(method, g):
    """A relabelled graph computes the same product up to the permutation:
    (P A P^T)(P B) = P (A B), checked with the FP64 oracle."""
    n, r, c, v = g
    G = cola.Graph.from_coo(n, r, c, v)
    H, perm = G.reorder(method)
    B = np.random.default_rng(n).uniform(-1, 1, (n, 8)).astype(np.float32)
    rp, ci, va = G.arrays()
    hp, hc, hv = H.arrays()
    AB = orc.spmm_f64(rp, ci, va, 0, n, B)
    HB = orc.spmm_f64(hp, hc, hv, 0, n, B[perm])
    assert np.allclose(HB, AB[perm], rtol=0, atol=1e-9)
