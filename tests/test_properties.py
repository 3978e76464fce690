"""Property-based checks of the host side against the oracle's restatements,
on random graphs drawn by hypothesis (host only, a few seconds).

- nnz_split (cuh:29-56): library == oracle restatement, and its invariants;
- row_partition (prep_torch_par.py:24-79): library == numpy restatement for
  random graphs, GPU counts and ranks, and the two halves rebuild the row
  panel exactly;
- MatrixMarket (prep_torch_par.py:14-20): random matrices written by scipy
  read back equal to scipy.io.mmread + tocsr(); the library's writer and
  reader round-trip every value;
- reorders keep the product up to the permutation;
- malformed inputs to the loader, nnz_split and the zip constructors are
  rejected with a status, and the plan builder runs on random partitions
  and options without faulting.
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


@pytest.mark.parametrize("method", ["cm", "rcm", "degree", "community"])
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


small_ints = st.lists(st.integers(-3, 12), min_size=0, max_size=7)


@settings(**dict(SETTINGS, max_examples=300))
@given(st.integers(-2, 24), st.integers(-1, 4), st.integers(-1, 4), small_ints, small_ints, small_ints, small_ints,
       small_ints, small_ints)
def test_partition_from_arrays_rejects_without_crashing(n, nGPU, rank, rp, cl, ri, cp, rr, cr):
    """Arbitrary (mostly malformed) partition arrays: either a valid partition
    or a ColaError, never a crash (the loader checks what the writer asserts
    plus every index range)."""
    try:
        part = cola.Partition.from_arrays(n, nGPU, rank, rp, cl, ri, cp, rr, cr)
    except cola.ColaError as e:
        assert e.code == cola._lib.COLA_ERR_INVALID
        return
    a = part.arrays()
    assert np.array_equal(a["csr_rowPtr_l"], np.asarray(rp, np.int32))
    assert np.array_equal(a["csc_colIdx_r"], np.asarray(cr, np.int32))


@settings(**dict(SETTINGS, max_examples=80))
@given(graphs(max_n=400, max_deg=30), st.integers(1, 6), st.data())
def test_host_plan_builder_on_random_partitions(g, nGPU, data):
    """The plan builder (degree bins, hub chunks, XCD ranges, panels) runs on
    random partitions and option sets; without a GPU the upload then fails
    with a HIP status, with one it succeeds — never a crash."""
    n, r, c, v = g
    G = cola.Graph.from_coo(n, r, c, v)
    rank = data.draw(st.integers(0, nGPU - 1))
    part = G.partition(nGPU, rank)
    opts = dict(hub_threshold=data.draw(st.sampled_from([0, 1, 4, 40])),
                chunk_nnz=data.draw(st.sampled_from([0, 1, 3, 16])),
                kahan_degree=data.draw(st.sampled_from([0, 2, 50])),
                short_degree=data.draw(st.sampled_from([0, 1, 5])),
                col_panels=data.draw(st.sampled_from([0, 2, 5])),
                row_order=data.draw(st.sampled_from([0, 1])))
    try:
        plan = cola.Plan(part, **opts)
    except cola.ColaError as e:
        assert e.code == cola._lib.COLA_ERR_HIP, str(e)
        return
    assert plan.info["nnz"] == part.info()["nnz_l"] + part.info()["nnz_r"]


@settings(**dict(SETTINGS, max_examples=200))
@given(small_ints, small_ints, st.integers(-1, 40))
def test_nnz_split_rejects_without_crashing(ri, rp, nnz):
    """Malformed (rowIdx, rowPtr, nnz): a ColaError or the oracle's answer."""
    try:
        t, p, gran = cola.nnz_split(np.asarray(ri, np.int32), np.asarray(rp, np.int32), nnz)
    except cola.ColaError as e:
        assert e.code == cola._lib.COLA_ERR_INVALID
        return
    t2, p2, g2 = orc.nnz_split(np.asarray(ri, np.int32), np.asarray(rp, np.int32), nnz)
    assert np.array_equal(t, t2) and np.array_equal(p, p2) and gran == g2


@settings(**dict(SETTINGS, max_examples=200))
@given(small_ints, small_ints, small_ints, st.booleans())
def test_zip_create_rejects_without_crashing(ptr, idx, zid, with_val):
    """cola_zip_create on arbitrary arrays: INVALID for malformed ones; well
    formed ones reach the device upload (a HIP status without a GPU)."""
    try:
        cola.d_csr_zip(ptr, idx, zid, np.ones(len(idx), np.float32) if with_val else None)
    except cola.ColaError as e:
        assert e.code in (cola._lib.COLA_ERR_INVALID, cola._lib.COLA_ERR_HIP), str(e)


tokens = st.sampled_from(["1", "2", "3", "0", "-1", "7", "1.5", "-2e3", "nan", "inf", "x", "+4", "1e400", "", " ",
                          "%", "%%", "99999999999", "2147483648", "\t"])


@settings(**dict(SETTINGS, max_examples=300))
@given(st.sampled_from(["real general", "pattern general", "integer symmetric", "real skew-symmetric",
                        "real symmetric", "complex general", "real hermitian", "bogus general"]),
       st.lists(tokens, min_size=0, max_size=4), st.lists(st.lists(tokens, min_size=0, max_size=4), max_size=12),
       st.booleans())
def test_mtx_reader_never_crashes(tmp_path_factory, header, size_line, lines, crlf):
    """Random and corrupted MatrixMarket text: a graph or a ColaError."""
    nl = "\r\n" if crlf else "\n"
    text = f"%%MatrixMarket matrix coordinate {header}{nl}" + " ".join(size_line) + nl
    text += nl.join(" ".join(t) for t in lines)
    path = tmp_path_factory.mktemp("fz") / "f.mtx"
    path.write_text(text)
    try:
        g = cola.Graph.from_mtx(str(path))
    except cola.ColaError as e:
        assert e.code in (cola._lib.COLA_ERR_INVALID, cola._lib.COLA_ERR_UNSUPPORTED, cola._lib.COLA_ERR_IO)
        return
    rp, ci, va = g.arrays()
    assert rp[0] == 0 and np.all(np.diff(rp) >= 0) and np.all((ci >= 0) & (ci < g.n))
