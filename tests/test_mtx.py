"""MatrixMarket ingestion (prep_torch_par.py:14-20 load_mtx) and graph
reorders (rabbit_order_write.py:46-77), host only.

The reader is pinned against scipy.io.mmread, the routine the reference
calls, followed by the reference's tocsr() conversion (duplicates summed,
columns sorted); the partition files written from an .mtx graph are pinned
byte for byte against a step-by-step scipy restatement of prep_torch_par.py
(oracle.par_files_from_mtx)."""
import numpy as np
import pytest
import scipy.io as sio
import scipy.sparse as sp

import cola_amd as cola
from cola_amd import _lib
from cola_amd._lib import ColaError
from oracle import oracle as orc


def write(path, text):
    path.write_bytes(text.encode())
    return str(path)


def scipy_csr(path):
    A = sp.csr_matrix(sio.mmread(path))
    A.sum_duplicates()
    A.sort_indices()
    return A


def assert_same_as_scipy(path, **kw):
    g = cola.Graph.from_mtx(path, **kw)
    rp, ci, va = g.arrays()
    A = scipy_csr(path)
    assert g.n == A.shape[0]
    assert np.array_equal(rp, A.indptr.astype(np.int64))
    assert np.array_equal(ci, A.indices.astype(np.int32))
    assert np.array_equal(va, A.data.astype(np.float32))
    return g


CASES = {
    "general_real_dups": """%%MatrixMarket matrix coordinate real general
% a comment
4 4 7
1 2 0.5
2 1 -1.25
1 2 0.25
4 4 3e-1
3 1 1.0E+2
2 3 -7
4 1 0.125
""",
    "symmetric_with_diagonal": """%%MatrixMarket matrix coordinate real symmetric
5 5 6
1 1 2.0
2 1 0.5
3 2 -0.75
5 1 1.5
5 5 4.0
4 3 0.25
""",
    "skew_symmetric": """%%MatrixMarket matrix coordinate real skew-symmetric
3 3 2
2 1 1.5
3 2 -2.5
""",
    "pattern_symmetric": """%%MatrixMarket matrix coordinate pattern symmetric
6 6 5
2 1
3 1
6 2
5 4
6 6
""",
    "integer_general": """%%MatrixMarket matrix coordinate integer general
3 3 4
1 3 7
3 1 -2
2 2 5
1 3 1
""",
    "comments_blank_lines_crlf": "%%MatrixMarket matrix coordinate real general\r\n%c1\r\n\r\n%c2\r\n"
                                 "3 3 3\r\n1 1 1.5\r\n\r\n2 3 2.5\r\n3 2   -0.5\r\n",
    "upper_case_banner_tabs": "%%MatrixMarket MATRIX Coordinate REAL General\n2 2 2\n1\t2\t0.5\n2\t1\t0.25",
    "isolated_nodes": """%%MatrixMarket matrix coordinate real general
8 8 2
2 7 1.0
7 2 1.0
""",
    "empty": """%%MatrixMarket matrix coordinate pattern general
4 4 0
""",
    "zero_nodes": "%%MatrixMarket matrix coordinate real general\n0 0 0\n",
    "no_final_newline": "%%MatrixMarket matrix coordinate real general\n2 2 1\n2 1 3.5",
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_small_files_match_scipy_mmread(tmp_path, name):
    assert_same_as_scipy(write(tmp_path / f"{name}.mtx", CASES[name]))


def test_symmetric_expansion_and_duplicate_sum_by_hand(tmp_path):
    g = cola.Graph.from_mtx(write(tmp_path / "s.mtx", CASES["symmetric_with_diagonal"]))
    rp, ci, va = g.arrays()
    D = np.zeros((5, 5), np.float32)
    for i in range(5):
        D[i, ci[rp[i]:rp[i + 1]]] = va[rp[i]:rp[i + 1]]
    E = np.zeros((5, 5), np.float32)
    E[0, 0], E[4, 4] = 2.0, 4.0
    for i, j, v in [(1, 0, 0.5), (2, 1, -0.75), (4, 0, 1.5), (3, 2, 0.25)]:
        E[i, j] = E[j, i] = v
    assert np.array_equal(D, E)
    g = cola.Graph.from_mtx(write(tmp_path / "d.mtx", CASES["general_real_dups"]))
    rp, ci, va = g.arrays()
    assert list(ci[rp[0]:rp[1]]) == [1] and va[rp[0]] == 0.75  # 0.5 + 0.25


@pytest.mark.parametrize("symmetry", ["general", "symmetric"])
def test_large_random_file_matches_scipy(tmp_path, symmetry):
    """A file large enough to be cut into one piece per host thread."""
    rng = np.random.default_rng(7)
    n, m = 30000, 400000
    r = rng.integers(0, n, m)
    c = rng.integers(0, n, m)
    v = rng.standard_normal(m)
    if symmetry == "symmetric":
        lo, hi = np.minimum(r, c), np.maximum(r, c)
        r, c = hi, lo
    A = sp.coo_matrix((v, (r, c)), shape=(n, n))
    path = str(tmp_path / "big.mtx")
    sio.mmwrite(path, A, symmetry=symmetry)
    assert_same_as_scipy(path)


def test_write_round_trip(tmp_path):
    g = cola.Graph.rmat(10, 8, seed=3)
    rp, ci, va = g.arrays()
    for with_values in (True, False):
        path = str(tmp_path / f"w{int(with_values)}.mtx")
        g.write_mtx(path, with_values=with_values)
        A = scipy_csr(path)
        assert np.array_equal(A.indptr, rp) and np.array_equal(A.indices, ci)
        h = cola.Graph.from_mtx(path)
        rp2, ci2, va2 = h.arrays()
        assert np.array_equal(rp2, rp) and np.array_equal(ci2, ci)
        if with_values:
            assert np.array_equal(A.data.astype(np.float32), va)
            assert np.array_equal(va2, va)  # 9 significant digits round-trip FP32
        else:
            assert np.all(va2 == 1.0)


def test_value_policies(tmp_path):
    path = write(tmp_path / "g.mtx", CASES["general_real_dups"])
    rp, ci, va = cola.Graph.from_mtx(path, values="one").arrays()
    assert np.all(va == 1.0)
    rp, ci, va = cola.Graph.from_mtx(path, values="gcn").arrays()
    deg = np.diff(rp).astype(np.float64)
    rows = np.repeat(np.arange(len(deg)), np.diff(rp))
    assert np.allclose(va, 1.0 / np.sqrt(deg[rows] * deg[ci]), rtol=1e-7, atol=0)


BAD = {
    "no_banner": ("3 3 1\n1 1 1\n", _lib.COLA_ERR_INVALID),
    "bad_object": ("%%MatrixMarket vector coordinate real general\n3 3 1\n1 1 1\n", _lib.COLA_ERR_INVALID),
    "array_format": ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", _lib.COLA_ERR_UNSUPPORTED),
    "complex": ("%%MatrixMarket matrix coordinate complex general\n2 2 1\n1 1 1 0\n", _lib.COLA_ERR_UNSUPPORTED),
    "hermitian": ("%%MatrixMarket matrix coordinate real hermitian\n2 2 1\n1 1 1\n", _lib.COLA_ERR_UNSUPPORTED),
    "unknown_field": ("%%MatrixMarket matrix coordinate quaternion general\n2 2 1\n1 1 1\n", _lib.COLA_ERR_INVALID),
    "non_square": ("%%MatrixMarket matrix coordinate real general\n3 4 1\n1 4 1\n", _lib.COLA_ERR_UNSUPPORTED),
    "index_zero": ("%%MatrixMarket matrix coordinate real general\n3 3 1\n0 1 1\n", _lib.COLA_ERR_INVALID),
    "entry_in_empty_matrix": ("%%MatrixMarket matrix coordinate real general\n0 0 1\n1 1 1\n", _lib.COLA_ERR_INVALID),
    "empty_file": ("", _lib.COLA_ERR_INVALID),
    "index_past_end": ("%%MatrixMarket matrix coordinate real general\n3 3 1\n1 4 1\n", _lib.COLA_ERR_INVALID),
    "too_few": ("%%MatrixMarket matrix coordinate real general\n3 3 3\n1 1 1\n2 2 1\n", _lib.COLA_ERR_INVALID),
    "too_many": ("%%MatrixMarket matrix coordinate real general\n3 3 1\n1 1 1\n2 2 1\n", _lib.COLA_ERR_INVALID),
    "bad_value": ("%%MatrixMarket matrix coordinate real general\n3 3 1\n1 1 x\n", _lib.COLA_ERR_INVALID),
    "missing_value": ("%%MatrixMarket matrix coordinate real general\n3 3 1\n1 1\n", _lib.COLA_ERR_INVALID),
    "extra_token": ("%%MatrixMarket matrix coordinate pattern general\n3 3 1\n1 1 5\n", _lib.COLA_ERR_INVALID),
    "bad_size_line": ("%%MatrixMarket matrix coordinate real general\n3 3\n1 1 1\n", _lib.COLA_ERR_INVALID),
    "no_size_line": ("%%MatrixMarket matrix coordinate real general\n% only a comment\n", _lib.COLA_ERR_IO),
}


@pytest.mark.parametrize("name", sorted(BAD))
def test_malformed_files_are_rejected(tmp_path, name):
    text, code = BAD[name]
    with pytest.raises(ColaError) as e:
        cola.Graph.from_mtx(write(tmp_path / "bad.mtx", text))
    assert e.value.code == code, str(e.value)


def test_missing_file_and_bad_policy(tmp_path):
    with pytest.raises(ColaError) as e:
        cola.Graph.from_mtx(str(tmp_path / "absent.mtx"))
    assert e.value.code == _lib.COLA_ERR_IO
    with pytest.raises(ColaError) as e:
        cola.Graph.from_mtx(str(tmp_path))  # a directory
    assert e.value.code == _lib.COLA_ERR_IO
    with pytest.raises(ColaError) as e:
        cola.Graph(None).write_mtx(str(tmp_path / "x.mtx"))
    assert e.value.code == _lib.COLA_ERR_INVALID


def test_partition_files_match_prep_torch_par(tmp_path):
    """mtx -> Graph.from_mtx -> serialize gives the same bytes as the reference
    writer's scipy steps on mmread's COO (duplicates and a symmetric file)."""
    rng = np.random.default_rng(11)
    n, m = 3001, 40000
    r = rng.integers(0, n, m)
    c = rng.integers(0, n, m)
    r[: n // 3] = 0  # a hub row
    A = sp.coo_matrix((np.ones(m), (r, c)), shape=(n, n))
    sym = sp.coo_matrix((np.ones(m), (np.maximum(r, c), np.minimum(r, c))), shape=(n, n))
    for name, mat, symmetry in [("dups", A, "general"), ("sym", sym, "symmetric")]:
        path = str(tmp_path / f"{name}.mtx")
        sio.mmwrite(path, mat, symmetry=symmetry)
        g = cola.Graph.from_mtx(path, values="one")
        cola.serialize(g, str(tmp_path / "SP"), name, [2, 4, 8], write_values=False)
        gdir = tmp_path / "SP" / name
        for nGPU in (2, 4, 8):
            ref = orc.par_files_from_mtx(path, nGPU)
            assert (gdir / "global_rowPtr.bin").read_bytes() == ref["global_rowPtr"].tobytes()
            assert (gdir / "global_colIdx.bin").read_bytes() == ref["global_colIdx"].tobytes()
            for (stem, rank), arr in ((k, v) for k, v in ref.items() if isinstance(k, tuple)):
                got = (gdir / f"GPU{nGPU}" / f"{stem}_{nGPU}_{rank}.bin").read_bytes()
                assert got == arr.tobytes(), (name, nGPU, rank, stem)


def test_par_restatement_agrees_with_row_partition_np(tmp_path):
    """The two restatements of prep_torch_par.py (scipy steps on the COO,
    numpy sorting on the deduplicated CSR) agree, and both satisfy the writer's
    asserted invariants."""
    rng = np.random.default_rng(5)
    n, m = 777, 9000
    A = sp.coo_matrix((np.ones(m), (rng.integers(0, n, m), rng.integers(0, n, m))), shape=(n, n))
    path = str(tmp_path / "a.mtx")
    sio.mmwrite(path, A)
    C = scipy_csr(path)
    for nGPU in (3, 4):
        ref = orc.par_files_from_mtx(path, nGPU)
        for rank in range(nGPU):
            np_part = orc.row_partition_np(C.indptr, C.indices, C.data, n, nGPU, rank)
            for stem in ("csr_rowPtr_l", "csr_colIdx_l", "csr_rowIdx_l", "csc_colPtr_r", "csc_rowIdx_r",
                         "csc_colIdx_r"):
                assert np.array_equal(ref[(stem, rank)], np_part[stem]), (nGPU, rank, stem)
            assert np.all(np.diff(ref[("csr_rowPtr_l", rank)]) > 0)   # par.py:53
            assert np.all(np.diff(ref[("csc_colPtr_r", rank)]) > 0)   # par.py:74


def band_graph(n, width, seed):
    """A banded symmetric graph under a random relabelling."""
    rng = np.random.default_rng(seed)
    r, c = [], []
    for d in range(1, width + 1):
        i = np.arange(n - d)
        r += [i, i + d]
        c += [i + d, i]
    r, c = np.concatenate(r), np.concatenate(c)
    perm = rng.permutation(n)
    return cola.Graph.from_coo(n, perm[r], perm[c], rng.uniform(0.5, 1.5, len(r)).astype(np.float32))


def bandwidth(g):
    rp, ci, _ = g.arrays()
    rows = np.repeat(np.arange(g.n), np.diff(rp))
    return int(np.max(np.abs(rows - ci))) if len(ci) else 0


@pytest.mark.parametrize("method", ["cm", "rcm", "degree", "community"])
def test_reorder_is_a_relabelling(method):
    g = cola.Graph.rmat(10, 6, seed=9)
    h, perm = g.reorder(method)
    n = g.n
    assert np.array_equal(np.sort(perm), np.arange(n))
    rp, ci, va = g.arrays()
    A = sp.csr_matrix((va, ci, rp), shape=(n, n))
    P = A[perm][:, perm]  # new i = old perm[i]
    P.sort_indices()
    rp2, ci2, va2 = h.arrays()
    assert np.array_equal(rp2, P.indptr) and np.array_equal(ci2, P.indices)
    assert np.array_equal(va2, P.data.astype(np.float32))
    if method == "degree":
        assert np.all(np.diff(np.diff(rp2)) <= 0)


def test_cuthill_mckee_recovers_a_band():
    g = band_graph(3000, 4, seed=2)
    assert bandwidth(g) > 1000
    h, perm = g.reorder("cm")
    assert bandwidth(h) <= 2 * 4 + 1  # CM levels of a width-4 band
    r, perm_r = g.reorder("rcm")
    assert np.array_equal(perm_r, perm[::-1])
    assert bandwidth(r) == bandwidth(h)


def cliques(sizes, seed):
    """Disjoint cliques of the given sizes under shuffled vertex ids."""
    n = sum(sizes)
    perm = np.random.default_rng(seed).permutation(n)
    r, c, groups, o = [], [], [], 0
    for s in sizes:
        ids = perm[o:o + s]
        o += s
        groups.append(ids)
        for i in ids:
            for j in ids:
                if i != j:
                    r.append(i)
                    c.append(j)
    return cola.Graph.from_coo(n, np.array(r), np.array(c), np.ones(len(r), np.float32)), groups


@pytest.mark.parametrize("sizes", [[5, 9, 3, 12, 7, 4], [2] * 30, [40, 3, 3, 3, 20], [6] * 10 + [1] * 5])
def test_community_order_keeps_cliques_contiguous(sizes):
    """Disjoint cliques: a vertex first joins a neighbouring singleton, since
    dQ = 2 (1/2m - d^2/(2m)^2) > 0 whenever d^2 < 2m (true for every clique
    here), and each later one joins its clique's community with the same
    sign of gain; no edge joins two cliques, so every clique is one merge
    tree and gets a contiguous id range (isolated vertices go last)."""
    g, groups = cliques(sizes, seed=len(sizes))
    h, perm = g.reorder("community")
    old2new = np.empty(g.n, np.int64)
    old2new[perm] = np.arange(g.n)
    for ids in groups:
        new = np.sort(old2new[ids])
        assert new[-1] - new[0] == len(ids) - 1, (len(ids), new)
    iso = [ids[0] for ids in groups if len(ids) == 1]
    assert sorted(old2new[iso].tolist()) == list(range(g.n - len(iso), g.n))


def test_community_order_is_deterministic_and_balances_rmat():
    """Same graph, same order; on an R-MAT graph the community order spreads
    the entries over the row partitions far more evenly than Cuthill-McKee,
    which puts the hubs' neighbourhood first."""
    g = cola.Graph.rmat(14, 16, seed=3)
    _, p1 = g.reorder("community")
    _, p2 = g.reorder("community")
    assert np.array_equal(p1, p2)
    def max_share(h):
        rp = h.arrays()[0]
        n = h.n
        cuts = [rp[min(n, k * n // 8)] for k in range(9)]
        return max(np.diff(cuts)) / rp[-1]
    assert max_share(g.reorder("community")[0]) < 0.5 * max_share(g.reorder("cm")[0])


def test_reorder_rejects_unknown_method():
    g = cola.Graph.rmat(6, 4, seed=1)
    with pytest.raises(ColaError):
        ph = _lib.ctypes.c_void_p()
        _lib.check(_lib.lib().cola_graph_reorder(g.handle, 9, None, _lib.ctypes.byref(ph)))


def test_process_graph_and_reorder_mtx_flow(tmp_path):
    """The reference's offline flow: <g>.mtx -> (reorder -> <g>_rabbit.mtx) ->
    per-GPU partition files, then a partition loaded back from them."""
    mtx = tmp_path / "mtx"
    (mtx / "rabbit").mkdir(parents=True)
    band = band_graph(2000, 3, seed=4)
    band.write_mtx(str(mtx / "band.mtx"), with_values=False)
    perm = cola.reorder_mtx(str(mtx / "band.mtx"), str(mtx / "rabbit" / "band_rabbit.mtx"), method="cm")
    assert np.array_equal(np.sort(perm), np.arange(2000))
    sp_dir = tmp_path / "SP"
    g = cola.process_graph("band_rabbit", str(mtx / "rabbit"), str(sp_dir), [2, 4])
    assert bandwidth(g) <= 7
    for nGPU in (2, 4):
        ref = orc.par_files_from_mtx(str(mtx / "rabbit" / "band_rabbit.mtx"), nGPU)
        for rank in range(nGPU):
            p = cola.Partition.load(str(sp_dir), "band_rabbit", nGPU, rank).arrays()
            for stem in ("csr_rowPtr_l", "csr_colIdx_l", "csr_rowIdx_l", "csc_colPtr_r", "csc_rowIdx_r",
                         "csc_colIdx_r"):
                assert np.array_equal(p[stem], ref[(stem, rank)]), (nGPU, rank, stem)
            # a locality order leaves few remote columns per partition
            assert len(p["csc_colIdx_r"]) <= 2 * 3 * 2


@pytest.mark.parametrize("method", ["cm", "rcm", "degree", "community"])
def test_reorder_directed_graph_with_sinks(method):
    """Vertices without out-edges that the breadth-first walk reaches appear
    once in the order (a directed graph, unlike the symmetric R-MAT ones)."""
    n = 6
    # 0 -> 1 -> 2 (2 is a sink), 3 -> 2, 4 isolated, 5 -> 0
    g = cola.Graph.from_coo(n, [0, 1, 3, 5], [1, 2, 2, 0], [1.0, 2.0, 3.0, 4.0])
    h, perm = g.reorder(method)
    assert sorted(perm.tolist()) == list(range(n))
    rp, ci, va = g.arrays()
    A = sp.csr_matrix((va, ci, rp), shape=(n, n))
    P = A[perm][:, perm]
    P.sort_indices()
    rp2, ci2, va2 = h.arrays()
    assert np.array_equal(rp2, P.indptr) and np.array_equal(ci2, P.indices) and np.array_equal(va2, P.data)


# ---------------------------------------------------------------------------
# Stage files of the pipelined path (prep_torch_ppl.py:24-84).
# ---------------------------------------------------------------------------
PPL_MTX = """%%MatrixMarket matrix coordinate real symmetric
% a hub row, duplicates, an explicit zero, isolated vertices at the end
13 13 16
1 1 1.0
2 1 0.5
5 2 1.0
5 2 2.0
7 3 0.0
4 4 1.0
9 1 1.0
9 2 1.0
9 3 1.0
9 4 1.0
9 5 1.0
9 6 1.0
9 7 1.0
9 8 1.0
10 6 1.0
11 10 1.0
"""


def structure(path):
    A = sio.mmread(path)
    S = np.zeros(A.shape, bool)
    S[np.asarray(A.row), np.asarray(A.col)] = True
    return S


@pytest.mark.parametrize("nGPU,gs", [(2, 1), (4, 1), (4, 2), (8, 2), (8, 3), (16, 1), (16, 8)])
def test_ppl_oracle_against_brute_force(tmp_path, nGPU, gs):
    """Each stage's columns are exactly the columns with an entry (explicit
    zeros included, as in the COO the reference reads) in the stage's rows."""
    path = write(tmp_path / "p.mtx", PPL_MTX)
    S = structure(path)
    n = S.shape[0]
    P = (n + nGPU - 1) // nGPU
    ref = orc.ppl_files_from_mtx(path, nGPU, gs)
    assert sorted(ref) == list(range(nGPU // gs))
    for s, cols in ref.items():
        lb, ub = min(n, s * gs * P), min(n, (s + 1) * gs * P)
        assert np.array_equal(cols, np.flatnonzero(S[lb:ub].any(axis=0))), s


@pytest.mark.parametrize("nGPU,gs", [(2, 1), (4, 1), (4, 2), (8, 2), (8, 3), (16, 1), (16, 8)])
def test_stage_files_match_the_reference_writer(tmp_path, nGPU, gs):
    """serialize_ppl on the library's reading of an .mtx file writes, byte for
    byte, what prep_torch_ppl.py writes (restated in oracle.ppl_files_from_mtx),
    including empty stages past the last row."""
    mtx = tmp_path / "mtx"
    mtx.mkdir()
    path = write(mtx / "p.mtx", PPL_MTX)
    sp_dir = tmp_path / "SP"
    cola.process_graph_ppl("p", str(mtx), str(sp_dir), [nGPU], [gs])
    S = nGPU // gs
    for s, cols in orc.ppl_files_from_mtx(path, nGPU, gs).items():
        f = sp_dir / "p" / f"ppl{S}" / f"colIdx_{nGPU}_{S}_{s}.bin"
        assert f.read_bytes() == cols.astype(np.int32).tobytes(), s
        assert np.array_equal(cola.stage_cols(str(sp_dir), "p", nGPU, gs, s), cols)


def test_stage_files_follow_the_reference_loop(tmp_path):
    """serialize_ppl writes only group sizes below nGPU (prep_torch_ppl.py:64);
    the C entry point rejects the others with a status."""
    g = cola.Graph.rmat(8, 4, seed=3)
    cola.serialize_ppl(g, str(tmp_path), "g", [2, 4], [1, 4, 8])
    names = sorted(p.relative_to(tmp_path).as_posix() for p in tmp_path.rglob("*.bin"))
    assert names == ["g/ppl2/colIdx_2_2_0.bin", "g/ppl2/colIdx_2_2_1.bin", "g/ppl4/colIdx_4_4_0.bin",
                     "g/ppl4/colIdx_4_4_1.bin", "g/ppl4/colIdx_4_4_2.bin", "g/ppl4/colIdx_4_4_3.bin"]
    for nGPU, gs in ((4, 4), (4, 0), (1, 1)):
        with pytest.raises(ColaError):
            _lib.check(_lib.lib().cola_serialize_ppl(g.handle, str(tmp_path).encode(), b"g", nGPU, gs))
