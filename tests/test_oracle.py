"""Pins the oracle before anything is checked against it.

The reference ships no tests or golden vectors; its one known answer is the
built-in check (distSpMM_ppl_batch.cu:83,99,221, utils.cuh:44-100): B = 1 and
values = 1 give C[i,:] = deg(i).  Everything else is pinned by closed forms,
brute force on tiny inputs and the format definitions, chosen so that a
dropped term, a wrong index or a transposed operand fails.
"""
import math

import numpy as np
import pytest

import cola_amd as cola
from oracle import oracle as orc


def tiny_graph(n=40, m=120, seed=0, weighted=True):
    rng = np.random.default_rng(seed)
    r = rng.integers(0, n, m)
    c = rng.integers(0, n, m)
    keep = r != c
    r, c = r[keep], c[keep]
    rows = np.concatenate([r, c])
    cols = np.concatenate([c, r])
    vals = rng.uniform(-2, 2, len(rows)).astype(np.float32) if weighted else None
    g = cola.Graph.from_coo(n, rows, cols, vals)
    return g


def test_spmm_f64_matches_dense_brute_force():
    g = tiny_graph()
    rp, ci, va = g.arrays()
    B = np.random.default_rng(1).uniform(-1, 1, (g.n, 7)).astype(np.float32)
    for lb, ub in [(0, 40), (5, 17), (39, 40)]:
        C = orc.spmm_f64(rp, ci, va, lb, ub, B)
        D = orc.dense_partition_product(rp, ci, va, g.n, lb, ub, B)
        assert np.allclose(C, D, rtol=1e-12, atol=1e-12)
    # a transposed operand or a dropped value would not match
    D_bad = orc.dense_partition_product(rp, ci, None, g.n, 0, 40, B)
    assert not np.allclose(orc.spmm_f64(rp, ci, va, 0, 40, B), D_bad)


def test_known_answer_degree():
    """The reference's own check: B = 1, values = 1 => C[i,:] = deg(i)."""
    g = cola.Graph.rmat(10, 8, seed=3, gcn=False)
    rp, ci, va = g.arrays()
    assert np.all(va == 1.0)
    B = np.ones((g.n, 5), np.float32)
    C = orc.spmm_f64(rp, ci, va, 0, g.n, B)
    assert np.array_equal(C, np.repeat(np.diff(rp)[:, None].astype(np.float64), 5, axis=1))
    S = orc.sag_unweighted(0, g.n, rp, ci, B)
    assert orc.check_result(S, C.astype(np.float32)) == 0


def test_gcn_closed_form():
    """B = 1 with GCN values: C[i,:] = sum_j 1/sqrt(d_i d_j)."""
    g = cola.Graph.rmat(9, 6, seed=4)
    rp, ci, va = g.arrays()
    deg = np.diff(rp).astype(np.float64)
    C = orc.spmm_f64(rp, ci, va, 0, g.n, np.ones((g.n, 2), np.float32))
    expect = np.zeros(g.n)
    for i in range(g.n):
        for k in range(rp[i], rp[i + 1]):
            expect[i] += 1.0 / math.sqrt(deg[i] * deg[ci[k]])
    assert np.allclose(C[:, 0], expect, rtol=1e-6)


def test_check_result_detects_a_planted_error():
    C = np.ones((4, 4), np.float32)
    R = C.copy()
    assert orc.check_result(C, R) == 0
    C[2, 3] = 1.02  # 2% off: above the 1e-2 relative tolerance
    assert orc.check_result(C, R) == 1
    C[2, 3] = 1.005
    assert orc.check_result(C, R) == 0


def test_nnz_split_hand_worked():
    # rows of 3, 70, 1 entries: avg = 74 // 3 = 24 -> gran 32
    rowIdx = np.array([2, 5, 9])
    rowPtr = np.array([0, 3, 73, 74])
    t, p, gran = orc.nnz_split(rowIdx, rowPtr, 74)
    assert gran == 32
    assert list(t) == [2, 5, 5, 5, 9]
    assert list(p) == [0, 3, 35, 67, 73, 74]
    # avg 100 -> gran 128
    t, p, gran = orc.nnz_split(np.array([0]), np.array([0, 100]), 100)
    assert gran == 128 and list(t) == [0] and list(p) == [0, 100]
    t, p, gran = orc.nnz_split(np.array([], np.int32), np.array([0]), 0)
    assert len(t) == 0 and len(p) == 0


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_nnz_split_invariants(seed):
    rng = np.random.default_rng(seed)
    deg = rng.integers(1, 400, 50)
    rowPtr = np.concatenate([[0], np.cumsum(deg)])
    rowIdx = np.sort(rng.choice(10000, 50, replace=False))
    t, p, gran = orc.nnz_split(rowIdx, rowPtr, rowPtr[-1])
    assert gran % 32 == 0 and gran >= 32
    assert gran == -(-(rowPtr[-1] // 50) // 32) * 32
    assert p[0] == 0 and p[-1] == rowPtr[-1] and np.all(np.diff(p) > 0)
    assert len(t) == len(p) - 1
    pieces = np.array([np.sum(t == r) for r in rowIdx])
    assert np.array_equal(pieces, -(-deg // gran))
    assert np.all(np.diff(p) <= gran)


def test_fp8_decode_all_codes():
    """Every E4M3 code against the format definition, plus anchors."""
    vals = orc.fp8_to_f32(np.arange(256))
    for c in range(256):
        s, e, m = c >> 7, (c >> 3) & 15, c & 7
        if e == 15 and m == 7:
            assert np.isnan(vals[c])
            continue
        mag = (m / 8.0) * 2.0 ** -6 if e == 0 else (1 + m / 8.0) * 2.0 ** (e - 7)
        assert vals[c] == (-mag if s else mag)
    assert vals[0x7E] == 448.0 and vals[0x08] == 2.0 ** -6 and vals[0x01] == 2.0 ** -9
    assert vals[0x38] == 1.0 and vals[0xB8] == -1.0


def test_fp8_encode_round_trip_and_ties():
    codes = np.array([c for c in range(256) if (c & 0x7F) != 0x7F], np.uint8)
    dec = orc.fp8_to_f32(codes)
    enc = orc.fp8_from_f32(dec)
    # +0 and -0 both decode to zero; every other code survives the round trip
    nz = dec != 0
    assert np.array_equal(enc[nz], codes[nz])
    # ties to even: 1 + 1/16 is halfway between 1 (m=0) and 1.125 (m=1)
    assert orc.fp8_from_f32(np.float32(1.0625)) == 0x38
    assert orc.fp8_from_f32(np.float32(1.1875)) == 0x3A  # between m=1 and m=2 -> m=2
    assert orc.fp8_from_f32(np.float32(1000.0)) == 0x7E  # saturates
    assert orc.fp8_from_f32(np.float32(-1000.0)) == 0xFE
    assert orc.fp8_from_f32(np.float32(2.0 ** -10)) == 0x00  # half of the smallest subnormal -> even (0)
    assert orc.fp8_from_f32(np.float32(3 * 2.0 ** -10)) == 0x02  # 1.5 quanta -> 2


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1.0 + 2 ** -9], np.float32)
    assert list(orc.bf16_from_f32(x)) == [0x3F80, 0x3F80, 0x3F82, 0xC020, 0x3F80]
    assert orc.bf16_from_f32(np.float32(np.nan)) & 0x7FC0 == 0x7FC0


def test_row_partition_np_invariants_and_reconstruction():
    g = tiny_graph(n=103, m=400, seed=7)
    rp, ci, va = g.arrays()
    for nGPU in (1, 2, 4, 5):
        P = -(-g.n // nGPU)
        for rank in range(nGPU):
            a = orc.row_partition_np(rp, ci, va, g.n, nGPU, rank)
            lb, ub = rank * P, min(g.n, (rank + 1) * P)
            M = max(ub - lb, 0)
            assert a["csr_rowPtr_l"][0] == 0 and np.all(np.diff(a["csr_rowPtr_l"]) > 0)
            assert len(a["csr_rowIdx_l"]) == len(a["csr_rowPtr_l"]) - 1
            assert np.all(np.diff(a["csr_rowIdx_l"]) > 0)
            assert np.all((a["csr_colIdx_l"] >= 0) & (a["csr_colIdx_l"] < P))
            assert np.all(np.diff(a["csc_colIdx_r"]) > 0)
            assert not np.any((a["csc_colIdx_r"] >= lb) & (a["csc_colIdx_r"] < lb + P))
            assert a["csc_colPtr_r"][0] == 0 and np.all(np.diff(a["csc_colPtr_r"]) > 0)
            assert len(a["csr_colIdx_l"]) + len(a["csc_rowIdx_r"]) == (rp[ub] - rp[lb] if M else 0)
            # rebuild the dense panel from both halves
            A = np.zeros((M, g.n))
            for z, r in enumerate(a["csr_rowIdx_l"]):
                for k in range(a["csr_rowPtr_l"][z], a["csr_rowPtr_l"][z + 1]):
                    A[r, a["csr_colIdx_l"][k] + rank * P] += a["csr_val_l"][k]
            for u, c in enumerate(a["csc_colIdx_r"]):
                for k in range(a["csc_colPtr_r"][u], a["csc_colPtr_r"][u + 1]):
                    A[a["csc_rowIdx_r"][k], c] += a["csc_val_r"][k]
            D = np.zeros((M, g.n))
            for i in range(lb, ub):
                for k in range(rp[i], rp[i + 1]):
                    D[i - lb, ci[k]] += va[k]
            assert np.array_equal(A, D)
            # rows sorted within each column of the CSC
            for u in range(len(a["csc_colIdx_r"])):
                seg = a["csc_rowIdx_r"][a["csc_colPtr_r"][u]:a["csc_colPtr_r"][u + 1]]
                assert np.all(np.diff(seg) > 0)


def test_ppl0_restatement_matches_fp64():
    g = cola.Graph.rmat(12, 10, seed=2)
    rp, ci, va = g.arrays()
    nGPU, dim = 4, 24
    P = -(-g.n // nGPU)
    B = np.random.default_rng(0).uniform(-1, 1, (P * nGPU, dim)).astype(np.float32)
    for rank in (0, 2):  # rank 2 walks remote columns in reverse
        a = orc.row_partition_np(rp, ci, va, g.n, nGPU, rank)
        tl, pl, _ = orc.nnz_split(a["csr_rowIdx_l"], a["csr_rowPtr_l"], len(a["csr_colIdx_l"]))
        tr, pr, _ = orc.nnz_split(a["csc_colIdx_r"], a["csc_colPtr_r"], len(a["csc_rowIdx_r"]))
        C = np.zeros((P, dim), np.float32)
        orc.ppl0_f32(C, B, dim, P, rank, pl, a["csr_colIdx_l"], tl, a["csr_val_l"], pr, a["csc_rowIdx_r"], tr,
                     a["csc_val_r"])
        M = min(g.n, (rank + 1) * P) - rank * P
        R = orc.spmm_f64(rp, ci, va, rank * P, rank * P + M, B[:g.n])
        assert np.linalg.norm(C[:M] - R) / np.linalg.norm(R) < 1e-6


def test_openmp_baseline_matches_fp64():
    g = cola.Graph.rmat(12, 10, seed=2)
    rp, ci, va = g.arrays()
    B = np.random.default_rng(0).uniform(-1, 1, (g.n, 32)).astype(np.float32)
    C, threads = orc.spmm_omp_f32(rp, ci, va, 100, 900, B)
    R = orc.spmm_f64(rp, ci, va, 100, 900, B)
    assert threads >= 1
    assert np.linalg.norm(C - R) / np.linalg.norm(R) < 1e-6
