"""The reference's own launch API (include/dist_spmm_ppl_batch.h: bin2vec,
nnz_split, ppl0_batch_SpMM, distSpMM_ppl_batch_wrapper,
run_distSpMM_ppl_batch) on the GPU.

1. tests/cpp/reference_driver.cpp, written like the reference's
   src/distSpMM_ppl_batch.cu:49-99,195-199,243-246, runs on partition files
   written by cola_serialize: the harness's own check (SAG_reference_host +
   check_result, 1e-2 relative) passes with B = 1 (the reference's
   fill_certain) and with fill_seq's row-dependent B; with a random B
   (signed, so sums can cancel below that check's relative bound) the dumped
   C matches the FP64 oracle (values 1.0, as the reference loads
   them) within 1e-5 relative Frobenius, with and without nnz_split on the
   remote side, through the cached plan and the literal atomic kernel, at
   every rank and at dims the reference cannot compute (96, 200).
2. The same launcher through the Python binding, with weighted (GCN) values
   and the remote rows read straight from B_global, against FP64.
"""
import os
import subprocess

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import cola_amd as cola  # noqa: E402
from oracle import oracle as orc  # noqa: E402

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a GPU")]

FP32_TOL = 1e-5


def rel_fro(C, R):
    d = np.linalg.norm(np.asarray(C, np.float64) - R)
    r = np.linalg.norm(R)
    return d / r if r > 0 else d


@pytest.fixture(scope="module")
def storage(tmp_path_factory):
    """A graph serialized the way prep_torch_par.py writes it (no value
    files), its driver, and its global CSR."""
    from tests.cpp_driver import build_driver
    d = tmp_path_factory.mktemp("storage")
    exe = build_driver(d, "reference_driver")
    g = cola.Graph.rmat(13, 12, seed=21, reorder=True, gcn=False)
    cola.serialize(g, str(d), "rmat13", [4, 16], write_values=False)
    rp, ci, _ = g.arrays()
    return dict(dir=str(d), exe=exe, g=g, rp=rp, ci=ci, n=g.n)


def run_driver(st, nGPU, dim, rank, fill="ones", out="-", split=1, literal=0, group_size=None, profile=0):
    args = [st["exe"], st["dir"], "rmat13", str(nGPU), str(dim), str(group_size or nGPU), str(split), str(rank),
            fill, out, str(literal), str(profile)]
    return subprocess.run(args, capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("dim", [32, 128, 96])
@pytest.mark.parametrize("literal", [0, 1])
def test_reference_driver_known_answer(storage, dim, literal):
    for rank in range(4):
        out = run_driver(storage, 4, dim, rank, "ones", literal=literal)
        assert out.returncode == 0, out.stdout + out.stderr
        assert f"PEid-{rank} passed" in out.stdout
        if rank == 0:  # nnz_split prints its ratio only for a non-empty part (cuh:32-33)
            assert "l_split_ratio" in out.stdout and "r_split_ratio" in out.stdout


def test_reference_driver_fill_seq(storage):
    """fill_seq (utils.cuh:38-42) makes B rows differ, so wrong remote-row
    addressing fails the harness's own check."""
    for rank in range(4):
        out = run_driver(storage, 4, 64, rank, "seq")
        assert out.returncode == 0, out.stdout + out.stderr
        assert f"PEid-{rank} passed" in out.stdout


@pytest.mark.parametrize("dim, split, literal", [(64, 1, 0), (64, 0, 0), (64, 1, 1), (64, 0, 1), (200, 1, 0),
                                                 (128, 1, 0)])
def test_reference_driver_random_B_matches_fp64(storage, tmp_path, dim, split, literal):
    nGPU = 4
    P = (storage["n"] + nGPU - 1) // nGPU
    rng = np.random.default_rng(dim * 10 + split * 2 + literal)
    B = rng.uniform(-1, 1, (P * nGPU, dim)).astype(np.float32)
    bfile = tmp_path / "B.bin"
    B.tofile(bfile)
    ones = np.ones(len(storage["ci"]), np.float32)
    for rank in range(nGPU):
        cfile = tmp_path / f"C{rank}.bin"
        out = run_driver(storage, nGPU, dim, rank, str(bfile), str(cfile), split=split, literal=literal)
        assert out.returncode == 0, out.stdout + out.stderr
        C = np.fromfile(cfile, np.float32).reshape(P, dim)
        lb = rank * P
        M = min(storage["n"], lb + P) - lb
        ref = orc.spmm_f64(storage["rp"], storage["ci"], ones, lb, lb + M, B[:storage["n"]])
        err = rel_fro(C[:M], ref)
        assert err <= FP32_TOL, (rank, err)
        assert not np.any(C[M:]), "rows past the partition must stay zero"


def test_reference_driver_timing_loop(storage):
    out = run_driver(storage, 4, 64, 0, "ones", profile=1)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "PEid-0 passed" in out.stdout and "PE-0 time (ms)" in out.stdout


def test_reference_driver_multi_gpu_forms_are_refused(storage):
    """More than one pipeline stage, and the buffered branch for more than 8
    ranks, need puts between GPUs: COLA_ERR_UNSUPPORTED (exit 4), not a
    wrong answer."""
    out = run_driver(storage, 4, 32, 0, "ones", group_size=2)
    assert out.returncode == 4, out.stdout + out.stderr
    assert "stage" in out.stderr
    out = run_driver(storage, 16, 32, 0, "ones")
    assert out.returncode == 4, out.stdout + out.stderr
    assert "8 ranks" in out.stderr


@pytest.mark.parametrize("literal", [False, True])
@pytest.mark.parametrize("rank", [0, 1, 2, 3])
def test_global_launcher_weighted_matches_fp64(rank, literal):
    """GCN values, nnz_split on both sides, remote rows read from B_global at
    their global row; accumulate into a prefilled C."""
    g = cola.Graph.rmat(14, 10, seed=5, reorder=True)
    rp, ci, va = g.arrays()
    nGPU, dim = 4, 64
    part = g.partition(nGPU, rank)
    info = part.info()
    P, M = info["nodesPerPE"], info["num_rows"]
    a = part.arrays()
    tl, pl, _ = cola.nnz_split(a["csr_rowIdx_l"], a["csr_rowPtr_l"], len(a["csr_colIdx_l"]))
    tr, pr, _ = cola.nnz_split(a["csc_colIdx_r"], a["csc_colPtr_r"], len(a["csc_rowIdx_r"]))
    zl = cola.d_csr_zip(pl, a["csr_colIdx_l"], tl, a["csr_val_l"])
    zr = cola.d_csc_zip(pr, a["csc_rowIdx_r"], tr, a["csc_val_r"])
    rng = np.random.default_rng(rank)
    Bh = rng.uniform(-1, 1, (P * nGPU, dim)).astype(np.float32)
    dev = torch.device("cuda:0")
    B = torch.from_numpy(Bh).to(dev)
    C = torch.full((P, dim), 0.5, dtype=torch.float32, device=dev)
    cola.ppl0_batch_SpMM_global(C, B[rank * P:], zl, zr, 0, 0, zl.num, zr.num, dim, P, rank, literal=literal)
    cola.ppl0_batch_SpMM_global(C, B[rank * P:], zl, zr, 0, 0, zl.num, zr.num, dim, P, rank, literal=literal)
    torch.cuda.synchronize()
    lb = rank * P
    ref = 0.5 + 2.0 * orc.spmm_f64(rp, ci, va, lb, lb + M, Bh[:g.n])
    assert rel_fro(C.cpu().numpy()[:M], ref) <= FP32_TOL
    cola.ppl0_plans_clear()


def test_global_launcher_offset_ranges():
    """Two calls over halves of the virtual rows / columns add up to the
    whole product (the offsets the multi-stage driver would pass)."""
    g = cola.Graph.rmat(13, 16, seed=9, reorder=True)
    rp, ci, va = g.arrays()
    nGPU, rank, dim = 2, 1, 32
    part = g.partition(nGPU, rank)
    P, M = part.info()["nodesPerPE"], part.info()["num_rows"]
    a = part.arrays()
    zl = cola.d_csr_zip(a["csr_rowPtr_l"], a["csr_colIdx_l"], a["csr_rowIdx_l"], a["csr_val_l"])
    zr = cola.d_csc_zip(a["csc_colPtr_r"], a["csc_rowIdx_r"], a["csc_colIdx_r"], a["csc_val_r"])
    rng = np.random.default_rng(3)
    Bh = rng.uniform(-1, 1, (P * nGPU, dim)).astype(np.float32)
    dev = torch.device("cuda:0")
    B = torch.from_numpy(Bh).to(dev)
    C = torch.zeros((P, dim), dtype=torch.float32, device=dev)
    hl, hr = zl.num // 2, zr.num // 3
    cola.ppl0_batch_SpMM_global(C, B[rank * P:], zl, zr, 0, 0, hl, hr, dim, P, rank)
    cola.ppl0_batch_SpMM_global(C, B[rank * P:], zl, zr, hl, hr, zl.num - hl, zr.num - hr, dim, P, rank)
    torch.cuda.synchronize()
    ref = orc.spmm_f64(rp, ci, va, rank * P, rank * P + M, Bh[:g.n])
    assert rel_fro(C.cpu().numpy()[:M], ref) <= FP32_TOL
    cola.ppl0_plans_clear()


def test_reference_header_is_what_the_driver_includes():
    src = open(os.path.join(os.path.dirname(__file__), "cpp", "reference_driver.cpp")).read()
    assert '#include "dist_spmm_ppl_batch.h"' in src
    for call in ("bin2vec<int>(", "nnz_split(", "new d_csc_zip(", "new d_csr_zip(", "run_distSpMM_ppl_batch<true"):
        assert call in src
