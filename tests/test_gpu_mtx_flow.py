"""The reference's whole offline-to-kernel flow on the GPU, starting from a
MatrixMarket file: load_mtx + serialize (prep_torch_par.py:14-20,93-118)
write the partition files, the partition load of distSpMM_ppl_batch.cu:49-99
reads them back, and the kernels run on every rank: the plan path, the
single-stage wrapper over nnz_split zips, and the C++ driver written like
the reference's (tests/cpp/reference_driver.cpp).  Checked against the FP64
oracle over scipy.io.mmread's CSR of the same file (values 1.0, the
reference driver's override, distSpMM_ppl_batch.cu:83,99).

Files: the committed tests/golden/par_small.mtx (real general) and a
generated symmetric file with repeated coordinates (summed by mmread) and a
diagonal, about 10^5 stored entries after mirroring."""
import os
import subprocess

import numpy as np
import pytest

torch = pytest.importorskip("torch")
sio = pytest.importorskip("scipy.io")

import cola_amd as cola  # noqa: E402
from oracle import oracle as orc  # noqa: E402

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a GPU")]

FP32_TOL = 1e-5
DEV = torch.device("cuda:0")
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rel_fro(C, R):
    d = np.linalg.norm(np.asarray(C, np.float64) - R)
    r = np.linalg.norm(R)
    return d / r if r > 0 else d


def write_symmetric_mtx(path, n, lines, seed):
    """Lower-triangle coordinates (diagonal included), a tenth of them
    repeated, real values."""
    rng = np.random.default_rng(seed)
    i = rng.integers(0, n, lines)
    j = np.minimum(i, (i - np.abs(rng.normal(0, n / 50, lines))).astype(np.int64).clip(0))
    j[::97] = i[::97]  # some diagonal entries
    rep = np.arange(0, lines, 10)
    i, j = np.concatenate([i, i[rep]]), np.concatenate([j, j[rep]])
    v = rng.uniform(-2, 2, len(i))
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real symmetric\n% generated\n")
        f.write(f"{n} {n} {len(i)}\n")
        for a, b, x in zip(i + 1, j + 1, v):
            f.write(f"{a} {b} {x:.6f}\n")


@pytest.fixture(scope="module", params=["par_small", "sym_dups"])
def flow(request, tmp_path_factory):
    d = tmp_path_factory.mktemp("mtxflow")
    mtx_dir = d / "mtx"
    mtx_dir.mkdir()
    name = request.param
    if name == "par_small":
        src = os.path.join(GOLDEN, "par_small.mtx")
        dst = mtx_dir / "par_small.mtx"
        dst.write_bytes(open(src, "rb").read())
    else:
        write_symmetric_mtx(mtx_dir / "sym_dups.mtx", 20000, 50000, seed=13)
    sp_dir = d / "SP"
    nGPUs = [2, 4]
    g = cola.process_graph(name, str(mtx_dir), str(sp_dir), nGPUs)
    A = sio.mmread(str(mtx_dir / f"{name}.mtx")).tocsr()
    A.sum_duplicates()
    rp, ci = A.indptr.astype(np.int64), A.indices.astype(np.int32)
    print(f"case mtx flow {name}: n={A.shape[0]} stored entries={A.nnz}", flush=True)
    assert g.n == A.shape[0] and g.info()[1] == A.nnz
    return dict(name=name, sp=str(sp_dir), mtx=str(mtx_dir), n=A.shape[0], rp=rp, ci=ci,
                ones=np.ones(A.nnz, np.float32), nGPUs=nGPUs, dir=d)


def test_mtx_flow_plan_and_wrapper_every_rank(flow):
    dim = 64
    for nGPU in flow["nGPUs"]:
        P = (flow["n"] + nGPU - 1) // nGPU
        Bh = np.random.default_rng(nGPU).uniform(-1, 1, (P * nGPU, dim)).astype(np.float32)
        B = torch.from_numpy(Bh).to(DEV)
        for rank in range(nGPU):
            part = cola.Partition.load(flow["sp"], flow["name"], nGPU, rank)
            M = part.info()["num_rows"]
            lb = rank * P
            ref = orc.spmm_f64(flow["rp"], flow["ci"], flow["ones"], lb, lb + M, Bh[:flow["n"]])
            # plan path
            plan = cola.Plan(part)
            packed = torch.empty((max(plan.U_r, 1), dim), dtype=torch.float32, device=DEV)
            C = torch.full((P, dim), 4.0, dtype=torch.float32, device=DEV)
            plan.pack(B, packed)
            plan.spmm(C, B[lb:lb + P], packed, flags=cola.OVERWRITE)
            # single-stage wrapper over nnz_split zips (accumulates)
            a = part.arrays()
            tl, pl, _ = cola.nnz_split(a["csr_rowIdx_l"], a["csr_rowPtr_l"], len(a["csr_colIdx_l"]))
            tr, pr, _ = cola.nnz_split(a["csc_colIdx_r"], a["csc_colPtr_r"], len(a["csc_rowIdx_r"]))
            C2 = torch.zeros((P, dim), dtype=torch.float32, device=DEV)
            if len(a["csr_colIdx_l"]) + len(a["csc_rowIdx_r"]) > 0:
                zl = cola.d_csr_zip(pl if len(pl) else [0], a["csr_colIdx_l"], tl, a["csr_val_l"])
                zr = cola.d_csc_zip(pr if len(pr) else [0], a["csc_rowIdx_r"], tr, a["csc_val_r"])
                packed2 = torch.empty((max(zr.n_unique, 1), dim), dtype=torch.float32, device=DEV)
                cola.dist_spmm_ppl_batch(C2, B, packed2, zl, zr, dim, P, rank)
            torch.cuda.synchronize()
            for got, fill in ((C, 4.0), (C2, 0.0)):
                got = got.cpu().numpy()
                if np.linalg.norm(ref) == 0:
                    assert not np.any(got[:M])
                else:
                    assert rel_fro(got[:M], ref) <= FP32_TOL, (nGPU, rank)
                assert np.all(got[M:P] == fill), "rows past the partition are not written"


def test_mtx_flow_cpp_driver(flow, tmp_path):
    from tests.cpp_driver import build_driver
    exe = build_driver(tmp_path, "reference_driver")
    dim = 32
    for nGPU in flow["nGPUs"]:
        P = (flow["n"] + nGPU - 1) // nGPU
        Bh = np.random.default_rng(10 + nGPU).uniform(-1, 1, (P * nGPU, dim)).astype(np.float32)
        bfile = tmp_path / "B.bin"
        Bh.tofile(bfile)
        for rank in range(nGPU):
            args = [exe, flow["sp"], flow["name"], str(nGPU), str(dim), str(nGPU), "1", str(rank)]
            out = subprocess.run(args + ["ones"], capture_output=True, text=True, timeout=120)
            assert out.returncode == 0, out.stdout + out.stderr
            assert f"PEid-{rank} passed" in out.stdout
            cfile = tmp_path / f"C{nGPU}_{rank}.bin"
            out = subprocess.run(args + [str(bfile), str(cfile)], capture_output=True, text=True, timeout=120)
            assert out.returncode == 0, out.stdout + out.stderr
            C = np.fromfile(cfile, np.float32).reshape(P, dim)
            lb = rank * P
            M = min(flow["n"], lb + P) - lb
            ref = orc.spmm_f64(flow["rp"], flow["ci"], flow["ones"], lb, lb + M, Bh[:flow["n"]])
            if np.linalg.norm(ref) > 0:
                assert rel_fro(C[:M], ref) <= FP32_TOL, (nGPU, rank)
