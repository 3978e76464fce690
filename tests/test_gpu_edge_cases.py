"""Edge cases of the device paths, against the FP64 oracle.

- Non-finite B: rows holding +Inf, -Inf and NaN give the IEEE result of the
  reference's loop over stored entries (only stored entries are multiplied,
  cuh:186-193), on the plan path (one-row-per-lane-group, whole-wave and hub
  items, with and without compensated sums), the literal virtual-row kernel
  and the batched kernel: padded broadcast slots contribute nothing.
- The column-panel schedule on a partition whose rows are not sorted by
  column (one item per row and panel, no racing read-add-write on C).
- An all-remote partition called with B_local = None.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import cola_amd as cola  # noqa: E402
from oracle import oracle as orc  # noqa: E402

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a GPU")]

FP32_TOL = 1e-5
DEV = torch.device("cuda:0")


def rel_fro(C, R):
    d = np.linalg.norm(np.asarray(C, np.float64) - R)
    r = np.linalg.norm(R)
    return d / r if r > 0 else d


def assert_ieee_match(C, ref):
    """NaN where the FP64 result is NaN, the same infinity where it is
    infinite, and 1e-5 relative Frobenius error over the finite elements."""
    C = np.asarray(C, np.float64)
    assert np.array_equal(np.isnan(C), np.isnan(ref)), "NaN pattern differs"
    inf = np.isinf(ref)
    assert np.array_equal(np.isinf(C), inf), "infinity pattern differs"
    assert np.array_equal(C[inf], ref[inf]), "sign of an infinity differs"
    fin = np.isfinite(ref)
    assert rel_fro(C[fin], ref[fin]) <= FP32_TOL


def nonfinite_B(rows, dim, bad_rows, seed):
    rng = np.random.default_rng(seed)
    B = rng.uniform(-1, 1, (rows, dim)).astype(np.float32)
    for k, r in enumerate(bad_rows):
        B[r, k % dim] = np.inf
        B[r, (k + 1) % dim] = -np.inf
        B[r, (k + 2) % dim] = np.nan
    return B


@pytest.mark.parametrize("opts, flags", [({}, cola.OVERWRITE),
                                         (dict(hub_threshold=40, chunk_nnz=16), cola.OVERWRITE),
                                         (dict(hub_threshold=40, chunk_nnz=16, kahan_degree=8),
                                          cola.OVERWRITE | cola.KAHAN),
                                         ({}, cola.ACCUMULATE)])
@pytest.mark.parametrize("rpw_short", [0, 1])
def test_nonfinite_B_plan_path(opts, flags, rpw_short):
    g = cola.Graph.rmat(12, 12, seed=3, reorder=True)
    rp, ci, va = g.arrays()
    nGPU, rank, dim = 2, 0, 64
    part = g.partition(nGPU, rank)
    P, M = part.info()["nodesPerPE"], part.info()["num_rows"]
    # Rows that end some item's entry list (the re-read padded slots load),
    # local and remote, and a few others.
    bad = sorted({int(ci[rp[i + 1] - 1]) for i in range(0, M, 97)} | {5, P + 7})
    Bh = nonfinite_B(P * nGPU, dim, bad, seed=1)
    ref = orc.spmm_f64(rp, ci, va, rank * P, rank * P + M, Bh[:g.n])
    B = torch.from_numpy(Bh).to(DEV)
    plan = cola.Plan(g.partition(nGPU, rank), **opts)
    plan.tuning = dict(rpw_short=rpw_short)
    packed = torch.empty((max(plan.U_r, 1), dim), dtype=torch.float32, device=DEV)
    C = torch.zeros((P, dim), dtype=torch.float32, device=DEV)
    plan.pack(B, packed)
    plan.spmm(C, B[rank * P:(rank + 1) * P], packed, flags=flags)
    torch.cuda.synchronize()
    assert_ieee_match(C.cpu().numpy()[:M], ref)


def test_nonfinite_B_literal_kernel():
    g = cola.Graph.rmat(12, 12, seed=4, reorder=True)
    rp, ci, va = g.arrays()
    nGPU, rank, dim = 4, 1, 32
    part = g.partition(nGPU, rank)
    P, M = part.info()["nodesPerPE"], part.info()["num_rows"]
    a = part.arrays()
    bad = sorted({int(a["csr_colIdx_l"][a["csr_rowPtr_l"][z + 1] - 1]) + rank * P
                  for z in range(0, len(a["csr_rowIdx_l"]), 13)})
    Bh = nonfinite_B(P * nGPU, dim, bad, seed=2)
    ref = orc.spmm_f64(rp, ci, va, rank * P, rank * P + M, Bh[:g.n])
    tl, pl, _ = cola.nnz_split(a["csr_rowIdx_l"], a["csr_rowPtr_l"], len(a["csr_colIdx_l"]))
    tr, pr, _ = cola.nnz_split(a["csc_colIdx_r"], a["csc_colPtr_r"], len(a["csc_rowIdx_r"]))
    zl = cola.d_csr_zip(pl, a["csr_colIdx_l"], tl, a["csr_val_l"])
    zr = cola.d_csc_zip(pr, a["csc_rowIdx_r"], tr, a["csc_val_r"])
    B = torch.from_numpy(Bh).to(DEV)
    C = torch.zeros((P, dim), dtype=torch.float32, device=DEV)
    cola.ppl0_batch_SpMM_global(C, B[rank * P:], zl, zr, 0, 0, zl.num, zr.num, dim, P, rank, literal=True)
    torch.cuda.synchronize()
    assert_ieee_match(C.cpu().numpy()[:M], ref)
    C.zero_()
    cola.ppl0_batch_SpMM_global(C, B[rank * P:], zl, zr, 0, 0, zl.num, zr.num, dim, P, rank, literal=False)
    torch.cuda.synchronize()
    assert_ieee_match(C.cpu().numpy()[:M], ref)
    cola.ppl0_plans_clear()


def test_nonfinite_B_batched_kernel():
    probs, refs = [], []
    for k, (log_n, deg, dim) in enumerate([(10, 8, 32), (11, 20, 64), (9, 40, 128)]):
        g = cola.Graph.rmat(log_n, deg, seed=30 + k)
        rp, ci, va = g.arrays()
        bad = sorted({int(ci[rp[i + 1] - 1]) for i in range(0, g.n, 37) if rp[i + 1] > rp[i]})
        Bh = nonfinite_B(g.n, dim, bad, seed=k)
        refs.append(orc.spmm_f64(rp, ci, va, 0, g.n, Bh))
        plan = cola.Plan(g.partition(1, 0), hub_threshold=48, chunk_nnz=16)
        probs.append(dict(plan=plan, B=torch.from_numpy(Bh).to(DEV),
                          C=torch.zeros((g.n, dim), dtype=torch.float32, device=DEV)))
    batch = cola.Batch(probs, dtype=cola.F32, flags=cola.OVERWRITE)
    batch.run()
    torch.cuda.synchronize()
    for p, ref in zip(probs, refs):
        assert_ieee_match(p["C"].cpu().numpy(), ref)


def test_column_panels_with_unsorted_rows():
    """Rows whose columns alternate between panels (local and remote entries
    shuffled within each half): one item per row and panel."""
    g = cola.Graph.rmat(12, 24, seed=8, reorder=True)
    rp, ci, va = g.arrays()
    nGPU, rank, dim = 2, 1, 64
    a = g.partition(nGPU, rank).arrays()
    rng = np.random.default_rng(0)
    col, val = a["csr_colIdx_l"].copy(), a["csr_val_l"].copy()
    for z in range(len(a["csr_rowIdx_l"])):
        s, e = a["csr_rowPtr_l"][z], a["csr_rowPtr_l"][z + 1]
        perm = rng.permutation(e - s)
        col[s:e], val[s:e] = col[s:e][perm], val[s:e][perm]
    part = cola.Partition.from_arrays(g.n, nGPU, rank, a["csr_rowPtr_l"], col, a["csr_rowIdx_l"],
                                      a["csc_colPtr_r"], a["csc_rowIdx_r"], a["csc_colIdx_r"], val, a["csc_val_r"])
    P, M = part.info()["nodesPerPE"], part.info()["num_rows"]
    Bh = np.random.default_rng(1).uniform(-1, 1, (P * nGPU, dim)).astype(np.float32)
    ref = orc.spmm_f64(rp, ci, va, rank * P, rank * P + M, Bh[:g.n])
    B = torch.from_numpy(Bh).to(DEV)
    for hub in (0, 40):
        plan = cola.Plan(part, col_panels=4, hub_threshold=hub, chunk_nnz=16)
        plan.tuning = dict(panels=1)
        packed = torch.empty((max(plan.U_r, 1), dim), dtype=torch.float32, device=DEV)
        plan.pack(B, packed)
        for _ in range(3):  # a lost update would show on some call
            C = torch.full((P, dim), 9.0, dtype=torch.float32, device=DEV)
            plan.spmm(C, B[rank * P:(rank + 1) * P], packed, flags=cola.OVERWRITE)
            torch.cuda.synchronize()
            assert rel_fro(C.cpu().numpy()[:M], ref) <= FP32_TOL


def test_all_remote_partition_without_B_local():
    """A partition with no local entries takes B_local = None."""
    n, nGPU, rank = 64, 2, 0
    P = n // nGPU
    rows = np.arange(P, dtype=np.int32)
    cols = (P + (rows * 7) % P).astype(np.int32)  # every entry remote
    # remote CSC over the sorted unique remote columns
    order = np.lexsort((rows, cols))
    uc, start = np.unique(cols[order], return_index=True)
    colPtr = np.append(start, len(order)).astype(np.int32)
    vals = np.linspace(0.5, 1.5, P).astype(np.float32)
    part = cola.Partition.from_arrays(n, nGPU, rank, [0], [], [], colPtr, rows[order], uc, [], vals[order])
    plan = cola.Plan(part)
    assert plan.U_l == 0
    dim = 16
    Bh = np.random.default_rng(2).uniform(-1, 1, (n, dim)).astype(np.float32)
    B = torch.from_numpy(Bh).to(DEV)
    packed = torch.empty((plan.U_r, dim), dtype=torch.float32, device=DEV)
    C = torch.zeros((P, dim), dtype=torch.float32, device=DEV)
    plan.pack(B, packed)
    plan.spmm(C, None, packed, dim=dim, flags=cola.OVERWRITE)
    plan.autotune(C, None, packed, dim=dim, iters=1)
    torch.cuda.synchronize()
    ref = vals[:, None].astype(np.float64) * Bh[cols].astype(np.float64)
    assert rel_fro(C.cpu().numpy(), ref) <= FP32_TOL


@pytest.mark.parametrize("dtype", [cola.F32, cola.BF16, cola.F16, cola.FP8_E4M3], ids=["f32", "bf16", "f16", "fp8"])
@pytest.mark.parametrize("dim", [32, 64, 128, 256])
def test_tiny_row_bundles(dtype, dim):
    """Rows of 1-4 entries run as bundles of floor(8/L) rows per lane group:
    bitwise equal to the short-row bin (one row per lane group, same entry
    order) and within the storage bound of FP64, in overwrite and accumulate
    mode, with and without XCD ranges, over whole rows and column tiles."""
    g = cola.Graph.rmat(14, 6, seed=12, reorder=True)  # many rows of 1-4 entries
    rp, ci, va = g.arrays()
    nGPU, rank = 2, 0
    part = g.partition(nGPU, rank)
    P, M = part.info()["nodesPerPE"], part.info()["num_rows"]
    plan = cola.Plan(part, row_order=1)
    assert plan.info["num_tiny_bundles"] > 0
    Bf = torch.from_numpy(np.random.default_rng(dim).uniform(-1, 1, (P * nGPU, dim)).astype(np.float32)).to(DEV)
    Bq, sc = (Bf, None) if dtype == cola.F32 else cola.quantize_rows(Bf, dtype)
    tdt = {cola.F32: torch.float32, cola.BF16: torch.uint16, cola.F16: torch.uint16, cola.FP8_E4M3: torch.uint8}[dtype]
    packed = torch.empty((max(plan.U_r, 1), dim), dtype=tdt, device=DEV)
    psc = torch.empty(max(plan.U_r, 1), dtype=torch.float32, device=DEV) if dtype == cola.FP8_E4M3 else None
    plan.pack(Bq, packed, dtype=dtype, scales_global=sc, packed_scales=psc)
    sl = None if sc is None else sc[:P]
    ref = orc.spmm_f64(rp, ci, va, 0, M, Bf.cpu().numpy()[:g.n])
    tol = {cola.F32: 1e-5, cola.BF16: 4e-3, cola.F16: 5e-4, cola.FP8_E4M3: 5e-2}[dtype]
    for tile in (0, -1):
        for xcd in (0, 1):
            outs = {}
            for tiny in (0, 1):
                plan.tuning = dict(col_tile=tile, rpw_short=0, xcd_remap=xcd, split_remote=0, tiny=tiny)
                for flags, init in ((cola.OVERWRITE, 5.0), (cola.ACCUMULATE, 0.25)):
                    C = torch.full((P, dim), init, dtype=torch.float32, device=DEV)
                    plan.spmm(C, Bq[:P], packed, dtype=dtype, flags=flags, scales_local=sl, packed_scales=psc)
                    torch.cuda.synchronize()
                    outs[(tiny, flags)] = C.cpu().numpy()
            for flags, init in ((cola.OVERWRITE, 0.0), (cola.ACCUMULATE, 0.25)):
                assert np.array_equal(outs[(0, flags)], outs[(1, flags)]), (tile, xcd, flags)
                got = outs[(1, flags)][:M] - init
                assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= tol
