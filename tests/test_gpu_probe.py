"""The gather-ceiling measurement kernels.  cola_probe_gather against the
FP64 oracle: mode 0 computes the partition product over the plan's items,
mode 1 the same product with every row index masked into the first
table_rows rows of each table; mode 2 runs; bad arguments (and tables the
buffers do not hold) are rejected before any launch.  cola_probe_row_stream
(the bare gather of a row stream) reads exactly the rows it is given."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import cola_amd as cola  # noqa: E402
from cola_amd import _lib  # noqa: E402
from oracle import oracle as orc  # noqa: E402

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a GPU")]

FP32_TOL = 1e-5
DEV = torch.device("cuda:0")


@pytest.fixture(scope="module")
def case():
    nGPU, rank, dim = 3, 2, 64  # the last rank: 1364 rows against nodesPerPE = 1366
    g = cola.Graph.rmat(12, 8, seed=21, reorder=False)
    rp, ci, va = g.arrays()
    part = g.partition(nGPU, rank)
    info = part.info()
    P, M = info["nodesPerPE"], info["num_rows"]
    rng = np.random.default_rng(5)
    Bh = rng.uniform(-1, 1, (P * nGPU, dim)).astype(np.float32)
    plan = cola.Plan(part)
    B = torch.from_numpy(Bh).to(DEV)
    B_local = B[rank * P:rank * P + M]  # exactly the partition's rows
    packed = torch.empty((max(plan.U_r, 1), dim), dtype=torch.float32, device=DEV)
    plan.pack(B, packed)
    print(f"case probe: n={g.n} rank={rank} M={M} P={P} U_r={plan.U_r} dim={dim}", flush=True)
    return dict(g=g, rp=rp, ci=ci, va=va, part=part, P=P, M=M, rank=rank, dim=dim, Bh=Bh, plan=plan,
                B=B, B_local=B_local, packed=packed, arrays=part.arrays())


def probe(c, C, mode, table_rows, tile=0, B_local=True):
    Bl = c["B_local"]
    return _lib.lib().cola_probe_gather(
        c["plan"].handle, ctypes.c_void_p(Bl.data_ptr() if B_local else 0), Bl.stride(0),
        ctypes.c_void_p(c["packed"].data_ptr()), c["packed"].stride(0), ctypes.c_void_p(C.data_ptr()), C.stride(0),
        c["dim"], tile, mode, table_rows, 0, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))


@pytest.mark.parametrize("tile", [0, 32, 16])
def test_real_rows_give_the_partition_product(case, tile):
    c = case
    assert c["plan"].info["num_hub_rows"] == 0  # every row's sum is stored (hub partials are not reduced)
    C = torch.zeros((c["M"], c["dim"]), dtype=torch.float32, device=DEV)
    assert probe(c, C, 0, 0, tile) == 0
    torch.cuda.synchronize()
    lb = c["rank"] * c["P"]
    ref = orc.spmm_f64(c["rp"], c["ci"], c["va"], lb, lb + c["M"], c["Bh"][:c["g"].n])
    err = np.linalg.norm(C.cpu().numpy() - ref) / np.linalg.norm(ref)
    assert err <= FP32_TOL, err


def test_masked_rows_give_the_masked_product(case):
    c = case
    T = 16
    a = c["arrays"]
    lb = c["rank"] * c["P"]
    # Expected: local column j reads B_local row (j & (T-1)); packed row u
    # reads packed row (u & (T-1)), which holds global row colIdx_r[u & (T-1)].
    ref = np.zeros((c["M"], c["dim"]))
    Bloc = c["Bh"][lb:lb + c["M"]].astype(np.float64)
    rows_l = np.repeat(a["csr_rowIdx_l"], np.diff(a["csr_rowPtr_l"]))
    np.add.at(ref, rows_l, a["csr_val_l"][:, None] * Bloc[a["csr_colIdx_l"] & (T - 1)])
    U = len(a["csc_colIdx_r"])
    u = np.repeat(np.arange(U), np.diff(a["csc_colPtr_r"]))
    Bp = c["Bh"][a["csc_colIdx_r"][u & (T - 1)]].astype(np.float64)
    np.add.at(ref, a["csc_rowIdx_r"], a["csc_val_r"][:, None] * Bp)
    C = torch.zeros((c["M"], c["dim"]), dtype=torch.float32, device=DEV)
    assert probe(c, C, 1, T) == 0
    torch.cuda.synchronize()
    err = np.linalg.norm(C.cpu().numpy() - ref) / np.linalg.norm(ref)
    assert err <= FP32_TOL, err


def test_streamed_mode_runs(case):
    c = case
    C = torch.zeros((c["M"], c["dim"]), dtype=torch.float32, device=DEV)
    assert probe(c, C, 2, c["M"]) == 0
    torch.cuda.synchronize()
    assert torch.isfinite(C).all()


@pytest.mark.parametrize("mode, table_rows, tile, with_B_local", [
    (3, 16, 0, True),          # unknown mode
    (1, 12, 0, True),          # table_rows not a power of two
    (1, 1 << 20, 0, True),     # larger than the partition's rows and U_r
    (2, 1 << 20, 0, True),     # larger than the partition's rows
    (2, 16, 0, False),         # mode 2 reads B_local
    (0, 0, 6, True),           # tile not a multiple of 4
])
def test_bad_arguments_are_rejected(case, mode, table_rows, tile, with_B_local):
    c = case
    C = torch.zeros((c["M"], c["dim"]), dtype=torch.float32, device=DEV)
    assert probe(c, C, mode, table_rows, tile, with_B_local) == _lib.COLA_ERR_INVALID
    assert _lib.lib().cola_last_error()


@pytest.mark.parametrize("xcd_ranges", [0, 1], ids=["strided", "xcd_ranges"])
@pytest.mark.parametrize("via_lds", [0, 1], ids=["vgpr", "lds"])
def test_row_stream_gather_sums_exactly_the_named_rows(via_lds, xcd_ranges):
    """cola_probe_row_stream loads the first 256 bytes of every named row once
    (a stream of whole 32-entry steps, so no entry is re-read at the end):
    with small-integer rows the per-block sums add up exactly, through VGPRs
    and through LDS (global_load_lds_dwordx4 lane layout)."""
    rng = np.random.default_rng(9)
    rows, width = 3000, 96  # 384-byte rows
    Bh = rng.integers(0, 4, (rows, width)).astype(np.float32)
    B = torch.from_numpy(Bh).to(DEV)
    n = 32 * 64
    idx = rng.integers(0, rows, n)
    offs = torch.from_numpy((idx * width * 4 // 16).astype(np.uint32).view(np.int32)).to(DEV)
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    out = torch.zeros(cus * 2, dtype=torch.float32, device=DEV)
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    st = _lib.lib().cola_probe_row_stream(ctypes.c_void_p(offs.data_ptr()), n, ctypes.c_void_p(B.data_ptr()),
                                          via_lds, xcd_ranges, 2, ctypes.c_void_p(out.data_ptr()), out.numel(), sp)
    assert st == 0, _lib.lib().cola_last_error()
    torch.cuda.synchronize()
    assert float(out.double().sum()) == float(Bh[idx, :64].astype(np.float64).sum())


def test_row_stream_rejects_bad_arguments():
    B = torch.zeros((4, 64), dtype=torch.float32, device=DEV)
    offs = torch.zeros(32, dtype=torch.int32, device=DEV)
    out = torch.zeros(8, dtype=torch.float32, device=DEV)  # fewer floats than blocks
    sp = ctypes.c_void_p(0)
    lib = _lib.lib()
    assert lib.cola_probe_row_stream(ctypes.c_void_p(offs.data_ptr()), 32, ctypes.c_void_p(B.data_ptr()), 0, 0, 2,
                                     ctypes.c_void_p(out.data_ptr()), out.numel(), sp) == _lib.COLA_ERR_INVALID
    big = torch.zeros(1 << 16, dtype=torch.float32, device=DEV)
    assert lib.cola_probe_row_stream(ctypes.c_void_p(offs.data_ptr()), 32, ctypes.c_void_p(B.data_ptr() + 4), 0, 0,
                                     2, ctypes.c_void_p(big.data_ptr()), big.numel(), sp) == _lib.COLA_ERR_INVALID
