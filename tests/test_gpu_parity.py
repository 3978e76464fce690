"""GPU parity tests: the HIP paths through the C ABI against the CPU oracle.

Bars: FP32 storage -> relative Frobenius error <= 1e-5 against the FP64 oracle
(the stated tolerance for the plain path); integer / byte work (packing,
conversions) -> bit-exact; reduced-precision storage -> the error bounds
stated per dtype below.
"""
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import cola_amd as cola  # noqa: E402
from oracle import oracle as orc  # noqa: E402

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a GPU")]

FP32_TOL = 1e-5
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rel_fro(C, R):
    d = np.linalg.norm(np.asarray(C, np.float64) - R)
    r = np.linalg.norm(R)
    return d / r if r > 0 else d


class Case:
    """A graph, one partition of it, a padded B and the FP64 oracle product."""

    def __init__(self, graph, nGPU, rank, dim, seed=0, b_ones=False):
        self.g = graph
        self.n = graph.n
        self.rp, self.ci, self.va = graph.arrays()
        self.part = graph.partition(nGPU, rank)
        self.info = self.part.info()
        self.P, self.M = self.info["nodesPerPE"], self.info["num_rows"]
        self.nGPU, self.rank, self.dim = nGPU, rank, dim
        rng = np.random.default_rng(seed)
        rows = self.P * nGPU
        self.Bh = np.ones((rows, dim), np.float32) if b_ones else rng.uniform(-1, 1, (rows, dim)).astype(np.float32)
        lb = rank * self.P
        self.ref = orc.spmm_f64(self.rp, self.ci, self.va, lb, lb + self.M, self.Bh[:self.n])
        self.dev = torch.device("cuda:0")
        self.B = torch.from_numpy(self.Bh).to(self.dev)

    def B_local(self):
        return self.B[self.rank * self.P:(self.rank + 1) * self.P]

    def run_plan(self, flags=cola.OVERWRITE, init=None, **opts):
        plan = cola.Plan(self.part, **opts)
        packed = torch.empty((max(plan.U_r, 1), self.dim), dtype=torch.float32, device=self.dev)
        C = (torch.zeros if init is None else torch.full)((self.P, self.dim), *(() if init is None else (init,)),
                                                          dtype=torch.float32, device=self.dev)
        plan.pack(self.B, packed)
        plan.spmm(C, self.B_local(), packed, flags=flags)
        torch.cuda.synchronize()
        return C.cpu().numpy(), plan

    def run_reference_path(self, split=True, literal=False, init=0.0, calls=1):
        """The reference API: zips (after nnz_split) + the single-stage wrapper,
        which runs through a plan cached from the zips; literal=True runs the
        ppl0_batch_SpMM restatement (virtual rows, FP32 atomics) instead."""
        a = self.part.arrays()
        if split:
            tl, pl, _ = cola.nnz_split(a["csr_rowIdx_l"], a["csr_rowPtr_l"], len(a["csr_colIdx_l"]))
            tr, pr, _ = cola.nnz_split(a["csc_colIdx_r"], a["csc_colPtr_r"], len(a["csc_rowIdx_r"]))
        else:
            tl, pl = a["csr_rowIdx_l"], a["csr_rowPtr_l"]
            tr, pr = a["csc_colIdx_r"], a["csc_colPtr_r"]
        zl = cola.d_csr_zip(pl, a["csr_colIdx_l"], tl, a["csr_val_l"])
        zr = cola.d_csc_zip(pr, a["csc_rowIdx_r"], tr, a["csc_val_r"])
        packed = torch.empty((max(zr.n_unique, 1), self.dim), dtype=torch.float32, device=self.dev)
        C = torch.full((self.P, self.dim), float(init), dtype=torch.float32, device=self.dev)
        for _ in range(calls):
            if literal:
                uniq = torch.from_numpy(np.unique(np.asarray(tr, np.int32))).to(self.dev)
                cola.pack_remote(self.B, uniq, packed, num_cols=zr.n_unique)
                cola.ppl0_batch_SpMM(C, self.B_local(), packed, zl, zr, 0, 0, zl.num, zr.num, self.dim, self.P,
                                     self.rank)
            else:
                cola.dist_spmm_ppl_batch(C, self.B, packed, zl, zr, self.dim, self.P, self.rank)
        torch.cuda.synchronize()
        return C.cpu().numpy()


@pytest.fixture(scope="module")
def c1():
    """The CPU-runnable configuration: R-MAT 2^16, degree 16, GCN values,
    partition 0 of 4, dim 64."""
    return Case(cola.Graph.rmat(16, 16, seed=1), 4, 0, 64, seed=11)


def test_plan_path_c1(c1):
    C, plan = c1.run_plan()
    err = rel_fro(C[:c1.M], c1.ref)
    print(f"plan path C1 rel_fro={err:.3e} items={plan.info['num_items']} hubs={plan.info['num_hub_rows']}")
    assert err <= FP32_TOL
    assert np.all(C[c1.M:] == 0)


@pytest.mark.parametrize("literal", [False, True])
def test_reference_path_c1(c1, literal):
    C = c1.run_reference_path(split=True, literal=literal)
    err = rel_fro(C[:c1.M], c1.ref)
    print(f"reference path (literal={literal}) C1 rel_fro={err:.3e}")
    assert err <= FP32_TOL
    C2 = c1.run_reference_path(split=False, literal=literal)
    assert rel_fro(C2[:c1.M], c1.ref) <= FP32_TOL


@pytest.mark.parametrize("literal", [False, True])
def test_reference_path_matches_fp32_restatement(c1, literal):
    """Both forms of the reference API against the oracle's FP32 restatement
    of ppl0_batch_SpMM_kernel on the same virtual rows/columns."""
    a = c1.part.arrays()
    tl, pl, _ = cola.nnz_split(a["csr_rowIdx_l"], a["csr_rowPtr_l"], len(a["csr_colIdx_l"]))
    tr, pr, _ = cola.nnz_split(a["csc_colIdx_r"], a["csc_colPtr_r"], len(a["csc_rowIdx_r"]))
    Cr = np.zeros((c1.P, c1.dim), np.float32)
    orc.ppl0_f32(Cr, c1.Bh, c1.dim, c1.P, c1.rank, pl, a["csr_colIdx_l"], tl, a["csr_val_l"], pr,
                 a["csc_rowIdx_r"], tr, a["csc_val_r"])
    C = c1.run_reference_path(split=True, literal=literal)
    assert rel_fro(C, Cr.astype(np.float64)) <= FP32_TOL


def test_reference_wrapper_follows_new_values_after_invalidate(c1):
    """Values written into the zips' device arrays reach the result once the
    cached plan is invalidated (2x the values: 2x the product)."""
    a = c1.part.arrays()
    tl, pl, _ = cola.nnz_split(a["csr_rowIdx_l"], a["csr_rowPtr_l"], len(a["csr_colIdx_l"]))
    tr, pr, _ = cola.nnz_split(a["csc_colIdx_r"], a["csc_colPtr_r"], len(a["csc_rowIdx_r"]))
    zl = cola.d_csr_zip(pl, a["csr_colIdx_l"], tl, a["csr_val_l"])
    zr = cola.d_csc_zip(pr, a["csc_rowIdx_r"], tr, a["csc_val_r"])
    packed = torch.empty((max(zr.n_unique, 1), c1.dim), dtype=torch.float32, device=c1.dev)
    C = torch.zeros((c1.P, c1.dim), dtype=torch.float32, device=c1.dev)
    cola.dist_spmm_ppl_batch(C, c1.B, packed, zl, zr, c1.dim, c1.P, c1.rank)
    torch.cuda.synchronize()
    assert rel_fro(C.cpu().numpy()[:c1.M], c1.ref) <= FP32_TOL
    # double every value in place through the exposed device pointers (a
    # device-to-device hipMemcpy; torch and the library share the runtime)
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")
    for ptr, vals in ((zl.d_csr_val, a["csr_val_l"]), (zr.d_csc_val, a["csc_val_r"])):
        src = torch.from_numpy(np.asarray(vals, np.float32) * 2).to(c1.dev)
        torch.cuda.synchronize()
        assert hip.hipMemcpy(ctypes.c_void_p(ptr), ctypes.c_void_p(src.data_ptr()),
                             ctypes.c_size_t(src.numel() * 4), 3) == 0
    zl.invalidate()
    C.zero_()
    cola.dist_spmm_ppl_batch(C, c1.B, packed, zl, zr, c1.dim, c1.P, c1.rank)
    torch.cuda.synchronize()
    assert rel_fro(C.cpu().numpy()[:c1.M] / 2, c1.ref) <= FP32_TOL


def test_reference_wrapper_accumulates_and_reuses_its_plan(c1):
    """The wrapper adds into C on every call (cuh:141,196), through the plan
    cached from the zips after the first call."""
    C = c1.run_reference_path(split=True, init=1.5, calls=3)
    assert rel_fro((C[:c1.M] - 1.5) / 3.0, c1.ref) <= FP32_TOL
    assert np.all(C[c1.M:] == 1.5)


def test_accumulate_semantics(c1):
    """C is accumulated into (cuh:141,196): C0 + A_p·B."""
    C, _ = c1.run_plan(flags=cola.ACCUMULATE, init=2.5)
    assert rel_fro(C[:c1.M] - 2.5, c1.ref) <= 2e-5
    assert np.all(C[c1.M:] == 2.5)


def test_plan_is_deterministic(c1):
    C1, _ = c1.run_plan()
    C2, _ = c1.run_plan()
    assert np.array_equal(C1, C2)


def test_known_answer_all_ones():
    """The reference's own check (distSpMM_ppl_batch.cu:83,99,221 + utils.cuh:44-100):
    B = 1 and all values 1 => C[i,:] = deg(i); exact in FP32 below 2^24."""
    g = cola.Graph.rmat(14, 8, seed=5, gcn=False)
    for rank in (0, 3):
        case = Case(g, 4, rank, 32, b_ones=True)
        lb = rank * case.P
        deg = np.diff(case.rp)[lb:lb + case.M].astype(np.float32)
        sag = orc.sag_unweighted(lb, lb + case.M, case.rp, case.ci, case.Bh[:case.n])
        for C in (case.run_plan()[0], case.run_reference_path(), case.run_reference_path(literal=True)):
            assert orc.check_result(C[:case.M], sag) == 0
            assert np.array_equal(C[:case.M], np.repeat(deg[:, None], 32, axis=1))


@pytest.mark.parametrize("dim", [1, 3, 8, 32, 33, 40, 64, 96, 128, 200, 256, 512])
def test_dims(dim):
    """Every dim is computed (the reference skips 65..96 and > 128)."""
    g = cola.Graph.rmat(11, 12, seed=dim)
    case = Case(g, 2, 1, dim, seed=dim)
    C, _ = case.run_plan(max_dim=max(dim, 256))
    assert rel_fro(C[:case.M], case.ref) <= FP32_TOL
    for literal in (False, True):
        Cr = case.run_reference_path(literal=literal)
        assert rel_fro(Cr[:case.M], case.ref) <= FP32_TOL, literal


@pytest.mark.parametrize("rank", [0, 1, 2, 3])
def test_every_rank_and_reverse_order(rank):
    """Ranks with peid & 2 walk remote columns in reverse (cuh:115-118)."""
    g = cola.Graph.rmat(13, 10, seed=9, reorder=True)
    case = Case(g, 4, rank, 64, seed=rank)
    if case.M == 0:
        pytest.skip("empty partition")
    assert rel_fro(case.run_plan()[0][:case.M], case.ref) <= FP32_TOL
    assert rel_fro(case.run_reference_path()[:case.M], case.ref) <= FP32_TOL
    assert rel_fro(case.run_reference_path(literal=True)[:case.M], case.ref) <= FP32_TOL


def test_identity_graph():
    """A = I: C = B_p exactly (all local, one entry per row)."""
    n = 1000
    g = cola.Graph.from_csr(np.arange(n + 1), np.arange(n), np.ones(n, np.float32))
    case = Case(g, 3, 1, 64, seed=1)
    C, _ = case.run_plan()
    assert np.array_equal(C[:case.M], case.Bh[case.P:case.P + case.M])
    assert np.array_equal(case.run_reference_path()[:case.M], case.Bh[case.P:case.P + case.M])
    assert np.array_equal(case.run_reference_path(literal=True)[:case.M], case.Bh[case.P:case.P + case.M])


def test_all_remote_partition():
    """Bipartite graph: partition 0's rows only reference the other half."""
    n = 512
    rows, cols = [], []
    rng = np.random.default_rng(0)
    for i in range(256):
        for j in rng.choice(np.arange(256, 512), 5, replace=False):
            rows += [i, j]
            cols += [j, i]
    g = cola.Graph.from_coo(n, rows, cols, rng.uniform(0.5, 1.5, len(rows)).astype(np.float32))
    case = Case(g, 2, 0, 64, seed=2)
    assert case.info["nnz_l"] == 0 and case.info["nnz_r"] > 0
    assert rel_fro(case.run_plan()[0][:case.M], case.ref) <= FP32_TOL
    assert rel_fro(case.run_reference_path()[:case.M], case.ref) <= FP32_TOL
    assert rel_fro(case.run_reference_path(literal=True)[:case.M], case.ref) <= FP32_TOL


def test_all_local_partition():
    """nGPU = 1: everything is local, no remote columns to pack."""
    g = cola.Graph.rmat(12, 8, seed=3)
    case = Case(g, 1, 0, 128, seed=3)
    assert case.info["U_r"] == 0
    assert rel_fro(case.run_plan()[0][:case.M], case.ref) <= FP32_TOL
    assert rel_fro(case.run_reference_path()[:case.M], case.ref) <= FP32_TOL
    assert rel_fro(case.run_reference_path(literal=True)[:case.M], case.ref) <= FP32_TOL


def test_empty_partition():
    """A partition without entries leaves C untouched (accumulate) / zero (overwrite)."""
    n = 100
    rows = np.arange(0, 40)
    g = cola.Graph.from_coo(n, np.concatenate([rows, rows + 1]), np.concatenate([rows + 1, rows]))
    case = Case(g, 4, 3, 32)  # rows 75..99 are isolated
    assert case.info["nnz_l"] + case.info["nnz_r"] == 0
    C, _ = case.run_plan(flags=cola.ACCUMULATE, init=1.0)
    assert np.all(C == 1.0)
    C, _ = case.run_plan(flags=cola.OVERWRITE, init=1.0)
    assert np.all(C == 0.0)
    assert np.all(case.run_reference_path() == 0.0)


@pytest.mark.parametrize("hub_threshold,chunk", [(0, 0), (64, 32), (100, 100)])
def test_hub_rows_split(hub_threshold, chunk):
    """A star (one hub row with thousands of entries) plus a ring; hub rows are
    cut into chunks and reduced in a fixed order."""
    n = 6000
    hub = np.zeros(n - 1, np.int64)
    leaves = np.arange(1, n)
    ring_a = np.arange(1, n)
    ring_b = np.concatenate([np.arange(2, n), [1]])
    rows = np.concatenate([hub, leaves, ring_a, ring_b])
    cols = np.concatenate([leaves, hub, ring_b, ring_a])
    vals = np.random.default_rng(1).uniform(-1, 1, len(rows)).astype(np.float32)
    g = cola.Graph.from_coo(n, rows, cols, vals)
    case = Case(g, 2, 0, 64, seed=4)
    C, plan = case.run_plan(hub_threshold=hub_threshold, chunk_nnz=chunk)
    assert plan.info["num_hub_rows"] >= 1
    assert rel_fro(C[:case.M], case.ref) <= FP32_TOL
    C2, _ = case.run_plan(flags=cola.ACCUMULATE | cola.KAHAN, init=0.0, hub_threshold=hub_threshold,
                          chunk_nnz=chunk, kahan_degree=16)
    assert rel_fro(C2[:case.M], case.ref) <= FP32_TOL


@pytest.mark.parametrize("tuning", [dict(col_tile=0, rpw_long=1, unroll=8), dict(col_tile=32, rpw_long=8, unroll=16),
                                    dict(col_tile=16, rpw_long=4, unroll=16, split_remote=1)])
def test_hub_rows_kahan_under_schedules(tuning):
    """Hub chunks, Kahan sums and the multi-pair staging of 16 loads in
    flight: partial rows and padding slots leave the result unchanged."""
    n = 5000
    rng = np.random.default_rng(3)
    hub = np.zeros(n - 1, np.int64)
    leaves = np.arange(1, n)
    r = rng.integers(1, n, 20000)
    c = rng.integers(1, n, 20000)
    rows = np.concatenate([hub, leaves, r, c])
    cols = np.concatenate([leaves, hub, c, r])
    vals = rng.uniform(-1, 1, len(rows)).astype(np.float32)
    g = cola.Graph.from_coo(n, rows, cols, vals)
    case = Case(g, 2, 0, 64, seed=5)
    plan = cola.Plan(case.part, hub_threshold=300, chunk_nnz=96, kahan_degree=8)
    assert plan.info["num_hub_rows"] >= 1
    full = dict(col_tile=0, rpw_short=0, rpw_long=1, blocks_per_cu=8, xcd_remap=0, split_remote=0, unroll=8, panels=0)
    full.update(tuning)
    plan.tuning = full
    packed = torch.empty((max(plan.U_r, 1), case.dim), dtype=torch.float32, device=case.dev)
    plan.pack(case.B, packed)
    for flags in (cola.OVERWRITE, cola.OVERWRITE | cola.KAHAN):
        C = torch.full((case.P, case.dim), 2.0, dtype=torch.float32, device=case.dev)
        plan.spmm(C, case.B_local(), packed, flags=flags)
        torch.cuda.synchronize()
        assert rel_fro(C.cpu().numpy()[:case.M], case.ref) <= FP32_TOL, flags


@pytest.mark.parametrize("row_order", [0, 1])
def test_panels_with_hub_rows_below_the_short_degree(row_order):
    """Column panels with hub_threshold < short_degree and with degree-ordered
    items: the panel pass must chunk exactly the rows the plan's hub table
    lists, in the same order (they share the hub reduction)."""
    g = cola.Graph.rmat(11, 12, seed=21)
    case = Case(g, 2, 0, 64, seed=6)
    plan = cola.Plan(case.part, hub_threshold=8, chunk_nnz=4, short_degree=24, col_panels=3, row_order=row_order)
    assert plan.info["num_hub_rows"] > 0
    plan.tuning = dict(col_tile=32, rpw_short=0, rpw_long=2, blocks_per_cu=4, xcd_remap=1, split_remote=0,
                       unroll=8, panels=1)
    packed = torch.empty((max(plan.U_r, 1), case.dim), dtype=torch.float32, device=case.dev)
    plan.pack(case.B, packed)
    for flags, init in ((cola.OVERWRITE, 3.0), (cola.ACCUMULATE, 0.5)):
        C = torch.full((case.P, case.dim), init, dtype=torch.float32, device=case.dev)
        plan.spmm(C, case.B_local(), packed, flags=flags)
        torch.cuda.synchronize()
        got = C.cpu().numpy()[:case.M] - (0.0 if flags == cola.OVERWRITE else init)
        assert rel_fro(got, case.ref) <= FP32_TOL, flags


@pytest.mark.parametrize("hot_cols", [1, 300, 1 << 30])
@pytest.mark.parametrize("tuning", [dict(col_tile=0, rpw_long=1, unroll=8), dict(col_tile=32, rpw_long=4, unroll=16),
                                    dict(col_tile=16, rpw_short=8, rpw_long=8, unroll=16, split_remote=1)])
def test_hot_cold_passes(c1, hot_cols, tuning):
    """Two-pass plans (hot entries write C, cold entries add onto it; hub
    chunks of both passes reduced together) give the one-pass result under
    every flag; split_remote is ignored for them."""
    plan = cola.Plan(c1.part, hot_cols=hot_cols, hub_threshold=200, chunk_nnz=64, kahan_degree=100, row_order=1)
    info = plan.info
    assert 0 < info["hot_nnz"] <= info["nnz"]
    assert (info["num_cold_items"] > 0) == (info["hot_nnz"] < info["nnz"])
    full = dict(col_tile=0, rpw_short=0, rpw_long=1, blocks_per_cu=4, xcd_remap=0, split_remote=0, unroll=8, panels=0)
    full.update(tuning)
    plan.tuning = full
    packed = torch.empty((plan.U_r, c1.dim), dtype=torch.float32, device=c1.dev)
    plan.pack(c1.B, packed)
    for flags, init in ((cola.OVERWRITE, 7.0), (cola.ACCUMULATE, 0.5), (cola.OVERWRITE | cola.KAHAN, 7.0)):
        C = torch.full((c1.P, c1.dim), init, dtype=torch.float32, device=c1.dev)
        plan.spmm(C, c1.B_local(), packed, flags=flags)
        torch.cuda.synchronize()
        got = C.cpu().numpy()
        assert rel_fro(got[:c1.M] - (init if flags == cola.ACCUMULATE else 0.0), c1.ref) <= FP32_TOL, flags
        assert np.all(got[c1.M:] == init)


def test_hot_cold_plans_are_not_batched(c1):
    plan = cola.Plan(c1.part, hot_cols=100)
    C = torch.zeros((c1.P, c1.dim), dtype=torch.float32, device=c1.dev)
    packed = torch.empty((plan.U_r, c1.dim), dtype=torch.float32, device=c1.dev)
    with pytest.raises(cola.ColaError) as e:
        cola.Batch([dict(plan=plan, B=c1.B_local(), packed=packed, C=C, dim=c1.dim)], dtype=cola.F32)
    assert e.value.code == cola._lib.COLA_ERR_UNSUPPORTED
    with pytest.raises(cola.ColaError):
        cola.Plan(c1.part, hot_cols=100, col_panels=2)


def test_tuning_ranges(c1):
    """col_tile -k means k x 128 bytes for k = 1..8; anything further out is
    rejected with a status, as are out-of-range knobs."""
    plan = cola.Plan(c1.part)
    plan.tuning = dict(col_tile=-8)
    assert plan.tuning["col_tile"] == -8
    for bad in (dict(col_tile=-9), dict(rpw_long=65), dict(unroll=12), dict(xcd_remap=2), dict(blocks_per_cu_short=65)):
        with pytest.raises(cola.ColaError) as e:
            plan.tuning = bad
        assert e.value.code == cola._lib.COLA_ERR_INVALID, bad


def test_linearity_in_B(c1):
    """A·(B1 + B2) = A·B1 + A·B2 to FP32 rounding."""
    plan = cola.Plan(c1.part)
    dev = c1.dev
    B1 = c1.B
    B2 = torch.flip(c1.B, dims=[1]).contiguous()
    outs = []
    for B in (B1, B2, B1 + B2):
        packed = torch.empty((plan.U_r, c1.dim), dtype=torch.float32, device=dev)
        C = torch.zeros((c1.P, c1.dim), dtype=torch.float32, device=dev)
        plan.pack(B, packed)
        plan.spmm(C, B[:c1.P], packed, flags=cola.OVERWRITE)
        outs.append(C.double())
    torch.cuda.synchronize()
    d = (outs[0] + outs[1] - outs[2]).norm() / outs[2].norm()
    assert d.item() <= 1e-5


def test_pack_is_bitexact(c1):
    plan = cola.Plan(c1.part)
    packed = torch.empty((plan.U_r, c1.dim), dtype=torch.float32, device=c1.dev)
    plan.pack(c1.B, packed)
    a = c1.part.arrays()
    assert np.array_equal(packed.cpu().numpy(), c1.Bh[a["csc_colIdx_r"]])
    # unaligned / element-sized path: a dim that is not a multiple of 4
    B3 = c1.B[:, :3].contiguous()
    p3 = torch.empty((plan.U_r, 3), dtype=torch.float32, device=c1.dev)
    plan.pack(B3, p3, dim=3)
    assert np.array_equal(p3.cpu().numpy(), c1.Bh[a["csc_colIdx_r"], :3])


def test_quantize_is_bitexact():
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, (64, 96)).astype(np.float32)
    x[0, :8] = [0.0, -0.0, 1e-30, -1e-30, 3.0e38, 1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-8]
    x[1] = 0.0
    xt = torch.from_numpy(x).cuda()
    bf, _ = cola.quantize_rows(xt, cola.BF16)
    assert np.array_equal(bf.cpu().numpy().view(np.uint16), orc.bf16_from_f32(x))
    f8, sc = cola.quantize_rows(xt, cola.FP8_E4M3)
    sc = sc.cpu().numpy()
    exp_sc = np.array([orc.fp8_row_scale(r) for r in x], np.float32)
    assert np.array_equal(sc, exp_sc)
    exp_codes = orc.fp8_from_f32((x / exp_sc[:, None]).astype(np.float32))
    assert np.array_equal(f8.cpu().numpy(), exp_codes)
    h, _ = cola.quantize_rows(xt, cola.F16)
    with np.errstate(over="ignore"):  # 3.0e38 overflows FP16 to inf on both sides
        want = x.astype(np.float16)
    assert np.array_equal(h.cpu().numpy().view(np.float16), want)


# Error bounds of the reduced-precision storage forms against the FP64 oracle
# (relative Frobenius): unit roundoff 2^-9 (bf16), 2^-12 (fp16), 2^-4 (e4m3)
# per stored element, averaged down by the row sums.
PREC_TOL = {cola.BF16: 4e-3, cola.F16: 5e-4, cola.FP8_E4M3: 5e-2}


@pytest.mark.parametrize("dtype", [cola.BF16, cola.F16, cola.FP8_E4M3])
@pytest.mark.parametrize("dim", [64, 128, 256])
def test_precision_forms(c1, dtype, dim):
    rng = np.random.default_rng(dim)
    Bh = rng.uniform(-1, 1, (c1.P * c1.nGPU, dim)).astype(np.float32)
    ref = orc.spmm_f64(c1.rp, c1.ci, c1.va, 0, c1.M, Bh[:c1.n])
    B = torch.from_numpy(Bh).cuda()
    Bq, sc = cola.quantize_rows(B, dtype)
    plan = cola.Plan(c1.part)
    eb = cola.DTYPE_BYTES[dtype]
    tdt = {1: torch.uint8, 2: torch.uint16}[eb]
    packed = torch.empty((plan.U_r, dim), dtype=tdt, device=B.device)
    psc = torch.empty(plan.U_r, dtype=torch.float32, device=B.device) if dtype == cola.FP8_E4M3 else None
    plan.pack(Bq, packed, dim=dim, dtype=dtype, scales_global=sc, packed_scales=psc)
    C = torch.zeros((c1.P, dim), dtype=torch.float32, device=B.device)
    plan.spmm(C, Bq[:c1.P], packed, dim=dim, dtype=dtype, flags=cola.OVERWRITE | cola.KAHAN,
              scales_local=None if sc is None else sc[:c1.P], packed_scales=psc)
    torch.cuda.synchronize()
    err = rel_fro(C.cpu().numpy()[:c1.M], ref)
    # Also: the device result equals the FP64 product of the dequantised B.
    if dtype == cola.FP8_E4M3:
        deq = orc.fp8_to_f32(Bq.cpu().numpy()) * sc.cpu().numpy()[:, None]
    elif dtype == cola.BF16:
        deq = (Bq.cpu().numpy().astype(np.uint32) << 16).view(np.float32)
    else:
        deq = Bq.cpu().numpy().view(np.float16).astype(np.float32)
    ref_q = orc.spmm_f64(c1.rp, c1.ci, c1.va, 0, c1.M, deq[:c1.n])
    err_q = rel_fro(C.cpu().numpy()[:c1.M], ref_q)
    print(f"dtype={cola.DTYPE_NAMES[dtype]} dim={dim} rel_fro_vs_fp64={err:.3e} vs_dequantised={err_q:.3e}")
    assert err <= PREC_TOL[dtype]
    assert err_q <= FP32_TOL


def test_python_wrapper_rejects_buffers_the_kernels_would_overrun(c1):
    """The C ABI takes raw pointers; the Python mirror checks every buffer it
    hands over (rows, used columns, element size) and raises before any
    launch."""
    plan = cola.Plan(c1.part)
    dim, dev = c1.dim, c1.dev
    packed = torch.empty((plan.U_r, dim), dtype=torch.float32, device=dev)
    C = torch.zeros((c1.P, dim), dtype=torch.float32, device=dev)
    Bl = c1.B_local()
    bad_calls = [
        dict(C=C[: c1.M - 1]),                                              # too few output rows
        dict(C=torch.zeros((c1.P, dim), dtype=torch.float16, device=dev)),  # C must be FP32
        dict(B_local=Bl[: c1.M - 1]),                                       # too few B rows
        dict(packed=packed[: plan.U_r - 1]),                                # too few packed rows
        dict(packed=packed[:, : dim - 1], dim=dim),                         # too few used columns
        dict(dtype=cola.BF16),                                              # FP32 tensors, BF16 storage
        dict(B_local=Bl.t()),                                               # column-major
    ]
    for bad in bad_calls:
        args = dict(C=C, B_local=Bl, packed=packed, dim=dim, dtype=cola.F32)
        args.update(bad)
        with pytest.raises(ValueError):
            plan.spmm(args["C"], args["B_local"], args["packed"], dim=args["dim"], dtype=args["dtype"])
    with pytest.raises(ValueError):
        plan.pack(c1.B, packed[: plan.U_r - 1])
    plan.pack(c1.B, packed)
    plan.spmm(C, Bl, packed, flags=cola.OVERWRITE)
    torch.cuda.synchronize()
    assert rel_fro(C.cpu().numpy()[: c1.M], c1.ref) <= FP32_TOL


def test_two_host_threads_with_their_own_plans_and_streams(c1):
    """Handles are re-entrant per handle: two host threads, each with its own
    plan, buffers and stream, call pack + SpMM concurrently (ctypes drops the
    GIL around each call); both keep matching the FP64 oracle, and a bad call
    on one thread leaves its own error text without disturbing the other."""
    import threading
    errors = []

    def worker(k):
        try:
            torch.cuda.set_device(0)
            s = torch.cuda.Stream()
            plan = cola.Plan(c1.part, row_order=k)
            packed = torch.empty((plan.U_r, c1.dim), dtype=torch.float32, device=c1.dev)
            C = torch.empty((c1.P, c1.dim), dtype=torch.float32, device=c1.dev)
            with torch.cuda.stream(s):
                for it in range(20):
                    plan.pack(c1.B, packed, stream=s)
                    plan.spmm(C, c1.B_local(), packed, flags=cola.OVERWRITE, stream=s)
                    if k == 1 and it == 10:
                        try:
                            plan.spmm(C, c1.B_local(), packed, dim=c1.dim, dtype=cola.F32, flags=64, stream=s)
                        except cola.ColaError as e:
                            assert "unknown flag" in str(e)
                        else:
                            raise AssertionError("a bad flag was accepted")
                s.synchronize()
            err = rel_fro(C.cpu().numpy()[: c1.M], c1.ref)
            if err > FP32_TOL:
                errors.append((k, err))
        except Exception as e:  # reported to the main thread
            errors.append((k, repr(e)))

    threads = [threading.Thread(target=worker, args=(k,)) for k in (0, 1)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors


def test_error_norms_on_device(c1):
    C, _ = c1.run_plan()
    Ct = torch.from_numpy(C[:c1.M]).cuda()
    Rt = torch.from_numpy(c1.ref).cuda()
    out = cola.error_norms(Ct, Rt).cpu().numpy()
    d = C[:c1.M].astype(np.float64) - c1.ref
    assert np.isclose(out[0], (d * d).sum(), rtol=1e-10, atol=1e-300)
    assert np.isclose(out[1], (c1.ref ** 2).sum(), rtol=1e-12)
    rn = (c1.ref ** 2).sum(1)
    nz = rn > 0
    assert np.isclose(out[2], ((d * d).sum(1)[nz] / rn[nz]).max(), rtol=1e-10)


@pytest.fixture(scope="module", params=["cm", "community"])
def full_size(request):
    """The 2^22-node R-MAT graph (degree 32, seed 2) in Cuthill-McKee or
    community order, partition 0 of 8, B of 2^22 x 128 FP32 uniform[-1,1]."""
    order = True if request.param == "cm" else "community"
    g = cola.Graph.rmat(22, 32, seed=2, reorder=order)
    part = g.partition(8, 0)
    gen = torch.Generator(device="cuda").manual_seed(5)
    P = part.info()["nodesPerPE"]
    B = torch.rand((P * 8, 128), generator=gen, device="cuda", dtype=torch.float32) * 2 - 1
    return dict(order=request.param, g=g, part=part, B=B)


# The large-graph default schedule and the benchmarked one (256-byte tiles,
# 16 blocks per CU on the short-row bin).
FULL_SIZE_TUNINGS = [None, dict(col_tile=64, rpw_short=0, rpw_long=4, blocks_per_cu=2, xcd_remap=1,
                                split_remote=0, unroll=8, panels=0, blocks_per_cu_short=16)]


@pytest.mark.parametrize("tuning", FULL_SIZE_TUNINGS, ids=["default", "bench"])
def test_full_size_partition_checksum(full_size, tuning):
    """Partition 0 of 8 of the 2^22-node graph at dim 128: row sums of C equal
    A_p·(B·1) from the FP64 oracle (a size-independent checksum), a sample of
    rows matches the FP64 product, and doubling B doubles C exactly (every
    product and sum scales by a power of two)."""
    g, part, B = full_size["g"], full_size["part"], full_size["B"]
    rp, ci, va = g.arrays()
    info = part.info()
    P, M, dim = info["nodesPerPE"], info["num_rows"], 128
    print(f"case full size order={full_size['order']} nnz={info['nnz_l'] + info['nnz_r']} tuning={tuning}",
          flush=True)
    plan = cola.Plan(part, row_order=1)
    if tuning is not None:
        plan.tuning = tuning
    packed = torch.empty((plan.U_r, dim), dtype=torch.float32, device="cuda")
    C = torch.zeros((P, dim), dtype=torch.float32, device="cuda")
    plan.pack(B, packed)
    plan.spmm(C, B[:P], packed, flags=cola.OVERWRITE)
    B2 = B * 2
    packed2 = torch.empty_like(packed)
    C2 = torch.zeros_like(C)
    plan.pack(B2, packed2)
    plan.spmm(C2, B2[:P], packed2, flags=cola.OVERWRITE)
    torch.cuda.synchronize()
    assert torch.equal(C2, C * 2)
    del B2, packed2, C2
    rowsum_B = B.double().sum(1).cpu().numpy()
    rowsum_C = C.double().sum(1).cpu().numpy()[:M]
    ref = orc.spmm_f64(rp, ci, va, 0, M, rowsum_B.astype(np.float32)[:, None])[:, 0]
    assert np.linalg.norm(rowsum_C - ref) / np.linalg.norm(ref) <= 1e-5
    sample = np.unique(np.concatenate([np.arange(0, M, 997), [0, 1, M - 1]]))
    Bh = B.cpu().numpy()
    Ch = C.cpu().numpy()
    refs = np.stack([orc.spmm_f64(rp, ci, va, int(i), int(i) + 1, Bh[:g.n])[0] for i in sample])
    assert rel_fro(Ch[sample], refs) <= FP32_TOL


def test_full_size_reference_entry_points(full_size):
    """The reference's entry points at full size (zips through nnz_split, as
    the shipped driver builds them): the single-stage wrapper rebuilds the
    default decreasing-degree plan from the zips, so it equals the plan path
    bit for bit; the literal virtual-row kernel with FP32 atomics matches the
    FP64 oracle on sampled rows."""
    g, part, B = full_size["g"], full_size["part"], full_size["B"]
    rp, ci, va = g.arrays()
    info = part.info()
    P, M, dim = info["nodesPerPE"], info["num_rows"], 128
    print(f"case full size reference entry points order={full_size['order']}", flush=True)
    a = part.arrays()
    tl, pl, _ = cola.nnz_split(a["csr_rowIdx_l"], a["csr_rowPtr_l"], len(a["csr_colIdx_l"]))
    tr, pr, _ = cola.nnz_split(a["csc_colIdx_r"], a["csc_colPtr_r"], len(a["csc_rowIdx_r"]))
    zl = cola.d_csr_zip(pl, a["csr_colIdx_l"], tl, a["csr_val_l"])
    zr = cola.d_csc_zip(pr, a["csc_rowIdx_r"], tr, a["csc_val_r"])
    packed = torch.empty((max(zr.n_unique, 1), dim), dtype=torch.float32, device="cuda")
    C_ref_api = torch.zeros((P, dim), dtype=torch.float32, device="cuda")
    cola.dist_spmm_ppl_batch(C_ref_api, B, packed, zl, zr, dim, P, 0)
    plan = cola.Plan(part, row_order=1)
    packed_p = torch.empty((plan.U_r, dim), dtype=torch.float32, device="cuda")
    C_plan = torch.zeros((P, dim), dtype=torch.float32, device="cuda")
    plan.pack(B, packed_p)
    plan.spmm(C_plan, B[:P], packed_p, flags=cola.OVERWRITE)
    torch.cuda.synchronize()
    assert torch.equal(C_ref_api, C_plan)
    del C_plan, packed_p
    C_lit = torch.zeros((P, dim), dtype=torch.float32, device="cuda")
    uniq = torch.from_numpy(np.unique(np.asarray(tr, np.int32))).cuda()
    cola.pack_remote(B, uniq, packed, num_cols=zr.n_unique)
    cola.ppl0_batch_SpMM(C_lit, B[:P], packed, zl, zr, 0, 0, zl.num, zr.num, dim, P, 0)
    torch.cuda.synchronize()
    sample = np.unique(np.concatenate([np.arange(0, M, 997), [0, 1, M - 1]]))
    Bh = B.cpu().numpy()
    refs = np.stack([orc.spmm_f64(rp, ci, va, int(i), int(i) + 1, Bh[:g.n])[0] for i in sample])
    assert rel_fro(C_lit.cpu().numpy()[sample], refs) <= FP32_TOL


def sampled_rows_csr(rp, ci, va, rows):
    """CSR of the given rows alone, columns renumbered over the B rows they
    reference (returned as `cols`), so an FP64 product over them needs only
    those B rows."""
    rp = np.asarray(rp, np.int64)
    idx = np.concatenate([np.arange(rp[i], rp[i + 1]) for i in rows])
    cols, inv = np.unique(np.asarray(ci)[idx], return_inverse=True)
    lens = np.array([rp[i + 1] - rp[i] for i in rows], np.int64)
    rp2 = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    return rp2, inv.astype(np.int32), np.asarray(va, np.float32)[idx], cols


@pytest.mark.parametrize("schedule", ["default", "tuned"])
@pytest.mark.parametrize("dim", [64, 128, 256])
@pytest.mark.parametrize("dtype", [cola.BF16, cola.F16, cola.FP8_E4M3], ids=["bf16", "f16", "fp8"])
def test_full_size_precision_forms(full_size, dtype, dim, schedule):
    """B stored in BF16 / FP16 / FP8-e4m3 (+ row scales) on the 2^22-node
    partition at each feature dim of the precision configuration (64, 128,
    256), FP32 accumulation with Kahan on rows above 1024 entries: on 527
    sampled rows the result equals the FP64 product of the dequantised B to
    FP32 accuracy (1e-5), and is within the storage format's error of the
    FP64 product of the FP32 B.  Each at the plan's default schedule and at
    a tuned one: the bench's committed table on the community order, the
    plan autotuner's pick on the Cuthill-McKee order."""
    g, part = full_size["g"], full_size["part"]
    rp, ci, va = g.arrays()
    info = part.info()
    P, M = info["nodesPerPE"], info["num_rows"]
    if dim == 128:
        B = full_size["B"]
    else:
        gen = torch.Generator(device="cuda").manual_seed(50 + dim)
        B = torch.rand((P * 8, dim), generator=gen, device="cuda", dtype=torch.float32) * 2 - 1
    print(f"case full size precision order={full_size['order']} dtype={cola.DTYPE_NAMES[dtype]} dim={dim}",
          flush=True)
    Bq, sc = cola.quantize_rows(B, dtype)
    plan = cola.Plan(part, row_order=1)
    tuned = schedule == "tuned" and full_size["order"] == "community"
    if tuned:
        # The schedule bench.py's precision sweep runs this form at (its
        # committed table, made on the community order).
        with open(os.path.join(ROOT, "profiles", "plan_rows_precision_rmat22_tuning.json")) as f:
            plan.tuning = json.load(f)["tunings"][str(dim)][cola.DTYPE_NAMES[dtype]]
    tdt = {1: torch.uint8, 2: torch.uint16}[cola.DTYPE_BYTES[dtype]]
    packed = torch.empty((plan.U_r, dim), dtype=tdt, device="cuda")
    psc = torch.empty(plan.U_r, dtype=torch.float32, device="cuda") if dtype == cola.FP8_E4M3 else None
    plan.pack(Bq, packed, dim=dim, dtype=dtype, scales_global=sc, packed_scales=psc)
    C = torch.zeros((P, dim), dtype=torch.float32, device="cuda")
    if schedule == "tuned" and not tuned:
        # Cuthill-McKee order: the plan autotuner's pick on these buffers.
        plan.autotune(C, Bq[:P], packed, dim=dim, dtype=dtype, flags=cola.OVERWRITE | cola.KAHAN, iters=1,
                      scales_local=None if sc is None else sc[:P], packed_scales=psc)
        C.zero_()
    plan.spmm(C, Bq[:P], packed, dim=dim, dtype=dtype, flags=cola.OVERWRITE | cola.KAHAN,
              scales_local=None if sc is None else sc[:P], packed_scales=psc)
    torch.cuda.synchronize()
    sample = np.unique(np.concatenate([np.arange(0, M, 997), [0, 1, M - 1]]))
    rp2, ci2, va2, cols = sampled_rows_csr(rp, ci, va, sample)
    cols_t = torch.from_numpy(cols.astype(np.int64)).cuda()
    Bf = B[cols_t].cpu().numpy()
    if Bq.dtype == torch.uint16:  # torch gathers no uint16 on the GPU: go through int16
        q = Bq.view(torch.int16)[cols_t].cpu().numpy().view(np.uint16)
    else:
        q = Bq[cols_t].cpu().numpy()
    if dtype == cola.FP8_E4M3:
        deq = orc.fp8_to_f32(q) * sc[cols_t].cpu().numpy()[:, None]
    elif dtype == cola.BF16:
        deq = (q.astype(np.uint32) << 16).view(np.float32)
    else:
        deq = q.view(np.float16).astype(np.float32)
    got = C.cpu().numpy()[sample]
    ref = orc.spmm_f64(rp2, ci2, va2, 0, len(sample), Bf)
    ref_q = orc.spmm_f64(rp2, ci2, va2, 0, len(sample), deq.astype(np.float32))
    err, err_q = rel_fro(got, ref), rel_fro(got, ref_q)
    print(f"rel_fro_vs_fp64={err:.3e} vs_dequantised={err_q:.3e}", flush=True)
    assert err_q <= FP32_TOL
    assert err <= PREC_TOL[dtype]



@pytest.mark.parametrize("dim", [64, 256])
def test_full_size_f32_other_dims(full_size, dim):
    """FP32 B at the precision configuration's other dims (64, 256) on the
    2^22-node partition, at the schedule bench.py's precision sweep uses for
    it (the committed table, community order) or the default schedule
    (Cuthill-McKee order): 527 sampled rows within 1e-5 of the FP64 product."""
    g, part = full_size["g"], full_size["part"]
    rp, ci, va = g.arrays()
    info = part.info()
    P, M = info["nodesPerPE"], info["num_rows"]
    gen = torch.Generator(device="cuda").manual_seed(70 + dim)
    B = torch.rand((P * 8, dim), generator=gen, device="cuda", dtype=torch.float32) * 2 - 1
    plan = cola.Plan(part, row_order=1)
    if full_size["order"] == "community":
        with open(os.path.join(ROOT, "profiles", "plan_rows_precision_rmat22_tuning.json")) as f:
            plan.tuning = json.load(f)["tunings"][str(dim)]["f32"]
    print(f"case full size f32 order={full_size['order']} dim={dim} tuning={plan.tuning}", flush=True)
    packed = torch.empty((plan.U_r, dim), dtype=torch.float32, device="cuda")
    C = torch.full((P, dim), 3.0, dtype=torch.float32, device="cuda")
    plan.pack(B, packed)
    plan.spmm(C, B[:P], packed, flags=cola.OVERWRITE)
    torch.cuda.synchronize()
    sample = np.unique(np.concatenate([np.arange(0, M, 997), [0, 1, M - 1]]))
    rp2, ci2, va2, cols = sampled_rows_csr(rp, ci, va, sample)
    Bf = B[torch.from_numpy(cols.astype(np.int64)).cuda()].cpu().numpy()
    ref = orc.spmm_f64(rp2, ci2, va2, 0, len(sample), Bf)
    assert rel_fro(C.cpu().numpy()[sample], ref) <= FP32_TOL

TUNINGS = [dict(col_tile=0, rpw_short=0, rpw_long=1, blocks_per_cu=8, xcd_remap=1, split_remote=0, unroll=8, panels=0),
           dict(col_tile=32, rpw_short=1, rpw_long=2, blocks_per_cu=2, xcd_remap=0, split_remote=1),
           dict(col_tile=16, rpw_short=2, rpw_long=4, blocks_per_cu=16, xcd_remap=1, split_remote=0),
           dict(col_tile=48, rpw_short=64, rpw_long=64, blocks_per_cu=1, xcd_remap=0, split_remote=1),
           dict(col_tile=0, rpw_short=4, rpw_long=1, blocks_per_cu=4, xcd_remap=1, split_remote=1, unroll=16),
           dict(col_tile=32, rpw_short=8, rpw_long=2, blocks_per_cu=8, xcd_remap=1, split_remote=0, unroll=16),
           dict(col_tile=0, rpw_short=0, rpw_long=1, blocks_per_cu=8, xcd_remap=1, split_remote=0, unroll=8, panels=1),
           dict(col_tile=32, rpw_short=0, rpw_long=2, blocks_per_cu=4, xcd_remap=0, split_remote=0, unroll=8,
                panels=1),
           # 16 loads in flight on groups narrower than 16 lanes: each lane
           # stages 2, 4 or 8 entries per chunk
           dict(col_tile=32, rpw_short=8, rpw_long=8, blocks_per_cu=2, xcd_remap=0, split_remote=0, unroll=16),
           # 128-byte tiles, 1 block per CU (the Cuthill-McKee partition's best)
           dict(col_tile=-1, rpw_short=0, rpw_long=8, blocks_per_cu=1, xcd_remap=0, split_remote=0, unroll=16),
           # the large-graph default: 256-byte tiles, XCD remap, short rows on a larger grid
           dict(col_tile=-2, rpw_short=0, rpw_long=4, blocks_per_cu=2, xcd_remap=1, split_remote=0, unroll=8,
                blocks_per_cu_short=4),
           dict(col_tile=0, rpw_short=2, rpw_long=1, blocks_per_cu=1, xcd_remap=0, split_remote=1, unroll=16,
                blocks_per_cu_short=16),
           dict(col_tile=16, rpw_short=16, rpw_long=8, blocks_per_cu=3, xcd_remap=1, split_remote=1, unroll=16),
           dict(col_tile=8, rpw_short=32, rpw_long=4, blocks_per_cu=2, xcd_remap=0, split_remote=0, unroll=16)]


@pytest.mark.parametrize("tuning", TUNINGS)
@pytest.mark.parametrize("dtype", [cola.F32, cola.BF16, cola.FP8_E4M3])
def test_tunings_give_the_same_result(c1, tuning, dtype):
    """Column tiles, rows per wave, grid size and XCD remap change the schedule,
    never the result: the same FP32-accumulated sums up to rounding order."""
    dim = 64
    plan = cola.Plan(c1.part, short_degree=24, hub_threshold=256, chunk_nnz=128, col_panels=4)
    B = c1.B
    Bq, sc = cola.quantize_rows(B, dtype)
    tdt = {cola.F32: torch.float32, cola.BF16: torch.uint16, cola.FP8_E4M3: torch.uint8}[dtype]
    packed = torch.empty((plan.U_r, dim), dtype=tdt, device=B.device)
    psc = torch.empty(plan.U_r, dtype=torch.float32, device=B.device) if dtype == cola.FP8_E4M3 else None
    plan.pack(Bq, packed, dtype=dtype, scales_global=sc, packed_scales=psc)
    outs = []
    for t in (TUNINGS[0], tuning):
        plan.tuning = t
        C = torch.full((c1.P, dim), 3.0, dtype=torch.float32, device=B.device)
        plan.spmm(C, Bq[:c1.P], packed, dim=dim, dtype=dtype, flags=cola.OVERWRITE,
                  scales_local=None if sc is None else sc[:c1.P], packed_scales=psc)
        outs.append(C.cpu().numpy())
    # accumulate mode on top of a non-zero C
    C = torch.full((c1.P, dim), 3.0, dtype=torch.float32, device=B.device)
    plan.spmm(C, Bq[:c1.P], packed, dim=dim, dtype=dtype, flags=cola.ACCUMULATE,
              scales_local=None if sc is None else sc[:c1.P], packed_scales=psc)
    acc = C.cpu().numpy()
    torch.cuda.synchronize()
    assert rel_fro(outs[1], outs[0].astype(np.float64)) <= 1e-6
    assert rel_fro(acc - 3.0, outs[0].astype(np.float64)) <= 2e-6
    if dtype == cola.F32:
        assert rel_fro(outs[1][:c1.M], c1.ref) <= FP32_TOL


def test_autotune_keeps_a_correct_schedule(c1):
    plan = cola.Plan(c1.part)
    packed = torch.empty((plan.U_r, c1.dim), dtype=torch.float32, device=c1.dev)
    plan.pack(c1.B, packed)
    C = torch.zeros((c1.P, c1.dim), dtype=torch.float32, device=c1.dev)
    best, ms = plan.autotune(C, c1.B_local(), packed, iters=2)
    assert ms > 0 and plan.tuning == best
    assert rel_fro(C.cpu().numpy()[:c1.M], c1.ref) <= FP32_TOL
    C2 = torch.zeros_like(C)
    plan.spmm(C2, c1.B_local(), packed, flags=cola.OVERWRITE)
    torch.cuda.synchronize()
    assert rel_fro(C2.cpu().numpy()[:c1.M], c1.ref) <= FP32_TOL
