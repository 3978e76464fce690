"""cola_spmm_partition_noalloc (hub partials in the library's fixed device
array) against cola_spmm_partition (the plan's own workspace), called
through the C ABI: the same kernels in the same order, so results are
bitwise equal.  Two calls in a row, the second on a smaller partition and
dim; hub-heavy plans, the column-panel schedule, BF16 storage with
compensated sums, a dim above the plan's max_dim, a plan too large for the
array (rejected, C untouched), and the benchmarked 2^22-node partition."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import cola_amd as cola  # noqa: E402
from cola_amd import _lib  # noqa: E402
from oracle import oracle as orc  # noqa: E402

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a GPU")]

DEV = torch.device("cuda:0")
SCRATCH_BYTES = 16 << 20


def vp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class Problem:
    def __init__(self, log_n, deg, nGPU, rank, dim, dtype=cola.F32, seed=0, **plan_opts):
        self.g = cola.Graph.rmat(log_n, deg, seed=seed, reorder=True)
        self.part = self.g.partition(nGPU, rank)
        info = self.part.info()
        self.P, self.M, self.rank, self.dim, self.dtype = info["nodesPerPE"], info["num_rows"], rank, dim, dtype
        self.plan = cola.Plan(self.part, **plan_opts)
        gen = torch.Generator(device=DEV).manual_seed(seed)
        self.Bf = torch.rand((self.P * nGPU, dim), generator=gen, device=DEV).mul_(2).sub_(1)
        self.B = self.Bf if dtype == cola.F32 else cola.quantize_rows(self.Bf, dtype)[0]
        tdt = torch.float32 if dtype == cola.F32 else torch.uint16
        self.packed = torch.empty((max(self.plan.U_r, 1), dim), dtype=tdt, device=DEV)
        self.plan.pack(self.B, self.packed, dtype=dtype)
        self.B_local = self.B[rank * self.P:(rank + 1) * self.P]

    def call(self, name, C, flags):
        fn = getattr(cola.lib(), name)
        return fn(self.plan.handle, vp(C), C.stride(0), vp(self.B_local), self.B_local.stride(0), None,
                  vp(self.packed), self.packed.stride(0), None, self.dim, self.dtype, flags,
                  ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))

    def both(self, flags):
        C_ref = torch.full((self.P, self.dim), 3.0, dtype=torch.float32, device=DEV)
        C = C_ref.clone()
        assert self.call("cola_spmm_partition", C_ref, flags) == 0, cola.lib().cola_last_error()
        assert self.call("cola_spmm_partition_noalloc", C, flags) == 0, cola.lib().cola_last_error()
        torch.cuda.synchronize()
        return C.cpu().numpy(), C_ref.cpu().numpy()


def test_noalloc_equals_plan_workspace_two_calls():
    big = Problem(15, 24, 2, 0, 128, hub_threshold=64, chunk_nnz=32, seed=1)
    assert big.plan.info["num_hub_chunks"] > 100
    for flags in (cola.OVERWRITE, cola.ACCUMULATE, cola.OVERWRITE | cola.KAHAN):
        C, C_ref = big.both(flags)
        assert np.array_equal(C, C_ref)
    # second call: a smaller partition and dim (the array keeps the first
    # call's partials; every partial is written before it is read)
    small = Problem(12, 10, 4, 1, 32, hub_threshold=40, chunk_nnz=16, seed=2)
    assert 0 < small.plan.info["num_hub_chunks"] < big.plan.info["num_hub_chunks"]
    C, C_ref = small.both(cola.OVERWRITE)
    assert np.array_equal(C, C_ref)
    # and against the FP64 oracle
    rp, ci, va = small.g.arrays()
    lb = small.rank * small.P
    ref = orc.spmm_f64(rp, ci, va, lb, lb + small.M, small.Bf.cpu().numpy()[:small.g.n])
    assert np.linalg.norm(C[:small.M] - ref) / np.linalg.norm(ref) <= 1e-5


def test_noalloc_bf16_kahan_and_panels():
    p = Problem(14, 20, 2, 1, 64, dtype=cola.BF16, hub_threshold=48, chunk_nnz=16, kahan_degree=40, seed=3)
    C, C_ref = p.both(cola.OVERWRITE | cola.KAHAN)
    assert np.array_equal(C, C_ref)
    q = Problem(13, 16, 2, 0, 64, hub_threshold=48, chunk_nnz=16, col_panels=3, seed=4)
    q.plan.tuning = dict(panels=1)
    C, C_ref = q.both(cola.OVERWRITE)
    assert np.array_equal(C, C_ref)


def test_noalloc_dim_above_plan_max_dim():
    """The fixed array is laid out per call at the call's dim, so a dim above
    the plan's max_dim (which its own workspace cannot hold) runs."""
    p = Problem(13, 16, 2, 0, 128, hub_threshold=48, chunk_nnz=16, max_dim=32, seed=5)
    C = torch.zeros((p.P, p.dim), dtype=torch.float32, device=DEV)
    assert p.call("cola_spmm_partition", C, cola.OVERWRITE) == _lib.COLA_ERR_INVALID
    assert p.call("cola_spmm_partition_noalloc", C, cola.OVERWRITE) == 0, cola.lib().cola_last_error()
    torch.cuda.synchronize()
    rp, ci, va = p.g.arrays()
    ref = orc.spmm_f64(rp, ci, va, 0, p.M, p.Bf.cpu().numpy()[:p.g.n])
    assert np.linalg.norm(C.cpu().numpy()[:p.M] - ref) / np.linalg.norm(ref) <= 1e-5


def test_noalloc_rejects_plans_beyond_the_array():
    p = Problem(15, 24, 1, 0, 256, hub_threshold=33, chunk_nnz=1, seed=6)
    need = p.plan.info["num_hub_chunks"] * 256 * 4
    assert need > SCRATCH_BYTES
    C = torch.full((p.P, p.dim), 5.0, dtype=torch.float32, device=DEV)
    st = p.call("cola_spmm_partition_noalloc", C, cola.OVERWRITE)
    assert st == _lib.COLA_ERR_INVALID
    assert b"workspace" in cola.lib().cola_last_error()
    torch.cuda.synchronize()
    assert torch.all(C == 5.0)


def test_noalloc_at_the_benchmarked_partition():
    """Partition 0 of 8 of the 2^22-node community-ordered graph at dim 128
    with the benchmarked schedule: its hub chunks fit the array (about 10 MB
    of 16 MiB) and two calls in a row equal cola_spmm_partition bitwise."""
    import json
    import os
    g = cola.Graph.rmat(22, 32, seed=2, reorder="community")
    part = g.partition(8, 0)
    P, dim = part.info()["nodesPerPE"], 128
    plan = cola.Plan(part, row_order=1)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    with open(os.path.join(root, "profiles", "plan_rows_f32_rmat22_traffic.json")) as f:
        plan.tuning = json.load(f)["tuning"]
    chunks = plan.info["num_hub_chunks"]
    print(f"case noalloc full size: {chunks} hub chunks", flush=True)
    assert 0 < chunks * dim * 4 <= SCRATCH_BYTES
    B = torch.rand((P * 8, dim), generator=torch.Generator(device=DEV).manual_seed(8), device=DEV).mul_(2).sub_(1)
    packed = torch.empty((plan.U_r, dim), dtype=torch.float32, device=DEV)
    plan.pack(B, packed)
    Bl = B[:P]
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def call(name, C):
        return getattr(cola.lib(), name)(plan.handle, vp(C), C.stride(0), vp(Bl), Bl.stride(0), None, vp(packed),
                                         packed.stride(0), None, dim, cola.F32, cola.OVERWRITE, sp)

    C_ref = torch.full((P, dim), 3.0, dtype=torch.float32, device=DEV)
    assert call("cola_spmm_partition", C_ref) == 0, cola.lib().cola_last_error()
    for _ in range(2):
        C = torch.full((P, dim), 4.0, dtype=torch.float32, device=DEV)
        assert call("cola_spmm_partition_noalloc", C) == 0, cola.lib().cola_last_error()
        torch.cuda.synchronize()
        assert torch.equal(C, C_ref)
