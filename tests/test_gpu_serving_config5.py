"""The serving configuration at its own workload: 200 device-resident graphs
and BF16 feature matrices drawn from the batched form's distribution (node
count log-uniform in [2^12, 2^18], degree uniform in [4, 64], dim from
{32, 64, 128, 256}; the bench's residents, workloads.batched_specs(200,
seed=11)), Poisson arrivals at 2000 req/s for 1 s, then 0.2 s at 20000
req/s (overload), continuous batching with Kahan sums on hub rows.

A fixed sample of 256 requests of each trace writes into caller-owned
outputs; each sampled result is
- bitwise equal to the same resident served alone (one request, one
  launch): the same result alone and in the mix, and
- within 1e-5 relative Frobenius error of the FP64 oracle over the
  dequantised BF16 B, on sampled rows (the max-degree row included).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import cola_amd as cola  # noqa: E402
from cola_amd import workloads  # noqa: E402
from oracle import oracle as orc  # noqa: E402

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a GPU")]

FP32_TOL = 1e-5
SAMPLE = 256
ROWS_PER_CHECK = 24


@pytest.fixture(scope="module")
def served():
    specs = workloads.batched_specs(200, seed=11)
    res = workloads.build_problems(specs, dtype=cola.BF16, plan_opts=dict(row_order=1))
    for p in res:
        del p["C"]
    sv = cola.Server(16 << 30, dtype=cola.BF16, max_batch=64, max_inflight=2, kahan=True)
    ids = [sv.add_resident(p["plan"], p["B"]) for p in res]
    print(f"case serving config: {len(res)} residents, {sum(p['nnz'] for p in res)} entries", flush=True)
    return dict(res=res, sv=sv, ids=np.asarray(ids, np.int32), alone={}, refs={})


def alone(served, r):
    """Resident r served by itself (one request, one launch)."""
    if r not in served["alone"]:
        p = served["res"][r]
        out = torch.full((p["n"], p["dim"]), -1.0, dtype=torch.float32, device="cuda")
        recs, _, _ = served["sv"].run_trace([0.0], [served["ids"][r]], outputs=[out])
        assert recs["batch_size"][0] == 1
        served["alone"][r] = out
    return served["alone"][r]


def fp64_rows(served, r):
    """Sampled rows of resident r and their FP64 product over the
    dequantised BF16 rows of B they reference."""
    if r not in served["refs"]:
        p = served["res"][r]
        rp, ci, va = p["graph"].arrays()
        deg = np.diff(rp)
        rng = np.random.default_rng(r)
        rows = np.unique(np.concatenate([rng.integers(0, p["n"], ROWS_PER_CHECK), [int(np.argmax(deg))]]))
        idx = np.concatenate([np.arange(rp[i], rp[i + 1]) for i in rows]).astype(np.int64)
        cols, inv = np.unique(ci[idx], return_inverse=True)
        rp2 = np.concatenate([[0], np.cumsum(deg[rows])]).astype(np.int64)
        bits = p["B"].view(torch.int16)[torch.from_numpy(cols.astype(np.int64)).cuda()].cpu().numpy()
        deq = (bits.view(np.uint16).astype(np.uint32) << 16).view(np.float32)
        ref = orc.spmm_f64(rp2, inv.astype(np.int32), va[idx], 0, len(rows), deq)
        served["refs"][r] = (torch.from_numpy(rows.astype(np.int64)).cuda(), ref)
    return served["refs"][r]


@pytest.mark.parametrize("rate, seconds", [(2000.0, 1.0), (20000.0, 0.2)], ids=["2k_req_s", "20k_req_s"])
def test_served_results_alone_in_the_mix_and_fp64(served, rate, seconds):
    res, ids = served["res"], served["ids"]
    arr, rid = workloads.poisson_trace(rate, seconds, len(res), seed=5)
    n = len(arr)
    assert n > SAMPLE
    pick = np.sort(np.random.default_rng(int(rate)).choice(n, SAMPLE, replace=False))
    outs = [None] * n
    for i in pick:
        p = res[rid[i]]
        outs[i] = torch.full((p["n"], p["dim"]), 7.0, dtype=torch.float32, device="cuda")
    recs, wall, peak = served["sv"].run_trace(arr, ids[rid], outputs=outs)
    torch.cuda.synchronize()
    stats = workloads.latency_stats(recs, wall)
    print(f"served {n} requests at {rate:.0f} req/s: {stats}", flush=True)
    assert np.all(recs["done_us"] >= recs["dispatch_us"]) and np.all(recs["dispatch_us"] >= recs["arrival_us"])
    assert stats["p99_ms"] > 0
    if rate >= 20000:
        assert stats["mean_batch"] > 1.5, "requests arriving together were not batched"
    for i in pick:
        r = int(rid[i])
        got = outs[i]
        assert torch.equal(got, alone(served, r)), f"request {i} (resident {r}) differs from the resident alone"
        rows, ref = fp64_rows(served, r)
        g = got.index_select(0, rows).cpu().numpy().astype(np.float64)
        nr = np.linalg.norm(ref)
        err = np.linalg.norm(g - ref) / nr if nr > 0 else np.linalg.norm(g)
        assert err <= FP32_TOL, (i, r, err)
