"""Serving: continuous batching over resident graphs, admission by memory
budget, overlap of launches, and results that are the same alone and in the
mix (bitwise: each item is reduced by one team in a fixed order whatever the
launch composition)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import cola_amd as cola  # noqa: E402
from cola_amd import workloads  # noqa: E402

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a GPU")]


@pytest.fixture(scope="module")
def residents():
    specs = workloads.batched_specs(12, seed=4, log_lo=9, log_hi=14)
    return workloads.build_problems(specs, dtype=cola.BF16, plan_opts=dict(hub_threshold=256, chunk_nnz=128))


def reference(p):
    C = torch.zeros_like(p["C"])
    p["plan"].spmm(C, p["B"], None, dim=p["dim"], dtype=cola.BF16, flags=cola.OVERWRITE)
    torch.cuda.synchronize()
    return C


def bursty_trace(n_ids, rate, seconds, seed, burst=24):
    """A burst of requests at t = 0 (contention for the launch slots) followed
    by Poisson arrivals."""
    arr, res = workloads.poisson_trace(rate, seconds, n_ids, seed=seed)
    rng = np.random.default_rng(seed + 100)
    arr = np.concatenate([np.zeros(burst), arr])
    res = np.concatenate([rng.integers(0, n_ids, burst).astype(np.int32), res])
    return arr, res


def make_server(residents, budget=1 << 30, **kw):
    sv = cola.Server(budget, dtype=cola.BF16, **kw)
    ids = [sv.add_resident(p["plan"], p["B"]) for p in residents]
    return sv, ids


def test_same_result_alone_and_in_the_mix(residents):
    sv, ids = make_server(residents, max_batch=16, max_inflight=2)
    x = 5
    alone = torch.full_like(residents[x]["C"], -3.0)
    recs, _, _ = sv.run_trace([0.0], [ids[x]], outputs=[alone])
    assert recs["batch_size"][0] == 1
    assert torch.allclose(alone, reference(residents[x]), rtol=0, atol=1e-5)
    # A Poisson mix in which resident x appears several times.
    arr, res = bursty_trace(len(ids), 4000, 0.05, seed=1)
    res[::7] = ids[x]
    outs = [torch.full_like(residents[r]["C"], 7.0) if r == ids[x] else None for r in res]
    recs, wall, peak = sv.run_trace(arr, res, outputs=outs)
    assert np.all(recs["done_us"] >= recs["arrival_us"])
    mixed = [o for o in outs if o is not None]
    assert len(mixed) >= 3
    assert recs["batch_size"].max() > 1, "requests were never batched together"
    for o in mixed:
        assert torch.equal(o, alone)


def test_every_request_is_correct_and_launches_overlap(residents):
    sv, ids = make_server(residents, max_batch=8, max_inflight=3)
    arr, res = bursty_trace(len(ids), 6000, 0.05, seed=2)
    outs = [torch.empty_like(residents[r]["C"]) for r in res]
    recs, wall, peak = sv.run_trace(arr, res, outputs=outs)
    refs = {r: reference(residents[r]) for r in set(res.tolist())}
    for o, r in zip(outs, res):
        assert torch.allclose(o, refs[r], rtol=0, atol=1e-5)
    # continuous batching: some launch was dispatched while an earlier one was still running
    order = np.argsort(recs["dispatch_us"])
    d, e = recs["dispatch_us"][order], recs["done_us"][order]
    assert np.any(d[1:] < np.maximum.accumulate(e)[:-1]), "no overlap of launches observed"


def test_admission_respects_the_memory_budget(residents):
    # room for the largest single request's output + workspace, not much more
    need = max(p["n"] * p["dim"] * 4 + p["plan"].info["num_hub_chunks"] * 256 * 4 for p in residents)
    budget = int(need * 1.5)
    sv, ids = make_server(residents, budget=budget, max_batch=32, max_inflight=2)
    arr, res = workloads.poisson_trace(20000, 0.02, len(ids), seed=3)
    recs, wall, peak = sv.run_trace(arr, res)
    assert peak <= budget
    assert np.all(recs["done_us"] > 0)


def test_plain_mode_matches(residents):
    sv, ids = make_server(residents, plain=True)
    arr, res = workloads.poisson_trace(3000, 0.03, len(ids), seed=4)
    outs = [torch.empty_like(residents[r]["C"]) for r in res]
    recs, wall, peak = sv.run_trace(arr, res, outputs=outs)
    assert np.all(recs["batch_size"] == 1)
    for o, r in zip(outs, res):
        assert torch.allclose(o, reference(residents[r]), rtol=0, atol=1e-5)


def test_request_larger_than_budget_is_rejected(residents):
    sv, ids = make_server(residents, budget=1 << 12)
    with pytest.raises(cola.ColaError):
        sv.run_trace([0.0], [ids[0]])


def test_budget_check_counts_the_rounding_admission_uses():
    """A request is sized as admission takes it (output and hub partials each
    on a 256-byte boundary): a budget that holds only its unrounded bytes is
    rejected up front instead of leaving the request waiting for memory, and
    a budget of the rounded size serves it."""
    r256 = lambda b: (b + 255) // 256 * 256  # noqa: E731
    n, dim, max_dim = 9, 8, 16
    rp = np.array([0, 8] + list(range(9, 17)), np.int64)  # row 0 -> rows 1..8, each of them -> row 0
    ci = np.array(list(range(1, 9)) + [0] * 8, np.int32)
    g = cola.Graph.from_csr(rp, ci, np.ones(16, np.float32))
    plan = cola.Plan(g.partition(1, 0), hub_threshold=4, chunk_nnz=4, max_dim=max_dim, short_degree=2)
    chunks = plan.info["num_hub_chunks"]
    assert chunks == 2
    out_bytes = n * dim * 4
    hub_bytes = chunks * ((max_dim + 15) // 16 * 16) * 4  # partial rows of max_dim rounded to 16 floats
    rounded = r256(r256(out_bytes) + hub_bytes)
    assert r256(out_bytes + hub_bytes) < rounded  # the case that tells the two sizings apart
    B = torch.rand((n, dim), device="cuda").to(torch.bfloat16)
    small = cola.Server(out_bytes + hub_bytes, dtype=cola.BF16)
    rid = small.add_resident(plan, B)
    with pytest.raises(cola.ColaError):
        small.run_trace([0.0], [rid])
    fits = cola.Server(rounded, dtype=cola.BF16)
    rid = fits.add_resident(plan, B)
    out = torch.full((n, dim), -1.0, device="cuda")
    recs, _, peak = fits.run_trace([0.0, 10.0], [rid, rid], outputs=[out, None])
    assert peak <= rounded and np.all(recs["done_us"] > 0)
    ref = torch.zeros_like(out)
    plan.spmm(ref, B, None, dim=dim, dtype=cola.BF16, flags=cola.OVERWRITE)
    torch.cuda.synchronize()
    assert torch.allclose(out, ref, rtol=0, atol=1e-5)
