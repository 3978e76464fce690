"""Batched form: many SpMMs in one persistent launch, hipGraph replay,
autotuning — checked against per-problem launches and the FP64 oracle."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import cola_amd as cola  # noqa: E402
from oracle import oracle as orc  # noqa: E402

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a GPU")]

TDT = {cola.F32: torch.float32, cola.BF16: torch.uint16, cola.F16: torch.uint16, cola.FP8_E4M3: torch.uint8}


def rel_fro(C, R):
    d = np.linalg.norm(np.asarray(C, np.float64) - R)
    r = np.linalg.norm(R)
    return d / r if r > 0 else d


def dequant(Bq, sc, dtype):
    h = Bq.cpu().numpy()
    if dtype == cola.F32:
        return h
    if dtype == cola.BF16:
        return (h.astype(np.uint32) << 16).view(np.float32)
    if dtype == cola.F16:
        return h.view(np.float16).astype(np.float32)
    return orc.fp8_to_f32(h) * sc.cpu().numpy()[:, None]


def make_problems(k, dtype, seed=0, log_lo=9, log_hi=13):
    """Whole-graph SpMMs drawn like the batched configuration (node count
    log-uniform, degree uniform in [4, 64], dim from {32, 64, 128, 256}),
    at smaller sizes."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        log_n = int(rng.integers(log_lo, log_hi + 1))
        deg = float(rng.uniform(4, 64))
        dim = int(rng.choice([32, 64, 128, 256]))
        g = cola.Graph.rmat(log_n, deg, seed=1000 + i + seed * 97)
        part = g.partition(1, 0)
        plan = cola.Plan(part, hub_threshold=256, chunk_nnz=128)
        B = torch.rand((g.n, dim), device="cuda", dtype=torch.float32) * 2 - 1
        Bq, sc = cola.quantize_rows(B, dtype)
        C = torch.full((g.n, dim), 9.0, device="cuda", dtype=torch.float32)
        out.append(dict(g=g, plan=plan, B=Bq, scales=sc, C=C, dim=dim, Bf=B))
    return out


def individual(pr, dtype, flags=cola.OVERWRITE):
    C = torch.full_like(pr["C"], 9.0)
    pr["plan"].spmm(C, pr["B"], None, dim=pr["dim"], dtype=dtype, flags=flags, scales_local=pr["scales"])
    return C


@pytest.mark.parametrize("dtype", [cola.F32, cola.BF16, cola.FP8_E4M3])
def test_batch_matches_individual_launches_and_oracle(dtype):
    probs = make_problems(10, dtype, seed=dtype)
    batch = cola.Batch(probs, dtype=dtype, flags=cola.OVERWRITE)
    batch.run()
    torch.cuda.synchronize()
    for pr in probs:
        Cb = pr["C"].cpu().numpy()
        Ci = individual(pr, dtype).cpu().numpy()
        assert rel_fro(Cb, Ci.astype(np.float64)) <= 1e-6
        rp, ci, va = pr["g"].arrays()
        ref = orc.spmm_f64(rp, ci, va, 0, pr["g"].n, dequant(pr["B"], pr["scales"], dtype))
        assert rel_fro(Cb, ref) <= 1e-5


def test_result_is_the_same_alone_and_in_the_mix():
    """A problem's result does not depend on the other problems in the launch."""
    probs = make_problems(8, cola.BF16, seed=3)
    alone = cola.Batch([probs[2]], dtype=cola.BF16)
    alone.run()
    torch.cuda.synchronize()
    solo = probs[2]["C"].clone()
    mix = cola.Batch(probs, dtype=cola.BF16)
    probs[2]["C"].fill_(-1.0)
    mix.run()
    torch.cuda.synchronize()
    assert torch.equal(solo, probs[2]["C"])
    mix2 = cola.Batch(probs[::-1], dtype=cola.BF16)
    mix2.run()
    torch.cuda.synchronize()
    assert torch.equal(solo, probs[2]["C"])


def test_graph_capture_and_replay():
    """Captured batch replayed several times with eager work in between; every
    replay equals the eager result and follows changes to B."""
    probs = make_problems(6, cola.BF16, seed=5)
    batch = cola.Batch(probs, dtype=cola.BF16)
    batch.run()
    torch.cuda.synchronize()
    eager = [p["C"].clone() for p in probs]
    s = torch.cuda.Stream()
    batch.capture(s)
    for rep in range(4):
        for p in probs:
            p["C"].fill_(float(rep))
        torch.cuda.synchronize()
        batch.replay(s)
        s.synchronize()
        for p, e in zip(probs, eager):
            assert torch.equal(p["C"], e), rep
        # eager work between replays
        x = torch.randn(1 << 20, device="cuda")
        (x * 2).sum().item()
    # the graph reads B at replay time
    probs[0]["B"].copy_(cola.quantize_rows(-probs[0]["Bf"], cola.BF16)[0])
    torch.cuda.synchronize()
    batch.replay(s)
    s.synchronize()
    assert torch.allclose(probs[0]["C"], -eager[0], rtol=0, atol=1e-6)


def test_back_to_back_launches_do_every_task_once():
    """Waves take tasks from per-launch counters that the kernel leaves at
    zero: five accumulating launches in a row (direct, then from a graph)
    must add each row's sum exactly five times, which a skipped or repeated
    task would break (row sums are bitwise the same in every launch)."""
    probs = make_problems(10, cola.BF16, seed=13)
    for p in probs:
        p["C"].zero_()
    one = cola.Batch(probs, dtype=cola.BF16)
    one.run()
    torch.cuda.synchronize()
    R = [p["C"].cpu().numpy().copy() for p in probs]
    expect = []
    for r in R:
        e = r.copy()
        for _ in range(4):
            e = (e + r).astype(np.float32)
        expect.append(e)
    acc = cola.Batch(probs, dtype=cola.BF16, flags=0)
    for p in probs:
        p["C"].zero_()
    for _ in range(5):
        acc.run()
    torch.cuda.synchronize()
    for p, e in zip(probs, expect):
        assert np.array_equal(p["C"].cpu().numpy(), e)
    s = torch.cuda.Stream()
    acc.capture(s)
    for p in probs:
        p["C"].zero_()
    torch.cuda.synchronize()
    for _ in range(5):
        acc.replay(s)
    s.synchronize()
    for p, e in zip(probs, expect):
        assert np.array_equal(p["C"].cpu().numpy(), e)


def test_batch_autotune_keeps_results():
    probs = make_problems(8, cola.BF16, seed=7)
    batch = cola.Batch(probs, dtype=cola.BF16)
    batch.run()
    torch.cuda.synchronize()
    ref = [p["C"].clone() for p in probs]
    ms = batch.autotune(iters=2)
    assert ms > 0
    for p in probs:
        p["C"].fill_(5.0)
    batch.run()
    torch.cuda.synchronize()
    for p, r in zip(probs, ref):
        assert rel_fro(p["C"].cpu().numpy(), r.cpu().numpy().astype(np.float64)) <= 1e-6


def test_batch_accumulate_and_partitions_with_remote_rows():
    g = cola.Graph.rmat(12, 12, seed=11, reorder=True)
    rp, ci, va = g.arrays()
    nG, dim = 4, 64
    P = -(-g.n // nG)
    B = torch.rand((P * nG, dim), device="cuda") * 2 - 1
    Bh = B.cpu().numpy()
    probs = []
    for rank in range(nG):
        part = g.partition(nG, rank)
        plan = cola.Plan(part)
        packed = torch.empty((max(plan.U_r, 1), dim), device="cuda")
        plan.pack(B, packed)
        C = torch.full((P, dim), 1.5, device="cuda")
        probs.append(dict(plan=plan, B=B[rank * P:(rank + 1) * P], packed=packed, C=C, dim=dim))
    batch = cola.Batch(probs, dtype=cola.F32, flags=cola.ACCUMULATE)
    batch.run()
    torch.cuda.synchronize()
    for rank, pr in enumerate(probs):
        M = min(g.n, (rank + 1) * P) - rank * P
        ref = orc.spmm_f64(rp, ci, va, rank * P, rank * P + M, Bh[:g.n])
        assert rel_fro(pr["C"].cpu().numpy()[:M] - 1.5, ref) <= 2e-5


def test_batch_rejects_bad_dims():
    probs = make_problems(1, cola.BF16, seed=1)
    probs[0]["dim"] = 36  # not a multiple of 8 bf16 elements
    with pytest.raises(cola.ColaError):
        cola.Batch(probs, dtype=cola.BF16)



def test_capture_errors_leave_direct_launches_working():
    """A capture that cannot happen (on the default stream) and a replay
    with nothing captured return errors, and direct launches keep working; a
    tuning change drops a captured graph, so replay errors until the batch is
    captured again."""
    probs = make_problems(4, cola.BF16, seed=9)
    batch = cola.Batch(probs, dtype=cola.BF16)
    with pytest.raises(cola.ColaError):
        batch.capture(0)
    with pytest.raises(cola.ColaError):
        batch.replay()
    batch.run()
    torch.cuda.synchronize()
    ref = [p["C"].clone() for p in probs]
    s = torch.cuda.Stream()
    batch.capture(s)
    batch.set_tuning(rpw_short=2)
    with pytest.raises(cola.ColaError):
        batch.replay(s)
    for p in probs:
        p["C"].fill_(1.0)
    batch.run()
    torch.cuda.synchronize()
    for p, r in zip(probs, ref):
        assert rel_fro(p["C"].cpu().numpy(), r.cpu().numpy().astype(np.float64)) <= 1e-6
    batch.capture(s)
    for p in probs:
        p["C"].fill_(2.0)
    torch.cuda.synchronize()
    batch.replay(s)
    s.synchronize()
    for p, r in zip(probs, ref):
        assert rel_fro(p["C"].cpu().numpy(), r.cpu().numpy().astype(np.float64)) <= 1e-6

def test_batched_configuration_64_problems():
    """The benchmarked batch itself (64 SpMMs, node count log-uniform in
    [2^12, 2^18], degree uniform in [4, 64], dim from {32, 64, 128, 256}, B in
    BF16, items by decreasing degree), autotuned and replayed from a hipGraph
    twice: every problem matches its own launch, and every eighth problem
    matches the FP64 product of its dequantised B on sampled rows."""
    from cola_amd import workloads
    specs = workloads.batched_specs(64, seed=7)
    probs = workloads.build_problems(specs, dtype=cola.BF16, plan_opts=dict(row_order=1))
    print(f"case batched configuration: {sum(p['nnz'] for p in probs)} entries", flush=True)
    refs = [individual(p, cola.BF16) for p in probs]
    batch = cola.Batch(probs, dtype=cola.BF16, flags=cola.OVERWRITE)
    batch.autotune(iters=2)
    s = torch.cuda.Stream()
    batch.capture(s)
    for rep in range(2):
        for p in probs:
            p["C"].fill_(float(rep) - 3.0)
        torch.cuda.synchronize()
        batch.replay(s)
        s.synchronize()
        for k, (p, r) in enumerate(zip(probs, refs)):
            assert rel_fro(p["C"].cpu().numpy(), r.cpu().numpy().astype(np.float64)) <= 1e-6, (rep, k)
    for p in probs[::8]:
        n = p["n"]
        rows = np.unique(np.concatenate([np.arange(0, n, 211), [n - 1]]))
        rp, ci, va = p["graph"].arrays()
        deq = dequant(p["B"], None, cola.BF16)
        ref = np.stack([orc.spmm_f64(rp, ci, va, int(i), int(i) + 1, deq)[0] for i in rows])
        assert rel_fro(p["C"].cpu().numpy()[rows], ref) <= 1e-5
