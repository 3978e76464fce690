"""Randomised GPU parity: hypothesis draws graphs (uniform or hub-heavy),
partitions, dims, storage dtypes, plan options, schedules and flags, and the
HIP result is checked against the FP64 oracle on the dequantised B (so only
FP32 accumulation error remains).  Derandomised: every run draws the same
examples, so a failure reproduces by rerunning the test."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
from hypothesis import HealthCheck, given, settings  # noqa: E402
from hypothesis import strategies as st  # noqa: E402

import cola_amd as cola  # noqa: E402
from oracle import oracle as orc  # noqa: E402

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a GPU")]

SETTINGS = dict(derandomize=True, deadline=None, database=None,
                suppress_health_check=[HealthCheck.too_slow, HealthCheck.data_too_large])
TOL = 1e-5
TDT = {cola.F32: torch.float32, cola.BF16: torch.uint16, cola.F16: torch.uint16, cola.FP8_E4M3: torch.uint8}


def rel_fro(C, R):
    d = np.linalg.norm(np.asarray(C, np.float64) - R)
    r = np.linalg.norm(R)
    return d / r if r > 0 else d


def assert_accumulated(got, fill, ref, scale=1, roundings=1):
    """got = fill + scale * ref, accumulated into an FP32 C prefilled with
    fill: the product within TOL relative Frobenius, plus half an ulp of
    |fill + scale * ref| per element for each FP32 add onto C (roundings of
    them), which no kernel avoids when the product is small next to what C
    already holds (0.75 + 1e-4 carries 3e-8 of rounding, 3e-4 of the 1e-4)."""
    got = np.asarray(got, np.float64)
    want = fill + scale * ref
    ulp = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
    err = np.linalg.norm(got - want)
    bound = TOL * scale * np.linalg.norm(ref) + roundings * 0.5 * np.linalg.norm(ulp)
    assert err <= bound, (err, bound)


def dequant(Bq, sc, dtype):
    h = Bq.cpu().numpy()
    if dtype == cola.F32:
        return h
    if dtype == cola.BF16:
        return (h.astype(np.uint32) << 16).view(np.float32)
    if dtype == cola.F16:
        return h.view(np.float16).astype(np.float32)
    return orc.fp8_to_f32(h) * sc.cpu().numpy()[:, None]


@st.composite
def graphs(draw, max_n=2500, max_deg=40):
    n = draw(st.integers(8, max_n))
    m = draw(st.integers(0, n * max_deg // 2))
    seed = draw(st.integers(0, 2**31 - 1))
    rng = np.random.default_rng(seed)
    skew = draw(st.sampled_from([0.0, 0.7, 0.95]))
    r = np.where(rng.random(m) < skew, rng.integers(0, max(1, n // 64), m), rng.integers(0, n, m))
    c = rng.integers(0, n, m)
    v = rng.uniform(-1, 1, m).astype(np.float32)
    # symmetric, as the reference's graphs are; hub rows come from the skew
    g = cola.Graph.from_coo(n, np.concatenate([r, c]), np.concatenate([c, r]), np.concatenate([v, v]))
    return g


plan_opts = st.fixed_dictionaries(dict(
    hub_threshold=st.sampled_from([0, 8, 64]), chunk_nnz=st.sampled_from([0, 4, 32]),
    kahan_degree=st.sampled_from([0, 4]), short_degree=st.sampled_from([0, 3, 16]),
    col_panels=st.sampled_from([0, 0, 3]), row_order=st.sampled_from([0, 1]),
    hot_cols=st.sampled_from([0, 0, 1, 50, 1 << 30])))
tunings = st.fixed_dictionaries(dict(
    col_tile=st.sampled_from([0, 8, 16, 32, 64]), rpw_short=st.sampled_from([0, 1, 2, 4, 8, 16, 32, 64]),
    rpw_long=st.sampled_from([0, 1, 2, 4, 8]), blocks_per_cu=st.sampled_from([1, 2, 4, 8]),
    xcd_remap=st.sampled_from([0, 1]), split_remote=st.sampled_from([0, 1]), unroll=st.sampled_from([8, 16]),
    panels=st.sampled_from([0, 1]), blocks_per_cu_short=st.sampled_from([0, 0, 4, 16])))


@settings(max_examples=150, **SETTINGS)
@given(graphs(), st.integers(1, 6), st.data(), st.sampled_from([1, 5, 16, 24, 32, 48, 64, 96, 128, 136, 256]),
       st.sampled_from([cola.F32, cola.F32, cola.BF16, cola.F16, cola.FP8_E4M3]), plan_opts, tunings,
       st.sampled_from([cola.OVERWRITE, cola.ACCUMULATE, cola.OVERWRITE | cola.KAHAN]))
def test_plan_path_random(g, nGPU, data, dim, dtype, opts, tuning, flags):
    n = g.n
    rank = data.draw(st.integers(0, nGPU - 1))
    part = g.partition(nGPU, rank)
    info = part.info()
    P, M = info["nodesPerPE"], info["num_rows"]
    print(f"case n={n} nnz={g.nnz} nGPU={nGPU} rank={rank} dim={dim} dtype={dtype} opts={opts} "
          f"tuning={tuning} flags={flags}", flush=True)
    if opts["col_panels"] > 1:
        opts = dict(opts, hot_cols=0)  # the two schedules are exclusive
    plan = cola.Plan(part, max_dim=max(dim, 256), **opts)
    plan.tuning = tuning
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(dim + 7 * nGPU)
    Bf = torch.from_numpy(rng.uniform(-1, 1, (P * nGPU, dim)).astype(np.float32)).to(dev)
    Bq, sc = cola.quantize_rows(Bf, dtype)
    packed = torch.empty((max(plan.U_r, 1), dim), dtype=TDT[dtype], device=dev)
    psc = torch.empty(max(plan.U_r, 1), dtype=torch.float32, device=dev) if dtype == cola.FP8_E4M3 else None
    plan.pack(Bq, packed, dim=dim, dtype=dtype, scales_global=sc, packed_scales=psc)
    overwrite = bool(flags & cola.OVERWRITE)
    fill = 3.0 if overwrite else 0.75
    C = torch.full((P, dim), fill, dtype=torch.float32, device=dev)
    plan.spmm(C, Bq[rank * P:(rank + 1) * P], packed, dim=dim, dtype=dtype, flags=flags,
              scales_local=None if sc is None else sc[rank * P:(rank + 1) * P], packed_scales=psc)
    torch.cuda.synchronize()
    Bd = dequant(Bq, sc, dtype)
    rp, ci, va = g.arrays()
    ref = orc.spmm_f64(rp, ci, va, rank * P, rank * P + M, Bd[:n])
    got = C.cpu().numpy()
    if overwrite:
        assert rel_fro(got[:M], ref) <= TOL
    else:  # at most three adds onto C per row: local, remote (split) and cold items
        assert_accumulated(got[:M], fill, ref, roundings=3)
    assert np.all(got[M:] == fill)  # rows past the partition are never touched


@settings(max_examples=50, **SETTINGS)
@given(st.lists(st.tuples(graphs(max_n=1500, max_deg=30), st.sampled_from([32, 64, 128, 256])), min_size=1,
                max_size=6),
       st.sampled_from([cola.F32, cola.BF16, cola.FP8_E4M3]), st.data())
def test_batch_random(problems, dtype, data):
    """A random batch: every problem equals its own plan-path launch (same
    sums, schedule-dependent rounding) and the FP64 oracle."""
    dev = torch.device("cuda:0")
    probs = []
    for k, (g, dim) in enumerate(problems):
        plan = cola.Plan(g.partition(1, 0), hub_threshold=data.draw(st.sampled_from([0, 16])),
                         chunk_nnz=data.draw(st.sampled_from([0, 8])), row_order=data.draw(st.sampled_from([0, 1])))
        rng = np.random.default_rng(k)
        Bf = torch.from_numpy(rng.uniform(-1, 1, (g.n, dim)).astype(np.float32)).to(dev)
        Bq, sc = cola.quantize_rows(Bf, dtype)
        C = torch.full((g.n, dim), 5.0, dtype=torch.float32, device=dev)
        probs.append(dict(g=g, plan=plan, B=Bq, scales=sc, C=C, dim=dim))
    print(f"case batch of {len(probs)} dtype={dtype} dims={[p['dim'] for p in probs]}", flush=True)
    batch = cola.Batch(probs, dtype=dtype, flags=cola.OVERWRITE)
    rl = data.draw(st.sampled_from([0, 1, 2, 4, 8]))
    tile = data.draw(st.sampled_from([0, 32, 64]))
    batch.set_tuning(-1, col_tile=tile, rpw_short=data.draw(st.sampled_from([0, 2, 8])), rpw_long=rl,
                     blocks_per_cu=data.draw(st.sampled_from([1, 2, 4, 8])))
    batch.run()
    torch.cuda.synchronize()
    for p in probs:
        Cb = p["C"].cpu().numpy()
        Ci = torch.full_like(p["C"], 9.0)
        p["plan"].spmm(Ci, p["B"], None, dim=p["dim"], dtype=dtype, flags=cola.OVERWRITE, scales_local=p["scales"])
        torch.cuda.synchronize()
        assert rel_fro(Cb, Ci.cpu().numpy().astype(np.float64)) <= 1e-6
        rp, ci, va = p["g"].arrays()
        ref = orc.spmm_f64(rp, ci, va, 0, p["g"].n, dequant(p["B"], p["scales"], dtype))
        assert rel_fro(Cb, ref) <= TOL


@settings(max_examples=60, **SETTINGS)
@given(graphs(), st.integers(1, 6), st.data(), st.sampled_from([1, 3, 32, 40, 64, 100, 128, 256]), st.booleans())
def test_reference_api_random(g, nGPU, data, dim, split):
    """The reference API on random partitions: zips (with or without
    nnz_split) through the single-stage wrapper (a plan rebuilt from the zips,
    called twice: C accumulates) and through the literal ppl0_batch_SpMM."""
    n = g.n
    rank = data.draw(st.integers(0, nGPU - 1))
    part = g.partition(nGPU, rank)
    info = part.info()
    P, M = info["nodesPerPE"], info["num_rows"]
    print(f"case reference api n={n} nnz={g.nnz} nGPU={nGPU} rank={rank} dim={dim} split={split}", flush=True)
    a = part.arrays()
    if split:
        tl, pl, _ = cola.nnz_split(a["csr_rowIdx_l"], a["csr_rowPtr_l"], len(a["csr_colIdx_l"]))
        tr, pr, _ = cola.nnz_split(a["csc_colIdx_r"], a["csc_colPtr_r"], len(a["csc_rowIdx_r"]))
    else:
        tl, pl = a["csr_rowIdx_l"], a["csr_rowPtr_l"]
        tr, pr = a["csc_colIdx_r"], a["csc_colPtr_r"]
    zl = cola.d_csr_zip(pl, a["csr_colIdx_l"], tl, a["csr_val_l"])
    zr = cola.d_csc_zip(pr, a["csc_rowIdx_r"], tr, a["csc_val_r"])
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(dim + nGPU)
    Bh = rng.uniform(-1, 1, (P * nGPU, dim)).astype(np.float32)
    B = torch.from_numpy(Bh).to(dev)
    packed = torch.empty((max(zr.n_unique, 1), dim), dtype=torch.float32, device=dev)
    rp, ci, va = g.arrays()
    ref = orc.spmm_f64(rp, ci, va, rank * P, rank * P + M, Bh[:n])
    C = torch.full((P, dim), 0.25, dtype=torch.float32, device=dev)
    cola.dist_spmm_ppl_batch(C, B, packed, zl, zr, dim, P, rank)
    cola.dist_spmm_ppl_batch(C, B, packed, zl, zr, dim, P, rank)
    torch.cuda.synchronize()
    got = C.cpu().numpy()
    assert_accumulated(got[:M], 0.25, ref, scale=2, roundings=2)  # two calls, each one add onto C
    assert np.all(got[M:] == 0.25)
    C.fill_(0.25)
    uniq = torch.from_numpy(np.unique(np.asarray(tr, np.int32))).to(dev)
    cola.pack_remote(B, uniq, packed, num_cols=zr.n_unique)
    cola.ppl0_batch_SpMM(C, B[rank * P:(rank + 1) * P], packed, zl, zr, 0, 0, zl.num, zr.num, dim, P, rank)
    torch.cuda.synchronize()
    got = C.cpu().numpy()
    # one atomic add onto C per (virtual) row segment of each CSR / CSC entry run
    assert_accumulated(got[:M], 0.25, ref, roundings=max(1, int(np.diff(rp).max(initial=0))))
