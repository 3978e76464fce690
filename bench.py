"""Benchmark: one row partition of the distributed SpMM C = A·B on one MI355X.

Workload (the real shape of one partition): R-MAT graph with 2^22 nodes,
average degree 32, a = 0.57, b = c = 0.19, symmetrised, self-loops and
duplicates dropped, GCN-normalised values, relabelled by the community
locality order (Rabbit Order's aggregation and depth-first numbering;
--order cm for Cuthill-McKee); 8 row partitions of 2^19 rows, partition 0;
B is 2^22 x 128 FP32 uniform[-1, 1],
resident in HBM.  One step = gather-pack of the unique remote B rows + the
degree-binned row kernel + the hub reduction, C overwritten.

value = algorithmic bytes of the partition SpMM / step time, with
  bytes_alg = 4(M+1) + nnz_p*(4+4) + (U_l + U_r)*N*s_B + M*N*4
(each distinct B row read once, C written once).

Extras (same partition unless stated): reduced-precision storage of B
(BF16 / FP16 / FP8-e4m3 + row scales) at dims 64 / 128 / 256 with relative
Frobenius and max-row error against the FP64 oracle (computed on the GPU in
FP64 over sampled rows); the reference-semantics path (nnz_split + CSC
scatter atomics); every row partition of the graph timed one after another
(FP32, default schedule); the batched form (64 SpMMs, one persistent launch);
the served form (Poisson arrivals over 200 resident graphs).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "SpMM HBM GB/s (% of peak) FP32 vs BF16 at stated error; served req/s + p99"


def bytes_alg(M, nnz, U, N, s_B, s_sc=0):
    return 4 * (M + 1) + nnz * 8 + U * (N * s_B + s_sc) + M * N * 4


def timed(torch, fn, reps, warm=1):
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s0.record()
    for _ in range(reps):
        fn()
    s1.record()
    torch.cuda.synchronize()
    return s0.elapsed_time(s1) / reps


class Sample:
    """FP64 oracle rows of the partition product for a fixed row sample, and
    on-device FP64 error norms of C over those rows."""

    def __init__(self, orc, torch, rp, ci, va, lb, M, Bh, n, count=512):
        self.rows = np.unique(np.concatenate([np.arange(0, M, max(1, M // count)), [M - 1]]))
        ref = np.stack([orc.spmm_f64(rp, ci, va, lb + int(i), lb + int(i) + 1, Bh[:n])[0] for i in self.rows])
        self.ref = torch.from_numpy(ref).cuda()
        self.idx = torch.from_numpy(self.rows.astype(np.int64)).cuda()

    def errors(self, cola, C):
        sub = C.index_select(0, self.idx).contiguous()
        out = cola.error_norms(sub, self.ref).cpu().numpy()
        return {"rel_fro": float(np.sqrt(out[0] / out[1])), "max_row_rel": float(np.sqrt(out[2]))}


def precision_section(cola, torch, orc, log, ctx, args, f32_tuning, table=None):
    """BF16 / FP16 / FP8-e4m3 storage of B at dims 64, 128, 256 (FP32
    accumulation, Kahan on rows above 1024 entries) against FP32 storage at
    the same dim, each at the schedule of `table` (the committed schedule
    table of this workload, made by cola_amd.tune_precision) or, where the
    table has none, autotuned."""
    plan, P, M, n, rp, ci, va, lb = ctx
    out = {}
    dev = torch.device("cuda:0")
    for dim in args.precision_dims:
        gen = torch.Generator(device=dev).manual_seed(4321 + dim)
        Bf = torch.rand((P * args.nparts, dim), generator=gen, device=dev, dtype=torch.float32).mul_(2).sub_(1)
        sample = Sample(orc, torch, rp, ci, va, lb, M, Bf.cpu().numpy(), n, count=256)
        per = {}
        for dt in (cola.F32, cola.BF16, cola.F16, cola.FP8_E4M3):
            name = cola.DTYPE_NAMES[dt]
            if dt == cola.F32:
                Bq, sc = Bf, None
            else:
                Bq, sc = cola.quantize_rows(Bf, dt)
            tdt = {cola.F32: torch.float32, cola.BF16: torch.uint16, cola.F16: torch.uint16,
                   cola.FP8_E4M3: torch.uint8}[dt]
            pq = torch.empty((max(plan.U_r, 1), dim), dtype=tdt, device=dev)
            psc = torch.empty(max(plan.U_r, 1), dtype=torch.float32, device=dev) if dt == cola.FP8_E4M3 else None
            Bl = Bq[args.rank * P:(args.rank + 1) * P]
            sl = None if sc is None else sc[args.rank * P:(args.rank + 1) * P]
            C = torch.empty((P, dim), dtype=torch.float32, device=dev)
            flags = cola.OVERWRITE | (cola.KAHAN if dt != cola.F32 else 0)
            plan.pack(Bq, pq, dtype=dt, scales_global=sc, packed_scales=psc)
            fixed = (table or {}).get(str(dim), {}).get(name)
            if fixed is not None:
                plan.tuning = fixed
                tq, source = plan.tuning, "profile"
            else:
                tq, _ = plan.autotune(C, Bl, pq, dtype=dt, flags=flags, iters=3, scales_local=sl,
                                      packed_scales=psc)
                source = "auto"

            def step():
                plan.pack(Bq, pq, dtype=dt, scales_global=sc, packed_scales=psc)
                plan.spmm(C, Bl, pq, dtype=dt, flags=flags, scales_local=sl, packed_scales=psc)

            ms = timed(torch, step, max(3, args.steps // 2))
            s_sc = 4 if dt == cola.FP8_E4M3 else 0
            rec = {"ms_per_step": ms,
                   "GBps_alg": bytes_alg(M, plan.nnz, plan.U_l + plan.U_r, dim, cola.DTYPE_BYTES[dt], s_sc) /
                   (ms * 1e-3) / 1e9,
                   "tuning": tq, "tuning_source": source}
            rec.update(sample.errors(cola, C))
            per[name] = rec
            del Bq, sc, pq, psc, C
        for name, rec in per.items():
            rec["speedup_vs_f32"] = per["f32"]["ms_per_step"] / rec["ms_per_step"]
        out[str(dim)] = per
        log("precision dim %d: " % dim + ", ".join(
            f"{k} {v['ms_per_step']:.3f} ms ({v['speedup_vs_f32']:.2f}x) err {v['rel_fro']:.1e}/{v['max_row_rel']:.1e}"
            for k, v in per.items()))
        del Bf, sample
    plan.tuning = f32_tuning
    return out


def partitions_section(cola, torch, log, g, B, args):
    """Every row partition of the benchmarked graph, one after another on this
    GPU: pack + cola_spmm_partition in FP32 at the plan's default schedule,
    so the shares of a sharded step can be compared (each figure is one
    partition on one GPU)."""
    out = []
    P = (g.n + args.nparts - 1) // args.nparts
    N = B.shape[1]
    for r in range(args.nparts):
        part = g.partition(args.nparts, r)
        info = part.info()
        rec = {"rank": r, "nnz": int(info["nnz_l"] + info["nnz_r"]), "unique_remote_rows": int(info["U_r"])}
        if rec["nnz"] == 0:
            rec["ms_per_step"] = 0.0
            out.append(rec)
            continue
        plan = cola.Plan(part, row_order=args.row_order)
        packed = torch.empty((max(plan.U_r, 1), N), dtype=torch.float32, device=B.device)
        C = torch.empty((P, N), dtype=torch.float32, device=B.device)
        B_local = B[r * P:(r + 1) * P]

        def step():
            plan.pack(B, packed)
            plan.spmm(C, B_local, packed, flags=cola.OVERWRITE)

        rec["ms_per_step"] = timed(torch, step, max(3, args.steps // 4), warm=2)
        out.append(rec)
        del plan, packed, C
    log("partitions: " + ", ".join(f"{x['rank']}: {x['nnz']} entries {x['ms_per_step']:.3f} ms" for x in out))
    return out


def batched_section(cola, torch, log):
    """64 independent SpMMs (node count log-uniform in [2^12, 2^18], degree
    uniform in [4, 64], dim from {32, 64, 128, 256}, B in BF16) in one
    persistent launch replayed from a hipGraph, against one launch per SpMM."""
    from cola_amd import workloads
    t0 = time.time()
    specs = workloads.batched_specs(64, seed=7)
    probs = workloads.build_problems(specs, dtype=cola.BF16, plan_opts=dict(row_order=1))
    balg = sum(workloads.bytes_alg(p, 2) for p in probs)
    nnz = sum(p["nnz"] for p in probs)
    log(f"batched: 64 problems, {nnz} entries, built in {time.time() - t0:.1f}s")

    def individual():
        for p in probs:
            p["plan"].spmm(p["C"], p["B"], None, dim=p["dim"], dtype=cola.BF16, flags=cola.OVERWRITE)

    ms_ind = timed(torch, individual, 5)
    batch = cola.Batch(probs, dtype=cola.BF16, flags=cola.OVERWRITE)
    ms_untuned = timed(torch, batch.run, 5)
    batch.autotune(iters=2)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        batch.capture(stream)
        ms_graph = timed(torch, lambda: batch.replay(stream), 10)
    torch.cuda.synchronize()
    tunings = {}
    for k, p in enumerate(probs):
        tunings.setdefault(p["dim"], batch.tuning(k))
    out = {"problems": 64, "entries": nnz, "dtype": "bf16", "bytes_alg": balg,
           "ms_one_launch_per_problem": ms_ind, "ms_persistent_untuned": ms_untuned,
           "ms_persistent_autotuned_graph": ms_graph,
           "GBps_alg_persistent": balg / (ms_graph * 1e-3) / 1e9,
           "speedup_vs_per_problem_launches": ms_ind / ms_graph,
           "tuning_per_dim": {str(d): t for d, t in tunings.items()}}
    log(f"batched: per-problem launches {ms_ind:.3f} ms, persistent {ms_untuned:.3f} ms, "
        f"autotuned + graph {ms_graph:.3f} ms ({out['GBps_alg_persistent']:.0f} GB/s alg)")
    del batch, probs
    return out, tunings


def serving_section(cola, torch, log, args, dim_tunings):
    """Poisson arrivals over resident graphs: continuous batching (BF16) vs
    plain per-request FP32 launches, same trace."""
    from cola_amd import workloads
    t0 = time.time()
    specs = workloads.batched_specs(args.serve_graphs, seed=11)
    res = workloads.build_problems(specs, dtype=cola.BF16, keep_f32=True, plan_opts=dict(row_order=1))
    for p in res:
        del p["C"]
    resident_bytes = sum(p["plan"].info["device_bytes"] + p["B"].numel() * 2 + p["Bf"].numel() * 4 for p in res)
    log(f"serving: {len(res)} resident graphs ({resident_bytes / 1e9:.1f} GB incl. FP32 copies), "
        f"built in {time.time() - t0:.1f}s")
    budget = 32 << 30
    out = {"resident_graphs": len(res), "rate_req_per_s": args.serve_rate, "trace_seconds": args.serve_seconds,
           "memory_budget_bytes": budget, "resident_bytes": resident_bytes}
    bpc = next(iter(dim_tunings.values()))["blocks_per_cu"] if dim_tunings else 0
    sv = cola.Server(budget, dtype=cola.BF16, max_batch=64, max_inflight=2, kahan=True, blocks_per_cu=bpc)
    ids_b = [sv.add_resident(p["plan"], p["B"]) for p in res]
    for d, t in dim_tunings.items():
        sv.set_tuning(d, col_tile=t["col_tile"], rpw_short=t["rpw_short"], rpw_long=t["rpw_long"])
    pl = cola.Server(budget, dtype=cola.F32, plain=True)
    ids_p = [pl.add_resident(p["plan"], p["Bf"]) for p in res]
    # The configured rate, then an overload point where requests arrive
    # faster than one launch per request can serve them.
    for rate, seconds, key in ((args.serve_rate, args.serve_seconds, ""),
                               (args.serve_stress_rate, 1.0, "_at_%d_req_per_s" % int(args.serve_stress_rate))):
        if rate <= 0:
            continue
        arr, rid = workloads.poisson_trace(rate, seconds, len(res), seed=5)
        sv.run_trace(arr[:50] - arr[0], np.asarray(ids_b, np.int32)[rid[:50]])  # warm-up
        recs, wall, peak = sv.run_trace(arr, np.asarray(ids_b, np.int32)[rid])
        out["continuous_batching_bf16" + key] = dict(workloads.latency_stats(recs, wall), peak_pool_bytes=peak)
        pl.run_trace(arr[:50] - arr[0], np.asarray(ids_p, np.int32)[rid[:50]])
        recs, wall, peak = pl.run_trace(arr, np.asarray(ids_p, np.int32)[rid])
        out["plain_per_request_f32" + key] = dict(workloads.latency_stats(recs, wall), peak_pool_bytes=peak)
        b, p = out["continuous_batching_bf16" + key], out["plain_per_request_f32" + key]
        log(f"serving at {rate:.0f} req/s: batched bf16 {b['req_per_s']:.0f} req/s p50 {b['p50_ms']:.2f} ms "
            f"p99 {b['p99_ms']:.2f} ms (mean batch {b['mean_batch']:.1f}); plain f32 {p['req_per_s']:.0f} req/s "
            f"p50 {p['p50_ms']:.2f} ms p99 {p['p99_ms']:.2f} ms")
    del sv, pl
    return out


def cpu_section(orc, log, rp, ci, va, lb, M, Bh, N, U_total, seconds):
    """OpenMP FP32 CSR SpMM (the oracle's CPU port) over the partition's rows:
    2 untimed passes, then the median of up to 10 timed passes within about
    `seconds` of CPU time.  If one pass would take longer than that, whole
    row blocks are timed until the budget is spent instead (a bounded sample,
    scaled by its row fraction)."""
    rows_per_block = max(1, M // 16)
    Cc = np.empty((rows_per_block, N), np.float32)

    def run_rows(r0, r1):
        tt = time.perf_counter()
        _, th = orc.spmm_omp_f32(rp, ci, va, r0, r1, Bh, Cc[:r1 - r0])
        return time.perf_counter() - tt, th

    # one timed pass over the blocks (stopping early if it exceeds the budget)
    done_rows = done_nnz = 0
    t_pass = 0.0
    threads = 1
    while done_rows < M and t_pass < seconds:
        r0 = lb + done_rows
        r1 = min(lb + M, r0 + rows_per_block)
        dt, threads = run_rows(r0, r1)
        t_pass += dt
        done_nnz += int(rp[r1] - rp[r0])
        done_rows = r1 - lb
    times = [t_pass]
    if done_rows == M:
        def full():
            return sum(run_rows(lb + k, min(lb + M, lb + k + rows_per_block))[0] for k in range(0, M, rows_per_block))
        full()  # second warm-up pass (the first timed one above also warmed up)
        times = []
        spent = 0.0
        while len(times) < 10 and spent < seconds:
            times.append(full())
            spent += times[-1]
    t_cpu = float(np.median(times))
    frac_rows = done_rows / M
    cb = 4 * (done_rows + 1) + done_nnz * 8 + int(U_total * frac_rows) * N * 4 + done_rows * N * 4
    cpu = {"value": cb / t_cpu / 1e9, "unit": "GB/s", "cores": int(threads), "kind": "port",
           "gflops": 2.0 * done_nnz * N / t_cpu / 1e9,
           "sample": f"OpenMP FP32 CSR SpMM (oracle; gcc -O3, row loop built for AVX-512 / AVX2+FMA / x86-64 "
                     f"and picked for the host CPU; schedule(dynamic,64); {threads} threads), "
                     f"rows [{lb},{lb + done_rows}) of the partition "
                     f"({done_nnz} entries, {frac_rows:.0%} of its rows) on {threads} threads; "
                     + (f"median of {len(times)} passes after 2 warm-up passes ({t_cpu:.2f} s each)"
                        if done_rows == M else f"one pass over that sample ({t_cpu:.1f} s), "
                                               f"algorithmic bytes scaled by the sampled row fraction")}
    log(f"cpu baseline {cpu['value']:.1f} GB/s on {threads} threads ({cpu['sample']})")
    return cpu


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--log-n", type=int, default=22)
    ap.add_argument("--degree", type=float, default=32)
    ap.add_argument("--nparts", type=int, default=8)
    ap.add_argument("--rank", type=int, default=0)
    ap.add_argument("--dim", type=int, default=128)
    ap.add_argument("--dtype", choices=["f32", "bf16", "f16", "fp8_e4m3"], default="f32",
                    help="storage dtype of B in the main step (the metric is quoted for f32)")
    ap.add_argument("--row-order", type=int, default=1, help="plan item order: 0 row order, 1 by degree")
    ap.add_argument("--order", choices=["cm", "community"], default="community",
                    help="locality order of the graph's vertex ids before partitioning")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip every extra section")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--precision-dims", type=int, nargs="*", default=[64, 128, 256])
    ap.add_argument("--no-reference-path", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-SpMM persistent-launch section")
    ap.add_argument("--no-partitions", action="store_true", help="skip timing every partition of the graph")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving section")
    ap.add_argument("--serve-graphs", type=int, default=200)
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--serve-stress-rate", type=float, default=20000.0,
                    help="second serving point, 1 s of arrivals at this rate (0: skip)")
    ap.add_argument("--tuning", default="profile",
                    help='"profile" (the schedule of the committed counter profile when it was taken on this '
                         'workload, else autotune), "auto" (autotune before timing), a JSON object of plan '
                         'tuning fields, or @file.json (its "tuning" object)')
    ap.add_argument("--save-tuning", default=None, help="write the FP32 tuning used to this JSON file")
    ap.add_argument("--precision-tuning", choices=["profile", "auto"], default="profile",
                    help='schedules of the precision section: "profile" (the committed table '
                         'profiles/plan_rows_precision_rmat22_tuning.json when made on this workload, else '
                         'autotune) or "auto"')
    ap.add_argument("--copy-calibrate-mib", type=int, default=0,
                    help="run one plain device copy of this many MiB first (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one partition on one GPU; runs on more than one GPU are not part of "
              "this build (the multi-GPU exchange is design-only).", file=sys.stderr)
        sys.exit(2)

    import torch

    import cola_amd as cola

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    t0 = time.time()
    g = cola.Graph.rmat(args.log_n, args.degree, seed=2, reorder=args.order)
    n, nnz_total, _ = g.info()
    part = g.partition(args.nparts, args.rank)
    info = part.info()
    P, M, N = info["nodesPerPE"], info["num_rows"], args.dim
    plan = cola.Plan(part, row_order=args.row_order)
    nnz_p = plan.nnz
    U_l, U_r = plan.U_l, plan.U_r
    lb = args.rank * P
    log(f"graph n={n} nnz={nnz_total} partition {args.rank}/{args.nparts}: M={M} nnz_p={nnz_p} "
        f"U_l={U_l} U_r={U_r} items={plan.info['num_items']} hubs={plan.info['num_hub_rows']} "
        f"setup {time.time() - t0:.1f}s")

    gen = torch.Generator(device=dev).manual_seed(1234)
    B = torch.rand((P * args.nparts, N), generator=gen, device=dev, dtype=torch.float32).mul_(2).sub_(1)
    # Storage dtype of B for the main step (FP32 unless --dtype says otherwise).
    dt = {"f32": cola.F32, "bf16": cola.BF16, "f16": cola.F16, "fp8_e4m3": cola.FP8_E4M3}[args.dtype]
    s_B = cola.DTYPE_BYTES[dt]
    s_sc = 4 if dt == cola.FP8_E4M3 else 0
    Bs, Bsc = (B, None) if dt == cola.F32 else cola.quantize_rows(B, dt)
    B_local = Bs[lb:lb + P]
    sc_local = None if Bsc is None else Bsc[lb:lb + P]
    tdt = {cola.F32: torch.float32, cola.BF16: torch.uint16, cola.F16: torch.uint16, cola.FP8_E4M3: torch.uint8}[dt]
    packed = torch.empty((max(U_r, 1), N), dtype=tdt, device=dev)
    psc = torch.empty(max(U_r, 1), dtype=torch.float32, device=dev) if dt == cola.FP8_E4M3 else None
    C = torch.empty((P, N), dtype=torch.float32, device=dev)

    def pack():
        plan.pack(Bs, packed, dtype=dt, scales_global=Bsc, packed_scales=psc)

    def spmm():
        plan.spmm(C, B_local, packed, dtype=dt, flags=cola.OVERWRITE, scales_local=sc_local, packed_scales=psc)

    if args.copy_calibrate_mib > 0:
        nel = args.copy_calibrate_mib * (1 << 20) // 4
        src = torch.ones(nel, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    pack()
    # The committed counter profile (HBM bytes per launch) holds only for the
    # workload and schedule it was measured with.
    tfile = os.path.join(ROOT, "profiles", {"community": "plan_rows_f32_rmat22_traffic.json",
                                            "cm": "plan_rows_f32_rmat22cm_traffic.json"}[args.order])
    prof = None
    if os.path.exists(tfile):
        with open(tfile) as f:
            prof = json.load(f)
        if not (prof.get("row_order", 0) == args.row_order and prof.get("order", "cm") == args.order
                and N == 128 and args.log_n == 22
                and args.degree == 32 and args.nparts == 8 and args.rank == 0 and dt == cola.F32):
            prof = None
    tuning_source = args.tuning
    if args.tuning == "profile" and prof is None:
        tuning_source = "auto"
    if tuning_source == "auto":
        tuning, tune_ms = plan.autotune(C, B_local, packed, dtype=dt, scales_local=sc_local, packed_scales=psc,
                                        iters=3)
        log(f"autotuned {args.dtype}: {tuning} ({tune_ms:.3f} ms)")
    else:
        if tuning_source == "profile":
            tj = prof["tuning"]
        elif args.tuning.startswith("@"):
            with open(os.path.join(ROOT, args.tuning[1:])) as f:
                tj = json.load(f)
            tj = tj.get("tuning", tj)
        else:
            tj = json.loads(args.tuning)
        plan.tuning = tj
        tuning = plan.tuning
        log(f"fixed {args.dtype} tuning ({tuning_source}): {tuning}")
    if args.save_tuning:
        with open(args.save_tuning, "w") as f:
            json.dump(tuning, f)

    for _ in range(args.warmup):
        pack()
        spmm()
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for k in range(args.steps):
        ev[k][0].record()
        pack()
        ev[k][1].record()
        spmm()
        ev[k][2].record()
    end.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / args.steps
    pack_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in ev]))
    spmm_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in ev]))
    balg = bytes_alg(M, nnz_p, U_l + U_r, N, s_B, s_sc)
    value = balg / (ms * 1e-3) / 1e9
    # Row kernel's own algorithmic bytes: entry lists + work items + distinct
    # local B rows + packed rows + C.
    kbytes = nnz_p * 8 + plan.info["num_items"] * 16 + (U_l + U_r) * (N * s_B + s_sc) + M * N * 4
    k_gbs = kbytes / (spmm_ms * 1e-3) / 1e9
    log(f"step {ms:.3f} ms (pack {pack_ms:.3f} + spmm {spmm_ms:.3f}); {value:.0f} GB/s alg; "
        f"spmm kernel {k_gbs:.0f} GB/s; gathered rows {nnz_p * N * s_B / (spmm_ms * 1e-3) / 1e9:.0f} GB/s")

    traffic = prof.get("hbm_bytes_per_launch") if prof is not None and prof.get("tuning") == tuning else None

    extras = {}
    if not (args.no_extras and args.no_cpu):
        from oracle import oracle as orc
        rp, ci, va = g.arrays()
    if not args.no_extras:
        Bh = B.cpu().numpy()
        sample = Sample(orc, torch, rp, ci, va, lb, M, Bh, n)
        extras["%s_errors_sampled_rows" % args.dtype] = sample.errors(cola, C)
        log(f"{args.dtype} errors {extras['%s_errors_sampled_rows' % args.dtype]}")
        del Bh, sample
        if not args.no_precision:
            # The committed schedule table holds only for the workload it was made on.
            ptable = None
            pfile = os.path.join(ROOT, "profiles", "plan_rows_precision_rmat22_tuning.json")
            if args.precision_tuning == "profile" and os.path.exists(pfile):
                with open(pfile) as f:
                    pj = json.load(f)
                if (pj.get("order") == args.order and pj.get("row_order") == args.row_order
                        and pj.get("log_n") == args.log_n and pj.get("degree") == args.degree
                        and pj.get("nparts") == args.nparts and pj.get("rank") == args.rank):
                    ptable = pj["tunings"]
            extras["precision"] = precision_section(cola, torch, orc, log, (plan, P, M, n, rp, ci, va, lb), args,
                                                    tuning, ptable)
            per = extras["precision"].get(str(N), {})
            # Reduced-precision storage against this FP32 step, with its
            # relative Frobenius error against FP64 of the FP32 B.
            extras["vs_f32_at_dim%d" % N] = {
                k: {"speedup": v["speedup_vs_f32"], "rel_fro": v["rel_fro"], "max_row_rel": v["max_row_rel"]}
                for k, v in per.items() if k != "f32"}
        if not args.no_reference_path:
            a = part.arrays()
            tl, pl, _ = cola.nnz_split(a["csr_rowIdx_l"], a["csr_rowPtr_l"], len(a["csr_colIdx_l"]))
            tr, pr, _ = cola.nnz_split(a["csc_colIdx_r"], a["csc_colPtr_r"], len(a["csc_rowIdx_r"]))
            zl = cola.d_csr_zip(pl, a["csr_colIdx_l"], tl, a["csr_val_l"])
            zr = cola.d_csc_zip(pr, a["csc_rowIdx_r"], tr, a["csc_val_r"])
            Cr = torch.zeros_like(C)
            # The reference's single-stage wrapper (plan cached from the zips)
            # and the literal ppl0_batch_SpMM restatement (virtual rows, atomics),
            # both FP32 as the reference is.
            packed32 = torch.empty((max(zr.n_unique, 1), N), dtype=torch.float32, device=dev)
            ms_ref = timed(torch, lambda: cola.dist_spmm_ppl_batch(Cr, B, packed32, zl, zr, N, P, args.rank),
                           max(3, args.steps // 4), warm=2)
            uniq = torch.from_numpy(np.unique(tr).astype(np.int32)).to(dev)

            def literal():
                cola.pack_remote(B, uniq, packed32, num_cols=zr.n_unique)
                cola.ppl0_batch_SpMM(Cr, B[lb:lb + P], packed32, zl, zr, 0, 0, zl.num, zr.num, N, P, args.rank)

            ms_lit = timed(torch, literal, max(3, args.steps // 4), warm=2)
            extras["reference_api_wrapper"] = {"ms_per_step": ms_ref, "GBps_alg": balg / (ms_ref * 1e-3) / 1e9,
                                               "path": "cola_dist_spmm_ppl_batch: plan rebuilt from the zips once"}
            extras["reference_kernel_restatement"] = {
                "ms_per_step": ms_lit, "GBps_alg": balg / (ms_lit * 1e-3) / 1e9,
                "path": "pack + cola_ppl0_batch_SpMM: nnz_split virtual rows, FP32 atomics"}
            log(f"reference API wrapper {ms_ref:.3f} ms; literal ppl0 restatement (atomics) {ms_lit:.3f} ms")
            del Cr, zl, zr, uniq, packed32
        if not args.no_partitions and args.dtype == "f32":
            extras["partitions"] = partitions_section(cola, torch, log, g, B, args)
        dim_tunings = {}
        if not args.no_batched:
            extras["batched"], dim_tunings = batched_section(cola, torch, log)
        if not args.no_serve:
            extras["serving"] = serving_section(cola, torch, log, args, dim_tunings)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_section(orc, log, rp, ci, va, lb, M, B.cpu().numpy()[:n], N, U_l + U_r, args.cpu_seconds)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic (seeded R-MAT graph, uniform[-1,1] B generated on the GPU)",
        "config": {"workload": f"rmat{args.log_n}_deg{int(args.degree)}_{args.order}_part{args.rank}of{args.nparts}"
                               f"_dim{N}_{args.dtype}",
                   "nodes": n, "nnz_graph": nnz_total, "rows": M, "nnz_partition": nnz_p,
                   "unique_local_cols": U_l, "unique_remote_cols": U_r, "dim": N,
                   "bytes_alg": balg, "step": "pack + plan row kernel + hub reduce, C overwritten",
                   "row_order": args.row_order, "tuning": tuning, "tuning_source": tuning_source},
        "roofline": {"bound": "hbm", "achieved": round(k_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(k_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     # The counters measure bytes that miss the L2 (TCC): HBM
                     # reads plus hits in the MALL (Infinity Cache), which
                     # the counters do not separate, so the rate below is an
                     # L2-miss rate, not an HBM rate.
                     "traffic_kind": "L2 (TCC) miss bytes per call, MALL hits included: FETCH_SIZE x calibration "
                                     "+ WRITE_SIZE, from rocprofv3 --pmc runs of this workload and schedule",
                     "traffic_source": (f"profiles/{os.path.basename(tfile)} (tuning {prof['tuning']})"
                                        if traffic else None),
                     "l2_miss_GBps": round(traffic / (spmm_ms * 1e-3) / 1e9, 1) if traffic else None,
                     "kernel": "plan_rows_kernel (+ hub_reduce_kernel), one cola_spmm_partition call",
                     "kernel_ms": round(spmm_ms, 4), "kernel_bytes_alg": kbytes},
        "cpu_baseline": cpu,
        "kernels_ms": {"pack": round(pack_ms, 4), "spmm": round(spmm_ms, 4)},
        # Context, not roofline: FLOP rate of the step, and the rate the
        # row kernel would need if every entry's B row came from HBM.
        "context": {"gflops_step": round(2 * nnz_p * N / (ms * 1e-3) / 1e9, 1),
                    "gather_no_reuse_GBps": round(nnz_p * N * s_B / (spmm_ms * 1e-3) / 1e9, 1)},
        "extras": extras,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
