"""Schedule table for the storage forms of B on one partition.

For every (dim, storage dtype) pair, runs the plan autotuner several times
(each run's pick is a candidate: one timing per configuration is noisy, so
runs disagree), then times every candidate over the whole step (pack + row
kernel) in interleaved rounds and keeps the fastest by its median round.
Writes a JSON table {dim: {dtype: tuning}} with the timings behind each pick,
which bench.py's precision section uses for the workload it was made on.

usage: python -m cola_amd.tune_precision --out profiles/plan_rows_precision_rmat22_tuning.json
"""
import argparse
import json
import statistics
import sys

DTYPES = ("f32", "bf16", "f16", "fp8_e4m3")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--log-n", type=int, default=22)
    ap.add_argument("--degree", type=float, default=32)
    ap.add_argument("--nparts", type=int, default=8)
    ap.add_argument("--rank", type=int, default=0)
    ap.add_argument("--order", choices=["cm", "community"], default="community")
    ap.add_argument("--row-order", type=int, default=1)
    ap.add_argument("--dims", type=int, nargs="*", default=[64, 128, 256])
    ap.add_argument("--dtypes", nargs="*", default=list(DTYPES), choices=DTYPES)
    ap.add_argument("--autotunes", type=int, default=3, help="autotuner runs per (dim, dtype)")
    ap.add_argument("--rounds", type=int, default=5)
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--out", required=True)
    args = ap.parse_args()

    import torch

    import cola_amd as cola

    def log(msg):
        print(f"[tune] {msg}", file=sys.stderr, flush=True)

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    g = cola.Graph.rmat(args.log_n, args.degree, seed=2, reorder=args.order)
    part = g.partition(args.nparts, args.rank)
    P = part.info()["nodesPerPE"]
    plan = cola.Plan(part, row_order=args.row_order)
    codes = {"f32": cola.F32, "bf16": cola.BF16, "f16": cola.F16, "fp8_e4m3": cola.FP8_E4M3}
    tdt = {cola.F32: torch.float32, cola.BF16: torch.uint16, cola.F16: torch.uint16, cola.FP8_E4M3: torch.uint8}

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.reps

    table, timings = {}, {}
    for dim in args.dims:
        # The same B as bench.py's precision section.
        gen = torch.Generator(device=dev).manual_seed(4321 + dim)
        Bf = torch.rand((P * args.nparts, dim), generator=gen, device=dev, dtype=torch.float32).mul_(2).sub_(1)
        for name in args.dtypes:
            dt = codes[name]
            Bq, sc = (Bf, None) if dt == cola.F32 else cola.quantize_rows(Bf, dt)
            pq = torch.empty((max(plan.U_r, 1), dim), dtype=tdt[dt], device=dev)
            psc = torch.empty(max(plan.U_r, 1), dtype=torch.float32, device=dev) if dt == cola.FP8_E4M3 else None
            Bl = Bq[args.rank * P:(args.rank + 1) * P]
            sl = None if sc is None else sc[args.rank * P:(args.rank + 1) * P]
            C = torch.empty((P, dim), dtype=torch.float32, device=dev)
            flags = cola.OVERWRITE | (cola.KAHAN if dt != cola.F32 else 0)
            plan.pack(Bq, pq, dtype=dt, scales_global=sc, packed_scales=psc)
            cands = []
            for _ in range(args.autotunes):
                t, _ = plan.autotune(C, Bl, pq, dtype=dt, flags=flags, iters=3, scales_local=sl, packed_scales=psc)
                if t not in cands:
                    cands.append(t)

            def step():
                plan.pack(Bq, pq, dtype=dt, scales_global=sc, packed_scales=psc)
                plan.spmm(C, Bl, pq, dtype=dt, flags=flags, scales_local=sl, packed_scales=psc)

            rounds = [[] for _ in cands]
            for _ in range(args.rounds):
                for i, t in enumerate(cands):
                    plan.tuning = t
                    rounds[i].append(timed(step))
            med = [statistics.median(r) for r in rounds]
            w = min(range(len(cands)), key=lambda i: med[i])
            table.setdefault(str(dim), {})[name] = cands[w]
            timings.setdefault(str(dim), {})[name] = {"candidates": len(cands), "median_ms": med,
                                                       "picked": w}
            log(f"dim {dim} {name}: {len(cands)} candidates, medians "
                + ", ".join(f"{m:.3f}" for m in med) + f" ms; picked #{w}")
            del Bq, sc, pq, psc, C
        del Bf
    out = {"workload": f"rmat{args.log_n}_deg{int(args.degree)}_{args.order}_part{args.rank}of{args.nparts}",
           "order": args.order, "row_order": args.row_order, "log_n": args.log_n, "degree": args.degree,
           "nparts": args.nparts, "rank": args.rank, "tunings": table, "step_ms": timings,
           "command": "python -m cola_amd.tune_precision " + " ".join(sys.argv[1:]),
           "note": "per (dim, storage dtype): the fastest of the plan autotuner's picks over "
                   f"{args.autotunes} runs, by the median of {args.rounds} interleaved rounds of "
                   f"{args.reps} steps (pack + row kernel, compensated sums for the 16- and 8-bit forms)"}
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    log(f"wrote {args.out}")


if __name__ == "__main__":
    main()
