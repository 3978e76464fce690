"""Gather ceiling of the benchmarked partition: the plan kernel beside
cola_probe_gather, the minimal kernel over the same items and entries
(include/cola_spmm.h), on the real B rows, on an L2-resident table and with
the same loads streamed in address order.

usage: python -m cola_amd.ceiling [--log-n 22] [--dim 128] [--tile 64] [--reps 10]
Prints one JSON object (times in ms, HIP events on the launching stream).
"""
import argparse
import ctypes
import json
import sys


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--log-n", type=int, default=22)
    ap.add_argument("--degree", type=float, default=32)
    ap.add_argument("--nparts", type=int, default=8)
    ap.add_argument("--dim", type=int, default=128)
    ap.add_argument("--tile", type=int, default=64, help="column tile of the probe (columns)")
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--tuning", default=None, help="JSON of plan tuning fields for the plan kernel")
    ap.add_argument("--only", default=None, choices=["plan", "real", "l2_resident", "streamed"],
                    help="run only this kernel (for counter runs)")
    ap.add_argument("--bpc", type=int, default=16, help="probe grid in blocks per CU with --only")
    ap.add_argument("--plan-opts", default=None, help="JSON of plan options (e.g. hot_cols, col_panels)")
    args = ap.parse_args()

    import torch

    import cola_amd as cola

    def log(msg):
        print(f"[ceiling] {msg}", file=sys.stderr, flush=True)

    dev = torch.device("cuda:0")
    g = cola.Graph.rmat(args.log_n, args.degree, seed=2, reorder="community")
    part = g.partition(args.nparts, 0)
    P = part.info()["nodesPerPE"]
    plan = cola.Plan(part, **dict(dict(row_order=1), **json.loads(args.plan_opts or "{}")))
    N = args.dim
    gen = torch.Generator(device=dev).manual_seed(1234)
    B = torch.rand((P * args.nparts, N), generator=gen, device=dev).mul_(2).sub_(1)
    packed = torch.empty((max(plan.U_r, 1), N), dtype=torch.float32, device=dev)
    C = torch.empty((P, N), dtype=torch.float32, device=dev)
    B_local = B[:P]
    plan.pack(B, packed)
    if args.tuning:
        plan.tuning = json.loads(args.tuning)
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)

    def timed(fn, reps=args.reps):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    def probe(mode, table_rows, bpc, tile=args.tile):
        def run():
            cola._lib.check(cola.lib().cola_probe_gather(
                plan.handle, ctypes.c_void_p(B_local.data_ptr()), B_local.stride(0),
                ctypes.c_void_p(packed.data_ptr()), packed.stride(0), ctypes.c_void_p(C.data_ptr()), C.stride(0),
                N, tile, mode, table_rows, bpc, sp))
        return timed(run)

    modes = {"real": (0, 0), "l2_resident": (1, 4096), "streamed": (2, P)}
    if args.only:
        if args.only == "plan":
            ms = timed(lambda: plan.spmm(C, B_local, packed, flags=cola.OVERWRITE), 3)
        else:
            ms = probe(*modes[args.only], args.bpc)
        print(json.dumps({"only": args.only, "bpc": args.bpc, "ms": ms}), flush=True)
        return
    out = {"workload": f"rmat{args.log_n}_deg{int(args.degree)}_community_part0of{args.nparts}_dim{N}_f32",
           "nnz": plan.nnz, "items": plan.info["num_items"], "tile_columns": args.tile,
           "row_loads_per_call": plan.nnz * ((N + args.tile - 1) // args.tile),
           "plan_kernel": {"tuning": plan.tuning,
                           "ms": timed(lambda: plan.spmm(C, B_local, packed, flags=cola.OVERWRITE))}}
    log(f"plan kernel {out['plan_kernel']['ms']:.3f} ms")
    # l2_resident: 4096 rows x the tile's bytes per table, which fits every XCD's L2
    for name, (mode, rows) in modes.items():
        per = {}
        for bpc in (1, 2, 4, 8, 16):
            per[str(bpc)] = probe(mode, rows, bpc)
        out["probe_" + name] = {"ms_by_blocks_per_cu": per, "best_ms": min(per.values()),
                                "table_rows": rows or None}
        log(f"probe {name}: " + ", ".join(f"{k}: {v:.3f}" for k, v in per.items()))
    bytes_gathered = plan.nnz * N * 4
    out["gathered_GBps_plan_kernel"] = bytes_gathered / (out["plan_kernel"]["ms"] * 1e-3) / 1e9
    out["plan_kernel_over_real_ceiling"] = out["plan_kernel"]["ms"] / out["probe_real"]["best_ms"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
