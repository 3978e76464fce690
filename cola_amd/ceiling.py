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
    ap.add_argument("--tuning", default=None,
                    help='JSON of plan tuning fields for the plan kernel, or @file.json (its "tuning" object)')
    ap.add_argument("--only", default=None, choices=["plan", "real", "l2_resident", "streamed"],
                    help="run only this kernel (for counter runs)")
    ap.add_argument("--bpc", type=int, default=16, help="probe grid in blocks per CU with --only")
    ap.add_argument("--plan-opts", default=None, help="JSON of plan options (e.g. hot_cols, col_panels)")
    ap.add_argument("--bare", action="store_true",
                    help="time the bare gather of the partition's row stream (cola_probe_row_stream, 256-byte "
                         "column passes, through VGPRs and through LDS) beside the plan kernel")
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
        if args.tuning.startswith("@"):
            with open(args.tuning[1:]) as f:
                tj = json.load(f)
            plan.tuning = tj.get("tuning", tj)
        else:
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

    if args.bare:
        bare(args, cola, torch, part, plan, B_local, packed, C, timed, log)
        return
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


def bare(args, cola, torch, part, plan, B_local, packed, C, timed, log):
    """The bare gather of the partition's row stream: entries ordered as the
    plan orders its items (rows by decreasing entry count, a row's local
    entries before its remote ones), each 256-byte column pass of every
    entry's B row loaded once, 8 loads in flight per 16-lane group, nothing
    else read or written.  Timed through VGPRs and through LDS at several
    grid sizes, beside the plan kernel on the same buffers."""
    import numpy as np
    N = args.dim
    if N * 4 % 256:
        raise SystemExit("--bare needs FP32 rows of a multiple of 256 bytes")
    a = part.arrays()
    U = len(a["csc_colIdx_r"])
    rows = np.concatenate([np.repeat(a["csr_rowIdx_l"].astype(np.int64), np.diff(a["csr_rowPtr_l"])),
                           a["csc_rowIdx_r"].astype(np.int64)])
    codes = np.concatenate([a["csr_colIdx_l"].astype(np.int64),
                            ~np.repeat(np.arange(U, dtype=np.int64), np.diff(a["csc_colPtr_r"]))])
    remote = np.concatenate([np.zeros(len(a["csr_colIdx_l"]), np.int8), np.ones(len(a["csc_rowIdx_r"]), np.int8)])
    deg = np.bincount(rows)
    order = np.lexsort((np.arange(len(rows)), remote, rows, -deg[rows]))
    codes = codes[order]
    lp, pp = B_local.data_ptr(), packed.data_ptr()
    base = min(lp, pp)
    addr = np.where(codes >= 0, lp + codes * B_local.stride(0) * 4, pp + (~codes) * packed.stride(0) * 4) - base
    if addr.min() < 0 or addr.max() >= 1 << 36 or np.any(addr % 16):
        raise SystemExit("rows do not share a 16-byte-aligned 64 GB span")
    offs = torch.from_numpy((addr >> 4).astype(np.uint32).view(np.int32)).cuda()
    out = torch.empty(torch.cuda.get_device_properties(0).multi_processor_count * 16, dtype=torch.float32,
                      device="cuda")
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    passes = N * 4 // 256

    def run(via_lds, xr, bpc):
        for t in range(passes):
            cola._lib.check(cola.lib().cola_probe_row_stream(
                ctypes.c_void_p(offs.data_ptr()), len(codes), ctypes.c_void_p(base + 256 * t), via_lds, xr, bpc,
                ctypes.c_void_p(out.data_ptr()), out.numel(), sp))

    res = {"workload": f"rmat{args.log_n}_deg{int(args.degree)}_community_part0of{args.nparts}_dim{N}_f32",
           "entries": len(codes), "column_passes": passes,
           "plan_kernel": {"tuning": plan.tuning,
                           "ms": timed(lambda: plan.spmm(C, B_local, packed, flags=cola.OVERWRITE))}}
    names = []
    for via_lds in (0, 1):
        for xr in (0, 1):
            name = "bare_gather_" + ("lds" if via_lds else "vgpr") + ("_xcd_ranges" if xr else "_strided")
            per = {str(bpc): timed(lambda: run(via_lds, xr, bpc)) for bpc in (1, 2, 4, 8, 16)}
            res[name] = {"ms_by_blocks_per_cu": per, "best_ms": min(per.values())}
            names.append(name)
            log(f"{name}: " + ", ".join(f"{k}: {v:.3f}" for k, v in per.items()))
    res["plan_kernel_over_best_bare_gather"] = res["plan_kernel"]["ms"] / min(res[k]["best_ms"] for k in names)
    log(f"plan kernel {res['plan_kernel']['ms']:.3f} ms")
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
