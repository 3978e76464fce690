"""Synthetic workloads of the batched and served forms.

Batched form: independent whole-graph SpMMs with R-MAT graphs whose node count
is log-uniform in [2^12, 2^18] (rounded to a power of two, as R-MAT needs),
average degree uniform in [4, 64], dense dim from {32, 64, 128, 256}, B
uniform[-1, 1] stored in BF16.  Served form: a Poisson arrival stream of
requests over a set of such resident graphs.
"""
import numpy as np

from . import _lib
from .partition import Graph
from .spmm import Plan, quantize_rows

DIMS = (32, 64, 128, 256)


def batched_specs(count, seed=0, log_lo=12, log_hi=18, deg_lo=4.0, deg_hi=64.0, dims=DIMS):
    """[(log_n, avg_degree, dim, graph_seed)] drawn with a fixed seed."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        log_n = int(round(rng.uniform(log_lo, log_hi)))
        deg = float(rng.uniform(deg_lo, deg_hi))
        dim = int(rng.choice(dims))
        out.append((log_n, deg, dim, int(seed * 100003 + i + 1)))
    return out


def build_problems(specs, dtype=_lib.BF16, device="cuda", keep_f32=False, plan_opts=None):
    """One resident whole-graph SpMM per spec: graph, plan, B (storage dtype),
    scales (FP8), C (FP32), and the FP32 B when keep_f32."""
    import torch
    probs = []
    gen = torch.Generator(device=device)
    for (log_n, deg, dim, gseed) in specs:
        g = Graph.rmat(log_n, deg, seed=gseed)
        plan = Plan(g.partition(1, 0), **(plan_opts or {}))
        gen.manual_seed(gseed)
        Bf = torch.rand((g.n, dim), generator=gen, device=device, dtype=torch.float32).mul_(2).sub_(1)
        Bq, sc = quantize_rows(Bf, dtype)
        C = torch.zeros((g.n, dim), device=device, dtype=torch.float32)
        p = dict(graph=g, plan=plan, B=Bq, scales=sc, C=C, dim=dim, n=g.n, nnz=plan.nnz, U=plan.U_l)
        if keep_f32:
            p["Bf"] = Bf
        probs.append(p)
    return probs


def bytes_alg(p, s_B):
    """Algorithmic bytes of one whole-graph SpMM: row pointer, entries,
    distinct B rows once, C once."""
    return 4 * (p["n"] + 1) + p["nnz"] * 8 + p["U"] * p["dim"] * s_B + p["n"] * p["dim"] * 4


def poisson_trace(rate_per_s, duration_s, n_resident, seed=0):
    """Arrival times (us, sorted) of a Poisson stream and uniformly drawn resident ids."""
    rng = np.random.default_rng(seed)
    n = int(rate_per_s * duration_s * 1.2) + 16
    gaps = rng.exponential(1e6 / rate_per_s, n)
    t = np.cumsum(gaps)
    t = t[t <= duration_s * 1e6]
    return t, rng.integers(0, n_resident, len(t)).astype(np.int32)


def latency_stats(recs, wall_us):
    lat = (recs["done_us"] - recs["arrival_us"]) / 1e3
    return {"requests": int(len(recs)), "wall_s": wall_us / 1e6,
            "req_per_s": len(recs) / (wall_us / 1e6) if wall_us > 0 else 0.0,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
            "mean_batch": float(recs["batch_size"].mean())}
