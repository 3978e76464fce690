"""Writes the golden fixtures of tests/test_golden.py (run from the repository
root: python tests/golden/make_fixtures.py).

- configs.json: the synthetic configurations as generated here (seeded R-MAT,
  symmetrised, GCN values): total nnz, max degree, a CRC32 of the CSR arrays
  and, per row partition, nnz_l, nnz_r, L, U_r, U_l, rows.  C1 = 2^16 nodes,
  degree 16, seed 1, 4 partitions; C2 = 2^22 nodes, degree 32, seed 2,
  community order, 8 partitions (the benchmarked workload); C2_cm = the same
  graph in Cuthill-McKee order.
- par_small.mtx + par_small.npz: a small MatrixMarket graph (duplicates, a
  hub row, a symmetric half) and the int32 partition files the reference's
  writer produces for it (prep_torch_par.py:14-101, restated with the same
  scipy routines in oracle.par_files_from_mtx) for 2 and 3 GPUs.
- ppl_small.npz: the stage files prep_torch_ppl.py:24-49 writes for the same
  graph (oracle.ppl_files_from_mtx) for (nGPU, group size) = (4, 1), (4, 2),
  (8, 2).
"""
import json
import os
import sys
import zlib

import numpy as np
import scipy.io as sio
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import cola_amd as cola  # noqa: E402
from oracle import oracle as orc  # noqa: E402


def describe(g, nGPU):
    rp, ci, va = g.arrays()
    n, nnz, md = g.info()
    crc = zlib.crc32(rp.astype(np.int64).tobytes())
    crc = zlib.crc32(ci.astype(np.int32).tobytes(), crc)
    parts = []
    for rank in range(nGPU):
        i = g.partition(nGPU, rank).info()
        parts.append({k: int(i[k]) for k in ("num_rows", "L", "U_r", "U_l", "nnz_l", "nnz_r", "max_row_nnz")})
    return {"nodes": int(n), "nnz": int(nnz), "max_degree": int(md), "csr_crc32": int(crc), "partitions": parts}


def main():
    configs = {
        "C1": dict(describe(cola.Graph.rmat(16, 16, seed=1), 4), log_n=16, degree=16, seed=1, reorder=False, nGPU=4),
        "C2": dict(describe(cola.Graph.rmat(22, 32, seed=2, reorder="community"), 8), log_n=22, degree=32,
                   seed=2, reorder="community", nGPU=8),
        "C2_cm": dict(describe(cola.Graph.rmat(22, 32, seed=2, reorder="cm"), 8), log_n=22, degree=32, seed=2,
                      reorder="cm", nGPU=8),
    }
    with open(os.path.join(HERE, "configs.json"), "w") as f:
        json.dump(configs, f, indent=1)
    rng = np.random.default_rng(20)
    n, m = 211, 1500
    r = rng.integers(0, n, m)
    c = rng.integers(0, n, m)
    r[:120] = 3  # a hub row
    A = sp.coo_matrix((np.ones(m), (r, c)), shape=(n, n))
    path = os.path.join(HERE, "par_small.mtx")
    sio.mmwrite(path, A)
    arrays = {}
    for nGPU in (2, 3):
        ref = orc.par_files_from_mtx(path, nGPU)
        for k, v in ref.items():
            name = k if isinstance(k, str) else f"{k[0]}_{nGPU}_{k[1]}"
            arrays[name] = v
    np.savez(os.path.join(HERE, "par_small.npz"), **arrays)
    stages = {}
    for nGPU, gs in ((4, 1), (4, 2), (8, 2)):
        for s, cols in orc.ppl_files_from_mtx(path, nGPU, gs).items():
            stages[f"colIdx_{nGPU}_{gs}_{s}"] = cols
    np.savez(os.path.join(HERE, "ppl_small.npz"), **stages)
    print("wrote configs.json, par_small.mtx, par_small.npz, ppl_small.npz")


if __name__ == "__main__":
    main()
