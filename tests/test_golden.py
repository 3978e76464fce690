"""Golden fixtures (tests/golden/, written by tests/golden/make_fixtures.py).

- The synthetic configurations regenerate bit-identically: the R-MAT
  generator, symmetrisation, GCN values, the community and Cuthill-McKee
  orders and the row partition are deterministic for a seed whatever the
  host thread count, so benchmark and profile figures refer to a fixed
  graph.
- A small MatrixMarket graph read and serialised by the library gives the
  int32 partition files that the reference's writer (prep_torch_par.py,
  restated with its own scipy routines) produced for it, byte for byte,
  without scipy at test time, and the stage files of prep_torch_ppl.py.
"""
import json
import os
import zlib

import numpy as np
import pytest

import cola_amd as cola

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def check_config(name):
    cfg = json.load(open(os.path.join(HERE, "configs.json")))[name]
    g = cola.Graph.rmat(cfg["log_n"], cfg["degree"], seed=cfg["seed"], reorder=cfg["reorder"])
    n, nnz, md = g.info()
    assert (n, nnz, md) == (cfg["nodes"], cfg["nnz"], cfg["max_degree"])
    rp, ci, _ = g.arrays()
    crc = zlib.crc32(ci.astype(np.int32).tobytes(), zlib.crc32(rp.astype(np.int64).tobytes()))
    assert crc == cfg["csr_crc32"]
    for rank, want in enumerate(cfg["partitions"]):
        info = g.partition(cfg["nGPU"], rank).info()
        assert {k: int(info[k]) for k in want} == want, rank


def test_c1_configuration_is_reproduced():
    check_config("C1")


def test_c2_configuration_is_reproduced():
    """The benchmarked 2^22-node graph in community order (about 30 s on 8
    host cores)."""
    check_config("C2")


def test_c2_cm_configuration_is_reproduced():
    """The same graph in Cuthill-McKee order (about 10 s)."""
    check_config("C2_cm")


@pytest.mark.parametrize("nGPU,gs", [(4, 1), (4, 2), (8, 2)])
def test_stage_files_of_an_mtx_graph(tmp_path, nGPU, gs):
    """The stage files (prep_torch_ppl.py:24-49) of the same small graph."""
    ref = np.load(os.path.join(HERE, "ppl_small.npz"))
    g = cola.load_mtx("par_small", HERE)
    cola.serialize_ppl(g, str(tmp_path), "par_small", [nGPU], [gs])
    S = nGPU // gs
    for s in range(S):
        got = (tmp_path / "par_small" / f"ppl{S}" / f"colIdx_{nGPU}_{S}_{s}.bin").read_bytes()
        assert got == ref[f"colIdx_{nGPU}_{gs}_{s}"].tobytes(), s


@pytest.mark.parametrize("nGPU", [2, 3])
def test_partition_files_of_an_mtx_graph(tmp_path, nGPU):
    ref = np.load(os.path.join(HERE, "par_small.npz"))
    g = cola.load_mtx("par_small", HERE)
    cola.serialize(g, str(tmp_path), "par_small", [nGPU], write_values=False)
    gdir = tmp_path / "par_small"
    assert (gdir / "global_rowPtr.bin").read_bytes() == ref["global_rowPtr"].tobytes()
    assert (gdir / "global_colIdx.bin").read_bytes() == ref["global_colIdx"].tobytes()
    for rank in range(nGPU):
        for stem in ("csr_rowPtr_l", "csr_colIdx_l", "csr_rowIdx_l", "csc_colPtr_r", "csc_rowIdx_r", "csc_colIdx_r"):
            got = (gdir / f"GPU{nGPU}" / f"{stem}_{nGPU}_{rank}.bin").read_bytes()
            assert got == ref[f"{stem}_{nGPU}_{rank}"].tobytes(), (rank, stem)
