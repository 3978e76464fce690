"""cola_amd — MI355X-native single-partition SpMM (one row partition of
CoLa's distributed C = A·B) behind a C ABI (include/cola_spmm.h).

Host objects (no GPU): Graph (R-MAT, CSR/COO, MatrixMarket, reorders),
Partition, serialize, load_mtx / process_graph / reorder_mtx, bin2vec,
nnz_split, the stage files serialize_ppl / process_graph_ppl / stage_cols.
Device path: d_csr_zip / d_csc_zip + ppl0_batch_SpMM / dist_spmm_ppl_batch
(reference semantics, FP32 atomics) and Plan (degree-binned, atomic-free,
FP32 / BF16 / FP16 / FP8-e4m3 storage of B).
"""
from ._lib import (ACCUMULATE, BF16, F16, F32, FP8_E4M3, KAHAN, OVERWRITE, ColaError,  # noqa: F401
                   LIB_PATH, lib)
from .partition import (Graph, Partition, bin2vec, load_mtx, nnz_split, process_graph,  # noqa: F401
                        process_graph_ppl, reorder_mtx, serialize, serialize_ppl, stage_cols)
from .spmm import (DTYPE_BYTES, DTYPE_NAMES, REQUEST_RECORD, Batch, Plan, Server, d_csc_zip, d_csr_zip, dist_spmm_ppl_batch,  # noqa: F401
                   error_norms, pack_remote, ppl0_batch_SpMM, ppl0_batch_SpMM_global, ppl0_plans_clear,
                   quantize_rows)

__version__ = "0.1.0"
