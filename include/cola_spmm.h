/*
 * cola_spmm.h — C ABI of the MI355X (gfx950) single-partition SpMM engine.
 *
 * One row partition p of a row-partitioned C = A·B: the partition owns rows
 * [p·P, min(n,(p+1)·P)) of A (P = nodesPerPE = ceil(n/nGPU)) and the same rows
 * of B.  A's row panel is stored as a local CSR ("zip": empty rows collapsed,
 * columns in [0,P) relative to the partition) plus a remote CSC over the
 * sorted unique remote global columns.  The device computes
 *
 *     C[0:M, :] (+)= A_local · B_p  +  A_remote · B
 *
 * Conventions (all entry points):
 *   - every function returns a cola_status (0 = COLA_OK); the text of the
 *     last failure on the calling thread is returned by cola_last_error().
 *     Nothing here calls exit().
 *   - device pointers are caller-owned unless a handle owns them; handles own
 *     what their create call allocated and free it in destroy.
 *   - hipStream_t is passed as a plain pointer (void*, NULL = default stream);
 *     every launch is stream-ordered and asynchronous; the caller synchronises.
 *   - dense matrices are row-major with a leading dimension (row stride) in
 *     elements; rows are addressed with 64-bit offsets.
 *   - a handle may be used from several threads for reads; concurrent
 *     launches that share a plan's hub workspace on different streams race.
 *
 * Reference interfaces each group replaces are cited per declaration
 * (paths relative to the CoLa source tree).
 */
#ifndef COLA_SPMM_H
#define COLA_SPMM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int cola_status;
enum {
  COLA_OK = 0,
  COLA_ERR_INVALID = 1,     /* bad argument, shape or inconsistent arrays   */
  COLA_ERR_IO = 2,          /* missing / truncated / mis-sized file          */
  COLA_ERR_HIP = 3,         /* HIP runtime error (launch, alloc, copy)       */
  COLA_ERR_UNSUPPORTED = 4, /* dim / dtype combination not provided          */
  COLA_ERR_NOMEM = 5        /* host allocation failed or budget exhausted    */
};

/* Storage type of the dense B rows (accumulation is always FP32). */
typedef enum {
  COLA_F32 = 0,
  COLA_BF16 = 1,
  COLA_F16 = 2,
  COLA_FP8_E4M3 = 3 /* OCP e4m3fn bytes + one FP32 scale per row */
} cola_dtype;

/* Flags of cola_spmm_partition / cola_dist_spmm. */
enum {
  COLA_ACCUMULATE = 0,   /* C += A_p·B   (reference semantics, cuh:141,196)  */
  COLA_OVERWRITE = 1,    /* C  = A_p·B   (rows without nnz are zeroed)       */
  COLA_KAHAN = 2         /* compensated FP32 sums on hub rows                */
};

const char *cola_last_error(void);
const char *cola_version(void);

/* ------------------------------------------------------------------------ */
/* Host-side graph / partition data (no GPU needed).                         */
/* ------------------------------------------------------------------------ */

/* Global CSR graph held in host memory (int32 indices, FP32 values). */
typedef struct cola_graph_s *cola_graph;

/* Seeded R-MAT graph: n = 2^log_n nodes, round(n·avg_degree/2) drawn edges
 * with quadrant probabilities (a, b, c, 1-a-b-c), each edge added in both
 * directions, self-loops and duplicates dropped, values 1/sqrt(d_i·d_j)
 * (GCN normalisation) when gcn_values != 0, else 1.0.
 * reorder: 0 = natural R-MAT ids, 1 = Cuthill-McKee locality order,
 * 2 = community order (COLA_ORDER_COMMUNITY).
 * Deterministic for a given seed whatever the host thread count. */
cola_status cola_rmat_graph(int log_n, double avg_degree, double a, double b,
                            double c, uint64_t seed, int reorder,
                            int gcn_values, cola_graph *out);

/* Global CSR from caller arrays (copied).  Columns within a row may come in
 * any order and may repeat: duplicates are summed, as scipy's tocsr() does
 * (prep_torch_par.py:93-101). */
cola_status cola_graph_from_csr(int64_t n, const int64_t *row_ptr,
                                const int32_t *col_idx, const float *val,
                                cola_graph *out);
cola_status cola_graph_info(cola_graph g, int64_t *n, int64_t *nnz,
                            int64_t *max_degree);
/* Borrowed pointers, valid until cola_graph_destroy. */
cola_status cola_graph_arrays(cola_graph g, const int64_t **row_ptr,
                              const int32_t **col_idx, const float **val);
cola_status cola_graph_destroy(cola_graph g);

/* Value policy when a graph is read from a MatrixMarket file. */
enum {
  COLA_MTX_VALUES_FILE = 0, /* the file's values (pattern files: 1.0)        */
  COLA_MTX_VALUES_ONE = 1,  /* every stored entry 1.0 (drv:83,99 override)    */
  COLA_MTX_VALUES_GCN = 2   /* 1/sqrt(d_i·d_j) over the deduplicated pattern */
};

/* Reads a MatrixMarket coordinate file into a global CSR graph: what
 * load_mtx (prep_torch_par.py:14-20, scipy.io.mmread) followed by the
 * writer's tocsr() conversions yields.  Fields real / double / integer /
 * pattern; symmetry general / symmetric (off-diagonal entries mirrored) /
 * skew-symmetric (mirrored negated).  Indices are 1-based in the file.
 * Duplicates are summed in FP64 and rounded once to FP32.  The matrix must
 * be square (rows = graph nodes).  Parsed in parallel on the host cores.
 * COLA_ERR_IO: missing / unreadable / truncated file;
 * COLA_ERR_UNSUPPORTED: array format, complex or hermitian fields, or a
 * non-square matrix; COLA_ERR_INVALID: malformed header, size line or entry,
 * index out of range, entry count different from the size line. */
cola_status cola_graph_from_mtx(const char *path, int values, cola_graph *out);

/* Writes the graph as "%%MatrixMarket matrix coordinate real general" (or
 * "pattern general" when with_values == 0), 1-based, row-major order, values
 * printed with 9 significant digits (FP32 round-trips exactly).  The writing
 * half of rabbit_order_write.py:60-77 (scipy.io.mmwrite of the reordered
 * graph). */
cola_status cola_graph_write_mtx(cola_graph g, const char *path,
                                 int with_values);

/* Locality orders (the reference reorders with Rabbit Order,
 * rabbit_order_write.py:46-55; this build's orders are its own). */
enum {
  COLA_ORDER_CUTHILL_MCKEE = 1, /* BFS from min-degree vertices, neighbours
                                   by (degree, id); isolated vertices last */
  COLA_ORDER_REVERSE_CM = 2,    /* the Cuthill-McKee order reversed         */
  COLA_ORDER_DEGREE = 3,        /* degree descending, ties by id           */
  COLA_ORDER_COMMUNITY = 4      /* Rabbit Order's two steps (incremental
                                   modularity aggregation in increasing
                                   degree order, then a depth-first walk of
                                   the merge trees): each community gets a
                                   contiguous id range; isolated last      */
};

/* Relabels the graph (values carried along): new vertex i = old vertex
 * new2old[i].  new2old (n entries) may be NULL. */
cola_status cola_graph_reorder(cola_graph g, int method, int32_t *new2old,
                               cola_graph *out);

/* One rank's partition in the prep_torch_par.py:24-79 layout, host memory.
 * Array lengths: csr_rowPtr_l L+1, csr_colIdx_l/csr_val_l nnz_l,
 * csr_rowIdx_l L, csc_colPtr_r U_r+1, csc_rowIdx_r/csc_val_r nnz_r,
 * csc_colIdx_r U_r. */
typedef struct cola_partition_s *cola_partition;
typedef struct {
  int64_t num_nodes;   /* n                                      */
  int32_t nGPU;        /* number of row partitions               */
  int32_t rank;        /* this partition                         */
  int32_t nodesPerPE;  /* P = ceil(n / nGPU)                     */
  int32_t num_rows;    /* M = min(n, (rank+1)·P) − rank·P        */
  int32_t L;           /* non-empty local rows                   */
  int32_t U_r;         /* unique remote columns                  */
  int64_t nnz_l;       /* local nnz                              */
  int64_t nnz_r;       /* remote nnz                             */
  int32_t U_l;         /* distinct local columns referenced      */
  int32_t max_row_nnz; /* max (local+remote) nnz of one row      */
} cola_partition_info;

typedef struct {
  const int32_t *csr_rowPtr_l, *csr_colIdx_l, *csr_rowIdx_l;
  const float *csr_val_l;
  const int32_t *csc_colPtr_r, *csc_rowIdx_r, *csc_colIdx_r;
  const float *csc_val_r;
} cola_partition_arrays;

/* prep_torch_par.py:24-79 row_partition for one rank, in memory. */
cola_status cola_row_partition(cola_graph g, int nGPU, int rank,
                               cola_partition *out);
/* Partition from caller arrays (copied, validated like a load). */
cola_status cola_partition_from_arrays(int64_t num_nodes, int nGPU, int rank,
                                       const cola_partition_arrays *arrs,
                                       int64_t L, int64_t nnz_l, int64_t U_r,
                                       int64_t nnz_r, cola_partition *out);
cola_status cola_partition_get_info(cola_partition p, cola_partition_info *info);
cola_status cola_partition_get_arrays(cola_partition p,
                                      cola_partition_arrays *arrs);
cola_status cola_partition_destroy(cola_partition p);

/* prep_torch_par.py:93-101 serialize: writes <bin_path>/<graph>/global_rowPtr.bin,
 * global_colIdx.bin and, for every rank, GPU<nGPU>/{csr_rowPtr_l,csr_colIdx_l,
 * csr_rowIdx_l,csc_colPtr_r,csc_rowIdx_r,csc_colIdx_r}_<nGPU>_<rank>.bin
 * (raw little-endian int32).  write_values != 0 adds csr_val_l_* / csc_val_r_*
 * (FP32) value files, which the reference does not have. */
cola_status cola_serialize(cola_graph g, const char *bin_path,
                           const char *graph_name, int nGPU, int write_values);

/* prep_torch_ppl.py:24-49 row_partition: the stage files of the pipelined
 * multi-stage path.  With S = nGPU / group_size stages and P =
 * ceil(n/nGPU), stage s covers rows [s*group_size*P, min((s+1)*group_size*P,
 * n)); <bin_path>/<graph_name>/ppl<S>/colIdx_<nGPU>_<S>_<s>.bin holds the
 * sorted unique columns of those rows (int32), the files
 * distSpMM_ppl_batch.cu:102-104 reads.  Needs 1 <= group_size < nGPU (the
 * reference writes nothing otherwise, prep_torch_ppl.py:64). */
cola_status cola_serialize_ppl(cola_graph g, const char *bin_path, const char *graph_name, int nGPU,
                               int group_size);

/* distSpMM_ppl_batch.cu:49-99 partition load: reads global_rowPtr.bin (for n)
 * and the six per-rank files with size and range checks (bin2vec,
 * dist_spmm_ppl_batch.cuh:7-27, without its 2 GiB int limit); value files are
 * used when present, else every value is 1.0 (distSpMM_ppl_batch.cu:83,99). */
cola_status cola_partition_load(const char *bin_path, const char *graph_name,
                                int nGPU, int rank, cola_partition *out);

/* dist_spmm_ppl_batch.cuh:29-56 nnz_split.  rowIdx (nrows) and rowPtr
 * (nrows+1) describe a zip CSR/CSC with nnz entries; rows are cut into
 * virtual rows of gran = ceil((nnz/nrows)/32)·32 entries.  Two-call protocol:
 * pass target=NULL/ptr=NULL to get *n_virtual, then arrays of n_virtual and
 * n_virtual+1 entries. */
cola_status cola_nnz_split(const int32_t *rowIdx, int64_t nrows,
                           const int32_t *rowPtr, int64_t nnz,
                           int32_t *target, int32_t *ptr, int64_t *n_virtual,
                           int32_t *gran);

/* ------------------------------------------------------------------------ */
/* Device data (sparse_data.h:6-77 d_csr / d_csr_zip / d_csc_zip).           */
/* ------------------------------------------------------------------------ */

/* Device copy of a zip CSR (d_csr_zip, sparse_data.h:53-77) or zip CSC
 * (d_csc_zip, sparse_data.h:28-52): hipMalloc + synchronous H2D of the four
 * arrays.  For the CSC, colIdx_zip holds global remote column ids and the
 * handle also stores, per (virtual) column, the row of the packed remote
 * buffer it reads (the position of its column among the unique ids). */
typedef struct cola_zip_s *cola_zip;
cola_status cola_zip_create(const int32_t *ptr, int64_t nptr,
                            const int32_t *idx, int64_t nnz,
                            const int32_t *zip, const float *val,
                            cola_zip *out);
/* Device pointers (borrowed): ptr, idx, zip, val and, for a CSC, pack_idx. */
cola_status cola_zip_device_arrays(cola_zip z, int32_t **ptr, int32_t **idx,
                                   int32_t **zip, float **val,
                                   int32_t **pack_idx, int64_t *nptr,
                                   int64_t *nnz, int64_t *n_unique);
cola_status cola_zip_destroy(cola_zip z);

/* ------------------------------------------------------------------------ */
/* Hot path.                                                                  */
/* ------------------------------------------------------------------------ */

/* Gather-pack of the unique remote B rows (replaces the per-warp NVSHMEM get,
 * dist_spmm_ppl_batch.cuh:120-125, and the request-coalescing pack kernel
 * batching_data_ppl, dist_spmm_ppl_batch.cuh:58-73):
 *   packed[u, 0:dim] = B_global[remote_cols[u], 0:dim],  u in [0, num_cols)
 * B_global holds every partition's slice back to back (row g of the graph is
 * row g of B_global).  For COLA_FP8_E4M3 the per-row scales travel too
 * (scales_global[g] -> packed_scales[u]); pass NULL for other dtypes. */
cola_status cola_pack_remote(const void *B_global, int64_t ld_B,
                             const float *scales_global,
                             const int32_t *remote_cols, int64_t num_cols,
                             int dim, cola_dtype dtype, void *packed,
                             int64_t ld_packed, float *packed_scales,
                             void *stream);

/* Reference-semantics partition SpMM (replaces ppl0_batch_SpMM<CF,res>,
 * dist_spmm_ppl_batch.cuh:202-245, and its kernel cuh:76-200): virtual local
 * rows [local_row_lb, local_row_lb+local_row_num) of a zip CSR and virtual
 * remote columns [remote_col_lb, remote_col_lb+remote_col_num) of a zip CSC
 * are accumulated into `output` (P × dim, row-major, ld = dim) with FP32
 * atomics.  Local columns index `input` (this partition's B slice, ld = dim);
 * remote column v reads packed row csc_packIdx_r[v] (filled by
 * cola_pack_remote over the unique remote columns).  Any dim >= 1 is
 * computed (the reference computes only dim <= 64 and 97..128). */
cola_status cola_ppl0_batch_spmm(
    float *output, const float *input, const float *packed,
    const int32_t *csr_rowPtr_l, const int32_t *csr_colIdx_l,
    const int32_t *csr_rowIdx_zip_l, const float *csr_val_l,
    const int32_t *csc_colPtr_r, const int32_t *csc_rowIdx_r,
    const int32_t *csc_packIdx_r, const float *csc_val_r, int local_row_lb,
    int remote_col_lb, int local_row_num, int remote_col_num, int dim,
    int nodesPerPE, int peid, void *stream);

/* Single-stage distSpMM_ppl_batch_wrapper (dist_spmm_ppl_batch.cuh:851-918
 * with group_size == nranks <= 8, cuh:837-848): output += A_p·B, with the
 * unique remote rows of `csc_zip_r` packed from B_global into `packed`
 * (n_unique x dim) and input = B_global + peid·nodesPerPE·dim.  The same sums
 * as ppl0_batch_SpMM over the virtual rows / columns, computed through a
 * degree-binned plan rebuilt from the two zips on the first call and cached
 * in csr_zip_l (no atomics, deterministic order); cola_ppl0_batch_spmm is the
 * literal virtual-row kernel.  The plan copies the zips' arrays at that
 * first call: after writing new values or indices into a zip's device arrays,
 * call cola_zip_invalidate on the local zip (or create new zips).  Calls on
 * one zip from several threads are serialised while the plan is built, and
 * each call holds its own reference to the plan until its launches are
 * enqueued (a concurrent rebuild or invalidate frees it only afterwards);
 * concurrent launches race on its hub workspace. */
cola_status cola_dist_spmm_ppl_batch(float *output, const float *B_global,
                                     float *packed, cola_zip csr_zip_l,
                                     cola_zip csc_zip_r, int dim,
                                     int nodesPerPE, int peid, void *stream);

/* ppl0_batch_SpMM<CF,is_residue> with the reference's own parameter list
 * (dist_spmm_ppl_batch.cuh:202-221; call sites cuh:841-846 and cuh:799-804):
 * the zips' device arrays (d_csr_zip / d_csc_zip members), the virtual-row
 * range [local_row_lb, +local_row_num) of the local zip CSR and the virtual-
 * column range [remote_col_lb, +remote_col_num) of the remote zip CSC, whose
 * csc_colIdx_zip_r holds global column ids.  `input` is this partition's
 * slice of a B_global that holds every partition back to back, so remote
 * global row g is read at input + (g - peid*nodesPerPE)*dim (on one GPU the
 * NVSHMEM get of cuh:120-125 becomes a plain load; no pack is needed).
 * output (nodesPerPE x dim, ld = dim) is accumulated into, as there.
 * literal = 0 (the default of the C++ header): the sums run through a
 * degree-binned plan rebuilt from the arrays on the first call with these
 * pointers and ranges, then cached (no atomics, deterministic order);
 * literal = 1: the virtual-row kernel with FP32 atomics (cuh:76-200).
 * Any dim >= 1 is computed (the reference computes only dim <= 64 and
 * 97..128).  The cache is keyed by the eight array pointers and the six
 * integer arguments; destroying a zip (cola_zip_destroy, the sparse_data.h
 * destructors) drops the plans built from its arrays.  After rewriting the
 * arrays behind cached pointers, or freeing arrays that no zip owns, call
 * cola_ppl0_plans_clear.  Remote columns must lie outside the partition's
 * own rows [peid*P, (peid+1)*P) (prep_torch_par.py:40-79 puts those in the
 * local CSR). */
cola_status cola_ppl0_batch_spmm_global(
    float *output, const float *input, const int32_t *csr_rowPtr_l,
    const int32_t *csr_colIdx_l, const int32_t *csr_rowIdx_zip_l,
    const float *csr_val_l, const int32_t *csc_colPtr_r,
    const int32_t *csc_rowIdx_r, const int32_t *csc_colIdx_zip_r,
    const float *csc_val_r, int local_row_lb, int remote_col_lb,
    int local_row_num, int remote_col_num, int dim, int nodesPerPE, int peid,
    int literal, void *stream);
/* Drops every plan cola_ppl0_batch_spmm_global cached; each is freed once
 * no call in flight holds it.  No counterpart in the reference (its launcher
 * keeps no state). */
cola_status cola_ppl0_plans_clear(void);
/* `launches` launches of an empty 1024 x 256 kernel (warmup_kernel,
 * utils.cuh:24-30, launched 50 times by the harness at cuh:966-969). */
cola_status cola_warmup(int launches, void *stream);
/* Drops the plan cola_dist_spmm_ppl_batch cached in a (local) zip; the next
 * call rebuilds it from the zips' current arrays.  No counterpart in the
 * reference (whose wrapper reads the zips on every call). */
cola_status cola_zip_invalidate(cola_zip csr_zip_l);

/* Degree-binned, atomic-free partition plan (replaces nnz_split's virtual
 * rows for the fast path): for every owned row, local nnz and remote nnz are
 * merged into one row list whose entries address either B_p (local column)
 * or the packed remote buffer (unique-column position).  Rows above
 * hub_threshold nnz are cut into chunks of chunk_nnz entries whose partial
 * sums are reduced in a fixed order by a second kernel, so results do not
 * depend on scheduling.  Work items are laid out so that each of the 8 XCDs
 * walks a contiguous row range. */
typedef struct cola_plan_s *cola_plan;
typedef struct {
  int32_t hub_threshold; /* rows with more nnz are split (0 = default 1024)  */
  int32_t chunk_nnz;     /* nnz per hub chunk (0 = default 512)              */
  int32_t kahan_degree;  /* rows above this degree use compensated sums when
                            COLA_KAHAN is passed (0 = default 1024)         */
  int32_t max_dim;       /* largest dim the hub workspace is sized for
                            (0 = default 256)                               */
  int32_t short_degree;  /* rows with at most this many entries form the
                            short-row bin (0 = default 32)                  */
  int32_t col_panels;    /* K > 1: also build a column-panel schedule (local
                            columns and packed rows each cut into K ranges,
                            one pass per range) for cache reuse; 0 = none   */
  int32_t row_order;     /* order of work items: 0 = row order, 1 = rows by
                            decreasing entry count                          */
  int32_t hot_cols;      /* H > 0: the H most referenced B rows (local or
                            packed) are "hot"; each row's hot entries run in
                            a first pass that writes C and its cold entries
                            in a second pass that adds onto it, so the hot
                            rows stay in L2 during the first pass.  Not with
                            col_panels or the batched form; 0 = one pass   */
} cola_plan_opts;
cola_status cola_plan_create(cola_partition p, const cola_plan_opts *opts,
                             cola_plan *out);
typedef struct {
  int64_t num_rows, num_items, num_hub_rows, num_hub_chunks, nnz, U_r, U_l;
  int64_t device_bytes;
  int64_t num_short_items, num_long_items;
  int64_t hot_nnz;       /* entries of the first (hot) pass; nnz if one pass */
  int64_t num_cold_items;/* items of the second pass                         */
  int64_t num_tiny_bundles; /* bundles of rows with at most 4 entries        */
} cola_plan_info;
cola_status cola_plan_get_info(cola_plan plan, cola_plan_info *info);

/* Launch parameters of the plan's kernels (the autotuner's search space).
 * A new plan with at least 2^22 entries starts with {-2, 0, 4, 2, 1, 0, 8,
 * 0, 4, 1}: 256-byte column tiles, 4 rows per wave on long rows, 2 blocks per
 * CU (4 for the short-row bin), XCD remap, 8 loads in flight, tiny-row
 * bundles (the tuned schedule of a partition of the community-ordered
 * 2^22-node R-MAT graph); a smaller one with {0, 0, 1, 8, 1, 0, 8, 0, 0, 0}:
 * all columns in one pass, one row per wave on long rows, 8 blocks per CU,
 * no bundles (on a small plan their extra launch costs more than it saves). */
typedef struct {
  int32_t col_tile;      /* dense columns per pass over B (0 = all of dim;
                            -k, k = 1..8: k x 128 bytes of the storage
                            dtype); rounded down to the 16-byte vector
                            width                                           */
  int32_t rpw_short;     /* short-row bin: rows in flight per wave (power of
                            two; 0 = one row per lane group)                */
  int32_t rpw_long;      /* long-row bin: rows in flight per wave (0 = 1)   */
  int32_t blocks_per_cu; /* grid size in 256-thread blocks per CU (0 = 8)   */
  int32_t xcd_remap;     /* 1: each XCD walks a contiguous row range; 0:
                            plain block order                               */
  int32_t split_remote;  /* 1: local entries first (B_p only), then remote
                            entries (packed only) added on top; 0: both in
                            one pass (default)                              */
  int32_t unroll;        /* row-gather loads in flight per lane group: 8
                            (also 0) or 16                                  */
  int32_t panels;        /* 1: run the column-panel schedule (plans built
                            with col_panels > 1): C is zeroed (overwrite),
                            then one accumulating pass per panel; 0: off    */
  int32_t blocks_per_cu_short; /* grid of the short-row bin's launches
                            (0 = blocks_per_cu): short rows are chains of
                            dependent loads and want more waves in flight
                            than long rows                                  */
  int32_t tiny;          /* 1: rows of at most 4 entries run as bundles of
                            floor(8/L) rows of L entries per lane group (one
                            batch of row loads, no padding slots), when the
                            schedule is one pass over 16-byte row vectors and
                            a lane group spans at least 8 lanes; 0: those
                            rows stay in the short-row bin                  */
} cola_plan_tuning;
cola_status cola_plan_set_tuning(cola_plan plan, const cola_plan_tuning *t);
cola_status cola_plan_get_tuning(cola_plan plan, cola_plan_tuning *t);

/* Project-written autotuner: times candidate launch parameters on the
 * caller's buffers with HIP events — a grid over the coupled knobs (column
 * tile x grid size x long-bin rows per wave x XCD remap x loads in
 * flight), then coordinate descent over every knob, then three interleaved
 * re-timings of the five fastest distinct candidates — keeps the fastest in
 * the plan and returns it (and its time).  C is used as scratch: on return it holds A_p·B
 * (overwrite).  Blocks the calling thread. */
cola_status cola_plan_autotune(cola_plan plan, float *C, int64_t ld_C,
                               const void *B_local, int64_t ld_B,
                               const float *scales_local, const void *packed,
                               int64_t ld_packed, const float *packed_scales,
                               int dim, cola_dtype dtype, int flags, int iters,
                               void *stream, cola_plan_tuning *best,
                               float *best_ms);
/* Device array of the unique remote column ids (U_r entries), for packing. */
cola_status cola_plan_remote_cols(cola_plan plan, const int32_t **d_cols);
cola_status cola_plan_destroy(cola_plan plan);

/* C[0:M,:] (+)= A_p·B through a plan.  B_local: this partition's slice
 * (P rows, ld_B); packed: the U_r packed remote rows (ld_packed).  For
 * COLA_FP8_E4M3, scales_local / packed_scales are the per-row FP32 scales.
 * C is FP32, ld_C >= dim. */
cola_status cola_spmm_partition(cola_plan plan, float *C, int64_t ld_C,
                                const void *B_local, int64_t ld_B,
                                const float *scales_local, const void *packed,
                                int64_t ld_packed, const float *packed_scales,
                                int dim, cola_dtype dtype, int flags,
                                void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): cola_spmm_partition with the hub rows' chunk partial sums in a
 * fixed device array of COLA_NOALLOC_SCRATCH_BYTES that the library holds on
 * every GPU it runs on until the process ends, instead of the plan's own
 * workspace.  Same kernels, same summation order, so the result is bitwise
 * that of cola_spmm_partition.  Calls no hipMalloc / hipFree or allocating
 * routine.  A plan whose hub chunks need more than the array at this dim
 * (chunks x round_up(dim, 16) x 4 bytes) is rejected with COLA_ERR_INVALID.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (different streams or threads) race. */
#define COLA_NOALLOC_SCRATCH_BYTES (16 << 20)
cola_status cola_spmm_partition_noalloc(cola_plan plan, float *C, int64_t ld_C,
                                        const void *B_local, int64_t ld_B,
                                        const float *scales_local,
                                        const void *packed, int64_t ld_packed,
                                        const float *packed_scales, int dim,
                                        cola_dtype dtype, int flags,
                                        void *stream);

/* ------------------------------------------------------------------------ */
/* Batched form: many independent SpMMs in one persistent launch.            */
/* ------------------------------------------------------------------------ */

/* One problem of a batch: C (+)= A·B for the plan's partition (a whole graph
 * is the partition of nGPU = 1), with its own dim, B and C.  The batch keeps
 * the pointers; the caller keeps the buffers alive and unchanged in size. */
typedef struct {
  cola_plan plan;
  int32_t dim;                /* a multiple of the 16-byte vector width
                                 (4 FP32, 8 BF16/FP16, 16 FP8), <= 64 vectors */
  const void *B;              /* local rows, ld_B elements per row           */
  int64_t ld_B;
  const float *scales;        /* FP8 row scales of B, else NULL              */
  const void *packed;         /* packed remote rows (partitions), or NULL    */
  int64_t ld_packed;
  const float *packed_scales;
  float *C;                   /* FP32 output, ld_C per row                   */
  int64_t ld_C;
} cola_batch_problem;
typedef struct cola_batch_s *cola_batch;
typedef struct {
  int32_t col_tile;  /* dense columns per pass (0 = all)                     */
  int32_t rpw_short; /* short-row bin: rows per wave (0 = one per lane group) */
  int32_t rpw_long;  /* long-row bin: rows per wave (0 = 1)                  */
} cola_batch_tuning;

/* Builds the batch's device work list (segments of every problem's degree
 * bins cut into tasks of rows-per-wave items x column tiles) for one storage
 * dtype; flags: COLA_OVERWRITE and/or COLA_KAHAN (else accumulate).  Hub-row
 * workspace is allocated per problem, so problems sharing a plan do not race. */
cola_status cola_batch_create(const cola_batch_problem *probs, int n,
                              cola_dtype dtype, int flags, cola_batch *out);
/* Sets one problem's schedule (problem = -1: all) and/or the grid size
 * (blocks_per_cu > 0); waits for the device, then rebuilds the work list and
 * drops a captured graph.  The batch's hub workspace is shared by its runs:
 * runs of one batch on different streams at the same time race. */
cola_status cola_batch_set_tuning(cola_batch b, int problem,
                                  const cola_batch_tuning *t, int blocks_per_cu);
cola_status cola_batch_get_tuning(cola_batch b, int problem,
                                  cola_batch_tuning *t, int *blocks_per_cu);
/* Autotuner: coordinate descent over each dim class's column tile and
 * rows-per-wave of both bins, then the grid size, timed with HIP events;
 * last, the XCD ranges are re-cut by per-XCD times that the kernel records
 * (kept only if the batch runs faster; a later set_tuning restores the
 * modelled cut).  blocks_per_cu counts 4-wave blocks: the kernel runs that
 * many waves per CU, a CU's waves in one block where registers allow.
 * The outputs hold the batch's result on return. */
cola_status cola_batch_autotune(cola_batch b, int iters, void *stream,
                                float *best_ms);
/* One persistent launch over every task (+ one launch that reduces the
 * hub rows of every problem). */
cola_status cola_batch_run(cola_batch b, void *stream);
/* Captures cola_batch_run into a hipGraph held by the batch (stream must not
 * be the default stream); cola_batch_replay launches it. */
cola_status cola_batch_capture(cola_batch b, void *stream);
cola_status cola_batch_replay(cola_batch b, void *stream);
cola_status cola_batch_destroy(cola_batch b);

/* ------------------------------------------------------------------------ */
/* Serving: continuous batching of SpMM requests over resident graphs.       */
/* ------------------------------------------------------------------------ */

typedef struct cola_server_s *cola_server;
typedef struct {
  int64_t memory_budget; /* device bytes for request outputs + hub workspace;
                            a request is admitted when its share fits      */
  int32_t max_batch;     /* requests per launch (0 = 64)                   */
  int32_t max_inflight;  /* launches in flight, one stream each (0 = 2)    */
  int32_t dtype;         /* storage dtype of every resident feature matrix */
  int32_t flags;         /* COLA_KAHAN or 0 (outputs are overwritten)      */
  int32_t plain;         /* 1: no batching, every request is its own launch
                            queued on one stream as it arrives (baseline)  */
  int32_t blocks_per_cu; /* grid of each launch (0 = 8); e.g. the value
                            cola_batch_get_tuning reports after autotuning */
} cola_server_opts;
typedef struct {
  double arrival_us, dispatch_us, done_us; /* from the start of the trace  */
  int32_t resident;                        /* which resident SpMM          */
  int32_t batch_size;                      /* requests in its launch       */
} cola_request_record;

cola_status cola_server_create(const cola_server_opts *opts, cola_server *out);
/* Registers a resident SpMM: a whole-graph plan and its feature matrix B
 * (device memory, kept by the caller).  Returns its id. */
cola_status cola_server_add_resident(cola_server s, cola_plan plan,
                                     const void *B, int64_t ld_B,
                                     const float *scales, int dim, int *id);
/* Schedule used for requests of one dim (e.g. from cola_batch_autotune). */
cola_status cola_server_set_tuning(cola_server s, int dim,
                                   const cola_batch_tuning *t);
/* Serves a request trace in real time: request i arrives arrival_us[i]
 * microseconds after the start (non-decreasing) and asks for C = A·B of
 * resident[i].  Arrived requests are admitted first-come first-served when
 * their output and workspace fit in the memory budget, and every free launch
 * slot takes all admitted requests (up to max_batch) into one persistent
 * launch, so requests arriving while others run join the next launch and
 * launches overlap on the device.  outputs[i] (or outputs NULL / entry NULL:
 * a pool buffer freed at completion) receives request i's C (ld = dim).
 * Fills records[i]; *wall_us = trace duration; *peak_bytes = pool peak. */
cola_status cola_server_run_trace(cola_server s, int64_t n,
                                  const double *arrival_us,
                                  const int32_t *resident,
                                  float *const *outputs,
                                  cola_request_record *records,
                                  double *wall_us, int64_t *peak_bytes);
cola_status cola_server_destroy(cola_server s);

/* Row-wise conversion of FP32 rows to a storage dtype (round to nearest
 * even; FP8 rows get scale = max|row|/448 and are stored as x/scale).
 * dst has ld_dst elements per row; scales (FP8 only) has rows entries. */
cola_status cola_quantize_rows(const float *src, int64_t ld_src, int64_t rows,
                               int dim, cola_dtype dtype, void *dst,
                               int64_t ld_dst, float *scales, void *stream);

/* On-device FP64 error norms of C (FP32, ld_C) against a reference R (FP64,
 * ld_R) over rows × dim: out[0] = Σ(C−R)², out[1] = ΣR², out[2] =
 * max over rows of ‖C_i−R_i‖²/‖R_i‖² (rows with R_i = 0 skipped).
 * out is 3 doubles of device memory. */
cola_status cola_error_norms(const float *C, int64_t ld_C, const double *R,
                             int64_t ld_R, int64_t rows, int dim, double *out,
                             void *stream);

/* ------------------------------------------------------------------------ */
/* Measurement.                                                              */
/* ------------------------------------------------------------------------ */

/* No counterpart in the reference.  The gather ceiling of a plan's entry
 * stream: a minimal FP32 kernel that reads each item's entries, loads the B
 * rows they name (16 bytes per lane, 8 rows in flight per lane group),
 * multiply-adds into registers and stores each row once, over the plan's
 * long-row then short-row items, one launch per column tile of col_tile
 * columns (0 = all; a multiple of 4, at most 256).  mode 0: the real rows;
 * 1: row indices masked into the first table_rows (a power of two, at most
 * the partition's rows and U_r) rows of each table, so the tables stay in
 * L2; 2: entry k reads local row k mod table_rows (table_rows at most the
 * partition's rows; B_local required): the same row loads in address
 * order, no gather.
 * Overwrites C (and the hub chunk partials, not reduced).  For timing only. */
cola_status cola_probe_gather(cola_plan plan, const float *B_local, int64_t ld_B,
                              const float *packed, int64_t ld_packed, float *C,
                              int64_t ld_C, int dim, int col_tile, int mode,
                              int64_t table_rows, int blocks_per_cu,
                              void *stream);

/* No counterpart in the reference.  The bare gather of a row stream: entry
 * k names a row starting at base + row_offs[k] * 16 bytes; each 16-lane group
 * loads the first 256 bytes of 8 consecutive entries' rows (16 bytes a lane,
 * all 8 loads in flight) and adds them up, its waves striding over the
 * whole stream or (xcd_ranges = 1) each XCD over one contiguous eighth of it,
 * through VGPRs (via_lds = 0) or
 * through LDS (via_lds = 1: global_load_lds_dwordx4).  Nothing else is read
 * or written but one float per block into out (out_len >= CUs x
 * blocks_per_cu, 0 = 2).  Every row must have 256 readable bytes and base
 * must be 16-byte aligned.  For timing only. */
cola_status cola_probe_row_stream(const uint32_t *row_offs, int64_t n, const void *base,
                                  int via_lds, int xcd_ranges, int blocks_per_cu,
                                  float *out, int64_t out_len, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* COLA_SPMM_H */
