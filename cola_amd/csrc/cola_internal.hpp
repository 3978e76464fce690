// Internal structures shared by the host library (host.cpp) and the device
// side (spmm.hip).  Not part of the public ABI.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/cola_spmm.h"

namespace cola {

// Records the failure text for cola_last_error() and returns `code`.
cola_status fail(cola_status code, const std::string &msg);

struct Graph {
  int64_t n = 0;
  std::vector<int64_t> row_ptr;  // n+1
  std::vector<int32_t> col_idx;  // nnz, sorted and unique within each row
  std::vector<float> val;        // nnz
};

struct Partition {
  int64_t num_nodes = 0;
  int32_t nGPU = 0, rank = 0, nodesPerPE = 0, num_rows = 0;
  // prep_torch_par.py:24-79 layout.
  std::vector<int32_t> csr_rowPtr_l, csr_colIdx_l, csr_rowIdx_l;
  std::vector<float> csr_val_l;
  std::vector<int32_t> csc_colPtr_r, csc_rowIdx_r, csc_colIdx_r;
  std::vector<float> csc_val_r;
};

// Graph helpers shared by the generator and the MatrixMarket reader.
// assign_values: 1/sqrt(d_i d_j) when gcn, else 1.0 for every entry.
void assign_values(Graph *g, bool gcn);
// Cuthill-McKee order, new -> old.
std::vector<int32_t> cuthill_mckee(const Graph &g);
// Community order (Rabbit Order's aggregation + depth-first numbering,
// community.cpp), new -> old.
std::vector<int32_t> community_order(const Graph &g);
// out = in relabelled (new vertex i = old new2old[i]); values carried when
// in.val is non-empty; columns sorted within each row.
void relabel(const Graph &in, const std::vector<int32_t> &new2old, Graph *out);

// Validates the structural invariants of a partition (what the reference's
// writer asserts, prep_torch_par.py:50,53,71,74, plus index ranges).
cola_status validate_partition(const Partition &p);

// Host form of the degree-binned plan (see cola_plan_create).
struct WorkItem {  // 16 bytes, loaded as one int4 by the row kernel
  int32_t out;     // >= 0: output row; < 0: -(hub chunk + 1)
  int32_t begin;   // first entry in the merged entry list
  int32_t end;     // one past the last entry
  int32_t flags;   // bit 0: compensated (Kahan) summation allowed;
                   // bits 1..: number of local entries (they come first)
};

struct HostPlan {
  int32_t num_rows = 0, nodesPerPE = 0, U_r = 0, U_l = 0;
  int32_t max_dim = 0;
  int64_t nnz = 0;
  // Merged row entries: col >= 0 -> local column of B_p; col < 0 -> packed
  // remote row ~col.
  std::vector<int32_t> cols;
  std::vector<float> vals;
  // Degree bins: short rows (at most short_degree entries) and long rows
  // (the rest, hub rows as chunks), each in row order with a contiguous item
  // range per XCD (9 entries).
  std::vector<WorkItem> items_short, items_long;
  std::vector<int32_t> xcd_short, xcd_long;
  // The same items restricted to those with remote entries (second phase of
  // the local-then-remote schedule).
  std::vector<WorkItem> items_short_r, items_long_r;
  std::vector<int32_t> xcd_short_r, xcd_long_r;
  // Second (cold) pass of a hot/cold plan (hot_cols > 0): items over each
  // row's cold entries, added onto what the first pass wrote; empty otherwise.
  std::vector<WorkItem> items_short2, items_long2;
  std::vector<int32_t> xcd_short2, xcd_long2;
  int64_t hot_nnz = 0;
  std::vector<int32_t> empty_rows;   // rows without entries (OVERWRITE)
  std::vector<int32_t> hub_rows;     // output row of each hub
  std::vector<int32_t> hub_chunk_ptr;// chunk range per hub (num_hubs+1)
  std::vector<int32_t> remote_cols;  // unique remote global columns (U_r)
  // Column panels (col_panels = K > 1): 2K passes, local columns cut into K
  // equal ranges then remote packed rows into K; panel k's items are the
  // runs of each row's entries whose column falls in range k (hub rows cut
  // into chunks numbered row by row, so a hub's chunks stay contiguous).
  int32_t panels = 0;
  std::vector<WorkItem> panel_items;    // all panels back to back
  std::vector<int64_t> panel_begin;     // 2K+1 offsets into panel_items
  std::vector<int32_t> panel_xcd;       // 2K x 9 item ranges (relative to the panel)
  std::vector<int32_t> panel_hub_ptr;   // chunk range per hub (num_hubs+1)
  int32_t panel_chunks = 0;
  // Tiny-row bundles (one-pass plans): rows of L <= 4 entries without the
  // compensated-sum flag, grouped by L into bundles of floor(8/L) rows whose
  // entries are copied back to back after the partition's entries, so one
  // lane group loads a whole bundle in one batch of row loads.  A bundle item
  // is (offset of its rows in tiny_rows, begin, end, L).  items_short_nt holds
  // the short-row items that are not in a bundle (the short bin of a
  // schedule that runs the bundles).
  std::vector<WorkItem> items_tiny, items_short_nt;
  std::vector<int32_t> xcd_tiny, xcd_short_nt;
  std::vector<int32_t> tiny_rows;
};

cola_status build_plan(const Partition &p, const cola_plan_opts &opts,
                       HostPlan *out);

}  // namespace cola

struct cola_graph_s {
  cola::Graph g;
};
struct cola_partition_s {
  cola::Partition p;
};
