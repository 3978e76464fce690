// Device-side handle structures and host helpers shared by spmm.hip (plan
// path) and batch.hip (batched / served path).  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "cola_internal.hpp"

namespace cola_devhost {
cola_status hip_check(hipError_t e, const char *what);
int device_cu_count();
size_t dtype_bytes(int dt);
int vec_elems(int dt);  // elements per 16-byte lane vector
bool aligned(const void *p, size_t bytes);
// Device address of the row kernels' zero row on the current device.
cola_status zero_row_address(uint64_t *out);

template <typename T>
cola_status upload(const std::vector<T> &v, T **d) {
  *d = nullptr;
  if (v.empty()) return COLA_OK;
  cola_status st = hip_check(hipMalloc((void **)d, v.size() * sizeof(T)), "hipMalloc");
  if (st) return st;
  return hip_check(hipMemcpy(*d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice),
                   "hipMemcpy H2D");
}
}  // namespace cola_devhost

struct cola_zip_s {
  int32_t *ptr = nullptr, *idx = nullptr, *zip = nullptr, *pack_idx = nullptr,
          *uniq = nullptr;
  float *val = nullptr;
  int64_t nptr = 0, nnz = 0, n_unique = 0;
  uint64_t serial = 0;  // unique per created zip (never reused)
  // Plan built from this (local) zip and the remote zip with serial
  // fast_peer by the single-stage wrapper, reused while its arguments match.
  // Shared: a call holds its own reference while it enqueues work, so a
  // concurrent rebuild or cola_zip_invalidate cannot free it under the call.
  std::shared_ptr<cola_plan_s> fast_plan;
  uint64_t fast_peer = 0;
  int fast_P = -1, fast_peid = -1;
};

struct PlanBin {
  int4 *items = nullptr;
  int32_t *xcd = nullptr;  // 9 entries: item range per XCD
  int64_t n = 0;
  // Host: prefix sums of each item's cost (entries + 8), n + 1 values; the
  // batched form splits its work list over the XCDs with it.
  std::vector<int64_t> cost;
};

struct cola_plan_s {
  int32_t num_rows = 0, nodesPerPE = 0, U_r = 0, U_l = 0, max_dim = 0;
  // Rows a remote entry can address in its table: U_r (packed buffer) or,
  // for a plan whose remote entries hold global row ids (the reference-
  // signature launcher reads B_global directly), the largest id + 1.
  int64_t remote_rows = 0;
  int64_t nnz = 0, num_empty = 0, num_hubs = 0, num_chunks = 0;
  PlanBin shortb, longb;      // every item
  PlanBin shortb_r, longb_r;  // items with remote entries (second phase)
  PlanBin emptyb;             // rows without entries as empty items (batched overwrite)
  PlanBin cshortb, clongb;    // second (cold) pass of a hot/cold plan, else empty
  PlanBin tinyb;              // tiny-row bundles (rows of <= 4 entries), else empty
  PlanBin shortb_nt;          // short rows outside the bundles
  int32_t *tiny_rows = nullptr;  // output rows of the bundles
  int64_t hot_nnz = 0;
  int32_t *cols = nullptr, *empty_rows = nullptr, *hub_rows = nullptr, *hub_ptr = nullptr,
          *remote_cols = nullptr;
  float *vals = nullptr, *partials = nullptr;
  int64_t pstride = 0;
  int64_t device_bytes = 0;
  // Column-panel schedule (see HostPlan): 2*panels item lists.
  int32_t panels = 0;
  int4 *panel_items = nullptr;
  int32_t *panel_xcd = nullptr, *panel_hub_ptr = nullptr;
  std::vector<int64_t> panel_begin;
  int64_t panel_chunks = 0;
  cola_plan_tuning tune{0, 0, 1, 8, 1, 0, 8, 0, 0, 1};  // set by size at creation
};
