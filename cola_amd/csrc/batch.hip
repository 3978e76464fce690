// Batched form: many independent SpMMs (each with its own plan, B, C, dim)
// in ONE persistent launch driven by a device work list, optionally captured
// in a hipGraph, with per-dim launch parameters chosen by an autotuner.
//
// Work list: the batch is a table of segments; a segment is one bin (long
// rows + hub chunks, short rows, or empty rows for overwrite) of one
// problem, cut into tasks of `rpw` consecutive items x `n_tiles` column
// tiles (tile-major).  Tasks are numbered consecutively across segments;
// each XCD walks a contiguous eighth of the task numbers and its waves stride
// through it, so one wave moves monotonically through the segment table.  A
// task runs the plan kernel's step (rows.hpp row_step) with the segment's
// row-group width G compiled in, selected per task by a wave-uniform switch
// (dims differ per problem).
// Every item is reduced by one team in a fixed order, so a problem's result
// does not depend on which other problems share the launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <thread>
#include <vector>

#include "device.hpp"
#include "rows.hpp"

using namespace cola_devhost;
using cola::fail;
using cola_dev::Vec;

namespace {

struct BatchSeg {
  const int4 *items;
  const int32_t *cols;
  const float *vals;
  const void *B;      // this problem's local rows (already offset by the tile at run time)
  int64_t ldB;
  const float *sl;
  const void *Bp;     // packed remote rows (partitions); NULL for whole graphs
  int64_t ldP;
  const float *sp;
  float *C;
  int64_t ldC;
  float *partials;
  int64_t pstride;
  int32_t n_items, dim, rpw, n_tiles, tile_w, g_log2, task_begin, task_end;
  int32_t n_grp;  // row groups of rpw items: tasks = n_grp x n_tiles
  uint64_t abase; // common base of B and Bp for 32-bit row offsets (RowArgs::abase), 0: none
  uint64_t zrow;  // the zero row (RowArgs::zrow)
};

struct HubSeg {
  const int32_t *hub_rows, *hub_ptr;
  const float *partials;
  int64_t pstride;
  float *C;
  int64_t ldC;
  int32_t n_hubs, dim;
  int64_t thread_begin;
};

// First task of each XCD's range (9 entries), chosen on the host so the
// ranges carry equal estimated work (entries x columns + a per-item share).
struct XcdBounds {
  int32_t v[9];
};

// Largest block of the persistent kernel: a whole CU's waves (16) in one
// block, except FP8 with compensated sums, whose 155 VGPRs do not fit 4
// waves per SIMD (it keeps blocks of 4 waves).
constexpr int batch_max_threads(int dt, bool kahan) { return dt == COLA_FP8_E4M3 && kahan ? 256 : 1024; }

template <int DT, int EPL, bool KAHAN>
__global__ __launch_bounds__(batch_max_threads(DT, KAHAN)) void batch_rows_kernel(const BatchSeg *__restrict__ segs,
                                                                                int n_segs, const XcdBounds xb,
                                                                                int accumulate, uint64_t *stamps) {
  using T = typename Vec<DT, EPL>::T;
  __shared__ int next_task;
  const int wpb = blockDim.x >> 6;
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int xcd = blockIdx.x & 7;
  const int blk = blockIdx.x >> 3;
  const int nblk = (gridDim.x - xcd + 7) >> 3;
  const int t0 = xb.v[xcd], t1 = xb.v[xcd + 1];
  const uint64_t start = stamps ? wall_clock64() : 0;
  if (threadIdx.x == 0) next_task = wpb;
  __syncthreads();
  // The range's blocks take its tasks in rounds of wpb consecutive tasks per
  // block, the blocks in serpentine order (block j takes the j-th chunk of
  // even rounds and the j-th from the end of odd ones: along a round tasks
  // get lighter, items being by decreasing degree).  Within a block the
  // waves draw the block's k-th task from an LDS counter, each asking for
  // its next before it works on the current one.  One block holds a CU's
  // waves: the CU issues from its oldest waves first, so waves with equal
  // static shares end hundreds of microseconds apart (measured on 64 BF16
  // problems: mean wave lifetime 89% of the launch), while waves that share
  // a counter end together (93%).
  auto task_of = [&](int k) {
    const int r = k / wpb;
    const int j = (r & 1) ? nblk - 1 - blk : blk;
    return t0 + (r * nblk + j) * wpb + (k - r * wpb);
  };
  int si = 0;
  int t = task_of(wid);
  while (t < t1) {
    int got = 0;
    if (lane == 0) got = __hip_atomic_fetch_add(&next_task, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    while (si < n_segs - 1 && t >= segs[si].task_end) ++si;
    const BatchSeg &sg = segs[si];
    // Tile-major within a segment: the waves sweep every row group of one
    // column tile before the next, as the plan path's tile passes do.
    const int lt = t - sg.task_begin;
    const int tile = lt / sg.n_grp, grp = lt - tile * sg.n_grp;
    const int c0 = tile * sg.tile_w;
    cola_dev::RowArgs a;
    a.items = sg.items;
    a.xcd_begin = nullptr;
    a.cols = sg.cols;
    a.vals = sg.vals;
    a.Bl = sg.B ? static_cast<const T *>(sg.B) + c0 : nullptr;
    a.ldB = sg.ldB;
    a.sl = sg.sl;
    a.Bp = sg.Bp ? static_cast<const T *>(sg.Bp) + c0 : nullptr;
    a.ldP = sg.ldP;
    a.sp = sg.sp;
    a.C = sg.C + c0;
    a.ldC = sg.ldC;
    a.partials = sg.partials ? sg.partials + c0 : nullptr;
    a.pstride = sg.pstride;
    a.dim = min(sg.tile_w, sg.dim - c0);
    a.accumulate = accumulate;
    a.kahan = KAHAN ? 1 : 0;
    a.rpw = sg.rpw;
    a.xcd_remap = 0;
    a.phase = 0;
    a.abase = sg.abase;
    a.zrow = sg.zrow;
    const int it0 = grp * sg.rpw;
    // The lane-group width is wave-uniform (one segment per task): each case
    // is the plan kernel's step with that width compiled in.  Two-lane groups
    // keep 4 loads in flight instead of 8: with 8 each lane stages 4 entries,
    // and that case alone set the kernel's register count (BF16 with
    // compensated sums 137 VGPRs, 3 waves per SIMD; with 4 loads 128, 4 waves).
    switch (sg.g_log2) {
      case 1: cola_dev::row_step<DT, EPL, 2, 1, 4, KAHAN, cola_dev::kBatchPipe<DT>>(a, it0, sg.n_items, sg.rpw); break;
      case 2: cola_dev::row_step<DT, EPL, 4, 1, 8, KAHAN, cola_dev::kBatchPipe<DT>>(a, it0, sg.n_items, sg.rpw); break;
      case 3: cola_dev::row_step<DT, EPL, 8, 1, 8, KAHAN, cola_dev::kBatchPipe<DT>>(a, it0, sg.n_items, sg.rpw); break;
      case 4: cola_dev::row_step<DT, EPL, 16, 1, 8, KAHAN, cola_dev::kBatchPipe<DT>>(a, it0, sg.n_items, sg.rpw); break;
      case 5: cola_dev::row_step<DT, EPL, 32, 1, 8, KAHAN, cola_dev::kBatchPipe<DT>>(a, it0, sg.n_items, sg.rpw); break;
      default: cola_dev::row_step<DT, EPL, 64, 1, 8, KAHAN, cola_dev::kBatchPipe<DT>>(a, it0, sg.n_items, sg.rpw); break;
    }
    t = task_of(__builtin_amdgcn_readfirstlane(got));
  }
  // Start and end of every wave, for the autotuner's per-XCD timing.
  if (stamps && lane == 0) {
    const int64_t w = (int64_t)blockIdx.x * wpb + wid;
    stamps[2 * w] = start;
    stamps[2 * w + 1] = wall_clock64();
  }
}

// Thread (hub segment, hub, column): partial rows of a hub added in chunk order.
__global__ __launch_bounds__(256) void batch_hub_reduce_kernel(const HubSeg *__restrict__ hs, int n_hs,
                                                               int64_t total, int accumulate, int kahan) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  int lo = 0, hi = n_hs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (hs[mid].thread_begin <= t) lo = mid;
    else hi = mid - 1;
  }
  const HubSeg &h = hs[lo];
  const int64_t r = t - h.thread_begin;
  const int hub = (int)(r / h.dim), col = (int)(r % h.dim);
  // The table's pointers are read from memory: gld / gst keep the accesses
  // global rather than flat.
  const int c0 = cola_dev::gld(h.hub_ptr + hub), c1 = cola_dev::gld(h.hub_ptr + hub + 1);
  const float s = cola_dev::hub_sum(h.partials + col, h.pstride, c0, c1, kahan != 0);
  float *dst = h.C + (int64_t)cola_dev::gld(h.hub_rows + hub) * h.ldC + col;
  cola_dev::gst(dst, accumulate ? cola_dev::gld(dst) + s : s);
}

// Grid of the persistent kernel: min(tasks, 4 x blocks_per_cu x CUs) waves
// (blocks_per_cu keeps its meaning of 4-wave blocks per CU), spread over the
// CUs in blocks of up to max_threads, a multiple of 8 blocks (one XCD range
// per blockIdx % 8).
struct Grid {
  int blocks, threads;
};
static Grid batch_grid(int64_t tasks, int blocks_per_cu, int max_threads) {
  const int64_t cus = device_cu_count();
  const int64_t waves = std::max<int64_t>(1, std::min<int64_t>(tasks, cus * 4 * blocks_per_cu));
  const int64_t wpb = std::min<int64_t>(max_threads / 64, (waves + cus - 1) / cus);
  int64_t blocks = (waves + wpb - 1) / wpb;
  blocks = std::max<int64_t>(8, (blocks + 7) / 8 * 8);
  return Grid{(int)blocks, (int)wpb * 64};
}

// Most waves one launch can have (the autotuner's stamp buffer).
static int64_t batch_max_waves() { return (int64_t)device_cu_count() * 4 * 64 + 8 * 16; }

template <int DT, int EPL>
void launch_batch(const BatchSeg *segs, int n_segs, const XcdBounds &xb, int accumulate, bool kahan, int64_t tasks,
                  int blocks_per_cu, uint64_t *stamps, hipStream_t s) {
  if (kahan) {
    const Grid g = batch_grid(tasks, blocks_per_cu, batch_max_threads(DT, true));
    hipLaunchKernelGGL((batch_rows_kernel<DT, EPL, true>), dim3(g.blocks), dim3(g.threads), 0, s, segs, n_segs, xb,
                       accumulate, stamps);
  } else {
    const Grid g = batch_grid(tasks, blocks_per_cu, batch_max_threads(DT, false));
    hipLaunchKernelGGL((batch_rows_kernel<DT, EPL, false>), dim3(g.blocks), dim3(g.threads), 0, s, segs, n_segs, xb,
                       accumulate, stamps);
  }
}

// Cost model of a segment for the XCD split: a task (one row group of one
// column tile) costs the entries of its items plus 8 per item, times the
// tile's columns; prefix = the bin's per-item cost prefix (n_items + 1).
struct SegCost {
  const std::vector<int64_t> *prefix;
  int64_t task_begin, task_end;
  int32_t n_grp, rpw, tile_w, dim, n_items;
};

// Cumulative cost of the first k tasks of a segment.
static double seg_cost_before(const SegCost &c, int64_t k) {
  const int64_t tile = k / c.n_grp, g = k % c.n_grp;
  const std::vector<int64_t> &P = *c.prefix;
  const double full = (double)P[c.n_items];
  const double cols_before = (double)std::min<int64_t>(tile * c.tile_w, c.dim);
  const double w = (double)std::min<int64_t>(c.tile_w, c.dim - tile * c.tile_w);
  const int64_t it = std::min<int64_t>(g * c.rpw, c.n_items);
  return full * cols_before + (tile * c.tile_w < c.dim ? w * (double)P[it] : 0.0);
}

// XCD boundaries at equal shares of the total estimated cost.
static XcdBounds xcd_bounds(const std::vector<SegCost> &costs, int64_t total_tasks) {
  XcdBounds xb{};
  std::vector<double> before(costs.size() + 1, 0.0);
  for (size_t i = 0; i < costs.size(); ++i)
    before[i + 1] = before[i] + seg_cost_before(costs[i], costs[i].task_end - costs[i].task_begin);
  const double total = before.back();
  xb.v[0] = 0;
  xb.v[8] = (int32_t)total_tasks;
  size_t si = 0;
  for (int x = 1; x < 8; ++x) {
    const double target = total * x / 8.0;
    while (si + 1 < costs.size() && before[si + 1] <= target) ++si;
    int64_t t = total_tasks;
    if (!costs.empty() && total > 0) {
      const SegCost &c = costs[si];
      int64_t lo = 0, hi = c.task_end - c.task_begin;  // first k with cost(k) >= target - before
      const double need = target - before[si];
      while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (seg_cost_before(c, mid) >= need) hi = mid;
        else lo = mid + 1;
      }
      t = c.task_begin + lo;
    }
    xb.v[x] = (int32_t)std::max<int64_t>(xb.v[x - 1], std::min<int64_t>(t, total_tasks));
  }
  return xb;
}

int ilog2(int v) {
  int r = 0;
  while ((1 << r) < v) ++r;
  return r;
}

}  // namespace

struct cola_batch_s {
  std::vector<cola_batch_problem> probs;
  std::vector<cola_batch_tuning> tune;  // per problem
  int dtype = COLA_F32;
  int flags = COLA_OVERWRITE;
  int blocks_per_cu = 8;
  // device tables
  BatchSeg *d_segs = nullptr;
  int n_segs = 0, total_tasks = 0, seg_cap = 0;
  XcdBounds xb{};            // task range of each XCD
  std::vector<SegCost> costs;     // cost model of the current segments
  HubSeg *d_hubs = nullptr;
  int n_hubs = 0, hub_cap = 0;
  int64_t hub_threads = 0;
  std::vector<float *> partials;  // per problem (NULL when the plan has no hubs)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static void drop_graph(cola_batch_s *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// Appends one problem's segments (degree bins cut into tasks) and hub table
// entry; task / hub-thread counters run across the whole work list.
static cola_status append_problem(const cola_batch_problem &pr, const cola_batch_tuning &t, float *partials,
                                  int dtype, int flags, std::vector<BatchSeg> *segs, std::vector<SegCost> *costs,
                                  std::vector<HubSeg> *hubs, int64_t *task, int64_t *hthreads) {
  const cola_plan_s *p = pr.plan;
  const int ev = vec_elems(dtype);
  int tile = pr.dim;
  if (t.col_tile > 0 && t.col_tile < pr.dim) tile = std::max(ev, t.col_tile / ev * ev);
  const int chunks = (tile + ev - 1) / ev;
  const int G = 1 << ilog2(std::max(chunks, 2));
  if (G > 64) return fail(COLA_ERR_UNSUPPORTED, "batch: a column tile wider than 64 lane vectors");
  const int S = 64 / G;
  const int n_tiles = (pr.dim + tile - 1) / tile;
  auto clamp = [&](int req) {
    int r = 1;
    while (r * 2 <= req && r * 2 <= S) r *= 2;
    return r;
  };
  // Common base of the problem's two row tables and the zero row when every
  // addressable row (local rows below nodesPerPE, packed rows below U_r, all
  // columns) lies within 2^36 bytes of it on 16-byte boundaries
  // (RowArgs::abase).
  uint64_t zrow = 0;
  if (cola_status zst = zero_row_address(&zrow)) return zst;
  uint64_t abase = 0;
  {
    const uint64_t eb = dtype_bytes(dtype);
    uint64_t lo = zrow, hi = zrow + cola_dev::kZeroRowBytes;
    bool ok = true;
    auto span = [&](const void *base, int64_t rows, int64_t ld) {
      if (!base || rows <= 0) return;
      const uint64_t bb = (uint64_t)(uintptr_t)base;
      ok = ok && bb % 16 == 0 && ((uint64_t)ld * eb) % 16 == 0;
      lo = std::min(lo, bb);
      hi = std::max(hi, bb + ((uint64_t)(rows - 1) * (uint64_t)ld + (uint64_t)pr.dim) * eb);
    };
    span(pr.B, p->nodesPerPE, pr.ld_B);
    span(pr.packed, p->remote_rows, pr.ld_packed);
    if (ok && hi - lo < (uint64_t(1) << 36)) abase = lo;
  }
  const PlanBin *bins[3] = {&p->longb, &p->shortb, (flags & COLA_OVERWRITE) ? &p->emptyb : nullptr};
  const int rpws[3] = {clamp(t.rpw_long > 0 ? t.rpw_long : 1), clamp(t.rpw_short > 0 ? t.rpw_short : S), S};
  for (int bi = 0; bi < 3; ++bi) {
    const PlanBin *bin = bins[bi];
    if (!bin || bin->n == 0) continue;
    BatchSeg s{};
    s.items = bin->items;
    s.cols = p->cols;
    s.vals = p->vals;
    s.B = pr.B;
    s.ldB = pr.ld_B;
    s.sl = pr.scales;
    s.Bp = pr.packed;
    s.ldP = pr.ld_packed;
    s.sp = pr.packed_scales;
    s.C = pr.C;
    s.ldC = pr.ld_C;
    s.partials = partials;
    s.pstride = p->pstride;
    s.n_items = (int32_t)bin->n;
    s.dim = pr.dim;
    s.rpw = rpws[bi];
    s.n_tiles = n_tiles;
    s.tile_w = tile;
    s.g_log2 = ilog2(G);
    s.n_grp = (int32_t)((bin->n + s.rpw - 1) / s.rpw);
    const int64_t tasks = (int64_t)s.n_grp * n_tiles;
    s.task_begin = (int32_t)*task;
    *task += tasks;
    if (*task > INT32_MAX - (1 << 24)) return fail(COLA_ERR_INVALID, "batch: too many tasks");
    s.task_end = (int32_t)*task;
    s.abase = abase;
    s.zrow = zrow;
    segs->push_back(s);
    costs->push_back(SegCost{&bin->cost, s.task_begin, s.task_end, s.n_grp, s.rpw, s.tile_w, s.dim, s.n_items});
  }
  if (p->num_hubs > 0) {
    HubSeg h{};
    h.hub_rows = p->hub_rows;
    h.hub_ptr = p->hub_ptr;
    h.partials = partials;
    h.pstride = p->pstride;
    h.C = pr.C;
    h.ldC = pr.ld_C;
    h.n_hubs = (int32_t)p->num_hubs;
    h.dim = pr.dim;
    h.thread_begin = *hthreads;
    *hthreads += (int64_t)p->num_hubs * pr.dim;
    hubs->push_back(h);
  }
  return COLA_OK;
}

// Launches the persistent row kernel over a work list and the hub reduction.
static cola_status launch_tables(const BatchSeg *d_segs, int n_segs, int total_tasks, const XcdBounds &xb,
                                 const HubSeg *d_hubs, int n_hubs, int64_t hub_threads, int dtype, int flags,
                                 int blocks_per_cu, hipStream_t s, uint64_t *stamps = nullptr) {
  const int accumulate = (flags & COLA_OVERWRITE) ? 0 : 1;
  const bool kahan = (flags & COLA_KAHAN) != 0;
  if (total_tasks > 0) {
    switch (dtype) {
      case COLA_F32:
        launch_batch<COLA_F32, 4>(d_segs, n_segs, xb, accumulate, kahan, total_tasks, blocks_per_cu, stamps, s);
        break;
      case COLA_BF16:
        launch_batch<COLA_BF16, 8>(d_segs, n_segs, xb, accumulate, kahan, total_tasks, blocks_per_cu, stamps, s);
        break;
      case COLA_F16:
        launch_batch<COLA_F16, 8>(d_segs, n_segs, xb, accumulate, kahan, total_tasks, blocks_per_cu, stamps, s);
        break;
      case COLA_FP8_E4M3:
        launch_batch<COLA_FP8_E4M3, 16>(d_segs, n_segs, xb, accumulate, kahan, total_tasks, blocks_per_cu, stamps, s);
        break;
    }
  }
  if (hub_threads > 0) {
    const int64_t blocks = (hub_threads + 255) / 256;
    hipLaunchKernelGGL(batch_hub_reduce_kernel, dim3((unsigned)blocks), dim3(256), 0, s, d_hubs, n_hubs, hub_threads,
                       accumulate, kahan ? 1 : 0);
  }
  return hip_check(hipGetLastError(), "batch launch");
}

// Rebuilds the segment / hub tables from the problems and their tunings.
// Tables already built may be in use by a launch still running on any
// stream (non-blocking streams do not order against hipMemcpy), so the device
// is drained before they are rewritten or the captured graph is dropped.
static cola_status build_tables(cola_batch_s *b) {
  if (b->n_segs > 0 || b->n_hubs > 0 || b->exec) {
    cola_status st = hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize before a table rebuild");
    if (st) return st;
  }
  drop_graph(b);
  std::vector<BatchSeg> segs;
  std::vector<SegCost> costs;
  std::vector<HubSeg> hubs;
  int64_t task = 0, hthreads = 0;
  for (size_t k = 0; k < b->probs.size(); ++k) {
    cola_status st = append_problem(b->probs[k], b->tune[k], b->partials[k], b->dtype, b->flags, &segs, &costs,
                                    &hubs, &task, &hthreads);
    if (st) return st;
  }
  b->xb = xcd_bounds(costs, task);
  b->costs = costs;
  if ((int)segs.size() > b->seg_cap) {
    (void)hipFree(b->d_segs);
    b->d_segs = nullptr;
    b->seg_cap = 0;
    cola_status st = hip_check(hipMalloc((void **)&b->d_segs, std::max<size_t>(1, segs.size()) * sizeof(BatchSeg)),
                               "hipMalloc batch segments");
    if (st) return st;
    b->seg_cap = (int)segs.size();
  }
  if ((int)hubs.size() > b->hub_cap) {
    (void)hipFree(b->d_hubs);
    b->d_hubs = nullptr;
    b->hub_cap = 0;
    cola_status st = hip_check(hipMalloc((void **)&b->d_hubs, std::max<size_t>(1, hubs.size()) * sizeof(HubSeg)),
                               "hipMalloc batch hubs");
    if (st) return st;
    b->hub_cap = (int)hubs.size();
  }
  cola_status st;
  if (!segs.empty() &&
      (st = hip_check(hipMemcpy(b->d_segs, segs.data(), segs.size() * sizeof(BatchSeg), hipMemcpyHostToDevice),
                      "hipMemcpy batch segments")))
    return st;
  if (!hubs.empty() &&
      (st = hip_check(hipMemcpy(b->d_hubs, hubs.data(), hubs.size() * sizeof(HubSeg), hipMemcpyHostToDevice),
                      "hipMemcpy batch hubs")))
    return st;
  b->n_segs = (int)segs.size();
  b->total_tasks = (int)task;
  b->n_hubs = (int)hubs.size();
  b->hub_threads = hthreads;
  return COLA_OK;
}

static cola_status run_batch(cola_batch_s *b, hipStream_t s, uint64_t *stamps = nullptr) {
  return launch_tables(b->d_segs, b->n_segs, b->total_tasks, b->xb, b->d_hubs, b->n_hubs, b->hub_threads, b->dtype,
                       b->flags, b->blocks_per_cu, s, stamps);
}

static void free_batch(cola_batch_s *b) {
  if (!b) return;
  drop_graph(b);
  (void)hipFree(b->d_segs);
  (void)hipFree(b->d_hubs);
  for (float *p : b->partials) (void)hipFree(p);
  delete b;
}

extern "C" {

cola_status cola_batch_create(const cola_batch_problem *probs, int n, cola_dtype dtype, int flags,
                              cola_batch *out) {
  if (!out) return fail(COLA_ERR_INVALID, "batch_create: out is NULL");
  *out = nullptr;
  if (n < 0 || (n > 0 && !probs)) return fail(COLA_ERR_INVALID, "batch_create: bad problem list");
  if (!dtype_bytes(dtype)) return fail(COLA_ERR_UNSUPPORTED, "batch_create: unknown dtype");
  if (flags & ~(COLA_OVERWRITE | COLA_KAHAN)) return fail(COLA_ERR_INVALID, "batch_create: unknown flag");
  const int ev = vec_elems(dtype);
  for (int k = 0; k < n; ++k) {
    const cola_batch_problem &pr = probs[k];
    if (!pr.plan) return fail(COLA_ERR_INVALID, "batch_create: problem without a plan");
    const cola_plan_s *p = pr.plan;
    if (pr.dim < 1 || pr.dim % ev != 0 || pr.dim / ev > 64)
      return fail(COLA_ERR_UNSUPPORTED, "batch_create: dim must be a multiple of the 16-byte vector width, "
                                        "at most 64 vectors");
    if (pr.ld_B < pr.dim || pr.ld_B % ev || pr.ld_C < pr.dim || pr.ld_C % 4 || !pr.C || !aligned(pr.C, 16) ||
        (p->U_l > 0 && (!pr.B || !aligned(pr.B, 16))))
      return fail(COLA_ERR_INVALID, "batch_create: B / C pointers or leading dimensions not 16-byte aligned");
    if (p->U_r > 0 && (!pr.packed || !aligned(pr.packed, 16) || pr.ld_packed < pr.dim || pr.ld_packed % ev))
      return fail(COLA_ERR_INVALID, "batch_create: partition with remote entries needs an aligned packed buffer");
    if (dtype == COLA_FP8_E4M3 && ((p->U_l > 0 && !pr.scales) || (p->U_r > 0 && !pr.packed_scales)))
      return fail(COLA_ERR_INVALID, "batch_create: FP8 needs row scales");
    if (p->num_chunks > 0 && pr.dim > p->max_dim) return fail(COLA_ERR_INVALID, "batch_create: dim exceeds max_dim");
    if (p->cshortb.n + p->clongb.n > 0)
      return fail(COLA_ERR_UNSUPPORTED, "batch_create: a hot/cold plan needs two ordered passes");
  }
  auto *b = new (std::nothrow) cola_batch_s;
  if (!b) return fail(COLA_ERR_NOMEM, "batch_create: out of host memory");
  b->probs.assign(probs, probs + n);
  b->tune.assign(n, cola_batch_tuning{0, 0, 1});
  b->dtype = dtype;
  b->flags = flags;
  b->partials.assign(n, nullptr);
  for (int k = 0; k < n; ++k) {
    const cola_plan_s *p = probs[k].plan;
    if (p->num_chunks > 0) {
      cola_status st = hip_check(
          hipMalloc((void **)&b->partials[k], (size_t)p->num_chunks * p->pstride * sizeof(float)),
          "hipMalloc batch partials");
      if (st) {
        free_batch(b);
        return st;
      }
    }
  }
  cola_status st = build_tables(b);
  if (st) {
    free_batch(b);
    return st;
  }
  *out = b;
  return COLA_OK;
}

cola_status cola_batch_set_tuning(cola_batch b, int problem, const cola_batch_tuning *t, int blocks_per_cu) {
  if (!b) return fail(COLA_ERR_INVALID, "batch_set_tuning: NULL batch");
  if (blocks_per_cu < 0 || blocks_per_cu > 64) return fail(COLA_ERR_INVALID, "batch_set_tuning: blocks_per_cu");
  if (blocks_per_cu > 0) b->blocks_per_cu = blocks_per_cu;
  if (t) {
    if (t->col_tile < 0 || t->rpw_short < 0 || t->rpw_short > 64 || t->rpw_long < 0 || t->rpw_long > 64)
      return fail(COLA_ERR_INVALID, "batch_set_tuning: parameter out of range");
    if (problem < -1 || problem >= (int)b->probs.size())
      return fail(COLA_ERR_INVALID, "batch_set_tuning: problem index out of range");
    for (size_t k = 0; k < b->probs.size(); ++k)
      if (problem == -1 || (int)k == problem) b->tune[k] = *t;
  }
  return build_tables(b);
}

cola_status cola_batch_get_tuning(cola_batch b, int problem, cola_batch_tuning *t, int *blocks_per_cu) {
  if (!b || problem < 0 || problem >= (int)b->probs.size())
    return fail(COLA_ERR_INVALID, "batch_get_tuning: bad batch or problem");
  if (t) *t = b->tune[problem];
  if (blocks_per_cu) *blocks_per_cu = b->blocks_per_cu;
  return COLA_OK;
}

cola_status cola_batch_run(cola_batch b, void *stream) {
  if (!b) return fail(COLA_ERR_INVALID, "batch_run: NULL batch");
  return run_batch(b, (hipStream_t)stream);
}

cola_status cola_batch_capture(cola_batch b, void *stream) {
  if (!b) return fail(COLA_ERR_INVALID, "batch_capture: NULL batch");
  if (!stream) return fail(COLA_ERR_INVALID, "batch_capture: needs a non-default stream");
  hipStream_t s = (hipStream_t)stream;
  drop_graph(b);
  cola_status st = hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  if (st) return st;
  cola_status run_st = run_batch(b, s);
  hipGraph_t g = nullptr;
  st = hip_check(hipStreamEndCapture(s, &g), "hipStreamEndCapture");
  if (run_st) {
    if (g) (void)hipGraphDestroy(g);
    return run_st;
  }
  if (st) return st;
  b->graph = g;
  st = hip_check(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "hipGraphInstantiate");
  if (st) {
    drop_graph(b);
    return st;
  }
  return COLA_OK;
}

cola_status cola_batch_replay(cola_batch b, void *stream) {
  if (!b) return fail(COLA_ERR_INVALID, "batch_replay: NULL batch");
  if (!b->exec) return fail(COLA_ERR_INVALID, "batch_replay: nothing captured (or the tuning changed since)");
  return hip_check(hipGraphLaunch(b->exec, (hipStream_t)stream), "hipGraphLaunch");
}

cola_status cola_batch_autotune(cola_batch b, int iters, void *stream, float *best_ms) {
  if (!b) return fail(COLA_ERR_INVALID, "batch_autotune: NULL batch");
  if (iters < 1) iters = 3;
  hipStream_t s = (hipStream_t)stream;
  const char *env = getenv("COLA_AUTOTUNE_VERBOSE");
  const bool verbose = env && env[0] == '1';
  hipEvent_t e0, e1;
  cola_status st;
  if ((st = hip_check(hipEventCreate(&e0), "hipEventCreate"))) return st;
  if ((st = hip_check(hipEventCreate(&e1), "hipEventCreate"))) {
    (void)hipEventDestroy(e0);
    return st;
  }
  auto measure = [&](float *ms) -> cola_status {
    cola_status r = run_batch(b, s);
    if (r) return r;
    if ((r = hip_check(hipEventRecord(e0, s), "hipEventRecord"))) return r;
    for (int i = 0; i < iters; ++i)
      if ((r = run_batch(b, s))) return r;
    if ((r = hip_check(hipEventRecord(e1, s), "hipEventRecord"))) return r;
    if ((r = hip_check(hipEventSynchronize(e1), "hipEventSynchronize"))) return r;
    float t = 0.f;
    if ((r = hip_check(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime"))) return r;
    *ms = t / iters;
    return COLA_OK;
  };
  float bms = 0.f;
  st = measure(&bms);
  // Problems sharing a dim share a schedule; tune each dim class in turn.
  std::vector<int> dims;
  for (const auto &p : b->probs)
    if (std::find(dims.begin(), dims.end(), p.dim) == dims.end()) dims.push_back(p.dim);
  const int ev = vec_elems(b->dtype);
  auto apply = [&](int dim, const cola_batch_tuning &t) {
    for (size_t k = 0; k < b->probs.size(); ++k)
      if (b->probs[k].dim == dim) b->tune[k] = t;
  };
  for (int pass = 0; pass < 2 && !st; ++pass) {
    for (int dim : dims) {
      cola_batch_tuning cur{0, 0, 1};
      for (size_t k = 0; k < b->probs.size(); ++k)
        if (b->probs[k].dim == dim) cur = b->tune[k];
      std::vector<cola_batch_tuning> cands;
      for (int tile : {0, dim / 2, dim / 4}) {
        if (tile != 0 && (tile < 2 * ev || tile % ev)) continue;
        cola_batch_tuning t = cur;
        t.col_tile = tile;
        cands.push_back(t);
      }
      for (int r : {1, 2, 4, 8, 16, 32}) {
        cola_batch_tuning t = cur;
        t.rpw_short = r;
        cands.push_back(t);
      }
      for (int r : {1, 2, 4, 8}) {
        cola_batch_tuning t = cur;
        t.rpw_long = r;
        cands.push_back(t);
      }
      for (const auto &t : cands) {
        apply(dim, t);
        if ((st = build_tables(b))) break;
        float ms = 0.f;
        if ((st = measure(&ms))) break;
        if (verbose)
          fprintf(stderr, "[batch autotune] dim=%d tile=%d rpw_short=%d rpw_long=%d: %.4f ms\n", dim, t.col_tile,
                  t.rpw_short, t.rpw_long, ms);
        if (ms < bms * 0.99f) {
          bms = ms;
          cur = t;
        }
      }
      apply(dim, cur);
      if (!st) st = build_tables(b);
      if (st) break;
    }
    if (st) break;
    const int bpc0 = b->blocks_per_cu;
    int best_bpc = bpc0;
    for (int bpc : {1, 2, 3, 4, 8, 16}) {
      b->blocks_per_cu = bpc;
      float ms = 0.f;
      if ((st = measure(&ms))) break;
      if (verbose) fprintf(stderr, "[batch autotune] blocks_per_cu=%d: %.4f ms\n", bpc, ms);
      if (ms < bms * 0.99f) {
        bms = ms;
        best_bpc = bpc;
      }
    }
    b->blocks_per_cu = best_bpc;
  }
  // Re-cut the XCD ranges by measured time: launches record every wave's
  // start and end, each XCD's time (its last wave's end after the launch's
  // first start, mean of 3 launches) is spread over its range's modelled
  // cost, and the ranges are cut again at equal shares of that time; three
  // rounds, each kept only if the whole batch runs faster.
  uint64_t *d_stamps = nullptr;
  if (!st && b->n_segs > 0 && b->total_tasks > 0)
    st = hip_check(hipMalloc((void **)&d_stamps, 2 * batch_max_waves() * sizeof(uint64_t)), "hipMalloc stamps");
  if (!st && b->n_segs > 0 && b->total_tasks > 0) {
    const bool kh = (b->flags & COLA_KAHAN) != 0;
    const int max_threads = batch_max_threads(b->dtype, kh);
    std::vector<uint64_t> h_st;
    int dev = 0, clock_khz = 0;
    if (!st) st = hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (!st) st = hip_check(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, dev), "wall clock rate");
    if (!st && clock_khz <= 0) clock_khz = 100000;
    auto xcd_times = [&](double t_x[8]) -> cola_status {
      const Grid g = batch_grid(b->total_tasks, b->blocks_per_cu, max_threads);
      const int wpb = g.threads / 64;
      const int64_t nw = (int64_t)g.blocks * wpb;
      h_st.resize(2 * nw);
      for (int x = 0; x < 8; ++x) t_x[x] = 0.0;
      for (int rep = 0; rep < 3; ++rep) {
        cola_status r = run_batch(b, s, d_stamps);
        if (!r) r = hip_check(hipMemcpyAsync(h_st.data(), d_stamps, h_st.size() * sizeof(uint64_t),
                                             hipMemcpyDeviceToHost, s), "hipMemcpyAsync stamps");
        if (!r) r = hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        if (r) return r;
        uint64_t t0 = UINT64_MAX, end[8] = {};
        for (int64_t w = 0; w < nw; ++w) t0 = std::min(t0, h_st[2 * w]);
        for (int64_t w = 0; w < nw; ++w) {
          const int x = (int)((w / wpb) & 7);
          end[x] = std::max(end[x], h_st[2 * w + 1]);
        }
        for (int x = 0; x < 8; ++x) t_x[x] += end[x] > t0 ? (double)(end[x] - t0) : 0.0;
      }
      for (int x = 0; x < 8; ++x) t_x[x] /= 3.0 * clock_khz;  // ms
      return COLA_OK;
    };
    // Modelled cost of the first k tasks of the whole list.
    std::vector<double> seg_before(b->costs.size() + 1, 0.0);
    for (size_t i = 0; i < b->costs.size(); ++i)
      seg_before[i + 1] = seg_before[i] + seg_cost_before(b->costs[i], b->costs[i].task_end - b->costs[i].task_begin);
    auto cost_to = [&](int64_t t) {
      size_t i = 0;
      while (i + 1 < b->costs.size() && b->costs[i].task_end <= t) ++i;
      const SegCost &c = b->costs[i];
      return seg_before[i] + seg_cost_before(c, std::max<int64_t>(0, std::min<int64_t>(t, c.task_end) - c.task_begin));
    };
    const XcdBounds model_xb = b->xb;
    XcdBounds cur = b->xb, best_xb = b->xb;
    for (int round = 0; round < 3 && !st; ++round) {
      double t_x[8];
      b->xb = cur;
      if ((st = xcd_times(t_x))) break;
      if (verbose) {
        fprintf(stderr, "[batch autotune] XCD range times (round %d):", round);
        for (int x = 0; x < 8; ++x) fprintf(stderr, " %.3f", t_x[x]);
        fprintf(stderr, " ms\n");
      }
      // measured time of the first t tasks: piecewise by range
      double c_x[9], m_x[9];
      for (int x = 0; x <= 8; ++x) c_x[x] = cost_to(cur.v[x]);
      m_x[0] = 0.0;
      for (int x = 0; x < 8; ++x) m_x[x + 1] = m_x[x] + t_x[x];
      auto time_to = [&](int64_t t) {
        int x = 0;
        while (x < 7 && cur.v[x + 1] <= t) ++x;
        const double span = c_x[x + 1] - c_x[x];
        const double f = span > 0 ? (cost_to(t) - c_x[x]) / span : 0.0;
        return m_x[x] + f * t_x[x];
      };
      XcdBounds nxt{};
      nxt.v[0] = 0;
      nxt.v[8] = b->total_tasks;
      for (int x = 1; x < 8; ++x) {
        const double target = m_x[8] * x / 8.0;
        int64_t lo = nxt.v[x - 1], hi = b->total_tasks;
        while (lo < hi) {
          const int64_t mid = (lo + hi) / 2;
          if (time_to(mid) >= target) hi = mid;
          else lo = mid + 1;
        }
        nxt.v[x] = (int32_t)lo;
      }
      b->xb = nxt;
      float ms = 0.f;
      if ((st = measure(&ms))) break;
      if (verbose) fprintf(stderr, "[batch autotune] XCD ranges re-cut by measured time: %.4f ms (best %.4f)\n", ms, bms);
      if (ms < bms * 0.995f) {
        bms = ms;
        best_xb = nxt;
      }
      cur = nxt;
    }
    b->xb = best_xb;
    if (verbose && std::memcmp(&best_xb, &model_xb, sizeof(XcdBounds)) != 0)
      fprintf(stderr, "[batch autotune] keeping the XCD ranges cut by measured time\n");
  }
  (void)hipFree(d_stamps);
  if (!st) st = run_batch(b, s);
  if (!st) st = hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st) return st;
  if (best_ms) *best_ms = bms;
  return COLA_OK;
}

cola_status cola_batch_destroy(cola_batch b) {
  free_batch(b);
  return COLA_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Serving: resident graphs + feature matrices, requests admitted against a
// device memory budget, continuous batching into the persistent kernel with
// up to max_inflight launches in flight (one stream each).
// ---------------------------------------------------------------------------
namespace {

struct LaunchSlot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool own_stream = false;
  bool busy = false;
  BatchSeg *h_segs = nullptr, *d_segs = nullptr;
  HubSeg *h_hubs = nullptr, *d_hubs = nullptr;
  int cap_segs = 0, cap_hubs = 0;
  std::vector<int64_t> reqs;
};

struct Resident {
  cola_batch_problem prob;  // C / ld_C filled per request
  int64_t out_bytes = 0, hub_bytes = 0;
};

struct ReqState {
  int64_t pool_off = -1, pool_bytes = 0;
  bool live = false;  // holds pool memory
  float *C = nullptr;
  float *partials = nullptr;
};

// Pool bytes of one request: its output (unless the caller gives one) and
// its hub partials, each part on a 256-byte boundary, as pool_alloc takes it.
int64_t round256(int64_t b) { return (b + 255) / 256 * 256; }
int64_t request_pool_bytes(int64_t out_bytes, int64_t hub_bytes, bool own_output) {
  return round256((own_output ? round256(out_bytes) : 0) + hub_bytes);
}

double now_us() {
  return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct cola_server_s {
  cola_server_opts o{};
  std::vector<Resident> res;
  std::map<int, cola_batch_tuning> tune;
  char *pool = nullptr;
  int64_t pool_size = 0;
  std::map<int64_t, int64_t> free_blocks;  // offset -> size
  std::vector<LaunchSlot> slots;
  int64_t peak_used = 0, used = 0;
};

static int64_t pool_alloc(cola_server_s *sv, int64_t bytes) {
  bytes = round256(bytes);
  for (auto it = sv->free_blocks.begin(); it != sv->free_blocks.end(); ++it) {
    if (it->second >= bytes) {
      const int64_t off = it->first, rest = it->second - bytes;
      sv->free_blocks.erase(it);
      if (rest > 0) sv->free_blocks[off + bytes] = rest;
      sv->used += bytes;
      sv->peak_used = std::max(sv->peak_used, sv->used);
      return off;
    }
  }
  return -1;
}

static void pool_free(cola_server_s *sv, int64_t off, int64_t bytes) {
  bytes = round256(bytes);
  sv->used -= bytes;
  auto it = sv->free_blocks.emplace(off, bytes).first;
  auto nx = std::next(it);
  if (nx != sv->free_blocks.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    sv->free_blocks.erase(nx);
  }
  if (it != sv->free_blocks.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      sv->free_blocks.erase(it);
    }
  }
}

static void free_server(cola_server_s *sv) {
  if (!sv) return;
  for (auto &sl : sv->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.own_stream && sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.h_segs) (void)hipHostFree(sl.h_segs);
    if (sl.h_hubs) (void)hipHostFree(sl.h_hubs);
    (void)hipFree(sl.d_segs);
    (void)hipFree(sl.d_hubs);
  }
  (void)hipFree(sv->pool);
  delete sv;
}

extern "C" {

cola_status cola_server_create(const cola_server_opts *opts, cola_server *out) {
  if (!opts || !out) return fail(COLA_ERR_INVALID, "server_create: NULL argument");
  *out = nullptr;
  cola_server_opts o = *opts;
  if (!dtype_bytes(o.dtype)) return fail(COLA_ERR_UNSUPPORTED, "server_create: unknown dtype");
  if (o.memory_budget <= 0) return fail(COLA_ERR_INVALID, "server_create: memory_budget must be positive");
  if (o.max_batch <= 0) o.max_batch = 64;
  if (o.max_inflight <= 0) o.max_inflight = 2;
  if (o.blocks_per_cu <= 0) o.blocks_per_cu = 8;
  if (o.max_inflight > 16 || o.max_batch > 4096 || o.blocks_per_cu > 64)
    return fail(COLA_ERR_INVALID, "server_create: limits too large");
  if (o.flags & ~COLA_KAHAN) return fail(COLA_ERR_INVALID, "server_create: unknown flag");
  auto *sv = new (std::nothrow) cola_server_s;
  if (!sv) return fail(COLA_ERR_NOMEM, "server_create: out of host memory");
  sv->o = o;
  sv->pool_size = (o.memory_budget + 255) / 256 * 256;
  cola_status st = hip_check(hipMalloc((void **)&sv->pool, sv->pool_size), "hipMalloc request pool");
  if (st) {
    free_server(sv);
    return st;
  }
  sv->free_blocks[0] = sv->pool_size;
  // Plain mode: per-request launches queued on one stream (ring of 64
  // contexts); batched mode: one context per stream.
  const int nslots = o.plain ? 64 : o.max_inflight;
  const int per_launch = o.plain ? 1 : o.max_batch;
  sv->slots.resize(nslots);
  hipStream_t shared = nullptr;
  for (int i = 0; i < nslots && !st; ++i) {
    LaunchSlot &sl = sv->slots[i];
    if (!o.plain || i == 0) {
      st = hip_check(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking), "hipStreamCreate");
      sl.own_stream = true;
      shared = sl.stream;
    } else {
      sl.stream = shared;
    }
    if (!st) st = hip_check(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming), "hipEventCreate");
    sl.cap_segs = 3 * per_launch;
    sl.cap_hubs = per_launch;
    if (!st) st = hip_check(hipHostMalloc((void **)&sl.h_segs, sl.cap_segs * sizeof(BatchSeg)), "hipHostMalloc");
    if (!st) st = hip_check(hipHostMalloc((void **)&sl.h_hubs, sl.cap_hubs * sizeof(HubSeg)), "hipHostMalloc");
    if (!st) st = hip_check(hipMalloc((void **)&sl.d_segs, sl.cap_segs * sizeof(BatchSeg)), "hipMalloc");
    if (!st) st = hip_check(hipMalloc((void **)&sl.d_hubs, sl.cap_hubs * sizeof(HubSeg)), "hipMalloc");
  }
  if (st) {
    free_server(sv);
    return st;
  }
  *out = sv;
  return COLA_OK;
}

cola_status cola_server_add_resident(cola_server sv, cola_plan plan, const void *B, int64_t ld_B,
                                     const float *scales, int dim, int *id) {
  if (!sv || !plan || !id) return fail(COLA_ERR_INVALID, "server_add_resident: NULL argument");
  const int ev = vec_elems(sv->o.dtype);
  if (dim < 1 || dim % ev || dim / ev > 64 || ld_B < dim || ld_B % ev)
    return fail(COLA_ERR_UNSUPPORTED, "server_add_resident: dim must be a multiple of the vector width, <= 64 vectors");
  if (plan->U_r > 0) return fail(COLA_ERR_INVALID, "server_add_resident: the plan must be a whole graph (no remote rows)");
  if (plan->U_l > 0 && (!B || !aligned(B, 16))) return fail(COLA_ERR_INVALID, "server_add_resident: B not 16-byte aligned");
  if (sv->o.dtype == COLA_FP8_E4M3 && plan->U_l > 0 && !scales)
    return fail(COLA_ERR_INVALID, "server_add_resident: FP8 needs row scales");
  if (plan->num_chunks > 0 && dim > plan->max_dim) return fail(COLA_ERR_INVALID, "server_add_resident: dim > max_dim");
  if (plan->cshortb.n + plan->clongb.n > 0)
    return fail(COLA_ERR_UNSUPPORTED, "server_add_resident: a hot/cold plan needs two ordered passes");
  Resident r;
  r.prob = cola_batch_problem{plan, dim, B, ld_B, scales, nullptr, dim, nullptr, nullptr, dim};
  r.out_bytes = (int64_t)plan->num_rows * dim * 4;
  r.hub_bytes = plan->num_chunks * plan->pstride * 4;
  sv->res.push_back(r);
  *id = (int)sv->res.size() - 1;
  return COLA_OK;
}

cola_status cola_server_set_tuning(cola_server sv, int dim, const cola_batch_tuning *t) {
  if (!sv || !t) return fail(COLA_ERR_INVALID, "server_set_tuning: NULL argument");
  sv->tune[dim] = *t;
  return COLA_OK;
}

cola_status cola_server_run_trace(cola_server sv, int64_t n, const double *arrival_us, const int32_t *resident,
                                  float *const *outputs, cola_request_record *records, double *wall_us,
                                  int64_t *peak_bytes) {
  if (!sv || n < 0 || (n > 0 && (!arrival_us || !resident || !records)))
    return fail(COLA_ERR_INVALID, "server_run_trace: NULL argument");
  for (int64_t i = 0; i < n; ++i) {
    if (resident[i] < 0 || resident[i] >= (int)sv->res.size())
      return fail(COLA_ERR_INVALID, "server_run_trace: unknown resident id");
    if (i && arrival_us[i] < arrival_us[i - 1]) return fail(COLA_ERR_INVALID, "server_run_trace: arrivals not sorted");
    // Checked as admission takes it: a request larger than the whole pool
    // would wait for memory forever.
    const Resident &r = sv->res[resident[i]];
    if (request_pool_bytes(r.out_bytes, r.hub_bytes, !(outputs && outputs[i])) > sv->pool_size)
      return fail(COLA_ERR_NOMEM, "server_run_trace: a request exceeds the memory budget");
  }
  for (int64_t i = 0; i < n; ++i) records[i] = cola_request_record{};
  const int flags = COLA_OVERWRITE | (sv->o.flags & COLA_KAHAN);
  std::vector<ReqState> rs(n);
  std::deque<int64_t> pending;
  int64_t next = 0, done = 0;
  sv->peak_used = sv->used;
  size_t ring = 0;
  cola_status st = COLA_OK;
  const double t0 = now_us();
  auto admit = [&](int64_t i) -> bool {
    const Resident &r = sv->res[resident[i]];
    ReqState &q = rs[i];
    const bool own = !(outputs && outputs[i]);
    q.pool_bytes = own || r.hub_bytes > 0 ? request_pool_bytes(r.out_bytes, r.hub_bytes, own) : 0;
    if (q.pool_bytes > 0) {
      q.pool_off = pool_alloc(sv, q.pool_bytes);
      if (q.pool_off < 0) return false;
      q.live = true;
    }
    q.C = own ? reinterpret_cast<float *>(sv->pool + q.pool_off) : outputs[i];
    q.partials = r.hub_bytes > 0
                     ? reinterpret_cast<float *>(sv->pool + q.pool_off + (own ? round256(r.out_bytes) : 0))
                     : nullptr;
    return true;
  };
  auto launch = [&](LaunchSlot &sl, const std::vector<int64_t> &batch, double tnow) -> cola_status {
    std::vector<BatchSeg> segs;
    std::vector<SegCost> costs;
    std::vector<HubSeg> hubs;
    int64_t task = 0, hth = 0;
    for (int64_t i : batch) {
      Resident &r = sv->res[resident[i]];
      cola_batch_problem pr = r.prob;
      pr.C = rs[i].C;
      pr.ld_C = pr.dim;
      auto it = sv->tune.find(pr.dim);
      const cola_batch_tuning t = it != sv->tune.end() ? it->second : cola_batch_tuning{0, 0, 1};
      cola_status e = append_problem(pr, t, rs[i].partials, sv->o.dtype, flags, &segs, &costs, &hubs, &task, &hth);
      if (e) return e;
      records[i].dispatch_us = tnow;
      records[i].batch_size = (int32_t)batch.size();
    }
    if ((int)segs.size() > sl.cap_segs || (int)hubs.size() > sl.cap_hubs)
      return fail(COLA_ERR_INVALID, "server: work list exceeds the launch context");
    std::copy(segs.begin(), segs.end(), sl.h_segs);
    std::copy(hubs.begin(), hubs.end(), sl.h_hubs);
    cola_status e;
    if (!segs.empty() && (e = hip_check(hipMemcpyAsync(sl.d_segs, sl.h_segs, segs.size() * sizeof(BatchSeg),
                                                       hipMemcpyHostToDevice, sl.stream), "hipMemcpyAsync")))
      return e;
    if (!hubs.empty() && (e = hip_check(hipMemcpyAsync(sl.d_hubs, sl.h_hubs, hubs.size() * sizeof(HubSeg),
                                                       hipMemcpyHostToDevice, sl.stream), "hipMemcpyAsync")))
      return e;
    if ((e = launch_tables(sl.d_segs, (int)segs.size(), (int)task, xcd_bounds(costs, task), sl.d_hubs,
                           (int)hubs.size(), hth, sv->o.dtype, flags, sv->o.blocks_per_cu, sl.stream)))
      return e;
    if ((e = hip_check(hipEventRecord(sl.done, sl.stream), "hipEventRecord"))) return e;
    sl.reqs = batch;
    sl.busy = true;
    return COLA_OK;
  };
  while (done < n && !st) {
    bool progress = false;
    const double tnow = now_us() - t0;
    while (next < n && arrival_us[next] <= tnow) {
      records[next].arrival_us = arrival_us[next];
      records[next].resident = resident[next];
      pending.push_back(next++);
      progress = true;
    }
    for (auto &sl : sv->slots) {
      if (!sl.busy) continue;
      const hipError_t q = hipEventQuery(sl.done);
      if (q == hipErrorNotReady) continue;
      if (q != hipSuccess) {
        st = hip_check(q, "hipEventQuery");
        break;
      }
      const double tdone = now_us() - t0;
      for (int64_t i : sl.reqs) {
        records[i].done_us = tdone;
        if (rs[i].live) pool_free(sv, rs[i].pool_off, rs[i].pool_bytes);
        rs[i].live = false;
        ++done;
      }
      sl.busy = false;
      sl.reqs.clear();
      progress = true;
    }
    if (st) break;
    if (sv->o.plain) {
      // Every arrived request gets its own launch, in arrival order.
      while (!pending.empty()) {
        LaunchSlot &sl = sv->slots[ring % sv->slots.size()];
        if (sl.busy) break;
        const int64_t i = pending.front();
        if (!admit(i)) break;
        pending.pop_front();
        if ((st = launch(sl, {i}, now_us() - t0))) break;
        ++ring;
        progress = true;
      }
    } else {
      for (auto &sl : sv->slots) {
        if (sl.busy || pending.empty()) continue;
        std::vector<int64_t> batch;
        while (!pending.empty() && (int)batch.size() < sv->o.max_batch) {
          if (!admit(pending.front())) break;  // first-come first-served: wait for memory
          batch.push_back(pending.front());
          pending.pop_front();
        }
        if (batch.empty()) break;
        if ((st = launch(sl, batch, now_us() - t0))) break;
        progress = true;
      }
    }
    if (!progress) {
      // Sleep only while nothing is on the device: with a launch in flight
      // the loop keeps polling its event, so a completion is seen as soon as
      // it happens rather than at the next arrival.
      bool inflight = false;
      for (const auto &sl : sv->slots) inflight = inflight || sl.busy;
      if (!inflight && next < n && pending.empty()) {
        const double wait = arrival_us[next] - (now_us() - t0);
        if (wait > 200.0) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)(wait - 100.0)));
      } else if (inflight) {
        std::this_thread::yield();
      }
    }
  }
  for (auto &sl : sv->slots)
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
  if (st) {
    for (int64_t i = 0; i < n; ++i)
      if (rs[i].live) pool_free(sv, rs[i].pool_off, rs[i].pool_bytes);
    for (auto &sl : sv->slots) {
      sl.busy = false;
      sl.reqs.clear();
    }
    return st;
  }
  if (wall_us) *wall_us = now_us() - t0;
  if (peak_bytes) *peak_bytes = sv->peak_used;
  return COLA_OK;
}

cola_status cola_server_destroy(cola_server sv) {
  free_server(sv);
  return COLA_OK;
}

}  // extern "C"
