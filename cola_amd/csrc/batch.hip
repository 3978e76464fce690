// Batched form: many independent SpMMs (each with its own plan, B, C, dim)
// in ONE persistent launch driven by a device work list, optionally captured
// in a hipGraph, with per-dim launch parameters chosen by an autotuner.
//
// Work list: the batch is a table of segments; a segment is one bin (long
// rows + hub chunks, short rows, or empty rows for overwrite) of one
// problem, cut into tasks of `rpw` consecutive items x `n_tiles` column
// tiles.  Tasks are numbered consecutively across segments; each XCD walks a
// contiguous eighth of the task numbers and its waves stride through it, so
// one wave moves monotonically through the segment table.  Within a task the
// lane layout is the plan path's (rows.hpp), with the row-group width G a
// run-time power of two taken from the segment (dims differ per problem).
// Every item is reduced by one team in a fixed order, so a problem's result
// does not depend on which other problems share the launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <new>
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

template <int DT, int EPL, int U, bool KAHAN>
__global__ __launch_bounds__(256) void batch_rows_kernel(const BatchSeg *__restrict__ segs, int n_segs,
                                                         int total, int accumulate) {
  using V = Vec<DT, EPL>;
  using T = typename V::T;
  using Raw = typename V::Raw;
  const int lane = threadIdx.x & 63;
  const int wpb = blockDim.x >> 6;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int xcd = blockIdx.x & 7;
  const int blk = blockIdx.x >> 3;
  const int nblk = (gridDim.x - xcd + 7) >> 3;
  const int t0 = (int)((int64_t)total * xcd / 8), t1 = (int)((int64_t)total * (xcd + 1) / 8);
  const int wave = blk * wpb + wid, nwave = nblk * wpb;
  int si = 0;
  for (int t = t0 + wave; t < t1; t += nwave) {
    while (si < n_segs - 1 && t >= segs[si].task_end) ++si;
    const BatchSeg &sg = segs[si];
    const int lt = t - sg.task_begin;
    const int tile = lt % sg.n_tiles, grp = lt / sg.n_tiles;
    const int glog = sg.g_log2, G = 1 << glog;
    const int S = 64 >> glog;
    const int rpw = sg.rpw, gpt = S / rpw, tsize = 64 / rpw;
    const int sub = lane >> glog, gl = lane & (G - 1);
    const int team = sub / gpt, tsub = sub % gpt;
    const int tbase = team * tsize, tl = lane - tbase;
    const int c0 = tile * sg.tile_w;
    const int w = min(sg.tile_w, sg.dim - c0);
    const int col = gl * EPL;
    const int colc = col < w ? col : 0;
    const T *Bl = static_cast<const T *>(sg.B) + c0;
    const T *Bp = sg.Bp ? static_cast<const T *>(sg.Bp) + c0 : nullptr;
    const int it = grp * rpw + team;
    const bool has = it < sg.n_items;
    int4 item = make_int4(0, 0, 0, 0);
    if (has) item = sg.items[it];
    const int b = item.y, e = item.z;
    const bool comp = KAHAN && (item.w & 1);
    float acc[EPL], err[EPL];
#pragma unroll
    for (int q = 0; q < EPL; ++q) acc[q] = err[q] = 0.f;
    int mc = 0, mc_next = 0;
    float mv = 0.f, mv_next = 0.f;
    if (b + tl < e) {
      mc = sg.cols[b + tl];
      mv = sg.vals[b + tl];
    }
    for (int base = b; base < e; base += tsize) {
      const int n = min(tsize, e - base);
      const int nb = base + tsize;
      if (nb + tl < e) {
        mc_next = sg.cols[nb + tl];
        mv_next = sg.vals[nb + tl];
      }
      for (int j = 0; j < n; j += gpt * U) {
        int cu[U];
        float vu[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int src = (tbase + min(j + u * gpt + tsub, n - 1)) << 2;
          cu[u] = __builtin_amdgcn_ds_bpermute(src, mc);
          vu[u] = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(mv)));
        }
        Raw raw[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int c = cu[u];
          const bool loc = c >= 0;
          const T *rp = (loc ? Bl : Bp) + (int64_t)(loc ? c : ~c) * (loc ? sg.ldB : sg.ldP);
          raw[u] = V::load(rp + colc);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          float v = vu[u];
          if constexpr (DT == COLA_FP8_E4M3) v *= (cu[u] >= 0 ? sg.sl[cu[u]] : sg.sp[~cu[u]]);
          v = (j + u * gpt + tsub < n) ? v : 0.f;
          float x[EPL];
          V::widen(raw[u], x);
          if (!comp) {
#pragma unroll
            for (int q = 0; q < EPL; ++q) acc[q] = fmaf(v, x[q], acc[q]);
          } else {
#pragma unroll
            for (int q = 0; q < EPL; ++q) {
              const float y = fmaf(v, x[q], -err[q]);
              const float tt = acc[q] + y;
              err[q] = (tt - acc[q]) - y;
              acc[q] = tt;
            }
          }
        }
      }
      mc = mc_next;
      mv = mv_next;
    }
    for (int off = G; off < tsize; off <<= 1) {
#pragma unroll
      for (int q = 0; q < EPL; ++q) {
        const float s2 = __shfl_xor(acc[q], off);
        const float e2 = __shfl_xor(err[q], off);
        if (!comp) {
          acc[q] += s2;
        } else {
          const float s1 = acc[q];
          const float tt = s1 + s2;
          const float lost = fabsf(s1) >= fabsf(s2) ? (s1 - tt) + s2 : (s2 - tt) + s1;
          acc[q] = tt;
          err[q] = err[q] + e2 - lost;
        }
      }
    }
    if (has && tsub == 0 && col < w) {
      float *dst;
      bool add_old;
      if (item.x >= 0) {
        dst = sg.C + (int64_t)item.x * sg.ldC + c0 + col;
        add_old = accumulate != 0;
      } else {
        dst = sg.partials + (int64_t)(-item.x - 1) * sg.pstride + c0 + col;
        add_old = false;
      }
#pragma unroll
      for (int q = 0; q < EPL; q += 4) {
        float4 o = make_float4(comp ? acc[q] - err[q] : acc[q], comp ? acc[q + 1] - err[q + 1] : acc[q + 1],
                               comp ? acc[q + 2] - err[q + 2] : acc[q + 2],
                               comp ? acc[q + 3] - err[q + 3] : acc[q + 3]);
        float4 *d4 = reinterpret_cast<float4 *>(dst + q);
        if (add_old) {
          const float4 old = *d4;
          o.x += old.x; o.y += old.y; o.z += old.z; o.w += old.w;
        }
        *d4 = o;
      }
    }
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
  const int c0 = h.hub_ptr[hub], c1 = h.hub_ptr[hub + 1];
  float s = 0.f, err = 0.f;
  for (int c = c0; c < c1; ++c) {
    const float x = h.partials[(int64_t)c * h.pstride + col];
    if (kahan) {
      const float y = x - err;
      const float tt = s + y;
      err = (tt - s) - y;
      s = tt;
    } else {
      s += x;
    }
  }
  float *dst = h.C + (int64_t)h.hub_rows[hub] * h.ldC + col;
  *dst = accumulate ? *dst + s : s;
}

template <int DT, int EPL>
void launch_batch(const BatchSeg *segs, int n_segs, int total, int accumulate, bool kahan, int blocks,
                  hipStream_t s) {
  if (kahan)
    hipLaunchKernelGGL((batch_rows_kernel<DT, EPL, 8, true>), dim3(blocks), dim3(256), 0, s, segs, n_segs,
                       total, accumulate);
  else
    hipLaunchKernelGGL((batch_rows_kernel<DT, EPL, 8, false>), dim3(blocks), dim3(256), 0, s, segs, n_segs,
                       total, accumulate);
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

// Rebuilds the segment / hub tables from the problems and their tunings.
static cola_status build_tables(cola_batch_s *b) {
  drop_graph(b);
  std::vector<BatchSeg> segs;
  std::vector<HubSeg> hubs;
  const size_t eb = dtype_bytes(b->dtype);
  const int ev = vec_elems(b->dtype);
  int64_t task = 0, hthreads = 0;
  for (size_t k = 0; k < b->probs.size(); ++k) {
    const cola_batch_problem &pr = b->probs[k];
    const cola_plan_s *p = pr.plan;
    const cola_batch_tuning &t = b->tune[k];
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
    const PlanBin *bins[3] = {&p->longb, &p->shortb, (b->flags & COLA_OVERWRITE) ? &p->emptyb : nullptr};
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
      s.partials = b->partials[k];
      s.pstride = p->pstride;
      s.n_items = (int32_t)bin->n;
      s.dim = pr.dim;
      s.rpw = rpws[bi];
      s.n_tiles = n_tiles;
      s.tile_w = tile;
      s.g_log2 = ilog2(G);
      const int64_t tasks = (bin->n + s.rpw - 1) / s.rpw * n_tiles;
      s.task_begin = (int32_t)task;
      task += tasks;
      if (task > INT32_MAX) return fail(COLA_ERR_INVALID, "batch: too many tasks");
      s.task_end = (int32_t)task;
      segs.push_back(s);
    }
    if (p->num_hubs > 0) {
      HubSeg h{};
      h.hub_rows = p->hub_rows;
      h.hub_ptr = p->hub_ptr;
      h.partials = b->partials[k];
      h.pstride = p->pstride;
      h.C = pr.C;
      h.ldC = pr.ld_C;
      h.n_hubs = (int32_t)p->num_hubs;
      h.dim = pr.dim;
      h.thread_begin = hthreads;
      hthreads += (int64_t)p->num_hubs * pr.dim;
      hubs.push_back(h);
    }
    (void)eb;
  }
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

static cola_status run_batch(cola_batch_s *b, hipStream_t s) {
  const int accumulate = (b->flags & COLA_OVERWRITE) ? 0 : 1;
  const bool kahan = (b->flags & COLA_KAHAN) != 0;
  if (b->total_tasks > 0) {
    const int64_t waves = b->total_tasks;
    int64_t blocks = std::min<int64_t>((waves + 3) / 4, (int64_t)device_cu_count() * b->blocks_per_cu);
    blocks = std::max<int64_t>(8, (blocks + 7) / 8 * 8);
    const int ev = vec_elems(b->dtype);
    switch (b->dtype) {
      case COLA_F32: launch_batch<COLA_F32, 4>(b->d_segs, b->n_segs, b->total_tasks, accumulate, kahan, (int)blocks, s); break;
      case COLA_BF16: launch_batch<COLA_BF16, 8>(b->d_segs, b->n_segs, b->total_tasks, accumulate, kahan, (int)blocks, s); break;
      case COLA_F16: launch_batch<COLA_F16, 8>(b->d_segs, b->n_segs, b->total_tasks, accumulate, kahan, (int)blocks, s); break;
      case COLA_FP8_E4M3: launch_batch<COLA_FP8_E4M3, 16>(b->d_segs, b->n_segs, b->total_tasks, accumulate, kahan, (int)blocks, s); break;
    }
    (void)ev;
  }
  if (b->hub_threads > 0) {
    const int64_t blocks = (b->hub_threads + 255) / 256;
    hipLaunchKernelGGL(batch_hub_reduce_kernel, dim3((unsigned)blocks), dim3(256), 0, s, b->d_hubs, b->n_hubs,
                       b->hub_threads, accumulate, kahan ? 1 : 0);
  }
  return hip_check(hipGetLastError(), "batch launch");
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
  const size_t eb = dtype_bytes(dtype);
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
    (void)eb;
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
      for (int r : {1, 2, 4}) {
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
    for (int bpc : {2, 4, 8, 16}) {
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
