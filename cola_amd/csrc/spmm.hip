// Device side of the single-partition SpMM engine for MI355X (gfx950, CDNA4,
// wave64).  Kernels:
//   pack_rows_kernel      gather-pack of unique remote B rows into one
//                         contiguous buffer (replaces the per-warp NVSHMEM get,
//                         dist_spmm_ppl_batch.cuh:120-125, and
//                         batching_data_ppl, cuh:58-73)
//   ppl0_kernel           reference-semantics partition SpMM: virtual local
//                         CSR rows + virtual remote CSC columns, FP32 atomics
//                         into C (ppl0_batch_SpMM_kernel, cuh:76-200)
//   plan_rows_kernel      degree-binned, atomic-free row kernel over the
//                         merged (local + packed-remote) row lists
//   hub_reduce_kernel     fixed-order reduction of hub-row chunk partials
//   zero_rows_kernel      zeroes rows without entries (overwrite mode)
//   quantize_rows_kernel  FP32 rows -> BF16 / FP16 / FP8-e4m3 (+ row scale)
//   error_norms_kernel    FP64 error norms of C against an FP64 reference
//
//   zero_matrix_kernel    zeroes C before the column-panel passes
//   fill_f64_kernel       device fill of the FP64 norm accumulators
//
// The plan also caches itself behind the reference's single-stage wrapper
// (cola_dist_spmm_ppl_batch): a plan rebuilt from the two zips on first use.
//
// SpMM is HBM/cache-bandwidth bound (about 2 FLOP per 8+ bytes), so nothing
// here uses MFMA.  The dense dimension is spread across the lanes of a row
// group (G lanes, 16 bytes per lane per vector), several row groups share a
// wave, and entries are staged per team lane with coalesced loads, turned
// into (B row address, value) pairs there and broadcast to the groups with
// ds_bpermute (rows.hpp).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "cola_internal.hpp"
#include "rows.hpp"
#include "device.hpp"

using namespace cola_devhost;

namespace cola_devhost {
using cola::fail;
bool aligned(const void *p, size_t bytes) { return ((uintptr_t)p % bytes) == 0; }

cola_status hip_check(hipError_t e, const char *what) {
  if (e == hipSuccess) return COLA_OK;
  return fail(COLA_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

int device_cu_count() {
  static std::atomic<int> cached[64] = {};  // per device; callers on several threads
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int n = cached[dev].load(std::memory_order_relaxed);
  if (n == 0) {
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
      n = 256;
    cached[dev].store(n, std::memory_order_relaxed);
  }
  return n;
}

size_t dtype_bytes(int dt) {
  switch (dt) {
    case COLA_F32: return 4;
    case COLA_BF16: case COLA_F16: return 2;
    case COLA_FP8_E4M3: return 1;
  }
  return 0;
}

int vec_elems(int dt) { return dt == COLA_F32 ? 4 : (dt == COLA_FP8_E4M3 ? 16 : 8); }

// The row kernels' zero row (RowArgs::zrow): a static device array, so zero
// from load and never written, which every GPU holds for the life of the
// process at its own address, looked up once per device.
__device__ __align__(256) uint8_t zero_row[cola_dev::kZeroRowBytes];
static std::atomic<uint64_t> g_zero_row_dev[64] = {};

cola_status zero_row_address(uint64_t *out) {
  int dev = -1;
  cola_status st = hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (st) return st;
  if (dev < 0 || dev >= 64) return fail(COLA_ERR_UNSUPPORTED, "device id above 63");
  uint64_t a = g_zero_row_dev[dev].load(std::memory_order_relaxed);
  if (!a) {
    void *ptr = nullptr;
    if ((st = hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(zero_row)), "hipGetSymbolAddress(zero_row)")))
      return st;
    if (!ptr) return fail(COLA_ERR_HIP, "hipGetSymbolAddress(zero_row) returned NULL");
    a = reinterpret_cast<uint64_t>(ptr);
    g_zero_row_dev[dev].store(a, std::memory_order_relaxed);
  }
  *out = a;
  return COLA_OK;
}
}  // namespace cola_devhost

using cola::fail;

namespace {

using cola_dev::Shape;
using cola_dev::Vec;
using cola_dev::fp8e4m3_to_f32;

// Round-to-nearest-even into bfloat16 (NaN -> quiet NaN with its sign).
__device__ __forceinline__ uint32_t f32_to_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return ((u >> 16) & 0x8000u) | 0x7fc0u;
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}
// Round-to-nearest-even into OCP FP8 E4M3 (e4m3fn), saturating at +-448.
__device__ __forceinline__ uint32_t f32_to_fp8e4m3_rne(float f) {
  const uint32_t sign = (__float_as_uint(f) >> 24) & 0x80u;
  float a = fabsf(f);
  if (!(a < 448.0f)) return sign | 0x7eu;  // also maps NaN to max finite
  if (a < 0x1p-6f) {                      // subnormal range: quantum 2^-9
    const uint32_t q = (uint32_t)__builtin_rintf(a * 512.0f);  // 0..8
    return sign | q;                                            // 8 == min normal
  }
  uint32_t u = __float_as_uint(a);
  u += 0x7ffffu + ((u >> 20) & 1u);  // RNE from 23 to 3 mantissa bits
  const int e = (int)(u >> 23) - 127 + 7;
  const uint32_t m = (u >> 20) & 7u;
  if (e > 15 || (e == 15 && m == 7)) return sign | 0x7eu;
  return sign | ((uint32_t)e << 3) | m;
}



// ---------------------------------------------------------------------------
// Gather-pack: packed[u] = B[cols[u]].  A group of lanes copies one row with
// 16-byte (or, when unaligned, element-sized) accesses.
// ---------------------------------------------------------------------------
template <typename V>
__global__ __launch_bounds__(256) void pack_rows_kernel(
    const char *__restrict__ B, int64_t ldB_bytes, const int32_t *__restrict__ cols,
    int64_t num_rows, int64_t row_vecs, char *__restrict__ out,
    int64_t ldO_bytes, const float *__restrict__ scales, float *__restrict__ out_scales,
    int lanes_per_row) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t groups_total = (int64_t)gridDim.x * blockDim.x / lanes_per_row;
  const int64_t g0 = tid / lanes_per_row;
  const int gl = (int)(tid % lanes_per_row);
  for (int64_t r = g0; r < num_rows; r += groups_total) {
    const int64_t src = cols[r];
    const V *s = reinterpret_cast<const V *>(B + src * ldB_bytes);
    V *d = reinterpret_cast<V *>(out + r * ldO_bytes);
    for (int64_t k = gl; k < row_vecs; k += lanes_per_row) d[k] = s[k];
    if (out_scales && gl == 0) out_scales[r] = scales[src];
  }
}

// ---------------------------------------------------------------------------
// Reference-semantics kernel (ppl0_batch_SpMM_kernel, cuh:76-200).  One group
// of G lanes per task: tasks [0, L) are virtual local rows, tasks [L, L+R)
// are virtual remote columns (visited in reverse when peid & 2, cuh:115-118).
// ---------------------------------------------------------------------------
template <int EPL, int G, int NV>
__global__ __launch_bounds__(256) void ppl0_kernel(
    float *__restrict__ out, const float *__restrict__ in, const float *__restrict__ packed,
    const int32_t *__restrict__ rowPtr_l, const int32_t *__restrict__ colIdx_l,
    const int32_t *__restrict__ rowIdx_l, const float *__restrict__ val_l,
    const int32_t *__restrict__ colPtr_r, const int32_t *__restrict__ rowIdx_r,
    const int32_t *__restrict__ packIdx_r, const float *__restrict__ val_r,
    int local_lb, int remote_lb, int L, int R, int dim, int peid) {
  using RowT = Vec<COLA_F32, EPL>;
  const int lane = threadIdx.x & 63;
  const int sub = lane / G, gl = lane % G;
  const int64_t T = (int64_t)L + R;
  const int64_t task = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  if (task >= T) return;
  const int gbase = sub * G;
  float acc[NV][EPL];
#pragma unroll
  for (int k = 0; k < NV; ++k)
#pragma unroll
    for (int q = 0; q < EPL; ++q) acc[k][q] = 0.f;

  if (task < L) {
    const int v = local_lb + (int)task;
    const int b = rowPtr_l[v], e = rowPtr_l[v + 1];
    for (int base = b; base < e; base += G) {
      const int cnt = min(G, e - base);
      int mc = 0;
      float mv = 0.f;
      if (gl < cnt) {
        mc = colIdx_l[base + gl];
        mv = val_l[base + gl];
      }
      for (int j = 0; j < cnt; j += 4) {
        float x[4][NV][EPL];
        float a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int jj = min(j + u, cnt - 1);
          const int c = __shfl(mc, gbase + jj);
          a[u] = (j + u < cnt) ? __shfl(mv, gbase + jj) : 0.f;
          const float *rp = in + (int64_t)c * dim;
#pragma unroll
          for (int k = 0; k < NV; ++k) {
            const int col = (k * G + gl) * EPL;
            // A slot past the row's end (j + u >= cnt) re-reads the last
            // entry's row; it is replaced by zeros so that an Inf or NaN
            // there never meets the slot's zero value.
            if (col < dim && j + u < cnt) RowT::widen(RowT::load(rp + col), x[u][k]);
            else
#pragma unroll
              for (int q = 0; q < EPL; ++q) x[u][k][q] = 0.f;
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int k = 0; k < NV; ++k)
#pragma unroll
            for (int q = 0; q < EPL; ++q) acc[k][q] += a[u] * x[u][k][q];
      }
    }
    const int row = rowIdx_l[v];
    float *op = out + (int64_t)row * dim;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const int col = (k * G + gl) * EPL;
      if (col < dim)
#pragma unroll
        for (int q = 0; q < EPL; ++q) unsafeAtomicAdd(op + col + q, acc[k][q]);
    }
    return;
  }

  // Remote virtual column: packed row held in registers, scattered with atomics.
  const int t = (int)(task - L);
  const int idx = remote_lb + ((peid & 2) == 0 ? t : (int)(R - t - 1));
  const float *rp = packed + (int64_t)packIdx_r[idx] * dim;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const int col = (k * G + gl) * EPL;
    if (col < dim) RowT::widen(RowT::load(rp + col), acc[k]);
  }
  const int b = colPtr_r[idx], e = colPtr_r[idx + 1];
  for (int base = b; base < e; base += G) {
    const int cnt = min(G, e - base);
    int mr = 0;
    float mv = 0.f;
    if (gl < cnt) {
      mr = rowIdx_r[base + gl];
      mv = val_r[base + gl];
    }
    for (int j = 0; j < cnt; ++j) {
      const int r = __shfl(mr, gbase + j);
      const float a = __shfl(mv, gbase + j);
      float *op = out + (int64_t)r * dim;
#pragma unroll
      for (int k = 0; k < NV; ++k) {
        const int col = (k * G + gl) * EPL;
        if (col < dim)
#pragma unroll
          for (int q = 0; q < EPL; ++q) unsafeAtomicAdd(op + col + q, a * acc[k][q]);
      }
    }
  }
}

// Hub rows: thread (h, col) adds the hub's chunk partials in chunk order.
__global__ __launch_bounds__(256) void hub_reduce_kernel(
    const int32_t *__restrict__ hub_rows, const int32_t *__restrict__ hub_ptr, int num_hubs,
    const float *__restrict__ partials, int64_t pstride, float *__restrict__ C, int64_t ldC,
    int dim, int accumulate, int kahan) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)num_hubs * dim) return;
  const int h = (int)(t / dim), col = (int)(t % dim);
  const float s = cola_dev::hub_sum(partials + col, pstride, hub_ptr[h], hub_ptr[h + 1], kahan != 0);
  float *dst = C + (int64_t)hub_rows[h] * ldC + col;
  *dst = accumulate ? *dst + s : s;
}

__global__ __launch_bounds__(256) void zero_rows_kernel(const int32_t *__restrict__ rows,
                                                        int num, float *__restrict__ C,
                                                        int64_t ldC, int dim) {
  const int r = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= num) return;
  float *dst = C + (int64_t)rows[r] * ldC;
  for (int col = lane; col < dim; col += 64) dst[col] = 0.f;
}

__global__ __launch_bounds__(256) void zero_matrix_kernel(float *__restrict__ C, int64_t ldC, int64_t rows,
                                                          int dim) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = rows * dim;
  for (int64_t i = t; i < total; i += (int64_t)gridDim.x * blockDim.x) C[(i / dim) * ldC + i % dim] = 0.f;
}

// One wave per row.  FP8 rows: scale = max|x| / 448 (1 for an all-zero row).
__global__ __launch_bounds__(256) void quantize_rows_kernel(const float *__restrict__ src,
                                                            int64_t lds, int64_t rows, int dim,
                                                            int dt, void *__restrict__ dst,
                                                            int64_t ldd, float *__restrict__ scales) {
  const int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (r >= rows) return;
  const float *s = src + r * lds;
  if (dt == COLA_F32) {
    float *d = (float *)dst + r * ldd;
    for (int c = lane; c < dim; c += 64) d[c] = s[c];
  } else if (dt == COLA_BF16) {
    uint16_t *d = (uint16_t *)dst + r * ldd;
    for (int c = lane; c < dim; c += 64) d[c] = (uint16_t)f32_to_bf16_rne(s[c]);
  } else if (dt == COLA_F16) {
    uint16_t *d = (uint16_t *)dst + r * ldd;
    for (int c = lane; c < dim; c += 64) d[c] = __half_as_ushort(__float2half(s[c]));
  } else {
    float m = 0.f;
    for (int c = lane; c < dim; c += 64) m = fmaxf(m, fabsf(s[c]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    const float scale = m > 0.f ? m / 448.0f : 1.0f;
    uint8_t *d = (uint8_t *)dst + r * ldd;
    for (int c = lane; c < dim; c += 64) d[c] = (uint8_t)f32_to_fp8e4m3_rne(s[c] / scale);
    if (lane == 0) scales[r] = scale;
  }
}

__global__ void fill_f64_kernel(double *p, int n) {
  const int i = threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// One wave per row: FP64 sums of squared error and of squared reference.
__global__ __launch_bounds__(256) void error_norms_kernel(const float *__restrict__ C, int64_t ldC,
                                                          const double *__restrict__ R,
                                                          int64_t ldR, int64_t rows, int dim,
                                                          double *__restrict__ out) {
  const int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (r >= rows) return;
  double se = 0.0, sr = 0.0;
  for (int c = lane; c < dim; c += 64) {
    const double ref = R[r * ldR + c];
    const double d = (double)C[r * ldC + c] - ref;
    se += d * d;
    sr += ref * ref;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    se += __shfl_xor(se, off);
    sr += __shfl_xor(sr, off);
  }
  if (lane == 0) {
    unsafeAtomicAdd(out + 0, se);
    unsafeAtomicAdd(out + 1, sr);
    if (sr > 0.0) {
      // Non-negative doubles order like their bit patterns.
      const double ratio = se / sr;
      atomicMax(reinterpret_cast<unsigned long long *>(out + 2),
                (unsigned long long)__double_as_longlong(ratio));
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Handles.
// ---------------------------------------------------------------------------
static void free_plan(cola_plan_s *p);
// Guards the plans the single-stage wrapper caches in zips.
static std::mutex g_zip_plan_mu;

// Drops the reference-signature launcher's cached plans built from any of
// these device arrays (their memory is about to be freed and may be reused).
static void drop_ref_plans_using(const void *a, const void *b, const void *c, const void *d);

static void free_zip(cola_zip_s *z) {
  if (!z) return;
  drop_ref_plans_using(z->ptr, z->idx, z->zip, z->val);
  {
    std::lock_guard<std::mutex> lock(g_zip_plan_mu);
    z->fast_plan.reset();
  }
  (void)hipFree(z->ptr);
  (void)hipFree(z->idx);
  (void)hipFree(z->zip);
  (void)hipFree(z->val);
  (void)hipFree(z->pack_idx);
  (void)hipFree(z->uniq);
  delete z;
}

static void free_plan(cola_plan_s *p) {
  if (!p) return;
  for (PlanBin *b : {&p->shortb, &p->longb, &p->shortb_r, &p->longb_r, &p->emptyb, &p->cshortb, &p->clongb,
                     &p->tinyb, &p->shortb_nt}) {
    (void)hipFree(b->items);
    (void)hipFree(b->xcd);
  }
  (void)hipFree(p->cols);
  (void)hipFree(p->vals);
  (void)hipFree(p->empty_rows);
  (void)hipFree(p->hub_rows);
  (void)hipFree(p->hub_ptr);
  (void)hipFree(p->remote_cols);
  (void)hipFree(p->partials);
  (void)hipFree(p->panel_items);
  (void)hipFree(p->panel_xcd);
  (void)hipFree(p->panel_hub_ptr);
  (void)hipFree(p->tiny_rows);
  delete p;
}

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
namespace {

template <int EPL, int G, int NV>
void launch_ppl0(float *out, const float *in, const float *packed, const int32_t *rp,
                 const int32_t *ci, const int32_t *ri, const float *vl, const int32_t *cp,
                 const int32_t *rr, const int32_t *pi, const float *vr, int llb, int rlb,
                 int L, int R, int dim, int peid, hipStream_t s) {
  const int64_t threads = ((int64_t)L + R) * G;
  const int64_t blocks = (threads + 255) / 256;
  hipLaunchKernelGGL((ppl0_kernel<EPL, G, NV>), dim3((unsigned)blocks), dim3(256), 0, s, out,
                     in, packed, rp, ci, ri, vl, cp, rr, pi, vr, llb, rlb, L, R, dim, peid);
}

bool launch_rows(int dt, const Shape &sh, const cola_dev::RowArgs &a, int blocks, hipStream_t s, bool u16) {
  switch (dt) {
    case COLA_F32: return cola_dev::launch_rows_f32(sh, a, blocks, 256, s, u16);
    case COLA_BF16: return cola_dev::launch_rows_bf16(sh, a, blocks, 256, s, u16);
    case COLA_F16: return cola_dev::launch_rows_f16(sh, a, blocks, 256, s, u16);
    case COLA_FP8_E4M3: return cola_dev::launch_rows_fp8(sh, a, blocks, 256, s, u16);
  }
  return false;
}

bool launch_tiny(int dt, const Shape &sh, const cola_dev::RowArgs &a, const int32_t *rowtab, int blocks,
                 hipStream_t s) {
  switch (dt) {
    case COLA_F32: return cola_dev::launch_tiny_f32(sh, a, rowtab, blocks, s);
    case COLA_BF16: return cola_dev::launch_tiny_bf16(sh, a, rowtab, blocks, s);
    case COLA_F16: return cola_dev::launch_tiny_f16(sh, a, rowtab, blocks, s);
    case COLA_FP8_E4M3: return cola_dev::launch_tiny_fp8(sh, a, rowtab, blocks, s);
  }
  return false;
}

int clamp_rpw(int req, int S) {
  int r = 1;
  while (r * 2 <= req && r * 2 <= S) r *= 2;
  return r;
}

struct SpmmCall {
  float *C;
  int64_t ldC;
  const void *Bl;
  int64_t ldB;
  const float *sl;
  const void *Bp;
  int64_t ldP;
  const float *sp;
  int dim;
  int dtype;
  int flags;
};

// Validates a call and runs it with the given tuning.  ws / ws_stride: the
// hub-chunk workspace (one row of ws_stride floats per chunk); NULL = the
// plan's own.
cola_status run_plan(const cola_plan_s *p, const SpmmCall &c, const cola_plan_tuning &t,
                     hipStream_t s, float *ws = nullptr, int64_t ws_stride = 0) {
  float *const partials = ws ? ws : p->partials;
  const int64_t pstride = ws ? ws_stride : p->pstride;
  const size_t eb = dtype_bytes(c.dtype);
  if (!eb) return fail(COLA_ERR_UNSUPPORTED, "spmm_partition: unknown dtype");
  if (c.flags & ~(COLA_OVERWRITE | COLA_KAHAN)) return fail(COLA_ERR_INVALID, "spmm_partition: unknown flag");
  if (c.dim < 1 || c.ldC < c.dim || c.ldB < c.dim || (p->U_r > 0 && c.ldP < c.dim))
    return fail(COLA_ERR_INVALID, "spmm_partition: bad dim or leading dimension");
  if (!ws && c.dim > p->max_dim && (p->num_chunks > 0 || p->panel_chunks > 0))
    return fail(COLA_ERR_INVALID, "spmm_partition: dim exceeds the plan's max_dim");
  if (p->num_rows > 0 && !c.C) return fail(COLA_ERR_INVALID, "spmm_partition: C is NULL");
  if (p->U_l > 0 && !c.Bl) return fail(COLA_ERR_INVALID, "spmm_partition: B_local is NULL");
  if (p->U_r > 0 && !c.Bp) return fail(COLA_ERR_INVALID, "spmm_partition: packed is NULL");
  if (c.dtype == COLA_FP8_E4M3 && ((p->U_l > 0 && !c.sl) || (p->U_r > 0 && !c.sp)))
    return fail(COLA_ERR_INVALID, "spmm_partition: FP8 needs row scales");
  const int ev = vec_elems(c.dtype);
  const bool vec_ok = c.ldC % 4 == 0 && c.ldB % ev == 0 && (p->U_r == 0 || c.ldP % ev == 0) &&
                      aligned(c.C, 16) && aligned(c.Bl, 16) && (p->U_r == 0 || aligned(c.Bp, 16));
  // Column tile: a multiple of the vector width (so tiles stay aligned).
  int tile = c.dim;
  const int want = t.col_tile < 0 ? -t.col_tile * 128 / dtype_bytes(c.dtype) : t.col_tile;
  if (want > 0 && want < c.dim && vec_ok && c.dim % ev == 0) {
    tile = std::max(ev, want / ev * ev);
  }
  const int accumulate = (c.flags & COLA_OVERWRITE) ? 0 : 1;
  const int kahan = (c.flags & COLA_KAHAN) ? 1 : 0;
  // Common base of the two row tables of a column tile starting at column
  // c0 (w columns): nonzero when every row an entry can address (local rows
  // below nodesPerPE, packed rows below U_r) lies within 2^36 bytes of it
  // and starts on a 16-byte boundary, so the row kernel can broadcast a
  // 32-bit offset in 16-byte units instead of a 64-bit address.
  uint64_t zrow = 0;
  if (cola_status zst = zero_row_address(&zrow)) return zst;
  auto table_base = [&](int c0, int w) -> uint64_t {
    if (!vec_ok) return 0;
    uint64_t lo = zrow, hi = zrow + cola_dev::kZeroRowBytes;
    auto span = [&](const void *base, int64_t rows, int64_t ld) {
      if (!base || rows <= 0) return;
      const uint64_t b = (uint64_t)(uintptr_t)base + (uint64_t)c0 * eb;
      lo = std::min(lo, b);
      hi = std::max(hi, b + ((uint64_t)(rows - 1) * (uint64_t)ld + (uint64_t)w) * eb);
    };
    span(c.Bl, p->nodesPerPE, c.ldB);
    span(c.Bp, p->remote_rows, c.ldP);
    if (lo % 16 != 0 || hi - lo >= (uint64_t(1) << 36)) return 0;
    return lo;
  };
  if (!accumulate && p->num_empty > 0) {
    const int64_t blocks = (p->num_empty * 64 + 255) / 256;
    hipLaunchKernelGGL(zero_rows_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p->empty_rows,
                       (int)p->num_empty, c.C, c.ldC, c.dim);
  }
  const int bpc = t.blocks_per_cu > 0 ? t.blocks_per_cu : 8;
  if (t.panels && p->panels > 1) {
    // Column-panel schedule: zero C (overwrite), then one accumulating pass
    // per (tile, panel), then the hub rows' partial sums added on top.
    if (!accumulate) {
      const int64_t total = (int64_t)p->num_rows * c.dim;
      const int64_t blocks = std::min<int64_t>((total + 255) / 256, (int64_t)device_cu_count() * 64);
      if (total > 0)
        hipLaunchKernelGGL(zero_matrix_kernel, dim3((unsigned)std::max<int64_t>(1, blocks)), dim3(256), 0, s, c.C,
                           c.ldC, (int64_t)p->num_rows, c.dim);
    }
    for (int c0 = 0; c0 < c.dim; c0 += tile) {
      const int w = std::min(tile, c.dim - c0);
      Shape sh;
      if (!cola_dev::pick_shape(c.dtype, w, vec_ok, &sh))
        return fail(COLA_ERR_UNSUPPORTED, "spmm_partition: dim " + std::to_string(c.dim) + " not supported");
      const int S = 64 / sh.g;
      for (int k = 0; k < 2 * p->panels; ++k) {
        const int64_t n = p->panel_begin[k + 1] - p->panel_begin[k];
        if (n == 0) continue;
        cola_dev::RowArgs a;
        a.items = p->panel_items + p->panel_begin[k];
        a.xcd_begin = p->panel_xcd + 9 * k;
        a.cols = p->cols;
        a.vals = p->vals;
        a.Bl = c.Bl ? (const char *)c.Bl + (size_t)c0 * eb : nullptr;
        a.ldB = c.ldB;
        a.sl = c.sl;
        a.Bp = c.Bp ? (const char *)c.Bp + (size_t)c0 * eb : nullptr;
        a.ldP = c.ldP;
        a.sp = c.sp;
        a.C = c.C + c0;
        a.ldC = c.ldC;
        a.partials = partials ? partials + c0 : nullptr;
        a.pstride = pstride;
        a.dim = w;
        a.accumulate = 1;
        a.kahan = kahan;
        a.rpw = clamp_rpw(t.rpw_long > 0 ? t.rpw_long : 1, S);
        a.xcd_remap = t.xcd_remap ? 1 : 0;
        a.phase = 0;
        a.abase = table_base(c0, w);
        a.zrow = zrow;
        const int64_t waves = (n + a.rpw - 1) / a.rpw;
        int64_t blocks = std::min<int64_t>((waves + 3) / 4, (int64_t)device_cu_count() * bpc);
        blocks = std::max<int64_t>(8, (blocks + 7) / 8 * 8);
        if (!launch_rows(c.dtype, sh, a, (int)blocks, s, t.unroll == 16))
          return fail(COLA_ERR_UNSUPPORTED, "spmm_partition: no kernel for this dim");
      }
    }
    if (p->num_hubs > 0) {
      const int64_t threads = p->num_hubs * (int64_t)c.dim;
      const int64_t blocks = (threads + 255) / 256;
      hipLaunchKernelGGL(hub_reduce_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p->hub_rows, p->panel_hub_ptr,
                         (int)p->num_hubs, partials, pstride, c.C, c.ldC, c.dim, 1, kahan);
    }
    return hip_check(hipGetLastError(), "spmm_partition launch");
  }
  for (int c0 = 0; c0 < c.dim; c0 += tile) {
    const int w = std::min(tile, c.dim - c0);
    Shape sh;
    if (!cola_dev::pick_shape(c.dtype, w, vec_ok, &sh))
      return fail(COLA_ERR_UNSUPPORTED, "spmm_partition: dim " + std::to_string(c.dim) + " not supported");
    const int S = 64 / sh.g;
    // phase 0: every entry at once; with split_remote, phase 1 (local
    // entries over this partition's B slice) then phase 2 (remote entries
    // over the packed buffer, added on top).  A hot/cold plan runs those over
    // its hot items, then phase 3 over the cold items (added on top), all
    // inside this column tile so the hot rows stay in L2.
    const bool cold = p->cshortb.n + p->clongb.n > 0;
    const bool split = t.split_remote && p->U_r > 0;
    // Tiny-row bundles replace their rows' short-bin items in a one-pass
    // schedule whose lane groups hold one 16-byte vector and span >= 8 lanes.
    const bool tiny = t.tiny && !split && p->tinyb.n > 0 && cola_dev::tiny_shape_ok(c.dtype, sh);
    for (int phase : {split ? 1 : 0, split ? 2 : -1, cold ? 3 : -1}) {
      if (phase < 0) continue;
      const PlanBin *bins[2] = {phase == 2 ? &p->longb_r : (phase == 3 ? &p->clongb : &p->longb),
                                phase == 2 ? &p->shortb_r
                                           : (phase == 3 ? &p->cshortb : (tiny ? &p->shortb_nt : &p->shortb))};
      for (int bi = 0; bi < 2; ++bi) {
        const PlanBin *bin = bins[bi];
        if (bin->n == 0) continue;
        cola_dev::RowArgs a;
        a.items = bin->items;
        a.xcd_begin = bin->xcd;
        a.cols = p->cols;
        a.vals = p->vals;
        a.Bl = c.Bl ? (const char *)c.Bl + (size_t)c0 * eb : nullptr;
        a.ldB = c.ldB;
        a.sl = c.sl;
        a.Bp = c.Bp ? (const char *)c.Bp + (size_t)c0 * eb : nullptr;
        a.ldP = c.ldP;
        a.sp = c.sp;
        a.C = c.C + c0;
        a.ldC = c.ldC;
        a.partials = partials ? partials + c0 : nullptr;
        a.pstride = pstride;
        a.dim = w;
        a.accumulate = accumulate;
        a.kahan = kahan;
        const int req = bi == 1 ? (t.rpw_short > 0 ? t.rpw_short : S) : (t.rpw_long > 0 ? t.rpw_long : 1);
        a.rpw = clamp_rpw(req, S);
        a.xcd_remap = t.xcd_remap ? 1 : 0;
        a.phase = phase;
        a.abase = table_base(c0, w);
        a.zrow = zrow;
        const int64_t waves = (bin->n + a.rpw - 1) / a.rpw;
        const int bin_bpc = bi == 1 && t.blocks_per_cu_short > 0 ? t.blocks_per_cu_short : bpc;
        int64_t blocks = std::min<int64_t>((waves + 3) / 4, (int64_t)device_cu_count() * bin_bpc);
        blocks = std::max<int64_t>(8, (blocks + 7) / 8 * 8);
        if (!launch_rows(c.dtype, sh, a, (int)blocks, s, t.unroll == 16))
          return fail(COLA_ERR_UNSUPPORTED, "spmm_partition: no kernel for this dim");
      }
      if (tiny && phase == 0) {
        cola_dev::RowArgs a;
        a.items = p->tinyb.items;
        a.xcd_begin = p->tinyb.xcd;
        a.cols = p->cols;
        a.vals = p->vals;
        a.Bl = c.Bl ? (const char *)c.Bl + (size_t)c0 * eb : nullptr;
        a.ldB = c.ldB;
        a.sl = c.sl;
        a.Bp = c.Bp ? (const char *)c.Bp + (size_t)c0 * eb : nullptr;
        a.ldP = c.ldP;
        a.sp = c.sp;
        a.C = c.C + c0;
        a.ldC = c.ldC;
        a.partials = nullptr;
        a.pstride = 0;
        a.dim = w;
        a.accumulate = accumulate;
        a.kahan = 0;
        a.rpw = S;
        a.xcd_remap = t.xcd_remap ? 1 : 0;
        a.phase = 0;
        a.abase = table_base(c0, w);
        a.zrow = zrow;
        const int64_t waves = (p->tinyb.n + S - 1) / S;
        const int bin_bpc = t.blocks_per_cu_short > 0 ? t.blocks_per_cu_short : bpc;
        int64_t blocks = std::min<int64_t>((waves + 3) / 4, (int64_t)device_cu_count() * bin_bpc);
        blocks = std::max<int64_t>(8, (blocks + 7) / 8 * 8);
        if (!launch_tiny(c.dtype, sh, a, p->tiny_rows, (int)blocks, s))
          return fail(COLA_ERR_UNSUPPORTED, "spmm_partition: no tiny-row kernel for this dim");
      }
    }
  }
  if (p->num_hubs > 0) {
    const int64_t threads = p->num_hubs * (int64_t)c.dim;
    const int64_t blocks = (threads + 255) / 256;
    hipLaunchKernelGGL(hub_reduce_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p->hub_rows,
                       p->hub_ptr, (int)p->num_hubs, partials, pstride, c.C, c.ldC, c.dim,
                       accumulate, kahan);
  }
  return hip_check(hipGetLastError(), "spmm_partition launch");
}

}  // namespace

// direct_remote: remote entries address global row g of B_global (~g)
// instead of packed row u (~u), for callers that pass B_global as the
// remote table.
static cola_status plan_from_partition(const cola::Partition &part, const cola_plan_opts &o, cola_plan_s **out,
                                       bool direct_remote = false);
static cola_status zip_pair_plan(cola_zip_s *l, const cola_zip_s *r, int P, int peid, int dim,
                                 std::shared_ptr<cola_plan_s> *out);

// ---------------------------------------------------------------------------
// C ABI (device part).
// ---------------------------------------------------------------------------
extern "C" {

cola_status cola_zip_create(const int32_t *ptr, int64_t nptr, const int32_t *idx, int64_t nnz,
                            const int32_t *zip, const float *val, cola_zip *out) {
  if (!out) return fail(COLA_ERR_INVALID, "zip_create: out is NULL");
  *out = nullptr;
  if (nptr < 0 || nnz < 0 || (nptr > 0 && !ptr) || (nnz > 0 && !idx) || (nptr > 1 && !zip))
    return fail(COLA_ERR_INVALID, "zip_create: NULL array or negative size");
  const int64_t nz = nptr > 0 ? nptr - 1 : 0;
  if (nptr > 0) {
    if (ptr[0] != 0 || ptr[nptr - 1] != nnz) return fail(COLA_ERR_INVALID, "zip_create: ptr must span [0, nnz]");
    for (int64_t i = 0; i < nz; ++i)
      if (ptr[i + 1] < ptr[i]) return fail(COLA_ERR_INVALID, "zip_create: ptr decreasing");
    for (int64_t i = 1; i < nz; ++i)
      if (zip[i] < zip[i - 1]) return fail(COLA_ERR_INVALID, "zip_create: zip ids must be sorted");
  }
  auto *z = new (std::nothrow) cola_zip_s;
  if (!z) return fail(COLA_ERR_NOMEM, "zip_create: out of host memory");
  static std::atomic<uint64_t> next_serial{1};
  z->serial = next_serial.fetch_add(1);
  z->nptr = nptr;
  z->nnz = nnz;
  std::vector<int32_t> vptr(ptr, ptr + nptr), vidx(idx, idx + nnz), vzip(zip, zip + nz);
  std::vector<float> vval(nnz, 1.0f);
  if (val) std::copy(val, val + nnz, vval.begin());
  // Packed-row position of every (virtual) entry of zip, and the unique ids.
  std::vector<int32_t> pack(nz), uniq;
  for (int64_t i = 0; i < nz; ++i) {
    if (i == 0 || vzip[i] != vzip[i - 1]) uniq.push_back(vzip[i]);
    pack[i] = (int32_t)uniq.size() - 1;
  }
  z->n_unique = (int64_t)uniq.size();
  cola_status st;
  if ((st = upload(vptr, &z->ptr)) || (st = upload(vidx, &z->idx)) || (st = upload(vzip, &z->zip)) ||
      (st = upload(vval, &z->val)) || (st = upload(pack, &z->pack_idx)) || (st = upload(uniq, &z->uniq))) {
    free_zip(z);
    return st;
  }
  *out = z;
  return COLA_OK;
}

cola_status cola_zip_device_arrays(cola_zip z, int32_t **ptr, int32_t **idx, int32_t **zip,
                                   float **val, int32_t **pack_idx, int64_t *nptr, int64_t *nnz,
                                   int64_t *n_unique) {
  if (!z) return fail(COLA_ERR_INVALID, "zip_device_arrays: NULL handle");
  if (ptr) *ptr = z->ptr;
  if (idx) *idx = z->idx;
  if (zip) *zip = z->zip;
  if (val) *val = z->val;
  if (pack_idx) *pack_idx = z->pack_idx;
  if (nptr) *nptr = z->nptr;
  if (nnz) *nnz = z->nnz;
  if (n_unique) *n_unique = z->n_unique;
  return COLA_OK;
}

cola_status cola_zip_destroy(cola_zip z) {
  free_zip(z);
  return COLA_OK;
}

cola_status cola_pack_remote(const void *B_global, int64_t ld_B, const float *scales_global,
                             const int32_t *remote_cols, int64_t num_cols, int dim,
                             cola_dtype dtype, void *packed, int64_t ld_packed,
                             float *packed_scales, void *stream) {
  const size_t eb = dtype_bytes(dtype);
  if (!eb) return fail(COLA_ERR_UNSUPPORTED, "pack_remote: unknown dtype");
  if (num_cols < 0 || dim < 1 || ld_B < dim || ld_packed < dim)
    return fail(COLA_ERR_INVALID, "pack_remote: bad sizes");
  if (num_cols == 0) return COLA_OK;
  if (!B_global || !remote_cols || !packed) return fail(COLA_ERR_INVALID, "pack_remote: NULL pointer");
  if (dtype == COLA_FP8_E4M3 && (!scales_global || !packed_scales))
    return fail(COLA_ERR_INVALID, "pack_remote: FP8 needs row scales");
  const int64_t row_bytes = (int64_t)dim * eb;
  const int64_t ldb = ld_B * eb, ldo = ld_packed * eb;
  hipStream_t s = (hipStream_t)stream;
  const bool v16 = row_bytes % 16 == 0 && ldb % 16 == 0 && ldo % 16 == 0 && aligned(B_global, 16) &&
                   aligned(packed, 16);
  const float *sc = dtype == COLA_FP8_E4M3 ? scales_global : nullptr;
  float *osc = dtype == COLA_FP8_E4M3 ? packed_scales : nullptr;
  const int64_t max_blocks = (int64_t)device_cu_count() * 16;
  if (v16) {
    const int64_t vecs = row_bytes / 16;
    int lpr = 1;
    while (lpr < vecs && lpr < 64) lpr <<= 1;
    const int64_t blocks = std::min<int64_t>(max_blocks, (num_cols * lpr + 255) / 256);
    hipLaunchKernelGGL((pack_rows_kernel<uint4>), dim3((unsigned)blocks), dim3(256), 0, s,
                       (const char *)B_global, ldb, remote_cols, num_cols, vecs, (char *)packed, ldo,
                       sc, osc, lpr);
  } else {
    const int64_t vecs = row_bytes / (int64_t)eb;
    int lpr = 1;
    while (lpr < vecs && lpr < 64) lpr <<= 1;
    const int64_t blocks = std::min<int64_t>(max_blocks, (num_cols * lpr + 255) / 256);
    if (eb == 4)
      hipLaunchKernelGGL((pack_rows_kernel<uint32_t>), dim3((unsigned)blocks), dim3(256), 0, s,
                         (const char *)B_global, ldb, remote_cols, num_cols, vecs, (char *)packed,
                         ldo, sc, osc, lpr);
    else if (eb == 2)
      hipLaunchKernelGGL((pack_rows_kernel<uint16_t>), dim3((unsigned)blocks), dim3(256), 0, s,
                         (const char *)B_global, ldb, remote_cols, num_cols, vecs, (char *)packed,
                         ldo, sc, osc, lpr);
    else
      hipLaunchKernelGGL((pack_rows_kernel<uint8_t>), dim3((unsigned)blocks), dim3(256), 0, s,
                         (const char *)B_global, ldb, remote_cols, num_cols, vecs, (char *)packed,
                         ldo, sc, osc, lpr);
  }
  return hip_check(hipGetLastError(), "pack_remote launch");
}

cola_status cola_ppl0_batch_spmm(float *output, const float *input, const float *packed,
                                 const int32_t *csr_rowPtr_l, const int32_t *csr_colIdx_l,
                                 const int32_t *csr_rowIdx_zip_l, const float *csr_val_l,
                                 const int32_t *csc_colPtr_r, const int32_t *csc_rowIdx_r,
                                 const int32_t *csc_packIdx_r, const float *csc_val_r,
                                 int local_row_lb, int remote_col_lb, int local_row_num,
                                 int remote_col_num, int dim, int nodesPerPE, int peid,
                                 void *stream) {
  if (local_row_num < 0 || remote_col_num < 0 || local_row_lb < 0 || remote_col_lb < 0 ||
      nodesPerPE < 0)
    return fail(COLA_ERR_INVALID, "ppl0_batch_spmm: negative range");
  if (local_row_num + (int64_t)remote_col_num == 0) return COLA_OK;  // cuh:232
  if (!output) return fail(COLA_ERR_INVALID, "ppl0_batch_spmm: output is NULL");
  if (local_row_num > 0 && (!input || !csr_rowPtr_l || !csr_colIdx_l || !csr_rowIdx_zip_l || !csr_val_l))
    return fail(COLA_ERR_INVALID, "ppl0_batch_spmm: NULL local array");
  if (remote_col_num > 0 && (!packed || !csc_colPtr_r || !csc_rowIdx_r || !csc_packIdx_r || !csc_val_r))
    return fail(COLA_ERR_INVALID, "ppl0_batch_spmm: NULL remote array");
  const bool vec_ok = aligned(output, 16) && (!input || aligned(input, 16)) &&
                      (!packed || aligned(packed, 16));
  Shape sh;
  if (!cola_dev::pick_shape(COLA_F32, dim, vec_ok, &sh))
    return fail(COLA_ERR_UNSUPPORTED, "ppl0_batch_spmm: dim " + std::to_string(dim) + " not supported");
  hipStream_t s = (hipStream_t)stream;
#define COLA_PPL0_CASE(E, G_, NV_)                                                             \
  if (sh.epl == E && sh.g == G_ && sh.nv == NV_) {                                             \
    launch_ppl0<E, G_, NV_>(output, input, packed, csr_rowPtr_l, csr_colIdx_l, csr_rowIdx_zip_l, \
                            csr_val_l, csc_colPtr_r, csc_rowIdx_r, csc_packIdx_r, csc_val_r,    \
                            local_row_lb, remote_col_lb, local_row_num, remote_col_num, dim,    \
                            peid, s);                                                           \
    return hip_check(hipGetLastError(), "ppl0_batch_spmm launch");                             \
  }
  COLA_PPL0_CASE(4, 2, 1)
  COLA_PPL0_CASE(4, 4, 1)
  COLA_PPL0_CASE(4, 8, 1)
  COLA_PPL0_CASE(4, 16, 1)
  COLA_PPL0_CASE(4, 32, 1)
  COLA_PPL0_CASE(4, 64, 1)
  COLA_PPL0_CASE(4, 64, 2)
  COLA_PPL0_CASE(4, 64, 4)
  COLA_PPL0_CASE(1, 64, 1)
  COLA_PPL0_CASE(1, 64, 2)
  COLA_PPL0_CASE(1, 64, 4)
  COLA_PPL0_CASE(1, 64, 8)
#undef COLA_PPL0_CASE
  return fail(COLA_ERR_UNSUPPORTED, "ppl0_batch_spmm: no kernel for this dim");
}

cola_status cola_dist_spmm_ppl_batch(float *output, const float *B_global, float *packed,
                                     cola_zip csr_zip_l, cola_zip csc_zip_r, int dim,
                                     int nodesPerPE, int peid, void *stream) {
  if (!csr_zip_l || !csc_zip_r) return fail(COLA_ERR_INVALID, "dist_spmm_ppl_batch: NULL handle");
  if (!B_global || dim < 1 || nodesPerPE < 1 || peid < 0)
    return fail(COLA_ERR_INVALID, "dist_spmm_ppl_batch: bad arguments");
  const int64_t L = csr_zip_l->nptr > 0 ? csr_zip_l->nptr - 1 : 0;
  const int64_t R = csc_zip_r->nptr > 0 ? csc_zip_r->nptr - 1 : 0;
  if (L > INT32_MAX || R > INT32_MAX) return fail(COLA_ERR_INVALID, "dist_spmm_ppl_batch: too many rows");
  // The same sums as ppl0_batch_SpMM over the virtual rows / columns, run
  // through a plan rebuilt from the two zips on the first call (cached in the
  // local zip): no atomics, deterministic order.
  // This call's own reference keeps the plan alive until its launches are
  // enqueued, whatever other threads do to the zip's cache meanwhile.
  std::shared_ptr<cola_plan_s> plan;
  cola_status st = zip_pair_plan(csr_zip_l, csc_zip_r, nodesPerPE, peid, dim, &plan);
  if (st) return st;
  if (plan->U_r > 0 && !packed) return fail(COLA_ERR_INVALID, "dist_spmm_ppl_batch: packed is NULL");
  st = cola_pack_remote(B_global, dim, nullptr, plan->remote_cols, plan->U_r, dim, COLA_F32, packed, dim, nullptr,
                        stream);
  if (st) return st;
  const float *input = B_global + (int64_t)peid * nodesPerPE * dim;
  return cola_spmm_partition(plan.get(), output, dim, input, dim, nullptr, packed, dim, nullptr, dim, COLA_F32,
                             COLA_ACCUMULATE, stream);
}

cola_status cola_zip_invalidate(cola_zip z) {
  if (!z) return fail(COLA_ERR_INVALID, "zip_invalidate: NULL handle");
  std::lock_guard<std::mutex> lock(g_zip_plan_mu);
  z->fast_plan.reset();  // freed once no call in flight holds it
  z->fast_peer = 0;
  return COLA_OK;
}

cola_status cola_plan_create(cola_partition part, const cola_plan_opts *opts, cola_plan *out) {
  if (!part || !out) return fail(COLA_ERR_INVALID, "plan_create: NULL argument");
  *out = nullptr;
  cola_plan_opts o{};
  if (opts) o = *opts;
  if (o.hub_threshold < 0 || o.chunk_nnz < 0 || o.kahan_degree < 0 || o.max_dim < 0 || o.short_degree < 0 ||
      o.col_panels < 0 || o.col_panels > 64 || o.row_order < 0 || o.row_order > 1 || o.hot_cols < 0)
    return fail(COLA_ERR_INVALID, "plan_create: negative option");
  return plan_from_partition(part->p, o, out);
}

}  // extern "C"

static cola_status plan_from_partition(const cola::Partition &part, const cola_plan_opts &o, cola_plan_s **out,
                                       bool direct_remote) {
  cola::HostPlan hp;
  cola_status st = cola::build_plan(part, o, &hp);
  if (st) return st;
  if (direct_remote)
    for (int32_t &c : hp.cols)
      if (c < 0) c = ~hp.remote_cols[~c];
  auto *p = new (std::nothrow) cola_plan_s;
  if (!p) return fail(COLA_ERR_NOMEM, "plan_create: out of host memory");
  p->num_rows = hp.num_rows;
  p->nodesPerPE = hp.nodesPerPE;
  p->U_r = hp.U_r;
  p->remote_rows = direct_remote ? (hp.remote_cols.empty() ? 0 : (int64_t)hp.remote_cols.back() + 1) : hp.U_r;
  p->U_l = hp.U_l;
  p->max_dim = hp.max_dim;
  p->pstride = (hp.max_dim + 15) / 16 * 16;
  p->nnz = hp.nnz;
  // Default schedule by size: large partitions gather from far more B rows
  // than the L2s hold and want 256-byte column tiles, 4 rows per wave at 2
  // blocks per CU and each XCD on its own item range; small ones want every
  // column in one pass and many waves.
  p->tune = hp.nnz >= (int64_t(1) << 22) ? cola_plan_tuning{-2, 0, 4, 2, 1, 0, 8, 0, 4, 1}
                                         : cola_plan_tuning{0, 0, 1, 8, 1, 0, 8, 0, 0, 0};
  p->num_empty = (int64_t)hp.empty_rows.size();
  p->num_hubs = (int64_t)hp.hub_rows.size();
  p->num_chunks = hp.hub_chunk_ptr.back();
  auto to_int4 = [](const std::vector<cola::WorkItem> &v) {
    std::vector<int4> r(v.size());
    for (size_t i = 0; i < v.size(); ++i) r[i] = make_int4(v[i].out, v[i].begin, v[i].end, v[i].flags);
    return r;
  };
  std::vector<int4> is = to_int4(hp.items_short), il = to_int4(hp.items_long);
  std::vector<int4> isr = to_int4(hp.items_short_r), ilr = to_int4(hp.items_long_r);
  p->shortb.n = (int64_t)is.size();
  p->longb.n = (int64_t)il.size();
  p->shortb_r.n = (int64_t)isr.size();
  p->longb_r.n = (int64_t)ilr.size();
  std::vector<int4> is2 = to_int4(hp.items_short2), il2 = to_int4(hp.items_long2);
  p->cshortb.n = (int64_t)is2.size();
  p->clongb.n = (int64_t)il2.size();
  p->hot_nnz = hp.hot_nnz;
  if ((st = upload(is, &p->shortb.items)) || (st = upload(hp.xcd_short, &p->shortb.xcd)) ||
      (st = upload(il, &p->longb.items)) || (st = upload(hp.xcd_long, &p->longb.xcd)) ||
      (st = upload(isr, &p->shortb_r.items)) || (st = upload(hp.xcd_short_r, &p->shortb_r.xcd)) ||
      (st = upload(ilr, &p->longb_r.items)) || (st = upload(hp.xcd_long_r, &p->longb_r.xcd)) ||
      (st = upload(is2, &p->cshortb.items)) || (st = upload(hp.xcd_short2, &p->cshortb.xcd)) ||
      (st = upload(il2, &p->clongb.items)) || (st = upload(hp.xcd_long2, &p->clongb.xcd)) ||
      (st = upload(hp.cols, &p->cols)) || (st = upload(hp.vals, &p->vals)) ||
      (st = upload(hp.empty_rows, &p->empty_rows)) || (st = upload(hp.hub_rows, &p->hub_rows)) ||
      (st = upload(hp.hub_chunk_ptr, &p->hub_ptr)) || (st = upload(hp.remote_cols, &p->remote_cols))) {
    free_plan(p);
    return st;
  }
  {
    std::vector<int4> it = to_int4(hp.items_tiny), isn = to_int4(hp.items_short_nt);
    p->tinyb.n = (int64_t)it.size();
    p->shortb_nt.n = (int64_t)isn.size();
    if ((st = upload(it, &p->tinyb.items)) || (st = upload(hp.xcd_tiny, &p->tinyb.xcd)) ||
        (st = upload(isn, &p->shortb_nt.items)) || (st = upload(hp.xcd_short_nt, &p->shortb_nt.xcd)) ||
        (st = upload(hp.tiny_rows, &p->tiny_rows))) {
      free_plan(p);
      return st;
    }
    p->device_bytes += (int64_t)((it.size() + isn.size()) * sizeof(int4) + 18 * 4 + hp.tiny_rows.size() * 4);
  }
  auto cost_prefix = [](const std::vector<int4> &v) {
    std::vector<int64_t> c(v.size() + 1, 0);
    for (size_t i = 0; i < v.size(); ++i) c[i + 1] = c[i] + (v[i].z - v[i].y) + 8;
    return c;
  };
  p->shortb.cost = cost_prefix(is);
  p->longb.cost = cost_prefix(il);
  p->tinyb.cost = cost_prefix(to_int4(hp.items_tiny));
  p->shortb_nt.cost = cost_prefix(to_int4(hp.items_short_nt));
  {
    std::vector<int4> ie(hp.empty_rows.size());
    for (size_t i = 0; i < ie.size(); ++i) ie[i] = make_int4(hp.empty_rows[i], 0, 0, 0);
    p->emptyb.cost = cost_prefix(ie);
    p->emptyb.n = (int64_t)ie.size();
    if ((st = upload(ie, &p->emptyb.items))) {
      free_plan(p);
      return st;
    }
  }
  if (hp.panels > 1) {
    std::vector<int4> pi = to_int4(hp.panel_items);
    p->panels = hp.panels;
    p->panel_begin = hp.panel_begin;
    p->panel_chunks = hp.panel_chunks;
    if ((st = upload(pi, &p->panel_items)) || (st = upload(hp.panel_xcd, &p->panel_xcd)) ||
        (st = upload(hp.panel_hub_ptr, &p->panel_hub_ptr))) {
      free_plan(p);
      return st;
    }
    p->device_bytes += (int64_t)(pi.size() * sizeof(int4) + hp.panel_xcd.size() * 4 + hp.panel_hub_ptr.size() * 4);
  }
  const int64_t chunks_alloc = std::max<int64_t>(p->num_chunks, p->panel_chunks);
  if (chunks_alloc > 0) {
    st = hip_check(hipMalloc((void **)&p->partials, (size_t)chunks_alloc * p->pstride * sizeof(float)),
                   "hipMalloc partials");
    if (st) {
      free_plan(p);
      return st;
    }
  }
  p->device_bytes += (int64_t)((is.size() + il.size() + isr.size() + ilr.size() + is2.size() + il2.size()) *
                                  sizeof(int4) + 54 * 4 + hp.cols.size() * 4 +
                              hp.vals.size() * 4 + hp.empty_rows.size() * 4 + hp.hub_rows.size() * 4 +
                              hp.hub_chunk_ptr.size() * 4 + hp.remote_cols.size() * 4 +
                              (size_t)chunks_alloc * p->pstride * 4);
  *out = p;
  return COLA_OK;
}

template <typename T>
static cola_status download(const T *d, int64_t n, std::vector<T> *h) {
  h->resize((size_t)std::max<int64_t>(n, 0));
  if (n <= 0) return COLA_OK;
  return hip_check(hipMemcpy(h->data(), d, (size_t)n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy zip arrays");
}

// Plan behind a (local zip, remote zip) pair, built on first use and cached in
// the local zip while the pair, nodesPerPE, peid and a dim within the plan's
// workspace stay the same.  The partition is recovered from the zips: the
// virtual rows (columns) that nnz_split cut from one row (column) are
// consecutive with the same id, so they are merged back; entries keep their
// order.  Items in decreasing-degree order; the plan's default schedule.
static cola_status zip_pair_plan(cola_zip_s *l, const cola_zip_s *r, int P, int peid, int dim,
                                 std::shared_ptr<cola_plan_s> *out) {
  std::lock_guard<std::mutex> lock(g_zip_plan_mu);
  if (l->fast_plan && l->fast_peer == r->serial && l->fast_P == P && l->fast_peid == peid &&
      dim <= l->fast_plan->max_dim) {
    *out = l->fast_plan;
    return COLA_OK;
  }
  l->fast_plan.reset();
  std::vector<int32_t> lptr, lidx, lzip, rptr, ridx, rzip;
  std::vector<float> lval, rval;
  const int64_t Lv = l->nptr > 0 ? l->nptr - 1 : 0, Rv = r->nptr > 0 ? r->nptr - 1 : 0;
  cola_status st;
  if ((st = download(l->ptr, l->nptr, &lptr)) || (st = download(l->idx, l->nnz, &lidx)) ||
      (st = download(l->zip, Lv, &lzip)) || (st = download(l->val, l->nnz, &lval)) ||
      (st = download(r->ptr, r->nptr, &rptr)) || (st = download(r->idx, r->nnz, &ridx)) ||
      (st = download(r->zip, Rv, &rzip)) || (st = download(r->val, r->nnz, &rval)))
    return st;
  cola::Partition part;
  part.nodesPerPE = P;
  part.rank = peid;
  part.num_rows = P;
  auto merge = [](const std::vector<int32_t> &ptr, const std::vector<int32_t> &zip, int64_t nv,
                  std::vector<int32_t> *ids, std::vector<int32_t> *mptr) {
    for (int64_t v = 0; v < nv; ++v)
      if (ids->empty() || ids->back() != zip[v]) {
        ids->push_back(zip[v]);
        mptr->push_back(ptr[v]);
      }
    mptr->push_back(nv > 0 ? ptr[nv] : 0);
  };
  merge(lptr, lzip, Lv, &part.csr_rowIdx_l, &part.csr_rowPtr_l);
  merge(rptr, rzip, Rv, &part.csc_colIdx_r, &part.csc_colPtr_r);
  part.csr_colIdx_l = std::move(lidx);
  part.csr_val_l = std::move(lval);
  part.csc_rowIdx_r = std::move(ridx);
  part.csc_val_r = std::move(rval);
  const int64_t max_col = part.csc_colIdx_r.empty() ? 0 : part.csc_colIdx_r.back();
  part.nGPU = (int32_t)std::max<int64_t>(peid + 1, max_col / P + 1);
  part.num_nodes = (int64_t)part.nGPU * P;
  if ((st = cola::validate_partition(part))) return st;
  cola_plan_opts o{};
  o.max_dim = std::max(dim, 256);
  o.row_order = 1;
  cola_plan_s *p = nullptr;
  if ((st = plan_from_partition(part, o, &p))) return st;
  l->fast_plan = std::shared_ptr<cola_plan_s>(p, free_plan);
  l->fast_peer = r->serial;
  l->fast_P = P;
  l->fast_peid = peid;
  *out = l->fast_plan;
  return COLA_OK;
}

extern "C" {

cola_status cola_plan_get_info(cola_plan p, cola_plan_info *info) {
  if (!p || !info) return fail(COLA_ERR_INVALID, "plan_info: NULL argument");
  info->num_rows = p->num_rows;
  info->num_items = p->shortb.n + p->longb.n + p->cshortb.n + p->clongb.n;
  info->num_hub_rows = p->num_hubs;
  info->num_hub_chunks = p->num_chunks;
  info->nnz = p->nnz;
  info->U_r = p->U_r;
  info->U_l = p->U_l;
  info->device_bytes = p->device_bytes;
  info->num_short_items = p->shortb.n;
  info->num_long_items = p->longb.n;
  info->hot_nnz = p->hot_nnz;
  info->num_cold_items = p->cshortb.n + p->clongb.n;
  info->num_tiny_bundles = p->tinyb.n;
  return COLA_OK;
}

cola_status cola_plan_set_tuning(cola_plan p, const cola_plan_tuning *t) {
  if (!p || !t) return fail(COLA_ERR_INVALID, "plan_set_tuning: NULL argument");
  if (t->col_tile < -8 || t->rpw_short < 0 || t->rpw_short > 64 || t->rpw_long < 0 || t->rpw_long > 64 ||
      t->blocks_per_cu < 0 || t->blocks_per_cu > 64 || t->blocks_per_cu_short < 0 || t->blocks_per_cu_short > 64 ||
      (t->xcd_remap != 0 && t->xcd_remap != 1) ||
      (t->split_remote != 0 && t->split_remote != 1) || (t->unroll != 0 && t->unroll != 8 && t->unroll != 16) ||
      (t->panels != 0 && t->panels != 1) || (t->tiny != 0 && t->tiny != 1))
    return fail(COLA_ERR_INVALID, "plan_set_tuning: parameter out of range");
  p->tune = *t;
  return COLA_OK;
}

cola_status cola_plan_get_tuning(cola_plan p, cola_plan_tuning *t) {
  if (!p || !t) return fail(COLA_ERR_INVALID, "plan_get_tuning: NULL argument");
  *t = p->tune;
  return COLA_OK;
}

cola_status cola_plan_remote_cols(cola_plan p, const int32_t **d_cols) {
  if (!p || !d_cols) return fail(COLA_ERR_INVALID, "plan_remote_cols: NULL argument");
  *d_cols = p->remote_cols;
  return COLA_OK;
}

cola_status cola_plan_destroy(cola_plan p) {
  free_plan(p);
  return COLA_OK;
}

cola_status cola_spmm_partition(cola_plan p, float *C, int64_t ld_C, const void *B_local,
                                int64_t ld_B, const float *scales_local, const void *packed,
                                int64_t ld_packed, const float *packed_scales, int dim,
                                cola_dtype dtype, int flags, void *stream) {
  if (!p) return fail(COLA_ERR_INVALID, "spmm_partition: NULL plan");
  SpmmCall c{C, ld_C, B_local, ld_B, scales_local, packed, ld_packed, packed_scales, dim, (int)dtype, flags};
  return run_plan(p, c, p->tune, (hipStream_t)stream);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free form of cola_spmm_partition: the hub-chunk partial sums go
// to `scratch`, a fixed device array every GPU holds for the life of the
// process, instead of the plan's workspace.
// ---------------------------------------------------------------------------
namespace {
constexpr int64_t kScratchBytes = COLA_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) float scratch[kScratchBytes / sizeof(float)];

std::mutex g_scratch_mu;
float *g_scratch_dev[64] = {};  // scratch's device address, per device

// The device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy of the array).  Host code never
// uses the array's name as an address: only this pointer.
cola_status scratch_address(float **out) {
  int dev = -1;
  cola_status st = hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (st) return st;
  if (dev < 0 || dev >= 64) return fail(COLA_ERR_UNSUPPORTED, "spmm_partition_noalloc: device id above 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *ptr = nullptr;
    if ((st = hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")))
      return st;
    if (!ptr) return fail(COLA_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_dev[dev] = static_cast<float *>(ptr);
  }
  *out = g_scratch_dev[dev];
  return COLA_OK;
}
}  // namespace

extern "C" {

cola_status cola_spmm_partition_noalloc(cola_plan p, float *C, int64_t ld_C, const void *B_local, int64_t ld_B,
                                        const float *scales_local, const void *packed, int64_t ld_packed,
                                        const float *packed_scales, int dim, cola_dtype dtype, int flags,
                                        void *stream) {
  if (!p) return fail(COLA_ERR_INVALID, "spmm_partition_noalloc: NULL plan");
  if (dim < 1) return fail(COLA_ERR_INVALID, "spmm_partition_noalloc: bad dim");
  // One row of the workspace per hub chunk (of whichever schedule runs),
  // rounded to 16 floats as the plan's own workspace is.
  const int64_t stride = ((int64_t)dim + 15) / 16 * 16;
  const int64_t chunks = (p->tune.panels && p->panels > 1) ? p->panel_chunks : p->num_chunks;
  const int64_t need = chunks * stride * (int64_t)sizeof(float);
  if (need > kScratchBytes)
    return fail(COLA_ERR_INVALID, "spmm_partition_noalloc: the plan's hub rows need " + std::to_string(need) +
                                      " bytes of workspace at dim " + std::to_string(dim) + "; the fixed array holds " +
                                      std::to_string(kScratchBytes) + " (use cola_spmm_partition)");
  float *ws = nullptr;
  if (chunks > 0) {
    cola_status st = scratch_address(&ws);
    if (st) return st;
  }
  SpmmCall c{C, ld_C, B_local, ld_B, scales_local, packed, ld_packed, packed_scales, dim, (int)dtype, flags};
  // Every partial is written before it is read (chunk items store, the hub
  // reduction then loads), so nothing in the array needs clearing.
  return run_plan(p, c, p->tune, (hipStream_t)stream, ws, stride);
}

cola_status cola_plan_autotune(cola_plan p, float *C, int64_t ld_C, const void *B_local,
                               int64_t ld_B, const float *scales_local, const void *packed,
                               int64_t ld_packed, const float *packed_scales, int dim,
                               cola_dtype dtype, int flags, int iters, void *stream,
                               cola_plan_tuning *best_out, float *best_ms) {
  if (!p) return fail(COLA_ERR_INVALID, "plan_autotune: NULL plan");
  if (iters < 1) iters = 3;
  hipStream_t s = (hipStream_t)stream;
  SpmmCall c{C, ld_C, B_local, ld_B, scales_local, packed, ld_packed, packed_scales, dim, (int)dtype,
             COLA_OVERWRITE | (flags & COLA_KAHAN)};
  hipEvent_t e0, e1;
  cola_status st;
  if ((st = hip_check(hipEventCreate(&e0), "hipEventCreate"))) return st;
  if ((st = hip_check(hipEventCreate(&e1), "hipEventCreate"))) {
    (void)hipEventDestroy(e0);
    return st;
  }
  auto time_of = [&](const cola_plan_tuning &t, float *ms) -> cola_status {
    cola_status r = run_plan(p, c, t, s);  // warm-up (and validation)
    if (r) return r;
    if ((r = hip_check(hipEventRecord(e0, s), "hipEventRecord"))) return r;
    for (int i = 0; i < iters; ++i)
      if ((r = run_plan(p, c, t, s))) return r;
    if ((r = hip_check(hipEventRecord(e1, s), "hipEventRecord"))) return r;
    if ((r = hip_check(hipEventSynchronize(e1), "hipEventSynchronize"))) return r;
    float t_ms = 0.f;
    if ((r = hip_check(hipEventElapsedTime(&t_ms, e0, e1), "hipEventElapsedTime"))) return r;
    *ms = t_ms / iters;
    return COLA_OK;
  };
  const char *env = getenv("COLA_AUTOTUNE_VERBOSE");
  const bool verbose = env && env[0] == '1';
  // Every distinct configuration timed so far, for the final re-timing.
  std::vector<std::pair<float, cola_plan_tuning>> seen;
  auto same = [](const cola_plan_tuning &x, const cola_plan_tuning &y) {
    return std::memcmp(&x, &y, sizeof(x)) == 0;
  };
  cola_plan_tuning best = p->tune;
  float bms = 0.f;
  auto consider = [&](const cola_plan_tuning &t, const char *tag) -> cola_status {
    for (const auto &e : seen)
      if (same(e.second, t)) return COLA_OK;
    float ms = 0.f;
    cola_status r = time_of(t, &ms);
    if (r) return r;
    seen.emplace_back(ms, t);
    if (verbose)
      fprintf(stderr,
              "[autotune %s] tile=%d rpw_short=%d rpw_long=%d blocks_per_cu=%d/%d xcd=%d split=%d unroll=%d "
              "panels=%d tiny=%d: %.4f ms\n",
              tag, t.col_tile, t.rpw_short, t.rpw_long, t.blocks_per_cu, t.blocks_per_cu_short, t.xcd_remap,
              t.split_remote, t.unroll, t.panels, t.tiny, ms);
    if (seen.size() == 1 || ms < bms * 0.99f) {
      bms = ms;
      best = t;
    }
    return COLA_OK;
  };
  st = consider(best, "start");
  const int ev = vec_elems(dtype);
  // Coupled knobs first: a grid over column tile x grid size x long-row
  // rows-per-wave x XCD remap x loads in flight.  (Narrow tiles pair with
  // large grids: the 16-bit forms at dim 256 run fastest at 64-column tiles
  // and 16 blocks per CU, which coordinate descent from a full-row winner
  // does not reach one knob at a time.)
  for (int tile : {0, dim / 2, dim / 4}) {
    if (tile != 0 && (tile < ev || tile % ev != 0)) continue;
    for (int bpc : {1, 2, 4, 8, 16})
      for (int rl : {1, 2, 4, 8})
        for (int x : {0, 1})
          for (int u : {8, 16}) {
            if (st) break;
            cola_plan_tuning t = best;
            t.col_tile = tile;
            t.blocks_per_cu = bpc;
            t.rpw_long = rl;
            t.xcd_remap = x;
            t.unroll = u;
            st = consider(t, "grid");
          }
  }
  // Then coordinate descent over one knob at a time.
  for (int pass = 0; pass < 2 && !st; ++pass) {
    std::vector<cola_plan_tuning> cands;
    for (int tile : {0, dim / 2, dim / 4, dim / 8}) {
      if (tile != 0 && (tile < ev || tile % ev != 0)) continue;
      cola_plan_tuning t = best;
      t.col_tile = tile;
      cands.push_back(t);
    }
    for (int r : {1, 2, 4, 8, 16, 32, 0}) {
      cola_plan_tuning t = best;
      t.rpw_short = r;
      cands.push_back(t);
    }
    for (int r : {1, 2, 4, 8}) {
      cola_plan_tuning t = best;
      t.rpw_long = r;
      cands.push_back(t);
    }
    for (int b : {1, 2, 3, 4, 8, 16}) {
      cola_plan_tuning t = best;
      t.blocks_per_cu = b;
      cands.push_back(t);
    }
    for (int b : {0, 2, 4, 8, 16}) {
      cola_plan_tuning t = best;
      t.blocks_per_cu_short = b;
      cands.push_back(t);
    }
    for (int x : {0, 1}) {
      cola_plan_tuning t = best;
      t.xcd_remap = x;
      cands.push_back(t);
    }
    for (int x : {0, 1}) {
      cola_plan_tuning t = best;
      t.split_remote = x;
      cands.push_back(t);
    }
    for (int u : {8, 16}) {
      cola_plan_tuning t = best;
      t.unroll = u;
      cands.push_back(t);
    }
    if (p->tinyb.n > 0)
      for (int x : {0, 1}) {
        cola_plan_tuning t = best;
        t.tiny = x;
        cands.push_back(t);
      }
    if (p->panels > 1)
      for (int x : {0, 1}) {
        cola_plan_tuning t = best;
        t.panels = x;
        cands.push_back(t);
      }
    for (const auto &t : cands)
      if ((st = consider(t, "descent"))) break;
  }
  // One timing per configuration is noisy: re-time the fastest few, three
  // interleaved rounds each, and keep the best minimum.
  if (!st) {
    std::stable_sort(seen.begin(), seen.end(),
                     [](const auto &x, const auto &y) { return x.first < y.first; });
    const size_t k = std::min<size_t>(seen.size(), 8);
    std::vector<float> score(k, 3.4e38f);
    for (int round = 0; round < 3 && !st; ++round)
      for (size_t i = 0; i < k && !st; ++i) {
        float ms = 0.f;
        if ((st = time_of(seen[i].second, &ms))) break;
        score[i] = std::min(score[i], ms);
      }
    if (!st) {
      size_t w = 0;
      for (size_t i = 1; i < k; ++i)
        if (score[i] < score[w]) w = i;
      best = seen[w].second;
      bms = score[w];
      if (verbose)
        for (size_t i = 0; i < k; ++i)
          fprintf(stderr,
                  "[autotune final] #%zu tile=%d rpw_short=%d rpw_long=%d blocks_per_cu=%d/%d xcd=%d unroll=%d: "
                  "%.4f ms%s\n",
                  i, seen[i].second.col_tile, seen[i].second.rpw_short, seen[i].second.rpw_long,
                  seen[i].second.blocks_per_cu, seen[i].second.blocks_per_cu_short, seen[i].second.xcd_remap,
                  seen[i].second.unroll, score[i],
                  i == w ? " <- best" : "");
    }
  }
  if (!st) st = run_plan(p, c, best, s);
  if (!st) st = hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st) return st;
  p->tune = best;
  if (best_out) *best_out = best;
  if (best_ms) *best_ms = bms;
  return COLA_OK;
}

cola_status cola_quantize_rows(const float *src, int64_t ld_src, int64_t rows, int dim,
                               cola_dtype dtype, void *dst, int64_t ld_dst, float *scales,
                               void *stream) {
  if (!dtype_bytes(dtype)) return fail(COLA_ERR_UNSUPPORTED, "quantize_rows: unknown dtype");
  if (rows < 0 || dim < 1 || ld_src < dim || ld_dst < dim)
    return fail(COLA_ERR_INVALID, "quantize_rows: bad sizes");
  if (rows == 0) return COLA_OK;
  if (!src || !dst || (dtype == COLA_FP8_E4M3 && !scales))
    return fail(COLA_ERR_INVALID, "quantize_rows: NULL pointer");
  const int64_t blocks = (rows * 64 + 255) / 256;
  if (blocks > INT32_MAX) return fail(COLA_ERR_INVALID, "quantize_rows: too many rows");
  hipLaunchKernelGGL(quantize_rows_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     src, ld_src, rows, dim, (int)dtype, dst, ld_dst, scales);
  return hip_check(hipGetLastError(), "quantize_rows launch");
}

cola_status cola_error_norms(const float *C, int64_t ld_C, const double *R, int64_t ld_R,
                             int64_t rows, int dim, double *out, void *stream) {
  if (!out || rows < 0 || dim < 1 || ld_C < dim || ld_R < dim)
    return fail(COLA_ERR_INVALID, "error_norms: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(fill_f64_kernel, dim3(1), dim3(64), 0, s, out, 3);
  if (rows > 0) {
    if (!C || !R) return fail(COLA_ERR_INVALID, "error_norms: NULL matrix");
    const int64_t blocks = (rows * 64 + 255) / 256;
    hipLaunchKernelGGL(error_norms_kernel, dim3((unsigned)blocks), dim3(256), 0, s, C, ld_C, R, ld_R,
                       rows, dim, out);
  }
  return hip_check(hipGetLastError(), "error_norms launch");
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Reference-signature launcher (ppl0_batch_SpMM's parameter list): the zips'
// device arrays and offsets as raw pointers, remote rows read from B_global
// at their global row (input is B_global + peid * nodesPerPE * dim).
// ---------------------------------------------------------------------------
namespace {

// Arguments that determine the work of one call; a plan is cached per key.
struct RefKey {
  const void *arr[8];
  int llb, rlb, lnum, rnum, P, peid;
  bool operator<(const RefKey &o) const {
    for (int i = 0; i < 8; ++i)
      if (arr[i] != o.arr[i]) return arr[i] < o.arr[i];
    const int a[6] = {llb, rlb, lnum, rnum, P, peid}, b[6] = {o.llb, o.rlb, o.lnum, o.rnum, o.P, o.peid};
    return std::lexicographical_compare(a, a + 6, b, b + 6);
  }
};
std::mutex g_ref_mu;
std::map<RefKey, std::shared_ptr<cola_plan_s>> g_ref_plans;

template <typename T>
cola_status download_range(const T *d, int64_t first, int64_t n, std::vector<T> *h) {
  h->assign((size_t)std::max<int64_t>(n, 0), T());
  if (n <= 0) return COLA_OK;
  return hip_check(hipMemcpy(h->data(), d + first, (size_t)n * sizeof(T), hipMemcpyDeviceToHost),
                   "hipMemcpy reference arrays");
}

// Rebuilds the partition that virtual rows [llb, llb+lnum) of the local zip
// CSR and virtual columns [rlb, rlb+rnum) of the remote zip CSC describe.
// Entries of one output row (column) are gathered in the order the virtual
// rows (columns) list them, whatever the order of those.
cola_status ref_partition(const RefKey &k, const int32_t *rowPtr_l, const int32_t *colIdx_l,
                          const int32_t *rowIdx_l, const float *val_l, const int32_t *colPtr_r,
                          const int32_t *rowIdx_r, const int32_t *colIdx_r, const float *val_r,
                          cola::Partition *part) {
  const int P = k.P;
  std::vector<int32_t> lptr, lrow, lcol, rptr, rcol, rrow;
  std::vector<float> lval, rval;
  cola_status st;
  if ((st = download_range(rowPtr_l, k.llb, k.lnum > 0 ? k.lnum + 1 : 0, &lptr)) ||
      (st = download_range(rowIdx_l, k.llb, k.lnum, &lrow)) ||
      (st = download_range(colPtr_r, k.rlb, k.rnum > 0 ? k.rnum + 1 : 0, &rptr)) ||
      (st = download_range(colIdx_r, k.rlb, k.rnum, &rcol)))
    return st;
  for (int v = 0; v < k.lnum; ++v)
    if (lptr[v + 1] < lptr[v]) return fail(COLA_ERR_INVALID, "ppl0_batch_spmm_global: csr_rowPtr_l decreasing");
  for (int v = 0; v < k.rnum; ++v)
    if (rptr[v + 1] < rptr[v]) return fail(COLA_ERR_INVALID, "ppl0_batch_spmm_global: csc_colPtr_r decreasing");
  const int64_t l0 = k.lnum > 0 ? lptr[0] : 0, l1 = k.lnum > 0 ? lptr[k.lnum] : 0;
  const int64_t r0 = k.rnum > 0 ? rptr[0] : 0, r1 = k.rnum > 0 ? rptr[k.rnum] : 0;
  if ((st = download_range(colIdx_l, l0, l1 - l0, &lcol)) || (st = download_range(val_l, l0, l1 - l0, &lval)) ||
      (st = download_range(rowIdx_r, r0, r1 - r0, &rrow)) || (st = download_range(val_r, r0, r1 - r0, &rval)))
    return st;
  for (int32_t r : lrow)
    if (r < 0 || r >= P) return fail(COLA_ERR_INVALID, "ppl0_batch_spmm_global: csr_rowIdx_zip_l out of [0,P)");
  // Local part: output rows in increasing order, each with its entries in
  // virtual-row order (a stable counting sort of the virtual rows).
  std::vector<int64_t> rcnt((size_t)P + 1, 0);
  for (int v = 0; v < k.lnum; ++v) rcnt[lrow[v] + 1] += lptr[v + 1] - lptr[v];
  for (int i = 0; i < P; ++i) rcnt[i + 1] += rcnt[i];
  part->csr_colIdx_l.resize(lcol.size());
  part->csr_val_l.resize(lcol.size());
  {
    std::vector<int64_t> cur(rcnt.begin(), rcnt.end() - 1);
    for (int v = 0; v < k.lnum; ++v)
      for (int64_t q = lptr[v]; q < lptr[v + 1]; ++q) {
        const int64_t w = cur[lrow[v]]++;
        part->csr_colIdx_l[w] = lcol[q - l0];
        part->csr_val_l[w] = lval[q - l0];
      }
  }
  part->csr_rowPtr_l.assign(1, 0);
  for (int i = 0; i < P; ++i)
    if (rcnt[i + 1] > rcnt[i]) {
      part->csr_rowIdx_l.push_back(i);
      part->csr_rowPtr_l.push_back((int32_t)rcnt[i + 1]);
    }
  // Remote part: distinct global columns in increasing order, each with its
  // entries in virtual-column order.
  std::vector<int32_t> order(k.rnum);
  for (int v = 0; v < k.rnum; ++v) order[v] = v;
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return rcol[a] < rcol[b]; });
  part->csc_colPtr_r.assign(1, 0);
  for (int oi = 0; oi < k.rnum; ++oi) {
    const int v = order[oi];
    if (rptr[v + 1] == rptr[v]) continue;
    if (part->csc_colIdx_r.empty() || part->csc_colIdx_r.back() != rcol[v]) {
      part->csc_colIdx_r.push_back(rcol[v]);
      part->csc_colPtr_r.push_back(part->csc_colPtr_r.back());
    }
    for (int64_t q = rptr[v]; q < rptr[v + 1]; ++q) {
      part->csc_rowIdx_r.push_back(rrow[q - r0]);
      part->csc_val_r.push_back(rval[q - r0]);
    }
    part->csc_colPtr_r.back() = (int32_t)part->csc_rowIdx_r.size();
  }
  part->nodesPerPE = P;
  part->rank = k.peid;
  part->num_rows = P;
  const int64_t max_col = part->csc_colIdx_r.empty() ? 0 : std::max<int64_t>(0, part->csc_colIdx_r.back());
  part->nGPU = (int32_t)std::max<int64_t>(k.peid + 1, max_col / P + 1);
  part->num_nodes = (int64_t)part->nGPU * P;
  return cola::validate_partition(*part);
}

}  // namespace

static void drop_ref_plans_using(const void *a, const void *b, const void *c, const void *d) {
  std::lock_guard<std::mutex> lock(g_ref_mu);
  for (auto it = g_ref_plans.begin(); it != g_ref_plans.end();) {
    bool uses = false;
    for (const void *x : it->first.arr) uses = uses || (x && (x == a || x == b || x == c || x == d));
    it = uses ? g_ref_plans.erase(it) : std::next(it);
  }
}

extern "C" {

cola_status cola_ppl0_batch_spmm_global(float *output, const float *input, const int32_t *csr_rowPtr_l,
                                        const int32_t *csr_colIdx_l, const int32_t *csr_rowIdx_zip_l,
                                        const float *csr_val_l, const int32_t *csc_colPtr_r,
                                        const int32_t *csc_rowIdx_r, const int32_t *csc_colIdx_zip_r,
                                        const float *csc_val_r, int local_row_lb, int remote_col_lb,
                                        int local_row_num, int remote_col_num, int dim, int nodesPerPE, int peid,
                                        int literal, void *stream) {
  if (local_row_num < 0 || remote_col_num < 0 || local_row_lb < 0 || remote_col_lb < 0 || nodesPerPE < 1 ||
      peid < 0 || dim < 1)
    return fail(COLA_ERR_INVALID, "ppl0_batch_spmm_global: negative range or bad sizes");
  if (local_row_num + (int64_t)remote_col_num == 0) return COLA_OK;  // cuh:232
  if (!output || !input) return fail(COLA_ERR_INVALID, "ppl0_batch_spmm_global: output or input is NULL");
  if (local_row_num > 0 && (!csr_rowPtr_l || !csr_colIdx_l || !csr_rowIdx_zip_l || !csr_val_l))
    return fail(COLA_ERR_INVALID, "ppl0_batch_spmm_global: NULL local array");
  if (remote_col_num > 0 && (!csc_colPtr_r || !csc_rowIdx_r || !csc_colIdx_zip_r || !csc_val_r))
    return fail(COLA_ERR_INVALID, "ppl0_batch_spmm_global: NULL remote array");
  // Remote global row g lives at input + (g - peid*P)*dim: B_global's base.
  const float *B_global = input - (int64_t)peid * nodesPerPE * dim;
  if (literal)
    return cola_ppl0_batch_spmm(output, input, B_global, csr_rowPtr_l, csr_colIdx_l, csr_rowIdx_zip_l, csr_val_l,
                                csc_colPtr_r, csc_rowIdx_r, csc_colIdx_zip_r, csc_val_r, local_row_lb,
                                remote_col_lb, local_row_num, remote_col_num, dim, nodesPerPE, peid, stream);
  RefKey key{{csr_rowPtr_l, csr_colIdx_l, csr_rowIdx_zip_l, csr_val_l, csc_colPtr_r, csc_rowIdx_r,
              csc_colIdx_zip_r, csc_val_r},
             local_row_lb, remote_col_lb, local_row_num, remote_col_num, nodesPerPE, peid};
  std::shared_ptr<cola_plan_s> plan;
  {
    std::lock_guard<std::mutex> lock(g_ref_mu);
    auto it = g_ref_plans.find(key);
    if (it != g_ref_plans.end() && dim <= it->second->max_dim) plan = it->second;
  }
  if (!plan) {
    cola::Partition part;
    cola_status st = ref_partition(key, csr_rowPtr_l, csr_colIdx_l, csr_rowIdx_zip_l, csr_val_l, csc_colPtr_r,
                                   csc_rowIdx_r, csc_colIdx_zip_r, csc_val_r, &part);
    if (st) return st;
    cola_plan_opts o{};
    o.max_dim = std::max(dim, 256);
    o.row_order = 1;
    cola_plan_s *p = nullptr;
    if ((st = plan_from_partition(part, o, &p, true))) return st;
    plan = std::shared_ptr<cola_plan_s>(p, free_plan);
    std::lock_guard<std::mutex> lock(g_ref_mu);
    g_ref_plans[key] = plan;
  }
  return cola_spmm_partition(plan.get(), output, dim, input, dim, nullptr, B_global, dim, nullptr, dim, COLA_F32,
                             COLA_ACCUMULATE, stream);
}

cola_status cola_ppl0_plans_clear(void) {
  std::lock_guard<std::mutex> lock(g_ref_mu);
  g_ref_plans.clear();  // each plan is freed once no call in flight holds it
  return COLA_OK;
}

}  // extern "C"

namespace {
__global__ void warmup_kernel() {}
}  // namespace

extern "C" cola_status cola_warmup(int launches, void *stream) {
  if (launches < 0) return fail(COLA_ERR_INVALID, "warmup: negative count");
  for (int i = 0; i < launches; ++i)
    hipLaunchKernelGGL(warmup_kernel, dim3(1024), dim3(256), 0, (hipStream_t)stream);
  return hip_check(hipGetLastError(), "warmup launch");
}
