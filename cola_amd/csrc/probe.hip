// Measurement kernel: the gather ceiling of a plan's entry stream.
//
// gather_ceiling_kernel does only what any kernel computing A_p·B over the
// plan's entries must do, in the simplest form: one group of G lanes per work
// item, each lane holding 16 bytes (4 FP32 columns) of the column tile; the
// group reads its entries (column, value) straight from memory (all G lanes
// read the same word, which the hardware broadcasts), loads 8 B rows ahead,
// multiply-adds into registers and stores the item's row once.  No staging,
// no cross-lane broadcast, no compensated sums.  Timing it beside the plan
// kernel on the same items and tiles bounds what any schedule of this entry
// stream can gain.  Modes:
//   0: the entries' real rows (local slice or packed buffer);
//   1: each row index masked into the first `table_rows` rows of its table,
//      a table that stays in L2 (same entry stream, same instructions);
//   2: entry k reads row k mod table_rows of the local slice: the same number
//      of row loads, streamed in address order (no gather at all).
// FP32 only; the column tile must be a multiple of 4 columns <= 256.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "device.hpp"
#include "rows.hpp"

using namespace cola_devhost;
using cola::fail;

namespace {

template <int G>
__global__ __launch_bounds__(256) void gather_ceiling_kernel(const int4 *__restrict__ items, int n_items,
                                                             const int32_t *__restrict__ cols,
                                                             const float *__restrict__ vals,
                                                             const float *__restrict__ Bl, int64_t ldB,
                                                             const float *__restrict__ Bp, int64_t ldP,
                                                             float *__restrict__ C, int64_t ldC,
                                                             float *__restrict__ partials, int64_t pstride, int mode,
                                                             int64_t table_rows) {
  constexpr int U = 8;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int gl = (int)(tid % G);
  const int64_t n_groups = (int64_t)gridDim.x * blockDim.x / G;
  const int lc = gl * 4;
  for (int64_t it = tid / G; it < n_items; it += n_groups) {
    const int4 item = items[it];
    const int b = item.y, e = item.z;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int k = b; k < e; k += U) {
      float4 x[U];
      float v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int kk = min(k + u, e - 1);
        const int c = cols[kk];
        v[u] = k + u < e ? vals[kk] : 0.f;
        const float *rp;
        if (mode == 2) {
          rp = Bl + (int64_t)(kk % table_rows) * ldB;
        } else if (c >= 0) {
          rp = Bl + (int64_t)(mode == 1 ? c & (table_rows - 1) : c) * ldB;
        } else {
          rp = Bp + (int64_t)(mode == 1 ? (~c) & (table_rows - 1) : ~c) * ldP;
        }
        x[u] = *reinterpret_cast<const float4 *>(rp + lc);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        acc[0] = fmaf(v[u], x[u].x, acc[0]);
        acc[1] = fmaf(v[u], x[u].y, acc[1]);
        acc[2] = fmaf(v[u], x[u].z, acc[2]);
        acc[3] = fmaf(v[u], x[u].w, acc[3]);
      }
    }
    float *dst = item.x >= 0 ? C + (int64_t)item.x * ldC : partials + (int64_t)(-item.x - 1) * pstride;
    *reinterpret_cast<float4 *>(dst + lc) = make_float4(acc[0], acc[1], acc[2], acc[3]);
  }
}

// The bare gather of a row stream: each 16-lane group takes 8 consecutive
// entries per step and loads the first 256 bytes of each entry's row (16
// bytes a lane), all 8 loads in flight, then adds them into registers.  The
// row of entry k starts at base + row_offs[k] * 16.  The waves stride over
// the whole stream, or with xcd_ranges block b works on the (b % 8)-th eighth
// of it, so each XCD walks one contiguous range as the row kernel's XCD item
// ranges do.  VIA_LDS: the loads go
// through LDS (global_load_lds_dwordx4 into an 8 KB area per wave, read back
// after the wave's loads have landed) instead of straight into VGPRs.  Each
// block stores the sum of its lanes' totals to out[blockIdx.x], which keeps
// the loads live.
template <bool VIA_LDS>
__global__ __launch_bounds__(256) void row_stream_kernel(const uint32_t *__restrict__ row_offs, int64_t n,
                                                         uint64_t base, int xcd_ranges, float *__restrict__ out) {
  typedef float f32x4 __attribute__((ext_vector_type(4)));
  __shared__ f32x4 slots[4][8][64];
  __shared__ float wsum[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, gl = lane & 15, grp = lane >> 4;
  const int xcd = blockIdx.x & 7;
  const int64_t lo = xcd_ranges ? n * xcd / 8 : 0, hi = xcd_ranges ? n * (xcd + 1) / 8 : n;
  const int64_t wave = xcd_ranges ? (int64_t)(blockIdx.x >> 3) * 4 + w : (int64_t)blockIdx.x * 4 + w;
  const int64_t nwave = xcd_ranges ? (int64_t)((gridDim.x - xcd + 7) >> 3) * 4 : (int64_t)gridDim.x * 4;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int64_t b = lo + wave * 32; b < hi; b += nwave * 32) {
    uint32_t o[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) o[u] = row_offs[min<int64_t>(b + grp * 8 + u, hi - 1)];
    if constexpr (!VIA_LDS) {
      f32x4 r[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) r[u] = cola_dev::load_global<f32x4>(base + ((uint64_t)o[u] << 4) + gl * 16);
#pragma unroll
      for (int u = 0; u < 8; ++u) acc += r[u];
    } else {
#pragma unroll
      for (int u = 0; u < 8; ++u)
        __builtin_amdgcn_global_load_lds(
            reinterpret_cast<const __attribute__((address_space(1))) void *>(base + ((uint64_t)o[u] << 4) + gl * 16),
            (__attribute__((address_space(3))) void *)&slots[w][u][0], 16, 0, 0);
      __builtin_amdgcn_s_waitcnt(0);  // every wait counter at zero: the LDS writes have landed
#pragma unroll
      for (int u = 0; u < 8; ++u) acc += slots[w][u][lane];
      __builtin_amdgcn_s_waitcnt(0);  // reads done before the next step refills the slots
    }
  }
  float s = acc.x + acc.y + acc.z + acc.w;
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) wsum[w] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

}  // namespace

extern "C" cola_status cola_probe_row_stream(const uint32_t *row_offs, int64_t n, const void *base, int via_lds,
                                             int xcd_ranges, int blocks_per_cu, float *out, int64_t out_len,
                                             void *stream) {
  if (n < 0 || (n > 0 && !row_offs) || !base || !out || !aligned(base, 16))
    return fail(COLA_ERR_INVALID, "probe_row_stream: bad arguments");
  const int64_t blocks = (int64_t)device_cu_count() * (blocks_per_cu > 0 ? blocks_per_cu : 2);
  if (out_len < blocks) return fail(COLA_ERR_INVALID, "probe_row_stream: out holds fewer floats than blocks");
  if (n == 0) return COLA_OK;
  hipStream_t s = (hipStream_t)stream;
  const uint64_t b = (uint64_t)(uintptr_t)base;
  if (via_lds)
    hipLaunchKernelGGL(row_stream_kernel<true>, dim3((unsigned)blocks), dim3(256), 0, s, row_offs, n, b, xcd_ranges, out);
  else
    hipLaunchKernelGGL(row_stream_kernel<false>, dim3((unsigned)blocks), dim3(256), 0, s, row_offs, n, b, xcd_ranges, out);
  return hip_check(hipGetLastError(), "probe_row_stream launch");
}

extern "C" cola_status cola_probe_gather(cola_plan p, const float *B_local, int64_t ld_B, const float *packed,
                                         int64_t ld_packed, float *C, int64_t ld_C, int dim, int col_tile, int mode,
                                         int64_t table_rows, int blocks_per_cu, void *stream) {
  if (!p || !C || dim < 4 || dim % 4 || ld_C < dim || ld_B < dim || (p->U_r > 0 && ld_packed < dim))
    return fail(COLA_ERR_INVALID, "probe_gather: bad arguments");
  if ((p->U_l > 0 && !B_local) || (p->U_r > 0 && !packed)) return fail(COLA_ERR_INVALID, "probe_gather: NULL B");
  if (!aligned(C, 16) || ld_C % 4 || (B_local && (!aligned(B_local, 16) || ld_B % 4)) ||
      (packed && (!aligned(packed, 16) || ld_packed % 4)))
    return fail(COLA_ERR_INVALID, "probe_gather: matrices must allow 16-byte row vectors");
  if (mode < 0 || mode > 2 || (mode > 0 && table_rows < 1) || (mode == 1 && (table_rows & (table_rows - 1))))
    return fail(COLA_ERR_INVALID, "probe_gather: mode 1 needs a power-of-two table_rows, mode 2 table_rows >= 1");
  // B_local holds the partition's num_rows rows (local columns lie below
  // it); the packed buffer holds U_r rows.
  if (mode == 1 && ((p->U_l > 0 && table_rows > p->num_rows) || (p->U_r > 0 && table_rows > p->U_r)))
    return fail(COLA_ERR_INVALID, "probe_gather: table_rows exceeds a table");
  if (mode == 2 && (!B_local || table_rows > p->num_rows))
    return fail(COLA_ERR_INVALID, "probe_gather: mode 2 reads B_local rows below min(table_rows, num_rows)");
  if (dim > p->max_dim && p->num_chunks > 0) return fail(COLA_ERR_INVALID, "probe_gather: dim above max_dim");
  const int tile = col_tile > 0 && col_tile < dim ? col_tile : dim;
  if (tile % 4 || tile > 256) return fail(COLA_ERR_INVALID, "probe_gather: tile must be a multiple of 4, <= 256");
  hipStream_t s = (hipStream_t)stream;
  const int bpc = blocks_per_cu > 0 ? blocks_per_cu : 8;
  for (int c0 = 0; c0 < dim; c0 += tile) {
    const int w = std::min(tile, dim - c0);
    if (w % 4) return fail(COLA_ERR_INVALID, "probe_gather: dim must be a multiple of the tile's vector width");
    const int G = w / 4;
    for (const PlanBin *bin : {&p->longb, &p->shortb}) {
      if (bin->n == 0) continue;
      const int64_t blocks = std::min<int64_t>((bin->n * G + 255) / 256, (int64_t)device_cu_count() * bpc);
#define COLA_PROBE_CASE(GG)                                                                                 \
  case GG:                                                                                                  \
    hipLaunchKernelGGL((gather_ceiling_kernel<GG>), dim3((unsigned)blocks), dim3(256), 0, s, bin->items,     \
                       (int)bin->n, p->cols, p->vals, B_local ? B_local + c0 : nullptr, ld_B,               \
                       packed ? packed + c0 : nullptr, ld_packed, C + c0, ld_C,                             \
                       p->partials ? p->partials + c0 : nullptr, p->pstride, mode, table_rows);             \
    break;
      switch (G) {
        COLA_PROBE_CASE(1)
        COLA_PROBE_CASE(2)
        COLA_PROBE_CASE(4)
        COLA_PROBE_CASE(8)
        COLA_PROBE_CASE(16)
        COLA_PROBE_CASE(32)
        COLA_PROBE_CASE(64)
        default:
          return fail(COLA_ERR_UNSUPPORTED, "probe_gather: tile width must be 4 x a power of two columns");
      }
#undef COLA_PROBE_CASE
    }
  }
  return hip_check(hipGetLastError(), "probe_gather launch");
}
