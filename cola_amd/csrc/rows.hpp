// Row kernel of the plan path, shared by the per-dtype translation units
// (rows_f32.hip, rows_bf16.hip, rows_f16.hip, rows_fp8.hip).
//
// A work item is one output row (or one chunk of a hub row) whose entries
// address either a row of this partition's B slice (entry column >= 0) or a
// row of the packed remote buffer (entry column < 0, packed row ~column).
//
// Lane layout (wave64): the dense dimension of one B row is spread over a
// row group of G lanes, EPL consecutive elements per lane per 16-byte vector
// (EPL = 4 FP32, 8 BF16/FP16, 16 FP8), NV vectors per lane.  The wave's
// S = 64/G groups form `rpw` teams; a team works on one item, its groups
// taking every (S/rpw)-th entry, U loads in flight per group, and the groups
// of a team are summed with cross-lane shuffles at the end.  rpw = 1 is one
// item per wave (long rows); rpw = S is one item per group (short rows).
// Entries are staged K per team lane (K = U/G when U > G, else 1) with
// coalesced loads, turned into (B row address, value) pairs there, and
// broadcast with ds_bpermute.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/cola_spmm.h"

namespace cola_dev {

__device__ __forceinline__ float bf16_to_f32(uint32_t h) { return __uint_as_float(h << 16); }
__device__ __forceinline__ float f16_to_f32(uint32_t h) {
  return __half2float(__ushort_as_half((unsigned short)h));
}
// OCP FP8 E4M3 (e4m3fn): 1 sign, 4 exponent (bias 7), 3 mantissa bits;
// exponent 0 is subnormal (m * 2^-9); 0x7f / 0xff are NaN (never written).
__device__ __forceinline__ float fp8e4m3_to_f32(uint32_t c) {
  const uint32_t e = (c >> 3) & 15u, m = c & 7u;
  const float mag = e == 0 ? (float)m * 0x1p-9f : __uint_as_float(((c & 0x7fu) << 20) + (120u << 23));
  return (c & 0x80u) ? -mag : mag;
}

// Load of type X from a global-memory address held as an integer (the row
// kernel stages row addresses as integers; the address space keeps the load
// a global_load rather than a flat one, which would also count against the
// LDS wait counter the broadcasts use).
template <typename X>
__device__ __forceinline__ X load_global(uint64_t addr) {
  return *reinterpret_cast<const __attribute__((address_space(1))) X *>(addr);
}
// 16-byte forms as clang vector types (HIP's float4 / uint4 structs do not
// copy from an address-space-qualified lvalue).
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 load_global_f4(uint64_t addr) {
  const f32x4_t v = load_global<f32x4_t>(addr);
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint4 load_global_u4(uint64_t addr) {
  const u32x4_t v = load_global<u32x4_t>(addr);
  return make_uint4(v.x, v.y, v.z, v.w);
}
// Pointer forms, for pointers the compiler cannot prove global (those the
// batched kernel reads from its segment table in memory).
template <typename X>
__device__ __forceinline__ X gld(const X *p) {
  return load_global<X>(reinterpret_cast<uint64_t>(p));
}
__device__ __forceinline__ int4 gld(const int4 *p) {
  typedef int32_t i32x4_t __attribute__((ext_vector_type(4)));
  const i32x4_t v = load_global<i32x4_t>(reinterpret_cast<uint64_t>(p));
  return make_int4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float4 gld(const float4 *p) { return load_global_f4(reinterpret_cast<uint64_t>(p)); }
__device__ __forceinline__ void gst(float *p, float v) {
  *reinterpret_cast<__attribute__((address_space(1))) float *>(reinterpret_cast<uint64_t>(p)) = v;
}
__device__ __forceinline__ void gst(float4 *p, const float4 &v) {
  const f32x4_t w = {v.x, v.y, v.z, v.w};
  *reinterpret_cast<__attribute__((address_space(1))) f32x4_t *>(reinterpret_cast<uint64_t>(p)) = w;
}

// Sum of a hub row's chunk partials p[c * pstride], c in [c0, c1), added in
// chunk order (Kahan-compensated when kahan).  32 loads are in flight per
// round (the last round's loads clamped to the range, its additions
// predicated), so the additions are exactly the sequential ones while a hub
// of a few hundred chunks costs about ten memory round trips rather than
// hundreds.
__device__ __forceinline__ float hub_sum(const float *p, int64_t pstride, int c0, int c1, bool kahan) {
  constexpr int kRound = 32;
  float s = 0.f, err = 0.f;
  for (int c = c0; c < c1; c += kRound) {
    float x[kRound];
#pragma unroll
    for (int i = 0; i < kRound; ++i) x[i] = gld(p + (int64_t)min(c + i, c1 - 1) * pstride);  // c < c1
#pragma unroll
    for (int i = 0; i < kRound; ++i) {
      if (c + i < c1) {
        if (kahan) {
          const float y = x[i] - err;
          const float t = s + y;
          // Once the sum is infinite or NaN there is nothing to compensate
          // (and (t - s) - y would turn an infinite sum into NaN).
          err = __builtin_isfinite(t) ? (t - s) - y : 0.f;
          s = t;
        } else {
          s += x[i];
        }
      }
    }
  }
  return s;
}

// Vec<DT, EPL>: element type T, the raw register form of one lane's vector,
// its load (from a pointer, or from a global address: gload) and its
// widening to FP32.
template <int DT, int EPL>
struct Vec;
template <>
struct Vec<COLA_F32, 4> {
  using T = float;
  using Raw = float4;
  static __device__ __forceinline__ Raw load(const T *p) { return *reinterpret_cast<const float4 *>(p); }
  static __device__ __forceinline__ Raw gload(uint64_t a) { return load_global_f4(a); }
  static __device__ __forceinline__ void widen(const Raw &r, float (&x)[4]) {
    x[0] = r.x; x[1] = r.y; x[2] = r.z; x[3] = r.w;
  }
};
template <>
struct Vec<COLA_F32, 1> {
  using T = float;
  using Raw = float;
  static __device__ __forceinline__ Raw load(const T *p) { return *p; }
  static __device__ __forceinline__ Raw gload(uint64_t a) { return load_global<float>(a); }
  static __device__ __forceinline__ void widen(const Raw &r, float (&x)[1]) { x[0] = r; }
};
template <int DT>
struct Vec16 {
  using T = uint16_t;
  using Raw = uint4;
  static __device__ __forceinline__ Raw load(const T *p) { return *reinterpret_cast<const uint4 *>(p); }
  static __device__ __forceinline__ Raw gload(uint64_t a) { return load_global_u4(a); }
  static __device__ __forceinline__ float cvt(uint32_t h) {
    if constexpr (DT == COLA_BF16) return bf16_to_f32(h);
    else return f16_to_f32(h);
  }
  static __device__ __forceinline__ void widen(const Raw &r, float (&x)[8]) {
    const uint32_t w[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      x[2 * i] = cvt(w[i] & 0xffffu);
      x[2 * i + 1] = cvt(w[i] >> 16);
    }
  }
};
template <>
struct Vec<COLA_BF16, 8> : Vec16<COLA_BF16> {};
template <>
struct Vec<COLA_F16, 8> : Vec16<COLA_F16> {};
template <int DT>
struct Vec16x1 {
  using T = uint16_t;
  using Raw = uint32_t;
  static __device__ __forceinline__ Raw load(const T *p) { return *p; }
  static __device__ __forceinline__ Raw gload(uint64_t a) { return load_global<uint16_t>(a); }
  static __device__ __forceinline__ void widen(const Raw &r, float (&x)[1]) { x[0] = Vec16<DT>::cvt(r); }
};
template <>
struct Vec<COLA_BF16, 1> : Vec16x1<COLA_BF16> {};
template <>
struct Vec<COLA_F16, 1> : Vec16x1<COLA_F16> {};
template <>
struct Vec<COLA_FP8_E4M3, 16> {
  using T = uint8_t;
  using Raw = uint4;
  static __device__ __forceinline__ Raw load(const T *p) { return *reinterpret_cast<const uint4 *>(p); }
  static __device__ __forceinline__ Raw gload(uint64_t a) { return load_global_u4(a); }
  // gfx950's v_cvt_pk_f32_fp8 decodes OCP E4M3 (checked against the format
  // definition for all 256 codes), two bytes per instruction.
  static __device__ __forceinline__ void widen(const Raw &r, float (&x)[16]) {
    const uint32_t w[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const auto lo = __builtin_amdgcn_cvt_pk_f32_fp8((int)w[i], false);
      const auto hi = __builtin_amdgcn_cvt_pk_f32_fp8((int)w[i], true);
      x[4 * i] = lo[0];
      x[4 * i + 1] = lo[1];
      x[4 * i + 2] = hi[0];
      x[4 * i + 3] = hi[1];
    }
  }
};
template <>
struct Vec<COLA_FP8_E4M3, 1> {
  using T = uint8_t;
  using Raw = uint32_t;
  static __device__ __forceinline__ Raw load(const T *p) { return *p; }
  static __device__ __forceinline__ Raw gload(uint64_t a) { return load_global<uint8_t>(a); }
  static __device__ __forceinline__ void widen(const Raw &r, float (&x)[1]) { x[0] = fp8e4m3_to_f32(r); }
};

// Bytes of the zero row: the widest column tile a row kernel reads (256
// 16-byte lane vectors, the limit of pick_shape).
inline constexpr int kZeroRowBytes = 4096;

struct RowArgs {
  const int4 *items;
  const int32_t *xcd_begin;  // 9 entries
  const int32_t *cols;
  const float *vals;
  const void *Bl;
  int64_t ldB;
  const float *sl;
  const void *Bp;
  int64_t ldP;
  const float *sp;
  float *C;
  int64_t ldC;
  float *partials;
  int64_t pstride;
  int dim;
  int accumulate;
  int kahan;
  int rpw;        // teams (items in flight) per wave, power of two <= 64/G
  int xcd_remap;  // 1: block b walks XCD (b % 8)'s item range
  int phase;      // 0: all entries; 1: local entries only; 2: remote entries
                  // only, added to what phase 1 wrote; 3: all entries, rows
                  // added to what the first pass of a hot/cold plan wrote
  uint64_t abase; // nonzero: every row address of Bl and Bp is abase + a
                  // 32-bit multiple of 16 bytes (the host checked the span
                  // and alignment), so one word per entry is broadcast
                  // instead of two; 0: full 64-bit addresses
  uint64_t zrow;  // kZeroRowBytes of zeros (with abase: inside its span), the
                  // row a staged slot past an item's end reads
};

// Two forms of the staging (PIPE): false turns the next chunk's staged
// entries into row addresses as soon as they arrive; true defers that (and
// FP8's scale loads) until this chunk's row loads are in flight and keeps the
// compiler from interleaving loads with their use.  Chosen per kernel by
// measurement: on the 2^22-node partition at dim 128 the plan kernel is
// 1.6% (FP32), 2.6% (BF16) and 3% (FP8) faster with PIPE, each at its best
// schedule, and far less sensitive to the schedule; the batched kernel (64
// BF16 problems) is 3% faster without.
template <int DT>
inline constexpr bool kPlanPipe = true;
template <int DT>
inline constexpr bool kBatchPipe = false;

// One step of a wave: the `rpw` items it0 .. it0+rpw-1 (those below ie), one
// per team, over the column range [0, a.dim) of the tile a's pointers start
// at.  Shared by the plan kernel (compile-time G per launch) and the batched
// persistent kernel (G chosen per segment).
template <int DT, int EPL, int G, int NV, int U, bool KAHAN, bool PIPE>
__device__ __forceinline__ void row_step(const RowArgs &a, int it0, int ie, int rpw) {
  using V = Vec<DT, EPL>;
  using T = typename V::T;
  using Raw = typename V::Raw;
  constexpr int S = 64 / G;
  const T *__restrict__ Bl = static_cast<const T *>(a.Bl);
  const T *__restrict__ Bp = static_cast<const T *>(a.Bp);
  const int lane = threadIdx.x & 63;
  const int sub = lane / G, gl = lane % G;
  const int gpt = S / rpw;     // groups per team
  const int tsize = 64 / rpw;  // lanes per team
  const int team = sub / gpt, tsub = sub % gpt;
  const int tbase = team * tsize, tl = lane - tbase;
  const int it = it0 + team;
  const bool has = it < ie;
  int4 item = make_int4(0, 0, 0, 0);
  if (has) item = gld(a.items + it);
  int b = item.y, e = item.z;
  if (a.phase == 1) e = b + (item.w >> 1);
  else if (a.phase == 2) b += item.w >> 1;
  // Compensated sums only in the KAHAN instantiation and on flagged items.
  const bool comp = KAHAN && (item.w & 1);
  float acc[NV][EPL], err[NV][EPL];
#pragma unroll
  for (int k = 0; k < NV; ++k)
#pragma unroll
    for (int q = 0; q < EPL; ++q) acc[k][q] = err[k][q] = 0.f;

  // Entries are taken in chunks of K*tsize: each team lane stages K entries
  // of the next chunk while this one is used, in two steps: the coalesced
  // (column, value) loads are issued before this chunk's row loads, and
  // turned into the address of their B row (local slice or packed buffer)
  // by stage_addr (with PIPE after this chunk's row loads are issued, so the
  // wait for the staged loads overlaps the row loads in flight).  The
  // address arithmetic thus runs once per entry rather than once per
  // broadcast slot.  FP8 rows also take their scale there, multiplied into
  // the value when the chunk becomes current (rotate).  A lane
  // whose entry lies past the item's end stages the zero row (a.zrow) with
  // value 0, so every broadcast slot holds a valid row address, no load or
  // FMA is predicated, and an Inf or NaN in a B row reaches a sum only
  // through a stored entry (the reference multiplies exactly the entries
  // of [rowPtr[v], rowPtr[v+1])).
  constexpr int K = U > G ? U / G : 1;  // entries per lane: one batch of U per group fills K*tsize
  const int csize = K * tsize;
  uint32_t mlo[K], mhi[K], mlo_next[K], mhi_next[K];
  float mv[K], mv_next[K], msc_next[K];
  int nc[K];    // staged columns of the next chunk
  float nx[K];  // staged values of the next chunk
  auto stage_load = [&](int base) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int cl = min(base + k * tsize + tl, e - 1);
      nc[k] = gld(a.cols + cl);
      nx[k] = gld(a.vals + cl);
    }
  };
  auto stage_addr = [&](int base) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int c = nc[k];
      const bool loc = c >= 0;
      const bool in = base + k * tsize + tl < e;
      const T *rp = loc ? Bl + (int64_t)c * a.ldB : Bp + (int64_t)(~c) * a.ldP;
      const uint64_t p = in ? reinterpret_cast<uint64_t>(rp) : a.zrow;
      if (a.abase) {
        mlo_next[k] = (uint32_t)((p - a.abase) >> 4);
        mhi_next[k] = 0;
      } else {
        mlo_next[k] = (uint32_t)p;
        mhi_next[k] = (uint32_t)(p >> 32);
      }
      mv_next[k] = in ? nx[k] : 0.f;
      if constexpr (DT == COLA_FP8_E4M3) msc_next[k] = in ? gld(loc ? a.sl + c : a.sp + ~c) : 1.f;
    }
  };
  auto rotate = [&]() {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      mlo[k] = mlo_next[k];
      mhi[k] = mhi_next[k];
      mv[k] = DT == COLA_FP8_E4M3 ? mv_next[k] * msc_next[k] : mv_next[k];
    }
  };
#pragma unroll
  for (int k = 0; k < K; ++k) {
    mlo[k] = mhi[k] = mlo_next[k] = mhi_next[k] = 0;
    mv[k] = mv_next[k] = 0.f;
    msc_next[k] = 1.f;
    nc[k] = 0;
    nx[k] = 0.f;
  }
  // This lane's offset into a B row, per vector (clamped to the row start
  // for the lanes past dim, whose results are never stored).
  int lcol[NV];
  uint64_t lbase[NV];  // with abase: this lane's column in the common table
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const int col = (k * G + gl) * EPL;
    lcol[k] = col < a.dim ? col : 0;
    lbase[k] = a.abase + (uint64_t)lcol[k] * sizeof(T);
  }
  if (b < e) {
    stage_load(b);
    stage_addr(b);
    rotate();
  }
  for (int base = b; base < e; base += csize) {
    const int n = min(csize, e - base);
    const bool more = base + csize < e;
    if (more) stage_load(base + csize);
    if constexpr (!PIPE)
      if (more) stage_addr(base + csize);
    for (int j = 0; j < n; j += gpt * U) {
      // Broadcast the batch's (row address, value) pairs first, then issue
      // every load.  With K > 1 the batch is the whole chunk (j = 0) and
      // slot u reads entry u / G of team lane (u % G) * gpt + tsub.
      float av[U];
      Raw raw[U][NV];
      if (a.abase) {
        // One word per entry: the row's 16-byte offset from abase.
        uint32_t ou[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int k = K > 1 ? u / G : 0;
          const int src = (tbase + (K > 1 ? (u % G) * gpt + tsub : j + u * gpt + tsub)) << 2;
          ou[u] = __builtin_amdgcn_ds_bpermute(src, (int)mlo[k]);
          av[u] = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(mv[k])));
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int k = 0; k < NV; ++k) raw[u][k] = V::gload(lbase[k] + ((uint64_t)ou[u] << 4));
      } else {
        uint64_t pu[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int k = K > 1 ? u / G : 0;
          const int src = (tbase + (K > 1 ? (u % G) * gpt + tsub : j + u * gpt + tsub)) << 2;
          const uint32_t lo = __builtin_amdgcn_ds_bpermute(src, (int)mlo[k]);
          const uint32_t hi = __builtin_amdgcn_ds_bpermute(src, (int)mhi[k]);
          pu[u] = ((uint64_t)hi << 32) | lo;
          av[u] = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(mv[k])));
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int k = 0; k < NV; ++k) raw[u][k] = V::gload(pu[u] + (uint64_t)lcol[k] * sizeof(T));
      }
      if constexpr (PIPE) {
        if (more && j == 0) {
          // Pin the staged registers here, after this batch's row loads (the
          // memory clobber keeps the loads above), so the address arithmetic
          // and its wait for the staged loads are not hoisted above them.
#pragma unroll
          for (int k = 0; k < K; ++k) asm volatile("" : "+v"(nc[k]), "+v"(nx[k])::"memory");
          stage_addr(base + csize);
        }
        // Keep every load of the batch ahead of its use: without this the
        // scheduler interleaves loads with the FP8 / 16-bit widening to save
        // registers and leaves one or two loads in flight instead of U.
        __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int k = 0; k < NV; ++k) {
          float x[EPL];
          V::widen(raw[u][k], x);
          if (!comp) {
#pragma unroll
            for (int q = 0; q < EPL; ++q) acc[k][q] = fmaf(av[u], x[q], acc[k][q]);
          } else {
            // Kahan: err carries the low-order part lost by acc.
#pragma unroll
            for (int q = 0; q < EPL; ++q) {
              const float y = fmaf(av[u], x[q], -err[k][q]);
              const float t = acc[k][q] + y;
              err[k][q] = __builtin_isfinite(t) ? (t - acc[k][q]) - y : 0.f;
              acc[k][q] = t;
            }
          }
        }
    }
    rotate();
  }
  // Sum the groups of each team (xor offsets below tsize stay in the team).
  for (int off = G; off < tsize; off <<= 1) {
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
      for (int q = 0; q < EPL; ++q) {
        const float s2 = __shfl_xor(acc[k][q], off);
        const float e2 = __shfl_xor(err[k][q], off);
        if (!comp) {
          acc[k][q] += s2;
        } else {
          // Neumaier combination of (acc, -err) pairs.
          const float s1 = acc[k][q];
          const float t = s1 + s2;
          const float lost = fabsf(s1) >= fabsf(s2) ? (s1 - t) + s2 : (s2 - t) + s1;
          acc[k][q] = t;
          err[k][q] = __builtin_isfinite(t) ? err[k][q] + e2 - lost : 0.f;
        }
      }
  }
  if (has && tsub == 0) {
    float *dst;
    bool add_old;
    if (item.x >= 0) {
      dst = a.C + (int64_t)item.x * a.ldC;
      add_old = a.accumulate != 0 || a.phase >= 2;
    } else {
      dst = a.partials + (int64_t)(-item.x - 1) * a.pstride;
      add_old = a.phase == 2;
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const int col = (k * G + gl) * EPL;
      if (col < a.dim) {
        float v[EPL];
#pragma unroll
        for (int q = 0; q < EPL; ++q) v[q] = comp ? acc[k][q] - err[k][q] : acc[k][q];
        if constexpr (EPL % 4 == 0) {
#pragma unroll
          for (int q = 0; q < EPL; q += 4) {
            float4 *d4 = reinterpret_cast<float4 *>(dst + col + q);
            float4 o = make_float4(v[q], v[q + 1], v[q + 2], v[q + 3]);
            if (add_old) {
              const float4 old = gld(d4);
              o.x += old.x; o.y += old.y; o.z += old.z; o.w += old.w;
            }
            gst(d4, o);
          }
        } else {
#pragma unroll
          for (int q = 0; q < EPL; ++q) gst(dst + col + q, add_old ? gld(dst + col + q) + v[q] : v[q]);
        }
      }
    }
  }
}

template <int DT, int EPL, int G, int NV, int U, bool KAHAN>
__global__ __launch_bounds__(256) void plan_rows_kernel(const RowArgs a) {
  const int rpw = a.rpw;
  const int wpb = blockDim.x >> 6;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int blk, nblk, ib, ie;
  if (a.xcd_remap) {
    const int xcd = blockIdx.x & 7;
    blk = blockIdx.x >> 3;
    nblk = (gridDim.x - xcd + 7) >> 3;
    ib = a.xcd_begin[xcd];
    ie = a.xcd_begin[xcd + 1];
  } else {
    blk = blockIdx.x;
    nblk = gridDim.x;
    ib = a.xcd_begin[0];
    ie = a.xcd_begin[8];
  }
  const int wave = blk * wpb + wid;
  const int nwave = nblk * wpb;
  for (int it0 = ib + wave * rpw; it0 < ie; it0 += nwave * rpw)
    row_step<DT, EPL, G, NV, U, KAHAN, kPlanPipe<DT>>(a, it0, ie, rpw);
}

// One bundle of tiny rows per lane group: R = 8 / L rows of L entries each
// (the last bundle of a length may hold fewer), entries back to back at
// [b, e).  Group lane u < R*L stages entry b + u (coalesced) and turns it
// into its B row's address and value, the group broadcasts the R*L pairs
// with ds_bpermute and issues all R*L row loads before any FMA, and row j
// sums slots j*L .. j*L+L-1 in entry order (the short bin's order for a row
// that one lane group sums alone).  One 16-byte vector per lane (NV = 1).
template <int DT, int EPL, int G, int L>
__device__ __forceinline__ void tiny_bundle(const RowArgs &a, const int32_t *__restrict__ rowtab, const int4 item,
                                            bool has, int gl, int tbase, uint64_t lbase, int lcol) {
  using V = Vec<DT, EPL>;
  using T = typename V::T;
  using Raw = typename V::Raw;
  constexpr int R = 8 / L, N = R * L;
  static_assert(N <= G, "a bundle's entries are staged one per group lane");
  const T *__restrict__ Bl = static_cast<const T *>(a.Bl);
  const T *__restrict__ Bp = static_cast<const T *>(a.Bp);
  const int b = item.y, n = has ? item.z - item.y : 0;  // n = rows * L <= N
  // Stage: lane gl < n takes entry b + gl; the others stage the zero row
  // with value 0, so every broadcast slot holds a valid row address and
  // only stored entries meet an Inf or NaN of B.
  uint32_t mlo = 0, mhi = 0;
  float mv = 0.f;
  if (has) {
    const int k = b + (gl < n ? gl : 0);
    const int c = gld(a.cols + k);
    const float x = gld(a.vals + k);
    const bool loc = c >= 0;
    const bool in = gl < n;
    const T *rp = loc ? Bl + (int64_t)c * a.ldB : Bp + (int64_t)(~c) * a.ldP;
    const uint64_t p = in ? reinterpret_cast<uint64_t>(rp) : a.zrow;
    if (a.abase) {
      mlo = (uint32_t)((p - a.abase) >> 4);
    } else {
      mlo = (uint32_t)p;
      mhi = (uint32_t)(p >> 32);
    }
    mv = in ? x : 0.f;
    if constexpr (DT == COLA_FP8_E4M3)
      if (in) mv *= gld(loc ? a.sl + c : a.sp + ~c);
  }
  if (n == 0) return;  // past the XCD range's last bundle (uniform over the group)
  float av[N];
  Raw raw[N];
  if (a.abase) {
    uint32_t ou[N];
#pragma unroll
    for (int u = 0; u < N; ++u) {
      const int src = (tbase + u) << 2;
      ou[u] = __builtin_amdgcn_ds_bpermute(src, (int)mlo);
      av[u] = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(mv)));
    }
#pragma unroll
    for (int u = 0; u < N; ++u) raw[u] = V::gload(lbase + ((uint64_t)ou[u] << 4));
  } else {
    uint64_t pu[N];
#pragma unroll
    for (int u = 0; u < N; ++u) {
      const int src = (tbase + u) << 2;
      const uint32_t lo = __builtin_amdgcn_ds_bpermute(src, (int)mlo);
      const uint32_t hi = __builtin_amdgcn_ds_bpermute(src, (int)mhi);
      pu[u] = ((uint64_t)hi << 32) | lo;
      av[u] = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(mv)));
    }
#pragma unroll
    for (int u = 0; u < N; ++u) raw[u] = V::gload(pu[u] + (uint64_t)lcol * sizeof(T));
  }
  // Output rows of the bundle: lane j < R reads row j, broadcast below.
  const int rows = n / L;
  const int rid = gl < rows ? gld(rowtab + item.x + gl) : 0;
  __builtin_amdgcn_sched_barrier(0);
  float acc[R][EPL];
#pragma unroll
  for (int j = 0; j < R; ++j)
#pragma unroll
    for (int q = 0; q < EPL; ++q) acc[j][q] = 0.f;
#pragma unroll
  for (int u = 0; u < N; ++u) {
    float x[EPL];
    V::widen(raw[u], x);
#pragma unroll
    for (int q = 0; q < EPL; ++q) acc[u / L][q] = fmaf(av[u], x[q], acc[u / L][q]);
  }
  const bool lane_ok = gl * EPL < a.dim;  // lanes past the tile's columns store nothing
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int row = __shfl(rid, tbase + j);
    if (j < rows && lane_ok) {
      float *dst = a.C + (int64_t)row * a.ldC + lcol;
#pragma unroll
      for (int q = 0; q < EPL; q += 4) {
        float4 o = make_float4(acc[j][q], acc[j][q + 1], acc[j][q + 2], acc[j][q + 3]);
        if (a.accumulate) {
          const float4 old = gld(reinterpret_cast<const float4 *>(dst + q));
          o.x += old.x; o.y += old.y; o.z += old.z; o.w += old.w;
        }
        gst(reinterpret_cast<float4 *>(dst + q), o);
      }
    }
  }
}

// One step of a wave over tiny-row bundles it0 .. it0+S-1 (those below ie),
// one per lane group (tiny_rows_kernel; the batched kernel runs these rows as
// short-bin items: the bundle's registers would make its other cases spill).
template <int DT, int EPL, int G>
__device__ __forceinline__ void tiny_step(const RowArgs &a, const int32_t *__restrict__ rowtab, int it0, int ie) {
  using T = typename Vec<DT, EPL>::T;
  const int lane = threadIdx.x & 63;
  const int sub = lane / G, gl = lane % G, tbase = sub * G;
  const int col = gl * EPL;
  const int lcol = col < a.dim ? col : 0;
  const uint64_t lbase = a.abase + (uint64_t)lcol * sizeof(T);
  const int it = it0 + sub;
  const bool has = it < ie;
  const int4 item = has ? gld(a.items + it) : make_int4(0, 0, 0, 4);
  // The length is uniform over a lane group; groups of one wave differ only
  // where the bundles of one length end and the next begin.
  switch (item.w) {
    case 1: tiny_bundle<DT, EPL, G, 1>(a, rowtab, item, has, gl, tbase, lbase, lcol); break;
    case 2: tiny_bundle<DT, EPL, G, 2>(a, rowtab, item, has, gl, tbase, lbase, lcol); break;
    case 3: tiny_bundle<DT, EPL, G, 3>(a, rowtab, item, has, gl, tbase, lbase, lcol); break;
    default: tiny_bundle<DT, EPL, G, 4>(a, rowtab, item, has, gl, tbase, lbase, lcol); break;
  }
}

// Tiny-row bundles (plan items (row offset, begin, end, L)): one bundle per
// lane group per step, items walked like plan_rows_kernel's (XCD ranges).
template <int DT, int EPL, int G>
__global__ __launch_bounds__(256) void tiny_rows_kernel(const RowArgs a, const int32_t *__restrict__ rowtab) {
  constexpr int S = 64 / G;
  const int wpb = blockDim.x >> 6;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int blk, nblk, ib, ie;
  if (a.xcd_remap) {
    const int xcd = blockIdx.x & 7;
    blk = blockIdx.x >> 3;
    nblk = (gridDim.x - xcd + 7) >> 3;
    ib = a.xcd_begin[xcd];
    ie = a.xcd_begin[xcd + 1];
  } else {
    blk = blockIdx.x;
    nblk = gridDim.x;
    ib = a.xcd_begin[0];
    ie = a.xcd_begin[8];
  }
  const int wave = blk * wpb + wid, nwave = nblk * wpb;
  for (int it0 = ib + wave * S; it0 < ie; it0 += nwave * S) tiny_step<DT, EPL, G>(a, rowtab, it0, ie);
}

// Lanes per row group and vectors per lane for a dim.  vec_ok: dim, leading
// dimensions and base pointers allow 16-byte vectors.
struct Shape {
  int epl, g, nv;
};

inline bool pick_shape(int dt, int dim, bool vec_ok, Shape *s) {
  if (dim < 1) return false;
  const int ev = dt == COLA_F32 ? 4 : (dt == COLA_FP8_E4M3 ? 16 : 8);
  if (vec_ok && dim % ev == 0) {
    const int chunks = dim / ev;
    s->epl = ev;
    s->nv = 1;
    if (chunks <= 2) s->g = 2;
    else if (chunks <= 4) s->g = 4;
    else if (chunks <= 8) s->g = 8;
    else if (chunks <= 16) s->g = 16;
    else if (chunks <= 32) s->g = 32;
    else if (chunks <= 64) s->g = 64;
    else if (chunks <= 128) { s->g = 64; s->nv = 2; }
    else if (chunks <= 256) { s->g = 64; s->nv = 4; }
    else return false;
    return true;
  }
  s->epl = 1;
  s->g = 64;
  if (dim <= 64) s->nv = 1;
  else if (dim <= 128) s->nv = 2;
  else if (dim <= 256) s->nv = 4;
  else if (dim <= 512) s->nv = 8;
  else return false;
  return true;
}

// Loads in flight per group: U8 (8 or 16), fewer when a lane already holds
// several vectors per entry.
template <int NV, int U8>
struct Unroll {
  static constexpr int value = NV >= U8 ? 1 : U8 / NV;
};

// Launches one row kernel; returns false if no instantiation matches.
bool launch_rows_f32(const Shape &sh, const RowArgs &a, int blocks, int threads, hipStream_t s,
                      bool unroll16);
bool launch_rows_bf16(const Shape &sh, const RowArgs &a, int blocks, int threads, hipStream_t s,
                      bool unroll16);
bool launch_rows_f16(const Shape &sh, const RowArgs &a, int blocks, int threads, hipStream_t s,
                      bool unroll16);
bool launch_rows_fp8(const Shape &sh, const RowArgs &a, int blocks, int threads, hipStream_t s,
                      bool unroll16);

// Launches the tiny-row bundle kernel; returns false when the shape has no
// bundle form (one 16-byte vector per lane and at least 8 lanes per group).
bool launch_tiny_f32(const Shape &sh, const RowArgs &a, const int32_t *rowtab, int blocks, hipStream_t s);
bool launch_tiny_bf16(const Shape &sh, const RowArgs &a, const int32_t *rowtab, int blocks, hipStream_t s);
bool launch_tiny_f16(const Shape &sh, const RowArgs &a, const int32_t *rowtab, int blocks, hipStream_t s);
bool launch_tiny_fp8(const Shape &sh, const RowArgs &a, const int32_t *rowtab, int blocks, hipStream_t s);

inline bool tiny_shape_ok(int dt, const Shape &sh) {
  const int ev = dt == COLA_F32 ? 4 : (dt == COLA_FP8_E4M3 ? 16 : 8);
  return sh.epl == ev && sh.nv == 1 && sh.g >= 8;
}

template <int DT, int EV>
bool launch_tiny_impl(const Shape &sh, const RowArgs &a, const int32_t *rowtab, int blocks, hipStream_t s) {
  if (sh.epl != EV || sh.nv != 1) return false;
  switch (sh.g) {
    case 8: hipLaunchKernelGGL((tiny_rows_kernel<DT, EV, 8>), dim3(blocks), dim3(256), 0, s, a, rowtab); return true;
    case 16: hipLaunchKernelGGL((tiny_rows_kernel<DT, EV, 16>), dim3(blocks), dim3(256), 0, s, a, rowtab); return true;
    case 32: hipLaunchKernelGGL((tiny_rows_kernel<DT, EV, 32>), dim3(blocks), dim3(256), 0, s, a, rowtab); return true;
    case 64: hipLaunchKernelGGL((tiny_rows_kernel<DT, EV, 64>), dim3(blocks), dim3(256), 0, s, a, rowtab); return true;
  }
  return false;
}

template <int DT, int EV>
bool launch_rows_impl(const Shape &sh, const RowArgs &a, int blocks, int threads, hipStream_t s,
                      bool unroll16) {
#define COLA_ROWS_LAUNCH(E, G_, NV_, U8)                                                          \
  do {                                                                                            \
    if (a.kahan)                                                                                  \
      hipLaunchKernelGGL((plan_rows_kernel<DT, E, G_, NV_, Unroll<NV_, U8>::value, true>),         \
                         dim3(blocks), dim3(threads), 0, s, a);                                   \
    else                                                                                          \
      hipLaunchKernelGGL((plan_rows_kernel<DT, E, G_, NV_, Unroll<NV_, U8>::value, false>),        \
                         dim3(blocks), dim3(threads), 0, s, a);                                   \
  } while (0)
#define COLA_ROWS_CASE(E, G_, NV_)                                                                \
  if (sh.epl == E && sh.g == G_ && sh.nv == NV_) {                                                \
    if (unroll16 && NV_ == 1 && E > 1) COLA_ROWS_LAUNCH(E, G_, NV_, 16);                          \
    else COLA_ROWS_LAUNCH(E, G_, NV_, 8);                                                         \
    return true;                                                                                  \
  }
  COLA_ROWS_CASE(EV, 2, 1)
  COLA_ROWS_CASE(EV, 4, 1)
  COLA_ROWS_CASE(EV, 8, 1)
  COLA_ROWS_CASE(EV, 16, 1)
  COLA_ROWS_CASE(EV, 32, 1)
  COLA_ROWS_CASE(EV, 64, 1)
  COLA_ROWS_CASE(EV, 64, 2)
  COLA_ROWS_CASE(EV, 64, 4)
  COLA_ROWS_CASE(1, 64, 1)
  COLA_ROWS_CASE(1, 64, 2)
  COLA_ROWS_CASE(1, 64, 4)
  COLA_ROWS_CASE(1, 64, 8)
#undef COLA_ROWS_CASE
#undef COLA_ROWS_LAUNCH
  return false;
}

}  // namespace cola_dev
