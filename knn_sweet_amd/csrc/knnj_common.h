// knnj_common.h -- types and wave64 helpers shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef unsigned long long u64;
typedef uint32_t u32;

#define KNNJ_WAVE 64
#define KEY_MAX 0xFFFFFFFFFFFFFFFFull
#define KNNJ_MAX_REPS 4096  // representatives per side (LDS-resident lists in the bounds kernel)
#define KNNJ_MAX_K 1024
#define KNNJ_BCAP 128       // per-query candidate buffer (two 64-candidate chunks)

// Candidate key: the squared distance (a non-negative binary32, whose bit
// pattern orders like its value) in the high word, a row id in the low word.
// Lexicographic u64 order is then the (squared distance, row) order.
__device__ __forceinline__ u64 make_key(float s, u32 row) {
  return ((u64)__float_as_uint(s) << 32) | (u64)row;
}
__device__ __forceinline__ float key_dist(u64 k) { return __uint_as_float((u32)(k >> 32)); }
__device__ __forceinline__ u32 key_row(u64 k) { return (u32)k; }

// Filter keys of the MFMA kernel carry, above a row field of `rowbits`
// bits, a code c of the entry's error bound E with 2^lgs steps per octave:
// decode(c) = 2^(c / 2^lgs + ebase) >= E, the top code meaning unbounded.
// (rowbits = 26 - lgs: a 6 + lgs bit code covers 64 octaves.)  A list
// entry's lower bound is key distance - 2 decode(c).
struct ECode {
  int ebase;    // codes start at 2^ebase
  int lgs;      // log2 of the steps per octave, 0..3
  int rowbits;  // width of the row field
};
__device__ __forceinline__ u32 ecode_top(const ECode &ec) { return (64u << ec.lgs) - 1u; }
__device__ __forceinline__ u32 key_row_of(u64 key, const ECode &ec) {
  return (u32)key & ((1u << ec.rowbits) - 1u);
}
__device__ __forceinline__ u32 ecode_encode(float E, const ECode &ec) {
  if (!(E > 0.f)) return 0u;
  const int e = ilogbf(E);
  if (e < ec.ebase) return 0u;  // E < 2^ebase = decode(0)
  const int S = 1 << ec.lgs;
  // j >= S log2(m), m = E / 2^e in [1, 2): the 1e-3 margin exceeds the
  // error of the hardware log2 by orders of magnitude
  const float t = (float)S * __log2f(ldexpf(E, -e));
  const int j = min(S, (int)ceilf(t + 1e-3f));
  const u32 c = ((u32)(e - ec.ebase) << ec.lgs) + (u32)j;
  return c >= ecode_top(ec) ? ecode_top(ec) : c;
}
// 2^(i/8) (1 + 2^-20) rounded up to binary32, i = 3 bits, as a select tree
// (the join's epilogues decode per B group: no division or exp2 there)
__device__ __forceinline__ float ecode_step(u32 i) {
  const float lo = (i & 2u) ? ((i & 1u) ? 0x1.4bfdc4p+0f : 0x1.306ff4p+0f)
                            : ((i & 1u) ? 0x1.172b96p+0f : 0x1.00001p+0f);
  const float hi = (i & 2u) ? ((i & 1u) ? 0x1.d581acp+0f : 0x1.ae8a16p+0f)
                            : ((i & 1u) ? 0x1.8ace6ep+0f : 0x1.6a09fep+0f);
  return (i & 4u) ? hi : lo;
}
__device__ __forceinline__ float ecode_decode(u32 c, const ECode &ec) {
  if (c >= ecode_top(ec)) return INFINITY;
  const u32 j = c & ((1u << ec.lgs) - 1u);  // steps of 1/2^lgs octave = j 2^(3-lgs) eighths
  return ldexpf(ecode_step(j << (3 - ec.lgs)), (int)(c >> ec.lgs) + ec.ebase);
}
__device__ __forceinline__ float key_lower(u64 key, const ECode &ec) {
  return key_dist(key) - 2.f * ecode_decode((u32)key >> ec.rowbits, ec);
}

// Orders this wave's LDS accesses across its lanes: a wave's LDS
// instructions execute in issue order, so compiler fences at wavefront scope
// around a wave barrier suffice (no workgroup barrier; the other waves of a
// multi-wave workgroup are not involved).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ u32 lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// number of set bits of `mask` below this lane
__device__ __forceinline__ u32 lanes_below(u64 mask) {
  return __builtin_amdgcn_mbcnt_hi((u32)(mask >> 32), __builtin_amdgcn_mbcnt_lo((u32)mask, 0u));
}

__device__ __forceinline__ u64 wave_ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ u64 shfl_xor_u64(u64 v, int m) {
  u32 lo = (u32)v, hi = (u32)(v >> 32);
  lo = __shfl_xor(lo, m);
  hi = __shfl_xor(hi, m);
  return ((u64)hi << 32) | lo;
}

// v of lane (lane ^ m), for a mask m known at compile time: DPP moves within
// rows of 16 lanes (quad permutes: ^1, ^2, ^3; half-row mirror: ^7; row
// mirror: ^15; two moves: ^4 = ^7^3, ^8 = ^15^7), the swizzle's xor mode
// within 32 lanes (^16, ^31), the swap of the two 32-lane halves
// (v_permlane32_swap: ^32; after a ^31 swizzle: ^63), a permute through LDS
// otherwise.  Only the permute makes an LDS round trip.  Callers run with
// all 64 lanes active (the DPP forms read 0 from a disabled source lane).
__device__ __forceinline__ u32 xor_lanes(u32 v, int m) {
  const int x = (int)v;
  switch (m) {
    case 1: return (u32)__builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, true);  // quad_perm [1,0,3,2]
    case 2: return (u32)__builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, true);  // quad_perm [2,3,0,1]
    case 3: return (u32)__builtin_amdgcn_mov_dpp(x, 0x1B, 0xF, 0xF, true);  // quad_perm [3,2,1,0]
    case 4: return (u32)__builtin_amdgcn_mov_dpp(__builtin_amdgcn_mov_dpp(x, 0x141, 0xF, 0xF, true), 0x1B, 0xF, 0xF,
                                                 true);
    case 7: return (u32)__builtin_amdgcn_mov_dpp(x, 0x141, 0xF, 0xF, true);  // row_half_mirror
    case 8: return (u32)__builtin_amdgcn_mov_dpp(__builtin_amdgcn_mov_dpp(x, 0x140, 0xF, 0xF, true), 0x141, 0xF, 0xF,
                                                 true);
    case 15: return (u32)__builtin_amdgcn_mov_dpp(x, 0x140, 0xF, 0xF, true);  // row_mirror
    case 16: return (u32)__builtin_amdgcn_ds_swizzle(x, 0x1F | (16 << 10));
    case 31: return (u32)__builtin_amdgcn_ds_swizzle(x, 0x1F | (31 << 10));
    case 32: {
      // lanes 32..63 of the first operand trade places with lanes 0..31 of
      // the second: the second then holds v[lane + 32] below lane 32, the
      // first v[lane - 32] above it
      const auto r = __builtin_amdgcn_permlane32_swap((u32)x, (u32)x, false, false);
      return lane_id() < 32 ? (u32)r[1] : (u32)r[0];
    }
    case 63: {
      const u32 y = (u32)__builtin_amdgcn_ds_swizzle(x, 0x1F | (31 << 10));
      const auto r = __builtin_amdgcn_permlane32_swap(y, y, false, false);
      return lane_id() < 32 ? (u32)r[1] : (u32)r[0];
    }
    default: return (u32)__shfl_xor(x, m);
  }
}
__device__ __forceinline__ u64 xor_lanes_u64(u64 v, int m) {
  return ((u64)xor_lanes((u32)(v >> 32), m) << 32) | xor_lanes((u32)v, m);
}
__device__ __forceinline__ float xor_lanes_f(float v, int m) {
  return __builtin_bit_cast(float, xor_lanes(__builtin_bit_cast(u32, v), m));
}
__device__ __forceinline__ int xor_lanes_i(int v, int m) { return (int)xor_lanes((u32)v, m); }

// Persistent launches: eight tile-slot counters q[0..8), one per XCD; slot s
// belongs to counter s % 8, the XCD the hardware's round-robin would give
// workgroup s.  A wave takes from its own XCD's counter (home), then from
// the others in turn (turn), and gets -1 once every counter has passed
// nslots.  Counters only grow, so every wave reaches -1.
__device__ __forceinline__ int queue_home() {
  // HW_REG_XCC_ID (register 20), bits [3:0]
  return (int)(__builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11)) & 7u);
}
__device__ __forceinline__ int queue_next(int32_t *q, int nslots, int home, int &turn) {
  while (turn < 8) {
    const int x = (home + turn) & 7;
    int t = 0;
    if ((threadIdx.x & 63) == 0) t = atomicAdd(q + x, 1);
    t = __shfl(t, 0);
    const int64_t s = 8 * (int64_t)t + x;
    if (s < nslots) return (int)s;
    ++turn;
  }
  return -1;
}

// butterfly reductions (the same lane pairs as __shfl_xor; the exchanges
// within 32 lanes as DPP moves and swizzles)
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += xor_lanes_f(v, m);
  return v;
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1)
    v += __builtin_bit_cast(double, xor_lanes_u64(__builtin_bit_cast(u64, v), m));
  return v;
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fminf(v, xor_lanes_f(v, m));
  return v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmaxf(v, xor_lanes_f(v, m));
  return v;
}

// Relative error bound of a binary32 Euclidean distance sqrt(sum_j (a_j-b_j)^2)
// over d dimensions (terms non-negative, any summation order, with or without
// FMA): (d+2) u for the squared sum, half that plus u after sqrt.  Doubled
// for margin.
__host__ __device__ __forceinline__ float dist_relerr(int d) { return (float)(d + 4) * 1.1920929e-7f; }

// The binary64 key of one pair, in the oracle's fixed operation sequence
// (ascending dimension, no fused multiply-add).
__device__ __forceinline__ double sq64_pair(const float *__restrict__ q, const float *__restrict__ r, int d) {
#pragma clang fp contract(off)
  double s = 0.0;
  for (int j = 0; j < d; ++j) {
    double t = (double)q[j] - (double)r[j];
    s = s + t * t;
  }
  return s;
}

// Per-side cluster index in device memory (all arrays device pointers).
struct SideIndex {
  int64_t n;          // rows
  int32_t m;          // clusters (representatives)
  int32_t dpad;       // padded row length of the permuted copy
  const float *reps;  // [m][dpad] representative rows (FP32, zero padded)
  const double *rad;  // [m] cluster radius (binary64, max member distance to rep)
  const int32_t *off; // [m+1] cluster start in permuted order
  const int32_t *perm;// [n] permuted position -> original row
  const float *pdist; // [n] distance to rep in permuted order, ascending per cluster
  const float *px;    // [n][dpad] permuted rows
};

// A query tile: `count` consecutive permuted queries of cluster `cluster`.
struct Tile {
  int32_t start;
  int32_t count;
  int32_t cluster;
  int32_t pad;
};
