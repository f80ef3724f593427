// k_join_mfma.h -- the join proper on the matrix cores (d >= 32).
//
// Same contract as k_join_valu (k_join.hip), with the distance tile in GEMM
// form on MFMA.  A workgroup of W waves owns 16 W consecutive queries of one
// query cluster; each wave owns 16 of them (the 16 rows of a 16x16 MFMA
// block) and streams the surviving reference clusters 16 candidates
// (columns) at a time.  The W waves walk the union of their cluster lists in
// one order with a workgroup barrier per cluster, so the reference rows one
// wave fetches are still in L2 when the others read them: W x fewer HBM
// reads of reference rows than independent 16-query tiles.  Everything else
// (thresholds, k-select lists, merges) is private to a wave and ordered with
// wave-level syncs only.  For reference cluster b with representative c:
//     s(q, r) = ||q-c||^2 + ||r-c||^2 - 2 (q-c).(r-c)
// q-c is formed per (tile, cluster) into the A operand registers, with
// ||q-c||^2 in binary32; r-c and ||r-c||^2 are precomputed (k_center.hip).
// The dot runs as v_mfma_f32_16x16x32_bf16 (BF16 filter) or
// v_mfma_f32_16x16x4_f32 (plain FP32), FP32 accumulation.  Lane (row c16,
// group g) holds dimensions 32 s + 8 g + j, j < 8, of every 32-wide step s;
// for the FP32 instruction the k-slot g of sub-step j stands for dimension
// 32 s + 8 g + j, so both instructions read the same 16-byte row slices.
//
// Before any of that, the reference clusters listed for the tile are screened
// 16 at a time: one FP32 MFMA block gives ||q - c_b|| for all 16 x 16
// (query, cluster) pairs, and only clusters whose lower bound
// d(q, c_b) - r_b can reach some query's threshold are visited.
//
// Rigorous error bound (Higham, Accuracy and Stability of Numerical
// Algorithms, 2nd ed., section 3.1; u = 2^-24, u_b = 2^-8, gamma_n =
// n u / (1 - n u)): with a = ||q-c||, R >= ||r-c|| (the cluster radius), the
// computed s_f satisfies
//     |s_f - s| <= E = c1 * a * R + c2 * (a^2 + R^2),
//     c1 = 2 cdot + 4u,  c2 = gamma_{d+3} + 8u  (norm errors, the final two
//                                  roundings, the binary32 centring),
// where |computed dot - (q-c).(r-c)| <= cdot ||q-c|| ||r-c|| and
//     fp32 operands:          cdot = gamma_d
//     one bf16 product:       cdot = 2 u_b + u_b^2 + gamma_d (1 + u_b)^2
//     bf16 hi/lo, 3 products: cdot = 3.1 u_b^2 + 1.02 gamma_{3d}
//     fp8 (e4m3), rows scaled:  cdot = 2 dl + dl^2 + gamma_d (1 + dl)^2,
//                               dl = 2^-4 + sqrt(d) 2^-16
// (hi = bf16(x), lo = bf16(x - hi) leaves |x - hi - lo| <= u_b^2 |x|; the
// dropped lo*lo product and the two residuals are <= 3.1 u_b^2 |q_j r_j|;
// every bf16 x bf16 product is exact in FP32.  FP8: a row x scaled by 2^e
// into [64, 128) rounds to nearest with 3 fraction bits, |error| <= 2^-4 |x_j|
// for normal values and <= 2^-10 (subnormals of either e4m3 encoding) in
// scaled units, i.e. 2^-10 2^-e <= 2^-16 max|x| <= 2^-16 ||x||; so
// ||x~ - x|| <= dl ||x||, and every fp8 x fp8 product is exact in FP32), all
// scaled by 1.01 on the host.  Integer inputs with 4 d (2 max|x|)^2 <
// 2^24 make every FP32 step exact (E = 0, exact mode).
//
// List keys are (U = s_f + E, row), where E is the value of the entry's
// error code (the bound rounded up to the code's grid, so the rerank can
// recover s_f = U - E and report the filter's own error); without codes
// E is the bound itself.  theta_q is the k-th smallest U; a
// candidate passes when L = s_f - E <= theta_q.  E_max(q) is the largest E of
// any candidate appended for the query (an entry in the list or evicted from
// it was appended, so its L >= U - 2 E_max); at the end every list entry with
// U <= theta_q + 2 E_max is re-scored in binary64 (emit_query), and a list
// whose last entry is still within that cutoff reports an overflow.
#pragma once

#include <type_traits>

#include "join_common.h"

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4_t;

// lower bound of an evicted filter key (per-entry error code)
struct DropLower {
  static constexpr bool tracks = true;
  ECode ec;
  __device__ float operator()(u64 key) const { return key_lower(key, ec); }
};
template <class D>
__device__ __forceinline__ D make_drop(const ECode &ec) {
  if constexpr (D::tracks)
    return D{ec};
  else
    return D{};
}

#define MF_TQ 16
#define MF_BC 64  // per-query candidate buffer

// Element ii (wave-uniform) of a lane's four per-row values.  The values pass
// through an empty asm statement so the compiler cannot fold the selects into
// one indexed load, which would move the per-row arrays to scratch memory.
template <class T>
__device__ __forceinline__ T pick4(T a0, T a1, T a2, T a3, int ii) {
  asm volatile("" : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3));
  return ii == 0 ? a0 : ii == 1 ? a1 : ii == 2 ? a2 : a3;
}

// LDS of one wave (u64 words): lists, buffers, merge scratch, then floats:
// row norms [16], screening distances [16][16], query rows [16][DK + 4]
__host__ __device__ constexpr size_t mfma_wave_words(int kcap, int dk) {
  return (size_t)MF_TQ * kcap + (size_t)MF_TQ * MF_BC + (size_t)kcap + (16 + 256 + (size_t)MF_TQ * (dk + 4)) / 2;
}
// waves per workgroup: the most of 4, 2, 1 whose LDS fits the 160 KiB of a
// gfx950 workgroup (plus the union slots)
__host__ __device__ constexpr int mfma_waves(int kcap, int dk) {
  return 4 * mfma_wave_words(kcap, dk) * 8 + 1024 <= 163840   ? 4
         : 2 * mfma_wave_words(kcap, dk) * 8 + 1024 <= 163840 ? 2
                                                               : 1;
}

// B-operand registers of one 16-column block and 32-dimension step
template <int PREC>
struct BOp;
template <>
struct BOp<0> {
  f32x4_t f0, f1;
};
template <>
struct BOp<1> {
  bf16x8_t h;
};
template <>
struct BOp<3> {
  bf16x8_t h, l;
};
template <>
struct BOp<8> {
  long h;  // 8 e4m3 values
};

template <int KCAP, bool EXACT, int PREC, int DK, int W>
__device__ __forceinline__ void mfma_tile(const JoinArgs &a, const Tile tile) {
  // W waves per workgroup; with W > 1 the B groups are double-buffered (the
  // next group's loads are in flight while this one's MFMAs run)
  constexpr bool DB = W > 1;
  // PREC: 0 fp32 operands, 1 bf16, 3 bf16 hi/lo (three products), 8 fp8
  // (e4m3) rows scaled per row by a power of two
  constexpr bool BF16 = PREC == 1 || PREC == 3;
  constexpr bool SPLIT = PREC == 3;
  constexpr bool FP8 = PREC == 8;
  constexpr int NS = DK / 32;  // 32-dimension steps
  // 16-column blocks loaded together (B registers: NB * NS * 4 or 8 VGPRs)
  constexpr int NB = DB ? (DK <= 128 ? 2 : 1) : (DK <= 128 ? 4 : 2);
  typedef BOp<PREC> BT;
  extern __shared__ u64 smem[];
  const int wv = W == 1 ? 0 : (int)(threadIdx.x >> 6);
  // workgroup-shared slots after the W wave regions (1 KiB)
  u32 *uslot = (u32 *)(smem + W * mfma_wave_words(KCAP, DK));  // [2][W] union slots
  u64 *sthr = (u64 *)(uslot + 64);                              // [16] split mode: shared thresholds
  float *semax = (float *)(sthr + 16);                          // [W][16] split mode: per-wave E_max
  float *sevl = semax + W * 16;                                 // [W][16] split mode: evicted bounds
  u64 *lists = smem + wv * mfma_wave_words(KCAP, DK);         // [16][KCAP]
  u64 *bufs = lists + MF_TQ * KCAP;         // [16][MF_BC]
  u64 *tmp = bufs + MF_TQ * MF_BC;          // [KCAP]
  float *fsh = (float *)(tmp + KCAP);       // [16] row norms
  float *dqs = fsh + 16;                    // [16 rows][16 cols] screening distances
  // [16 rows][DK + 4] the tile's query rows; the 4-float pad spreads the 16
  // rows a 16-byte read touches over different LDS banks
  float *qsh = dqs + 256;
  const int lane = threadIdx.x & 63, g = lane >> 4, c16 = lane & 15;
  const u64 clk0 = a.tile_prof ? wall_clock64() : 0;
  u64 nvisit = 0, nseedb = 0, nsweepb = 0;  // diagnostics: visits, screened blocks
  u64 ngroups = 0, mticks = 0;              // diagnostics: B groups computed, ticks in sweep merges
  // Split mode (a.split > 1, W > 1): every wave serves the same (up to) 16
  // queries and takes every W-th block of 16 listed clusters, lowering its
  // thresholds from the others' published k-th keys (each is an upper bound
  // of the final one); at the end wave 0 merges the W lists and emits.  It
  // cuts a tile's latency W-fold where tiles are too few to fill the chip.
  const bool splitm = W > 1 && a.split > 1;
  // this wave's queries: permuted positions q0 .. q0 + nq - 1
  const int A = tile.cluster, q0 = tile.start + (splitm ? 0 : 16 * wv);
  const int nq = max(0, min(16, tile.count - (splitm ? 0 : 16 * wv)));
  // workgroup-wide OR of a wave-uniform value (double-buffered slots: a
  // slot is rewritten only after the next union's barrier, which every wave
  // reaches after reading it)
  int uphase = 0;
  auto wg_or = [&](u32 v) -> u32 {
    if constexpr (W == 1) {
      return v;
    } else {
      u32 *sl = uslot + uphase * W;
      uphase ^= 1;
      if (lane == 0) sl[wv] = v;
      __syncthreads();
      u32 r = 0;
#pragma unroll
      for (int w = 0; w < W; ++w) r |= sl[w];
      return r;
    }
  };
  // the operand planes are DK wide (zero beyond d): every 32-dimension step
  // runs, with no per-step branch between the MFMAs
  constexpr int dk = DK;
  const int d = a.d, k = a.k, dpQ = a.Q.dpad;
  const float eps = dist_relerr(d);

  for (int e = lane; e < MF_TQ * KCAP; e += 64) lists[e] = KEY_MAX;

  const double th = a.theta[A];
  const float T0 = (float)(th * th * (1.0 + 1e-6));
  if (splitm) {
    if (threadIdx.x < 16) sthr[threadIdx.x] = ((u64)__float_as_uint(T0) << 32) | 0xFFFFFFFFull;
    __syncthreads();  // workgroup-uniform: a.split and W are
  }
  // per-lane state of this lane's four C rows r = 4g + i
  u64 thr[4];
  int cnt[4];
  float emax[4], evl[4];
  bool live[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    thr[i] = ((u64)__float_as_uint(T0) << 32) | 0xFFFFFFFFull;
    cnt[i] = 0;
    emax[i] = 0.f;
    evl[i] = INFINITY;
    live[i] = (4 * g + i) < nq;
  }
  const ECode ec = a.ec;
  typedef typename std::conditional<EXACT, NoDrop, DropLower>::type DropL;
  // this lane's slice of query row c16: dimensions 32 s + 8 g + j, in FP32
  const bool alive = c16 < nq;
  // q_slice(s): this lane's 8 dimensions 32 s + 8 g + j of query row c16
  // (re-read from LDS when needed rather than held in registers)
  for (int e = lane; e < MF_TQ * DK; e += 64) {
    const int r = e / DK, dim = e % DK;
    qsh[r * (DK + 4) + dim] = (r < nq && dim < d) ? a.Q.px[(int64_t)(q0 + r) * dpQ + dim] : 0.f;
  }
  wave_sync();
  const float *qrow_sh = qsh + c16 * (DK + 4) + 8 * g;
  auto q_slice = [&](int s, float (&v)[8]) {
    const f32x4_t v0 = *(const f32x4_t *)(qrow_sh + 32 * s);
    const f32x4_t v1 = *(const f32x4_t *)(qrow_sh + 32 * s + 4);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      v[j] = v0[j];
      v[4 + j] = v1[j];
    }
  };
  float qn = 0.f;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    float v[8];
    q_slice(s, v);
#pragma unroll
    for (int j = 0; j < 8; ++j) qn = __builtin_fmaf(v[j], v[j], qn);
  }
  qn += xor_lanes_f(qn, 16);
  qn += xor_lanes_f(qn, 32);
  if (g == 0) fsh[c16] = qn;
  wave_sync();
  float qnr[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) qnr[i] = fsh[4 * g + i];
  wave_sync();
  const float gE = 1.01f * ((float)(d + 3) * 5.97e-8f + 8.f * 5.97e-8f);  // gamma_{d+3} + 8u, with margin
  u64 evals = 0, nmerge = 0, npass = 0;

  // merge of query r (wave-uniform r)
  // after list r changed: its k-th key bounds row r's threshold
  auto absorb = [&](int r, float dropped) {
    const int gi = r >> 2, ii = r & 3;
    const u64 kth = lists[r * KCAP + k - 1];
    u64 t = KEY_MAX;
    if (kth != KEY_MAX) t = EXACT ? kth - 1 : (((u64)__float_as_uint(key_dist(kth)) << 32) | 0xFFFFFFFFull);
    // every element written with a static index (no indexed store)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const bool me = g == gi && j == ii;
      cnt[j] = me ? 0 : cnt[j];
      thr[j] = (me && t < thr[j]) ? t : thr[j];
      evl[j] = me ? fminf(evl[j], dropped) : evl[j];
    }
    if (splitm && lane == 16 * gi && t != KEY_MAX) atomicMin(sthr + r, t);
  };
  auto merge = [&](int r) {
    const int gi = r >> 2, ii = r & 3;
    const int mine = pick4(cnt[0], cnt[1], cnt[2], cnt[3], ii);
    const int cr = __builtin_amdgcn_readlane(mine, 16 * gi);  // r is wave-uniform
    const float dropped = wave_merge<KCAP, MF_BC>(lists + r * KCAP, bufs + r * MF_BC, cr, tmp, make_drop<DropL>(ec));
    ++nmerge;
    absorb(r, dropped);
  };

  const int np = a.pcnt[A];
  const int32_t *pl = a.plist + (int64_t)A * a.R.m;
  // The list ascends by lb_b <= d(a, s) for every s of cluster b, so a row
  // with d(q, a) has d(q, s) >= lb_b - d(q, a) there: once lb_b exceeds every
  // live row's d(q, a) + search radius, no later cluster can hold a candidate.
  const float *pll = a.plb + (int64_t)A * a.R.m;
  float dqa[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) dqa[i] = live[i] ? a.Q.pdist[q0 + 4 * g + i] * (1.f + 2.f * eps) : 0.f;
  // wave-uniform: true when listed cluster pb and all after it are out of reach
  auto out_of_reach = [&](const int pb) {
    float reach = 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (live[i]) reach = fmaxf(reach, sqrtf(key_dist(thr[i])) * (1.f + eps) + dqa[i]);
    reach = wave_max(reach);
    return pll[pb] * (1.f - 2.f * eps) > reach;
  };
  // What the screening of one block needs from memory, for lane c16's
  // cluster: its representative's slice (dimensions 32 s + 8 g + j), squared
  // norm, member range and radius.  The seed search and the sweep load the
  // next block's while the current one computes (one list entry and one
  // representative row per lane: a dependent pair of loads whose latency a
  // single wave per SIMD cannot otherwise hide).
  struct Pre {
    f32x4_t c[2 * NS];
    float cn, Rb;
    int s0, s1;
  };
  auto load_pre = [&](const int bj, Pre &P) {
    const float *cr = a.reps32 + (int64_t)bj * dk + 8 * g;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      P.c[2 * s] = *(const f32x4_t *)(cr + 32 * s);
      P.c[2 * s + 1] = *(const f32x4_t *)(cr + 32 * s + 4);
    }
    P.cn = a.rn32[bj];
    P.s0 = a.R.off[bj];
    P.s1 = a.R.off[bj + 1];
    P.Rb = (float)(a.R.rad[bj] * (1.0 + 1e-6));
  };
  // list entry pj, clamped into the list (0 for an empty list: never used)
  auto list_at = [&](const int pj) { return np > 0 ? pl[min(pj, np - 1)] : 0; };
  // q.c for this lane's four rows and the representative of lane c16's
  // cluster (columns), on v_mfma_f32_16x16x4_f32; two accumulation chains
  // (even and odd k-steps) summed at the end
  auto center_dots = [&](const Pre &P) {
    f32x4_t acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      float qv[8];
      q_slice(s, qv);
#pragma unroll
      for (int j = 0; j < 4; j += 2) {
        acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(qv[j], P.c[2 * s][j], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(qv[j + 1], P.c[2 * s][j + 1], acc1, 0, 0, 0);
      }
#pragma unroll
      for (int j = 0; j < 4; j += 2) {
        acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(qv[4 + j], P.c[2 * s + 1][j], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(qv[5 + j], P.c[2 * s + 1][j + 1], acc1, 0, 0, 0);
      }
    }
    return f32x4_t{acc0[0] + acc1[0], acc0[1] + acc1[1], acc0[2] + acc1[2], acc0[3] + acc1[3]};
  };

  // Screening and visits of one block of up to 16 clusters: lane c16's
  // cluster bj, taking part if lv.  lock: the multi-wave lockstep applies.
  // cslot: the seed slot of lane c16's cluster (-1: not a seed).  The seed
  // block (seedmode) may serve any row; the sweep skips the seeds (served).
  auto process_block = [&](const int bj, const bool lv, const bool lock, const int cslot, const bool seedmode,
                           const Pre &P) {
      // ---- screening of 16 listed clusters: ||q - c||^2 = ||q||^2 + ||c||^2 - 2 q.c
      // on v_mfma_f32_16x16x4_f32 (rows = queries, columns = representatives),
      // error <= gE (||q|| + ||c||)^2 (any summation order of the dot is
      // covered); a cluster is kept if d(q,c) - r_b can be <= theta_q for some
      // query.  This lane's cluster metadata (P) is broadcast per visit.
      const int s0j = P.s0, s1j = P.s1;
      const float Rbj = P.Rb;
      {
        const f32x4_t acc = center_dots(P);
        const float cn = P.cn;
        const float rbj = Rbj;
        bool need = false;
  #pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float sqc = (qnr[i] + cn) - 2.f * acc[i];
          const float nsum = sqrtf(qnr[i]) + sqrtf(cn);
          const float dlo = sqrtf(fmaxf(sqc - gE * nsum * nsum, 0.f));
          const float thi = sqrtf(key_dist(thr[i])) * (1.f + eps);
          need |= lv && live[i] && (seedmode || cslot < 0) && !(dlo * (1.f - eps) - rbj > thi);
          dqs[(4 * g + i) * 16 + c16] = dlo;
        }
        need |= xor_lanes_i((int)need, 16) != 0;
        need |= xor_lanes_i((int)need, 32) != 0;
        wave_sync();
        // clusters any wave of the workgroup needs, visited in one order
        const u32 own = (u32)wave_ballot(need) & 0xFFFFu;
        u32 pending = (lock && a.lockstep && !splitm) ? wg_or(own) : own;
        bool first = true;
        for (; pending; pending &= pending - 1) {
          const int jl = __builtin_ctz(pending);
          // lockstep: every wave finishes the previous cluster before any
          // starts this one (no-op for W = 1; `first` is workgroup-uniform)
          if constexpr (W > 1) {
            if (!first && lock && a.lockstep && !splitm) __syncthreads();
          }
          first = false;
          // a cluster this wave visited as a seed is not visited again: the
          // union of a lockstep workgroup can hold it because another wave
          // needs it (a second visit would append the same keys twice)
          if (!seedmode && __builtin_amdgcn_readlane(cslot, jl) >= 0) continue;
          const int b = __builtin_amdgcn_readlane(bj, jl);
          const int s0 = __builtin_amdgcn_readlane(s0j, jl), s1 = __builtin_amdgcn_readlane(s1j, jl);
          const float Rb = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, Rbj), jl));
          // first 64 chunk-start distances of the cluster, requested now so the
          // load overlaps the A-operand build below
          const int nch = (s1 - s0 + 63) >> 6;
          const float cmin0 = lane < nch ? a.R.pdist[s0 + 64 * lane] : INFINITY;
          const float cnext0 = (lane == 63 && nch > 64) ? a.R.pdist[s0 + 64 * 64] : INFINITY;
          // re-check with the thresholds as they stand now
          bool use[4];
          bool anyl = false;
  #pragma unroll
          for (int i = 0; i < 4; ++i) {
            const float dlo = dqs[(4 * g + i) * 16 + jl];
            const float thi = sqrtf(key_dist(thr[i])) * (1.f + eps);
            use[i] = live[i] && !(dlo * (1.f - eps) - Rb > thi);
            anyl |= use[i];
          }
          if (wave_ballot(anyl) == 0) continue;
          ++nvisit;
          // A operand: q - c for row c16, dimensions of k-group g
          const float *rep = a.reps32 + (int64_t)b * dk + 8 * g;
          typename std::conditional<BF16, bf16x8_t, float[8]>::type av[NS];
          bf16x8_t avl[SPLIT ? NS : 1];
          long av8[FP8 ? NS : 1];
          float part = 0.f;
  #pragma unroll
          for (int s = 0; s < NS; ++s) {
            const f32x4_t c0 = *(const f32x4_t *)(rep + 32 * s);
            const f32x4_t c1 = *(const f32x4_t *)(rep + 32 * s + 4);
            float qv[8];
            q_slice(s, qv);
  #pragma unroll
            for (int j = 0; j < 8; ++j) {
              const float v = alive ? qv[j] - (j < 4 ? c0[j] : c1[j - 4]) : 0.f;
              part = __builtin_fmaf(v, v, part);
              if constexpr (BF16) {
                const __bf16 h = (__bf16)v;
                av[s][j] = h;
                if constexpr (SPLIT) avl[s][j] = (__bf16)(v - (float)h);
              } else {
                av[s][j] = v;
              }
            }
          }
          part += xor_lanes_f(part, 16);
          part += xor_lanes_f(part, 32);
          if (g == 0) fsh[c16] = part;
          wave_sync();
          float nqr[4];
  #pragma unroll
          for (int i = 0; i < 4; ++i) nqr[i] = fsh[4 * g + i];
          wave_sync();
          // FP8: row c16 of q - c scaled by 2^e so that its largest magnitude
          // lies in [64, 128), rounded to e4m3; iqr[i] = 2^-e of row 4g + i
          float iqr[4] = {1.f, 1.f, 1.f, 1.f};
          if constexpr (FP8) {
            float mx = 0.f;
  #pragma unroll
            for (int s2 = 0; s2 < NS; ++s2)
  #pragma unroll
              for (int j = 0; j < 8; ++j) mx = fmaxf(mx, fabsf(av[s2][j]));
            mx = fmaxf(mx, xor_lanes_f(mx, 16));
            mx = fmaxf(mx, xor_lanes_f(mx, 32));
            const int e = mx > 0.f ? 6 - ilogbf(mx) : 0;
  #pragma unroll
            for (int s2 = 0; s2 < NS; ++s2) {
              int lo = __builtin_amdgcn_cvt_pk_fp8_f32(ldexpf(av[s2][0], e), ldexpf(av[s2][1], e), 0, false);
              lo = __builtin_amdgcn_cvt_pk_fp8_f32(ldexpf(av[s2][2], e), ldexpf(av[s2][3], e), lo, true);
              int hi = __builtin_amdgcn_cvt_pk_fp8_f32(ldexpf(av[s2][4], e), ldexpf(av[s2][5], e), 0, false);
              hi = __builtin_amdgcn_cvt_pk_fp8_f32(ldexpf(av[s2][6], e), ldexpf(av[s2][7], e), hi, true);
              av8[s2] = (long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
            }
            const float iq = ldexpf(1.f, -e);
  #pragma unroll
            for (int i = 0; i < 4; ++i) iqr[i] = __shfl(iq, 4 * g + i);
          }

          // Ei: the rigorous bound; Ed: the bound the key carries (its code's
          // decoded value, >= Ei, with per-entry codes; Ei without)
          float Ei[4], Ed[4], dqbv[4];
          u32 ecd[4];
          float lo = INFINITY, hi = -INFINITY;
  #pragma unroll
          for (int i = 0; i < 4; ++i) {
            const float dqb = sqrtf(nqr[i]);
            dqbv[i] = dqb;
            const float thi = sqrtf(key_dist(thr[i])) * (1.f + eps);
            use[i] = use[i] && !(dqb * (1.f - 2.f * eps) - Rb > thi);
            Ei[i] = EXACT ? 0.f : a.c1 * dqb * (1.f + eps) * Rb + a.c2 * (nqr[i] + Rb * Rb);
            const u32 code = (!EXACT && a.ecode) ? ecode_encode(Ei[i], ec) : 0u;
            ecd[i] = code << ec.rowbits;
            Ed[i] = (!EXACT && a.ecode) ? ecode_decode(code, ec) : Ei[i];
            if (use[i]) {
              lo = fminf(lo, dqb * (1.f - 2.f * eps) - thi);
              hi = fmaxf(hi, dqb * (1.f + 2.f * eps) + thi);
            }
          }
          lo = wave_min(lo);
          hi = wave_max(hi);
          if (!(lo <= hi)) continue;
          lo = lo > 0.f ? lo * (1.f - eps) : -INFINITY;
          hi = hi * (1.f + eps);

          for (int cg = 0; cg < nch; cg += 64) {
            const int c = cg + lane;
            const float cmin = cg == 0 ? cmin0 : (c < nch ? a.R.pdist[s0 + 64 * c] : INFINITY);
            float nmin = __shfl_down(cmin, 1);
            if (lane == 63)
              nmin = cg == 0 ? cnext0 : ((c + 1 < nch) ? a.R.pdist[s0 + 64 * (c + 1)] : INFINITY);
            const bool needc = (c < nch) && (cmin <= hi) && (nmin >= lo);
            const u64 m = wave_ballot(needc);
            if (m == 0) {
              if (__builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, cmin), 0)) > hi) break;
              continue;
            }
            const int pbeg = s0 + 64 * (cg + __builtin_ctzll(m));
            const int pend = min(s1, s0 + 64 * (cg + 64 - __builtin_clzll(m)));
            // B operands, norms and rows of NB 16-column blocks, loaded
            // together so one memory latency covers NB blocks of MFMA work
            struct Group {
              BT bo[NB][NS];
              float nrv[NB];
              float isv[FP8 ? NB : 1];  // FP8: the rows' inverse scales
              u32 rowv[NB];
              bool validv[NB];
            };
            auto load_group = [&](int g0, Group &G) {
  #pragma unroll
              for (int nb = 0; nb < NB; ++nb) {
                const int p = g0 + 16 * nb + c16;
                G.validv[nb] = p < pend;
                const int pp = G.validv[nb] ? p : g0;
  #pragma unroll
                for (int s = 0; s < NS; ++s) {
                  if constexpr (SPLIT) {
                    const __bf16 *rr = (const __bf16 *)a.rc + (int64_t)pp * 2 * dk + 8 * g + 32 * s;
                    G.bo[nb][s].h = *(const bf16x8_t *)rr;
                    G.bo[nb][s].l = *(const bf16x8_t *)(rr + dk);
                  } else if constexpr (BF16) {
                    G.bo[nb][s].h = *(const bf16x8_t *)((const __bf16 *)a.rc + (int64_t)pp * dk + 8 * g + 32 * s);
                  } else if constexpr (FP8) {
                    G.bo[nb][s].h = *(const long *)((const unsigned char *)a.rc + (int64_t)pp * dk + 8 * g + 32 * s);
                  } else {
                    const float *rr = (const float *)a.rc + (int64_t)pp * dk + 8 * g + 32 * s;
                    G.bo[nb][s].f0 = *(const f32x4_t *)rr;
                    G.bo[nb][s].f1 = *(const f32x4_t *)(rr + 4);
                  }
                }
                G.nrv[nb] = a.nr[pp];
                if constexpr (FP8) G.isv[nb] = a.rinv[pp];
                G.rowv[nb] = EXACT ? (u32)a.R.perm[pp] : (u32)pp;
              }
            };
            auto compute_group = [&](int g0, Group &G) {
              ++ngroups;
              if constexpr (!EXACT) {
                // rows are sorted by distance to the representative, so the
                // group's last row bounds ||r - c|| for all of it: a tighter
                // R in the error bound than the cluster radius
                const float pd = a.R.pdist[min(g0 + 16 * NB, pend) - 1];
                const float Rg = fminf(Rb, pd * (1.f + 1e-6f));
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                  Ei[i] = a.c1 * dqbv[i] * (1.f + eps) * Rg + a.c2 * (nqr[i] + Rg * Rg);
                  const u32 code = a.ecode ? ecode_encode(Ei[i], ec) : 0u;
                  ecd[i] = code << ec.rowbits;
                  Ed[i] = a.ecode ? ecode_decode(code, ec) : Ei[i];
                }
              }
              // MFMA stage: the NB blocks are independent accumulation
              // chains, issued interleaved (blocks past the group's end hold
              // clamped in-bounds rows; their results are not used)
              f32x4_t accv[NB];
  #pragma unroll
              for (int nb = 0; nb < NB; ++nb) accv[nb] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  #pragma unroll
              for (int s = 0; s < NS; ++s) {
                if constexpr (SPLIT) {
  #pragma unroll
                  for (int nb = 0; nb < NB; ++nb)
                    accv[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(avl[s], G.bo[nb][s].h, accv[nb], 0, 0, 0);
  #pragma unroll
                  for (int nb = 0; nb < NB; ++nb)
                    accv[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[s], G.bo[nb][s].l, accv[nb], 0, 0, 0);
  #pragma unroll
                  for (int nb = 0; nb < NB; ++nb)
                    accv[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[s], G.bo[nb][s].h, accv[nb], 0, 0, 0);
                } else if constexpr (BF16) {
  #pragma unroll
                  for (int nb = 0; nb < NB; ++nb)
                    accv[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[s], G.bo[nb][s].h, accv[nb], 0, 0, 0);
                } else if constexpr (FP8) {
  #pragma unroll
                  for (int nb = 0; nb < NB; ++nb)
                    accv[nb] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(av8[s], G.bo[nb][s].h, accv[nb], 0, 0, 0);
                } else {
  #pragma unroll
                  for (int j = 0; j < 8; ++j)
  #pragma unroll
                    for (int nb = 0; nb < NB; ++nb)
                      accv[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                          av[s][j], j < 4 ? G.bo[nb][s].f0[j] : G.bo[nb][s].f1[j - 4], accv[nb], 0, 0, 0);
                }
              }
              // epilogue per block
  #pragma unroll
              for (int nb = 0; nb < NB; ++nb) {
                if (g0 + 16 * nb >= pend) break;
                const bool valid = G.validv[nb];
                const f32x4_t acc = accv[nb];
                const float nrb = G.nrv[nb];
                const u32 row = G.rowv[nb];
                bool pass[4];
                u64 key[4];
                bool anyp = false;
  #pragma unroll
                for (int i = 0; i < 4; ++i) {
                  // FP8: the dot of the scaled rows, times both inverse scales
                  // (powers of two: exact)
                  const float dotv = FP8 ? acc[i] * (iqr[i] * G.isv[FP8 ? nb : 0]) : acc[i];
                  const float sf = fmaxf((nqr[i] + nrb) - 2.f * dotv, 0.f);
                  if (EXACT) {
                    key[i] = make_key(sf, row);
                    pass[i] = valid && use[i] && key[i] <= thr[i];
                  } else {
                    key[i] = make_key(sf + Ed[i], row | ecd[i]);
                    pass[i] = valid && use[i] && (sf - Ei[i]) <= key_dist(thr[i]);
                  }
                  anyp |= pass[i];
                  evals += __popcll(wave_ballot(valid && use[i]));
                }
                if (wave_ballot(anyp) != 0) {
                  const u32 below = (1u << c16) - 1u;
  #pragma unroll
                  for (int i = 0; i < 4; ++i) {
                    const u64 mk = wave_ballot(pass[i]);
                    const u32 gm = (u32)(mk >> (16 * g)) & 0xFFFFu;
                    if (pass[i]) bufs[(4 * g + i) * MF_BC + cnt[i] + __popc(gm & below)] = key[i];
                    cnt[i] += __popc(gm);
                    npass += __popcll(mk);
                    // every entry ever appended for this row has E <= emax
                    if (gm) emax[i] = fmaxf(emax[i], Ed[i]);
                  }
                  // a 16-column block appends at most 16 keys per query, so a
                  // buffer below MF_BC - 16 has room for the next block
                  constexpr int WM = MF_BC - 16;
                  const bool hot = cnt[0] >= WM || cnt[1] >= WM || cnt[2] >= WM || cnt[3] >= WM;
                  if (wave_ballot(hot) != 0) {
                    const u64 cm0 = a.tile_prof ? wall_clock64() : 0;
                    wave_sync();
                    // rows 4g + i with a full buffer: lanes 16 g hold their counts
                    u32 rows = 0;
  #pragma unroll
                    for (int i = 0; i < 4; ++i) {
                      const u64 hb = wave_ballot(cnt[i] >= WM);
  #pragma unroll
                      for (int gg = 0; gg < 4; ++gg) rows |= (u32)((hb >> (16 * gg)) & 1u) << (4 * gg + i);
                    }
                    for (; rows; rows &= rows - 1) merge(__builtin_ctz(rows));
                    if (a.tile_prof) mticks += wall_clock64() - cm0;
                  }
                }
              }
            };
            constexpr int GS = 16 * NB;
            if constexpr (DB) {
              Group GA, GB;
              load_group(pbeg, GA);
              for (int g0 = pbeg; g0 < pend; g0 += 2 * GS) {
                const int g1 = g0 + GS;
                if (g1 < pend) load_group(g1, GB);
                compute_group(g0, GA);
                if (g1 >= pend) break;
                if (g1 + GS < pend) load_group(g1 + GS, GA);
                compute_group(g1, GB);
              }
            } else {
              Group G0;
              for (int g0 = pbeg; g0 < pend; g0 += GS) {
                load_group(g0, G0);
                compute_group(g0, G0);
              }
            }
          }
        }
      }
  
  };

  // ---- seed: each query's nearest listed cluster (by centre distance) is
  // visited first, so thresholds are tight before the sweep; a query cluster
  // mixing several dense regions would otherwise start from its loose
  // cluster bound and visit everything its sweep meets first.  The search
  // walks the whole list, or its first seed_blocks x 16 entries when
  // seed_blocks > 0 (measured: a capped search loses most of the seed's gain).
  int seedv = -1;  // lane c16: seed cluster of row c16, -1 if none or a duplicate
  u64 clk_search = 0;
  {
    float bestd[4];
    int bestj[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      bestd[i] = INFINITY;
      bestj[i] = -1;
    }
    const int np_seed = a.qseed ? 0 : a.seed_blocks > 0 ? min(np, 16 * a.seed_blocks) : np;
    if (a.qseed && np > 0 && nq > 0) {
      // one block: lane c16's column is the nearest reference representative
      // of row c16 (from the index build), and each row takes the nearest of
      // the tile's 16 -- a representative outside the list is still a valid
      // start (its members are real candidates and the bound below holds for
      // any cluster)
      ++nseedb;
      const bool lv = c16 < nq;
      const int bj = a.qseed[a.Q.perm[q0 + (lv ? c16 : 0)]];
      Pre cur;
      load_pre(bj, cur);
      const f32x4_t acc = center_dots(cur);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float sqc = (qnr[i] + cur.cn) - 2.f * acc[i];
        if (lv && sqc < bestd[i]) {
          bestd[i] = sqc;
          bestj[i] = bj;
        }
      }
    }
    if (np_seed > 0) {
      int bjc = list_at(c16), bjn = list_at(16 + c16);
      Pre cur;
      load_pre(bjc, cur);
      for (int pb = 0; pb < np_seed; pb += 16) {
        ++nseedb;
        // next block's loads in flight while this one computes
        Pre nxt;
        load_pre(bjn, nxt);
        const int bjnn = list_at(pb + 32 + c16);
        const bool lv = pb + c16 < np;
        const f32x4_t acc = center_dots(cur);
        const float cn = cur.cn;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float sqc = (qnr[i] + cn) - 2.f * acc[i];
          if (lv && sqc < bestd[i]) {
            bestd[i] = sqc;
            bestj[i] = bjc;
          }
        }
        bjc = bjn;
        bjn = bjnn;
        cur = nxt;
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) {
        const float od = xor_lanes_f(bestd[i], off);
        const int oj = xor_lanes_i(bestj[i], off);
        if (od < bestd[i] || (od == bestd[i] && oj < bestj[i])) {
          bestd[i] = od;
          bestj[i] = oj;
        }
      }
    // Each row's threshold from its nearest centre c_b: the k members of b
    // nearest to c_b lie within d(q, c_b) + pdist[off_b + k - 1] of q (rows
    // of a cluster ascend by pdist), so that bounds the k-th neighbour
    // distance -- usually far below the query cluster's theta_A, which cuts
    // the candidates the seed visit appends before its list holds k keys.
    // d(q, c_b) is taken from above through the screening's error bound.
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int b = bestj[i];
      if (live[i] && b >= 0) {
        const int o = a.R.off[b];
        if (a.R.off[b + 1] - o >= k) {
          const float pk = a.R.pdist[o + k - 1] * (1.f + 1e-6f);
          const float nsum = sqrtf(qnr[i]) + sqrtf(a.rn32[b]);
          const float dhi = sqrtf(fmaxf(bestd[i] + gE * nsum * nsum, 0.f)) * (1.f + eps);
          const float ub = (dhi + pk) * (1.f + eps);
          const u64 t = ((u64)__float_as_uint(ub * ub * (1.f + 2.f * eps)) << 32) | 0xFFFFFFFFull;
          if (t < thr[i]) thr[i] = t;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int v = __shfl(bestj[i], 16 * (c16 >> 2));
      if ((c16 & 3) == i) seedv = v;
    }
    if (c16 >= nq) seedv = -1;
    // one slot per distinct seed: its first row's lane
    int first = c16;
    for (int j = 0; j < 15; ++j) {
      const int vj = __builtin_amdgcn_readlane(seedv, j);
      if (j < first && vj == seedv) first = j;
    }
    if (first != c16) seedv = -1;
    // split mode: wave 0 visits the seeds and publishes its thresholds
    // (cluster 0 stands in for lanes without a seed: never visited)
    clk_search = a.tile_prof ? wall_clock64() : 0;
    if ((!splitm || wv == 0) && np > 0) {
      const int sb = seedv >= 0 ? seedv : 0;
      Pre sp;
      load_pre(sb, sp);
      process_block(sb, seedv >= 0, false, c16, true, sp);
      if (splitm) {
        wave_sync();
        for (int r = 0; r < MF_TQ; ++r) {
          const int gi = r >> 2, ii = r & 3;
          const int mine = pick4(cnt[0], cnt[1], cnt[2], cnt[3], ii);
          if (__builtin_amdgcn_readlane(mine, 16 * gi) > 0) merge(r);
        }
      }
    }
    if (splitm) __syncthreads();
  }
  const u64 clk_seed = a.tile_prof ? wall_clock64() : 0;
  // seed slot of cluster bj, -1 if it is no row's seed
  auto seed_slot = [&](const int bj) {
    int cs = -1;
    for (int j = 0; j < 16; ++j)
      if (__builtin_amdgcn_readlane(seedv, j) == bj) cs = j;
    return cs;
  };

  {
    const int pstep = splitm ? 16 * W : 16;
    const int pb0 = splitm ? 16 * wv : 0;
    int bjc = list_at(pb0 + c16), bjn = list_at(pb0 + pstep + c16);
    Pre cur;
    if (pb0 < np) load_pre(bjc, cur);
    for (int pb = pb0; pb < np; pb += pstep) {
      if (splitm) {
        // thresholds the other waves have reached
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const u64 t = __atomic_load_n(sthr + 4 * g + i, __ATOMIC_RELAXED);
          if (t < thr[i]) thr[i] = t;
        }
      }
      // the waves of a lockstep workgroup leave together (the visits hold
      // workgroup barriers); split-mode and single waves leave on their own
      if constexpr (W > 1) {
        if (a.lockstep && !splitm) {
          if (wg_or(out_of_reach(pb) ? 0u : 1u) == 0u) break;
        } else if (out_of_reach(pb)) {
          break;
        }
      } else {
        if (out_of_reach(pb)) break;
      }
      ++nsweepb;
      // next block's loads in flight while this one is screened and visited
      Pre nxt;
      if (pb + pstep < np) load_pre(bjn, nxt);
      const int bjnn = list_at(pb + 2 * pstep + c16);
      process_block(bjc, pb + c16 < np, true, seed_slot(bjc), false, cur);
      bjc = bjn;
      bjn = bjnn;
      cur = nxt;
    }
  }
  const u64 clk_sweep = a.tile_prof ? wall_clock64() : 0;
  wave_sync();
  for (int r = 0; r < MF_TQ; ++r) {
    const int gi = r >> 2, ii = r & 3;
    const int mine = pick4(cnt[0], cnt[1], cnt[2], cnt[3], ii);
    if (__builtin_amdgcn_readlane(mine, 16 * gi) > 0) merge(r);
  }
  wave_sync();
  if (splitm) {
    // wave 0 takes over the other waves' lists, E_max and evicted bounds
    if (c16 == 0) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        semax[wv * 16 + 4 * g + i] = emax[i];
        sevl[wv * 16 + 4 * g + i] = evl[i];
      }
    }
    __syncthreads();
    if (wv == 0) {
      for (int w = 1; w < W; ++w) {
        u64 *lw = smem + w * mfma_wave_words(KCAP, DK);
        for (int r = 0; r < nq; ++r) {
          const float dropped =
              wave_merge<KCAP, KCAP>(lists + r * KCAP, lw + r * KCAP, KCAP, tmp, make_drop<DropL>(ec));
          absorb(r, dropped);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          emax[i] = fmaxf(emax[i], semax[w * 16 + 4 * g + i]);
          evl[i] = fminf(evl[i], sevl[w * 16 + 4 * g + i]);
        }
      }
      wave_sync();
    }
  }
  // in split mode only wave 0 holds the merged lists
  const int nemit = (splitm && wv != 0) ? 0 : nq;
  const u64 clk_merge = a.tile_prof ? wall_clock64() : 0;

  // row r's final threshold, error maximum and smallest evicted lower bound
  // (wave-uniform r), and the cutoff of its rerank
  auto row_final = [&](const int r, float &Tr, float &Vr, float &cutoff) {
    const int gi = r >> 2, ii = r & 3;
    const u64 t = pick4(thr[0], thr[1], thr[2], thr[3], ii);
    const float em = pick4(emax[0], emax[1], emax[2], emax[3], ii);
    const float ev = pick4(evl[0], evl[1], evl[2], evl[3], ii);
    Tr = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, key_dist(t)), 16 * gi));
    const float Er = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, em), 16 * gi));
    Vr = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, ev), 16 * gi));
    cutoff = EXACT ? Tr : (Tr + 2.f * Er) * (1.f + 2.4e-7f);
  };
  // several rows at a time (their rerank loads in flight together): eight up
  // to capacity 64, four up to 128 (measured on the d = 128 mixture, k = 100:
  // join 10.16 ms with two, 9.92 ms with four), one at a time above
  constexpr int NQE = KCAP <= 64 ? 8 : KCAP <= 128 ? 4 : 1;
  int r = 0;
  for (; r + NQE <= nemit; r += NQE) {
    const u64 *Ls[NQE];
    int qp[NQE];
    float cu[NQE], th[NQE], ev[NQE];
    const float *qr[NQE];
#pragma unroll
    for (int n = 0; n < NQE; ++n) {
      row_final(r + n, th[n], ev[n], cu[n]);
      Ls[n] = lists + (r + n) * KCAP;
      qp[n] = q0 + r + n;
      qr[n] = a.Q.px + (int64_t)(q0 + r + n) * dpQ;
    }
    emit_queries<KCAP, EXACT, DK / 4, NQE>(a, Ls, qp, A, cu, qr, th, ev);
  }
  for (; r < nemit; ++r) {
    float Tr, Vr, cutoff;
    row_final(r, Tr, Vr, cutoff);
    emit_query<KCAP, EXACT, DK / 4>(a, lists + r * KCAP, q0 + r, A, cutoff, a.Q.px + (int64_t)(q0 + r) * dpQ, Tr,
                                    Vr);
  }
  if (a.tile_prof && wv == 0) {
    // diagnostic record: clocks at the phase boundaries, B groups and ticks
    // spent in merges during visits, visits and screened blocks
    if (threadIdx.x == 0) {
      u64 *tp = a.tile_prof + 8 * (size_t)blockIdx.x;
      tp[0] = clk0;
      tp[1] = wall_clock64();
      tp[4] = clk_seed;
      tp[5] = clk_sweep;
      tp[6] = clk_merge;
      tp[7] = clk_search;
      tp[2] = (ngroups << 32) | min(mticks, (u64)0xFFFFFFFFull);
      tp[3] = (min(nvisit, (u64)0xFFFF) << 48) | (min(nseedb, (u64)0xFFFF) << 32) | (min(nsweepb, (u64)0xFFFF) << 16) |
              (u64)min(np, 0xFFFF);
    }
  }
  if (lane == 0) {
    atomicAdd(a.n_evals, evals);
    atomicAdd(a.n_evals + 1, nmerge);
    atomicAdd(a.n_evals + 2, npass);
  }
}

template <int KCAP, bool EXACT, int PREC, int DK, int W>
__global__ __launch_bounds__(64 * W) void k_join_mfma(JoinArgs a) {
  if ((int)blockIdx.x >= *a.ntiles) return;
  const Tile tile = a.tiles[blockIdx.x];
  if (tile.count <= 0) return;
  mfma_tile<KCAP, EXACT, PREC, DK, W>(a, tile);
}

// Many joins in one launch: slot b holds a tile of instance tiles[b].pad,
// whose arguments are argv[tiles[b].pad].
template <int KCAP, bool EXACT, int PREC, int DK, int W>
__global__ __launch_bounds__(64 * W) void k_join_mfma_multi(const JoinArgs *__restrict__ argv,
                                                            const Tile *__restrict__ tiles, int nslots) {
  if ((int)blockIdx.x >= nslots) return;
  const Tile tile = tiles[blockIdx.x];
  if (tile.count <= 0) return;
  mfma_tile<KCAP, EXACT, PREC, DK, W>(argv[tile.pad], tile);
}

// Persistent form of k_join_mfma_multi for single-wave workgroups (see
// k_join_valu_queue): tile slots from the eight device counters queue[0..8).  Multi-wave
// workgroups are not run this way: their waves meet at barriers inside a
// tile, and a wave that ends a tile early must not go on to the next slot's.
template <int KCAP, bool EXACT, int PREC, int DK>
__global__ __launch_bounds__(64) void k_join_mfma_queue(const JoinArgs *__restrict__ argv,
                                                        const Tile *__restrict__ tiles, int nslots, int32_t *queue) {
  const int home = queue_home();
  int turn = 0;
  for (;;) {
    const int slot = queue_next(queue, nslots, home, turn);
    if (slot < 0) break;
    const Tile tile = tiles[slot];
    if (tile.count > 0) mfma_tile<KCAP, EXACT, PREC, DK, 1>(argv[tile.pad], tile);
  }
}

// waves per workgroup actually launched: 1, or the most that fit when the
// arguments ask for more than one
inline int mfma_launch_waves(int kcap, int dkmax, int want) { return want > 1 ? mfma_waves(kcap, dkmax) : 1; }

inline size_t mfma_lds(int kcap, int dkmax, int waves) {
  return waves * mfma_wave_words(kcap, dkmax) * sizeof(u64) + 1024;
}

// Launch of one operand width DK (a.dk == DK): the capacity, exact-mode and
// waves-per-workgroup instantiations.  Each width is instantiated in its own
// translation unit (k_join_mfma_d<DK>.hip); k_join_mfma.hip dispatches.
template <int PREC, int DK>
hipError_t launch_mfma_dk(const JoinArgs &a, int64_t grid, hipStream_t s, const JoinArgs *argv,
                          const Tile *tiles, int32_t *queue) {
  const int waves = mfma_launch_waves(a.kcap, DK, a.waves);
  const size_t lds = mfma_lds(a.kcap, DK, waves);
#define KJ_LAUNCH_W(KC, EX, WV)                                                                                \
  do {                                                                                                          \
    if (argv && queue && WV == 1) {                                                                             \
      const void *kq_ = (const void *)k_join_mfma_queue<KC, EX, PREC, DK>;                                      \
      const hipError_t e_ = hipFuncSetAttribute(kq_, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);     \
      if (e_ != hipSuccess) return e_;                                                                          \
      const int64_t g_ = queue_grid(kq_, 64, lds, grid);                                                        \
      hipLaunchKernelGGL((k_join_mfma_queue<KC, EX, PREC, DK>), dim3((unsigned)g_), dim3(64), lds, s, argv,     \
                         tiles, (int)grid, queue);                                                              \
    } else if (argv) {                                                                                          \
      const hipError_t e_ = hipFuncSetAttribute((const void *)k_join_mfma_multi<KC, EX, PREC, DK, WV>,          \
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);          \
      if (e_ != hipSuccess) return e_;                                                                          \
      hipLaunchKernelGGL((k_join_mfma_multi<KC, EX, PREC, DK, WV>), dim3((unsigned)grid), dim3(64 * WV), lds, s, \
                         argv, tiles, (int)grid);                                                               \
    } else {                                                                                                    \
      const hipError_t e_ = hipFuncSetAttribute((const void *)k_join_mfma<KC, EX, PREC, DK, WV>,                \
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);          \
      if (e_ != hipSuccess) return e_;                                                                          \
      hipLaunchKernelGGL((k_join_mfma<KC, EX, PREC, DK, WV>), dim3((unsigned)grid), dim3(64 * WV), lds, s, a);   \
    }                                                                                                           \
  } while (0)
#define KJ_LAUNCH(KC, EX)                                        \
  do {                                                           \
    constexpr int WM = mfma_waves(KC, DK);                       \
    if constexpr (WM > 1) {                                      \
      if (waves > 1) {                                           \
        KJ_LAUNCH_W(KC, EX, WM);                                 \
        break;                                                   \
      }                                                          \
    }                                                            \
    KJ_LAUNCH_W(KC, EX, 1);                                      \
  } while (0)
#define KJ_EX(KC)                                    \
  if (a.exact && PREC == 0) {                        \
    if constexpr (PREC == 0) { KJ_LAUNCH(KC, true); } \
  } else {                                           \
    KJ_LAUNCH(KC, false);                            \
  }
  switch (a.kcap) {
    case 64: KJ_EX(64) break;
    case 128: KJ_EX(128) break;
    case 256: KJ_EX(256) break;
    case 512: KJ_EX(512) break;
    default: return hipErrorInvalidValue;
  }
#undef KJ_EX
#undef KJ_LAUNCH
#undef KJ_LAUNCH_W
  return hipGetLastError();
}

// operand widths with a compiled kernel (dk = mfma_dk(d))
#define KJ_FOR_WIDTHS(X) X(32) X(64) X(96) X(128) X(256)
#define KJ_DK_INST(P, DKV)                                                                                 \
  template hipError_t launch_mfma_dk<P, DKV>(const JoinArgs &, int64_t, hipStream_t, const JoinArgs *, const Tile *, \
                                             int32_t *);
#define KJ_DK_EXTERN(P, DKV)                                                                                      \
  extern template hipError_t launch_mfma_dk<P, DKV>(const JoinArgs &, int64_t, hipStream_t, const JoinArgs *,     \
                                                    const Tile *, int32_t *);
