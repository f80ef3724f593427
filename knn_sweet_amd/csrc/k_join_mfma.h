// k_join_mfma.h -- the join proper on the matrix cores (d >= 32).
//
// Same contract as k_join_valu (k_join.hip), with the distance tile in GEMM
// form on MFMA.  One wave owns 16 queries of one query cluster (the 16 rows
// of a 16x16 MFMA block) and streams the surviving reference clusters 16
// candidates (columns) at a time.  For reference cluster b with
// representative c:
//     s(q, r) = ||q-c||^2 + ||r-c||^2 - 2 (q-c).(r-c)
// q-c is formed per (tile, cluster) into the A operand registers, with
// ||q-c||^2 in binary32; r-c and ||r-c||^2 are precomputed (k_center.hip).
// The dot runs as v_mfma_f32_16x16x32_bf16 (BF16 filter) or
// v_mfma_f32_16x16x4_f32 (plain FP32), FP32 accumulation.
//
// Rigorous error bound (Higham, Accuracy and Stability of Numerical
// Algorithms, 2nd ed., section 3.1; u = 2^-24, u_in = 2^-8 for bf16 operands
// and 0 for fp32, gamma_n = n u / (1 - n u)): with a = ||q-c||, R >= ||r-c||
// (the cluster radius), the computed s_f satisfies
//     |s_f - s| <= E = c1 * a * R + c2 * (a^2 + R^2),
//     c1 = 2 (2 u_in + u_in^2 + gamma_d (1 + u_in)^2) + 4u,
//     c2 = gamma_{d+3} + 8u      (norm errors, the final two roundings, the
//                                  binary32 centring q-c and r-c),
// both scaled by 1.01 on the host.  Integer inputs with 4 d (2 max|x|)^2 <
// 2^24 make every FP32 step exact (E = 0, exact mode).
//
// List keys are (U = s_f + E, row).  theta_q is the k-th smallest U; a
// candidate passes when L = s_f - E <= theta_q.  E_max(q) is the largest E the
// query used; at the end every list entry with U <= theta_q + 2 E_max is
// re-scored in binary64 (emit_query), and a list whose last entry is still
// within that cutoff reports an overflow.
#pragma once

#include <type_traits>

#include "join_common.h"

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4_t;

#define MF_TQ 16
#define MF_BC 64  // per-query candidate buffer
#define MF_WM 48  // merge watermark: a 16-column block adds at most 16

template <bool BF16, int DK>
struct AOps {
  static constexpr int KS = BF16 ? 32 : 16;  // dimensions per step
  static constexpr int NS = DK / KS;
  typedef typename std::conditional<BF16, bf16x8_t, f32x4_t>::type T;
  T v[NS];
};

template <int KCAP, bool EXACT, bool BF16, int DK>
__global__ __launch_bounds__(64) void k_join_mfma(JoinArgs a) {
  extern __shared__ u64 smem[];
  u64 *lists = smem;                        // [16][KCAP]
  u64 *bufs = lists + MF_TQ * KCAP;         // [16][MF_BC]
  u64 *tmp = bufs + MF_TQ * MF_BC;          // [KCAP]
  float *fsh = (float *)(tmp + KCAP);       // [16]
  const int lane = threadIdx.x, g = lane >> 4, c16 = lane & 15;
  if ((int)blockIdx.x >= *a.ntiles) return;
  const Tile tile = a.tiles[blockIdx.x];
  const int A = tile.cluster, nq = tile.count;
  const int d = a.d, k = a.k, dk = a.dk, dpQ = a.Q.dpad, dpR = a.R.dpad;
  const float eps = dist_relerr(d);
  typedef AOps<BF16, DK> Ops;
  const int nst = dk / Ops::KS;

  for (int e = lane; e < MF_TQ * KCAP; e += 64) lists[e] = KEY_MAX;
  __syncthreads();

  const double th = a.theta[A];
  const float T0 = (float)(th * th * (1.0 + 1e-6));
  // per-lane state of this lane's four C rows r = 4g + i
  u64 thr[4];
  int cnt[4];
  float emax[4];
  bool live[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    thr[i] = ((u64)__float_as_uint(T0) << 32) | 0xFFFFFFFFull;
    cnt[i] = 0;
    emax[i] = 0.f;
    live[i] = (4 * g + i) < nq;
  }
  const bool alive = c16 < nq;  // this lane's A-operand row
  const float *qa = a.Q.px + (int64_t)(tile.start + (alive ? c16 : 0)) * dpQ;
  u64 evals = 0, nmerge = 0, npass = 0;

  // merge of query r (wave-uniform r)
  auto merge = [&](int r) {
    const int gi = r >> 2, ii = r & 3;
    int mine = cnt[0];
#pragma unroll
    for (int j = 1; j < 4; ++j)
      if (j == ii) mine = cnt[j];
    const int cr = __shfl(mine, 16 * gi);
    wave_merge<KCAP, MF_BC>(lists + r * KCAP, bufs + r * MF_BC, cr, tmp);
    ++nmerge;
    const u64 kth = lists[r * KCAP + k - 1];
    u64 t = KEY_MAX;
    if (kth != KEY_MAX) t = EXACT ? kth - 1 : (((u64)__float_as_uint(key_dist(kth)) << 32) | 0xFFFFFFFFull);
    if (g == gi) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (j == ii) {
          cnt[j] = 0;
          if (t < thr[j]) thr[j] = t;
        }
    }
  };

  const int np = a.pcnt[A];
  const int32_t *pl = a.plist + (int64_t)A * a.R.m;
  for (int pi = 0; pi < np; ++pi) {
    const int b = pl[pi];
    const int s0 = a.R.off[b], s1 = a.R.off[b + 1];
    const float Rb = (float)(a.R.rad[b] * (1.0 + 1e-6));
    const float *rep = a.R.reps + (int64_t)b * dpR;

    // A operand: q - c for row c16, dimensions of k-group g
    Ops av;
    float part = 0.f;
#pragma unroll
    for (int s = 0; s < Ops::NS; ++s) {
      constexpr int W = BF16 ? 8 : 4;
#pragma unroll
      for (int j = 0; j < W; ++j) {
        const int dim = s * Ops::KS + W * g + j;
        const float v = (alive && s < nst && dim < d) ? qa[dim] - rep[dim] : 0.f;
        part = __builtin_fmaf(v, v, part);
        if constexpr (BF16)
          av.v[s][j] = (__bf16)v;
        else
          av.v[s][j] = v;
      }
    }
    part += __shfl_xor(part, 16);
    part += __shfl_xor(part, 32);
    if (g == 0) fsh[c16] = part;
    __syncthreads();
    float nqr[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) nqr[i] = fsh[4 * g + i];
    __syncthreads();

    bool use[4];
    float Ei[4];
    float lo = INFINITY, hi = -INFINITY;
    bool anyl = false;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float dqb = sqrtf(nqr[i]);
      const float thi = sqrtf(key_dist(thr[i])) * (1.f + eps);
      const float lb = dqb * (1.f - 2.f * eps) - Rb;
      use[i] = live[i] && !(lb > thi);
      Ei[i] = EXACT ? 0.f : a.c1 * dqb * (1.f + eps) * Rb + a.c2 * (nqr[i] + Rb * Rb);
      if (use[i]) {
        anyl = true;
        lo = fminf(lo, dqb * (1.f - 2.f * eps) - thi);
        hi = fmaxf(hi, dqb * (1.f + 2.f * eps) + thi);
        emax[i] = fmaxf(emax[i], Ei[i]);
      }
    }
    if (wave_ballot(anyl) == 0) continue;
    lo = wave_min(lo);
    hi = wave_max(hi);
    lo = lo > 0.f ? lo * (1.f - eps) : -INFINITY;
    hi = hi * (1.f + eps);

    const int nch = (s1 - s0 + 63) >> 6;
    for (int cg = 0; cg < nch; cg += 64) {
      const int c = cg + lane;
      const float cmin = c < nch ? a.R.pdist[s0 + 64 * c] : INFINITY;
      float nmin = __shfl_down(cmin, 1);
      if (lane == 63) nmin = (c + 1 < nch) ? a.R.pdist[s0 + 64 * (c + 1)] : INFINITY;
      const bool need = (c < nch) && (cmin <= hi) && (nmin >= lo);
      const u64 m = wave_ballot(need);
      if (m == 0) {
        if (__shfl(cmin, 0) > hi) break;
        continue;
      }
      const int pbeg = s0 + 64 * (cg + __builtin_ctzll(m));
      const int pend = min(s1, s0 + 64 * (cg + 64 - __builtin_clzll(m)));
      for (int p0 = pbeg; p0 < pend; p0 += 16) {
        const int p = p0 + c16;
        const bool valid = p < s1;
        const int pp = valid ? p : p0;
        f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
        if constexpr (BF16) {
          const __bf16 *rr = (const __bf16 *)a.rc + (int64_t)pp * dk + 8 * g;
#pragma unroll
          for (int s = 0; s < Ops::NS; ++s)
            if (s < nst) {
              const bf16x8_t bv = *(const bf16x8_t *)(rr + 32 * s);
              acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av.v[s], bv, acc, 0, 0, 0);
            }
        } else {
          const float *rr = (const float *)a.rc + (int64_t)pp * dk + 4 * g;
#pragma unroll
          for (int s = 0; s < Ops::NS; ++s)
            if (s < nst) {
              const f32x4_t bv = *(const f32x4_t *)(rr + 16 * s);
              acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.v[s][0], bv[0], acc, 0, 0, 0);
              acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.v[s][1], bv[1], acc, 0, 0, 0);
              acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.v[s][2], bv[2], acc, 0, 0, 0);
              acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.v[s][3], bv[3], acc, 0, 0, 0);
            }
        }
        const float nrv = a.nr[pp];
        const u32 row = EXACT ? (u32)a.R.perm[pp] : (u32)pp;
        bool pass[4];
        u64 key[4];
        bool anyp = false;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float sf = fmaxf((nqr[i] + nrv) - 2.f * acc[i], 0.f);
          if (EXACT) {
            key[i] = make_key(sf, row);
            pass[i] = valid && use[i] && key[i] <= thr[i];
          } else {
            key[i] = make_key(sf + Ei[i], row);
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
          }
          const bool hot = cnt[0] >= MF_WM || cnt[1] >= MF_WM || cnt[2] >= MF_WM || cnt[3] >= MF_WM;
          if (wave_ballot(hot) != 0) {
            __syncthreads();
            for (int r = 0; r < MF_TQ; ++r) {
              const int gi = r >> 2, ii = r & 3;
              int mine = cnt[0];
#pragma unroll
              for (int j = 1; j < 4; ++j)
                if (j == ii) mine = cnt[j];
              if (__shfl(mine, 16 * gi) >= MF_WM) merge(r);
            }
          }
        }
      }
    }
  }
  __syncthreads();
  for (int r = 0; r < MF_TQ; ++r) {
    const int gi = r >> 2, ii = r & 3;
    int mine = cnt[0];
#pragma unroll
    for (int j = 1; j < 4; ++j)
      if (j == ii) mine = cnt[j];
    if (__shfl(mine, 16 * gi) > 0) merge(r);
  }
  __syncthreads();

  for (int r = 0; r < nq; ++r) {
    const int gi = r >> 2, ii = r & 3;
    u64 t = thr[0];
    float em = emax[0];
#pragma unroll
    for (int j = 1; j < 4; ++j)
      if (j == ii) {
        t = thr[j];
        em = emax[j];
      }
    const float Tr = __shfl(key_dist(t), 16 * gi);
    const float Er = __shfl(em, 16 * gi);
    const float cutoff = EXACT ? Tr : (Tr + 2.f * Er) * (1.f + 2.4e-7f);
    emit_query<KCAP, EXACT>(a, lists + r * KCAP, tile.start + r, A, cutoff,
                            a.Q.px + (int64_t)(tile.start + r) * dpQ);
  }
  if (lane == 0) {
    atomicAdd(a.n_evals, evals);
    atomicAdd(a.n_evals + 1, nmerge);
    atomicAdd(a.n_evals + 2, npass);
  }
}

inline size_t mfma_lds(int kcap) {
  return sizeof(u64) * ((size_t)MF_TQ * kcap + (size_t)MF_TQ * MF_BC + (size_t)kcap) + 16 * sizeof(float);
}

template <bool BF16>
hipError_t launch_mfma_any(const JoinArgs &a, int64_t grid, hipStream_t s) {
  const size_t lds = mfma_lds(a.kcap);
  const bool big = a.dk > 128;
#define KJ_LAUNCH(KC, EX, DKV) \
  hipLaunchKernelGGL((k_join_mfma<KC, EX, BF16, DKV>), dim3((unsigned)grid), dim3(64), lds, s, a)
#define KJ_DK(KC, EX)     \
  if (big)                \
    KJ_LAUNCH(KC, EX, 256); \
  else                    \
    KJ_LAUNCH(KC, EX, 128);
#define KJ_EX(KC)                               \
  if (a.exact && !BF16) {                       \
    if constexpr (!BF16) { KJ_DK(KC, true) }    \
  } else {                                      \
    KJ_DK(KC, false)                            \
  }
  switch (a.kcap) {
    case 64: KJ_EX(64) break;
    case 128: KJ_EX(128) break;
    case 256: KJ_EX(256) break;
    case 512: KJ_EX(512) break;
    case 1024: KJ_EX(1024) break;
    default: return hipErrorInvalidValue;
  }
#undef KJ_EX
#undef KJ_DK
#undef KJ_LAUNCH
  return hipGetLastError();
}
