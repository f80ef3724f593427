// join_common.h -- result emission shared by the join kernels.
#pragma once

#include "knnj_common.h"
#include "knnj_kernels.h"
#include "wave_sort.h"

// Writes row `qo` of the answer from a query's sorted candidate list L.
//
// Exact mode: the keys are (exact binary32 squared distance, original row);
// the first k are the answer.
// Filter mode: the keys are (upper bound of the squared distance, permuted
// reference position).  Entries that may belong to the answer are re-scored
// with the oracle's binary64 sequence and sorted by (binary64 squared
// distance, original row):
//  - with per-entry error codes (a.ecode): entries whose lower bound is
//    <= theta; an overflow is an evicted entry whose lower bound (evl) is
//    <= theta;
//  - otherwise: entries with key distance <= cutoff; an overflow is a last
//    list entry within the cutoff.
// An overflowed query is appended to the overflow list and nothing is written.
// NV > 0: rows have at most NV 16-byte steps (compile time), so the rerank
// is straight-line code whose row loads all issue before the arithmetic.
template <int KCAP, bool EXACT, int NV = 0>
__device__ __forceinline__ void emit_query(const JoinArgs &a, const u64 *L, int qpos, int cluster,
                                           float cutoff, const float *qrow, float theta = 0.f,
                                           float evl = INFINITY) {
  const int lane = threadIdx.x & 63;
  const int k = a.k;
  const int64_t qo = a.Q.perm[qpos];
  if (EXACT) {
    for (int j = lane; j < k; j += 64) {
      const u64 key = L[j];
      const float s = key_dist(key);
      const int32_t row = (int32_t)key_row(key);
      const float dist = a.out_mode == 1 ? s : (float)sqrt((double)s);
      if (a.out_mode == 2) {
        a.out_idx[(qo * k + j) * 2] = row;
        a.out_idx[(qo * k + j) * 2 + 1] = __float_as_int(dist);
      } else {
        a.out_idx[qo * k + j] = row;
        a.out_dist[qo * k + j] = dist;
      }
    }
    return;
  }
  const u64 last = L[KCAP - 1];
  const float thr_l = theta * (1.f + 0x1p-20f);
  const bool ovf = a.ecode ? (evl <= thr_l) : (last != KEY_MAX && key_dist(last) <= cutoff);
  if (ovf) {
    if (lane == 0) {
      const int slot = atomicAdd(a.ovf_count, 1);
      Tile t;
      t.start = qpos;
      t.count = 1;
      t.cluster = cluster;
      t.pad = 0;
      a.ovf_tiles[slot] = t;
      if (a.ovf_diag && slot < 64) {
        // diagnostic record: cutoff, k-th key, last key, first key
        float *dg = a.ovf_diag + 4 * slot;
        dg[0] = cutoff;
        dg[1] = key_dist(L[k - 1]);
        dg[2] = key_dist(last);
        dg[3] = key_dist(L[0]);
      }
    }
    return;
  }
  constexpr int E = KCAP / 64;
  u64 hs[E];
  u32 rs[E];
  float err = 0.f;
  const int d = a.d, dpR = a.R.dpad;
  // The E entries of this lane are re-scored together: 16-byte row loads
  // (rows are dpad = roundup(d, 4) floats, 16-byte aligned) and E independent
  // binary64 chains, each in the oracle's order (sq64_pair: t_j = (q_j -
  // r_j)^2, s = s + t_j for j = 0 .. d-1, no contraction).  A lane without an
  // entry reads the query row itself (one broadcast line) and discards it.
  typedef __attribute__((__vector_size__(4 * sizeof(float)))) float ef32x4;
  bool okv[E];
  int pv[E];
  const ef32x4 *r4[E];
  double s64v[E];
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const u64 key = L[lane + 64 * u];
    okv[u] = key != KEY_MAX && (a.ecode ? key_lower(key, a.ec) <= thr_l : key_dist(key) <= cutoff);
    pv[u] = okv[u] ? (int)(a.ecode ? key_row_of(key, a.ec) : key_row(key)) : 0;
    r4[u] = okv[u] ? (const ef32x4 *)(a.R.px + (int64_t)pv[u] * dpR) : (const ef32x4 *)qrow;
    s64v[u] = 0.0;
  }
  {
#pragma clang fp contract(off)
    const ef32x4 *q4 = (const ef32x4 *)qrow;
    const int nv = (d + 3) >> 2;
    auto step = [&](const ef32x4 qv, const ef32x4 (&rv)[E], int t) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (4 * t + c < d) {
#pragma unroll
          for (int u = 0; u < E; ++u) {
            const double x = (double)qv[c] - (double)rv[u][c];
            s64v[u] = s64v[u] + x * x;
          }
        }
      }
    };
    if constexpr (NV > 0) {
#pragma unroll
      for (int t = 0; t < NV; ++t) {
        if (t < nv) {
          ef32x4 rv[E];
#pragma unroll
          for (int u = 0; u < E; ++u) rv[u] = r4[u][t];
          step(q4[t], rv, t);
        }
      }
    } else {
#pragma unroll 4
      for (int t = 0; t < nv; ++t) {
        ef32x4 rv[E];
#pragma unroll
        for (int u = 0; u < E; ++u) rv[u] = r4[u][t];
        step(q4[t], rv, t);
      }
    }
  }
#pragma unroll
  for (int u = 0; u < E; ++u) {
    hs[u] = KEY_MAX;
    rs[u] = 0xFFFFFFFFu;
    if (okv[u]) {
      const u64 key = L[lane + 64 * u];
      hs[u] = (u64)__double_as_longlong(s64v[u]);
      rs[u] = (u32)a.R.perm[pv[u]];
      err = fmaxf(err, (float)fabs((double)key_dist(key) - s64v[u]));
    }
  }
  wave_sort_pair<E>(hs, rs);
  err = wave_max(err);
  if (lane == 0 && err > 0.f) atomicMax(a.max_err_bits, __float_as_uint(err));
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int j = lane + 64 * u;
    if (j < k) {
      const double s64 = __longlong_as_double((long long)hs[u]);
      const float dist = a.out_mode == 1 ? (float)s64 : (float)sqrt(s64);
      if (a.out_mode == 2) {
        a.out_idx[(qo * k + j) * 2] = (int32_t)rs[u];
        a.out_idx[(qo * k + j) * 2 + 1] = __float_as_int(dist);
      } else {
        a.out_idx[qo * k + j] = (int32_t)rs[u];
        a.out_dist[qo * k + j] = dist;
      }
    }
  }
}
