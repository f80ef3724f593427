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
// emit_queries writes NQ queries' rows at once: their re-scoring loads are in
// flight together (each entry's arithmetic is the same as alone).
// Diagnostic (a.ferr_bits set): the filter's own error on every entry of a
// query's list that the rerank re-scores, recomputed from scratch after the
// answer is written so the rerank itself is unchanged.  A coded key is
// U = fl(s_f + E) with E the decoded bound, so s_f = U - E up to U's
// rounding (half an ulp, taken off the error so that a ratio above 1 proves
// a violated bound); a direct-form key is s_f itself, |s_f - s| <= delta s.
template <int KCAP>
__device__ __forceinline__ void filter_stats_query(const JoinArgs &a, const u64 *L, const float *qrow, float cutoff,
                                                float thr_l) {
  const int lane = threadIdx.x & 63;
  float ferr = 0.f, fbnd = 0.f, frat = 0.f;
  for (int j = lane; j < KCAP; j += 64) {
    const u64 key = L[j];
    if (key == KEY_MAX) continue;
    if (!(a.ecode ? key_lower(key, a.ec) <= thr_l : key_dist(key) <= cutoff)) continue;
    const int pos = (int)(a.ecode ? key_row_of(key, a.ec) : key_row(key));
    const double s64 = sq64_pair(qrow, a.R.px + (int64_t)pos * a.R.dpad, a.d);
    const double U = (double)key_dist(key);
    double sf = U, eb = 0.0, slack = 0.0;
    if (a.ecode) {
      eb = (double)ecode_decode((u32)key >> a.ec.rowbits, a.ec);
      sf = U - eb;
      slack = U * 0x1p-24;
    } else if (a.relbound) {
      eb = a.delta * s64;
    }
    if (!(eb < INFINITY)) continue;
    const double fe = fmax(fabs(sf - s64) - slack, 0.0);
    ferr = fmaxf(ferr, (float)fe);
    fbnd = fmaxf(fbnd, (float)eb);
    if (eb > 0.0) frat = fmaxf(frat, (float)(fe / eb));
  }
  ferr = wave_max(ferr);
  fbnd = wave_max(fbnd);
  frat = wave_max(frat);
  if (lane == 0) {
    if (ferr > 0.f) atomicMax(a.ferr_bits, __float_as_uint(ferr));
    if (fbnd > 0.f) atomicMax(a.ferr_bits + 1, __float_as_uint(fbnd));
    if (frat > 0.f) atomicMax(a.ferr_bits + 2, __float_as_uint(frat));
  }
}

template <int KCAP, bool EXACT, int NV, int NQ>
__device__ __forceinline__ void emit_queries(const JoinArgs &a, const u64 *const (&L)[NQ], const int (&qpos)[NQ],
                                             int cluster, const float (&cutoff)[NQ], const float *const (&qrow)[NQ],
                                             const float (&theta)[NQ], const float (&evl)[NQ]) {
  const int lane = threadIdx.x & 63;
  const int k = a.k;
  if (EXACT) {
#pragma unroll
    for (int n = 0; n < NQ; ++n) {
      const int64_t qo = a.Q.perm[qpos[n]];
      for (int j = lane; j < k; j += 64) {
        const u64 key = L[n][j];
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
    }
    return;
  }
  bool ovf[NQ];
  float thr_l[NQ];
#pragma unroll
  for (int n = 0; n < NQ; ++n) {
    const u64 last = L[n][KCAP - 1];
    thr_l[n] = theta[n] * (1.f + 0x1p-20f);
    ovf[n] = a.ecode ? (evl[n] <= thr_l[n]) : (last != KEY_MAX && key_dist(last) <= cutoff[n]);
    if (ovf[n] && lane == 0) {
      const int slot = atomicAdd(a.ovf_count, 1);
      Tile t;
      t.start = qpos[n];
      t.count = 1;
      t.cluster = cluster;
      t.pad = 0;
      a.ovf_tiles[slot] = t;
      if (a.ovf_diag && slot < 64) {
        // diagnostic record: cutoff, k-th key, last key, first key
        float *dg = a.ovf_diag + 4 * slot;
        dg[0] = cutoff[n];
        dg[1] = key_dist(L[n][k - 1]);
        dg[2] = key_dist(last);
        dg[3] = key_dist(L[n][0]);
      }
    }
  }
  constexpr int E = KCAP / 64;
  const int d = a.d, dpR = a.R.dpad;
  // The E entries of each query on this lane are re-scored together: 16-byte
  // row loads (rows are dpad = roundup(d, 4) floats, 16-byte aligned) and one
  // independent binary64 chain per entry, each in the oracle's order
  // (sq64_pair: t_j = (q_j - r_j)^2, s = s + t_j for j = 0 .. d-1, no
  // contraction).  A lane without an entry reads its query row (one
  // broadcast line) and discards it; an overflowed query has no entries.
  typedef __attribute__((__vector_size__(4 * sizeof(float)))) float ef32x4;
  bool okv[NQ][E];
  int pv[NQ][E];
  const ef32x4 *r4[NQ][E];
  double s64v[NQ][E];
#pragma unroll
  for (int n = 0; n < NQ; ++n)
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const u64 key = L[n][lane + 64 * u];
      okv[n][u] = !ovf[n] && key != KEY_MAX &&
                  (a.ecode ? key_lower(key, a.ec) <= thr_l[n] : key_dist(key) <= cutoff[n]);
      pv[n][u] = okv[n][u] ? (int)(a.ecode ? key_row_of(key, a.ec) : key_row(key)) : 0;
      r4[n][u] = okv[n][u] ? (const ef32x4 *)(a.R.px + (int64_t)pv[n][u] * dpR) : (const ef32x4 *)qrow[n];
      s64v[n][u] = 0.0;
    }
  {
#pragma clang fp contract(off)
    const int nv = (d + 3) >> 2;
    auto step = [&](const ef32x4 (&qv)[NQ], const ef32x4 (&rv)[NQ][E], int t) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (4 * t + c < d) {
#pragma unroll
          for (int n = 0; n < NQ; ++n)
#pragma unroll
            for (int u = 0; u < E; ++u) {
              const double x = (double)qv[n][c] - (double)rv[n][u][c];
              s64v[n][u] = s64v[n][u] + x * x;
            }
        }
      }
    };
    auto load_step = [&](int t, ef32x4 (&qv)[NQ], ef32x4 (&rv)[NQ][E]) {
#pragma unroll
      for (int n = 0; n < NQ; ++n) {
        qv[n] = ((const ef32x4 *)qrow[n])[t];
#pragma unroll
        for (int u = 0; u < E; ++u) rv[n][u] = r4[n][u][t];
      }
    };
    if constexpr (NV > 0) {
#pragma unroll
      for (int t = 0; t < NV; ++t) {
        if (t < nv) {
          ef32x4 qv[NQ], rv[NQ][E];
          load_step(t, qv, rv);
          step(qv, rv, t);
        }
      }
    } else {
#pragma unroll 4
      for (int t = 0; t < nv; ++t) {
        ef32x4 qv[NQ], rv[NQ][E];
        load_step(t, qv, rv);
        step(qv, rv, t);
      }
    }
  }
#pragma unroll
  for (int n = 0; n < NQ; ++n) {
    if (ovf[n]) continue;  // wave-uniform
    const int64_t qo = a.Q.perm[qpos[n]];
    u64 hs[E];
    u32 rs[E];
    float err = 0.f;
#pragma unroll
    for (int u = 0; u < E; ++u) {
      hs[u] = KEY_MAX;
      rs[u] = 0xFFFFFFFFu;
      if (okv[n][u]) {
        const u64 key = L[n][lane + 64 * u];
        hs[u] = (u64)__double_as_longlong(s64v[n][u]);
        rs[u] = (u32)a.R.perm[pv[n][u]];
        err = fmaxf(err, (float)fabs((double)key_dist(key) - s64v[n][u]));
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
    if (a.ferr_bits) filter_stats_query<KCAP>(a, L[n], qrow[n], cutoff[n], thr_l[n]);
  }
}

// One query (see emit_queries).
template <int KCAP, bool EXACT, int NV = 0>
__device__ __forceinline__ void emit_query(const JoinArgs &a, const u64 *L, int qpos, int cluster,
                                           float cutoff, const float *qrow, float theta = 0.f,
                                           float evl = INFINITY) {
  const u64 *const Ls[1] = {L};
  const int qp[1] = {qpos};
  const float cu[1] = {cutoff}, th[1] = {theta}, ev[1] = {evl};
  const float *const qr[1] = {qrow};
  emit_queries<KCAP, EXACT, NV, 1>(a, Ls, qp, cluster, cu, qr, th, ev);
}
