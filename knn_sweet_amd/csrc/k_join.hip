// k_join.hip -- the join proper: VALU distance tiles, per-query wave k-select,
// exact rerank.
//
// Replaces KNNQuery_base (reference src/knnjoin.cu:762-858) and the
// small-batch KNNQuery_theta / KNNQuery / final path (knnjoin.cu:859-979).
// The reference runs one thread per query over a random gather of R rows,
// with an insertion-sorted K-list in per-thread scratch (IndexDist
// knearest[1000]).  Here one wave64 owns a tile of TQ queries of one query
// cluster and streams the surviving reference clusters (contiguous,
// permuted rows) 64 candidates at a time, one candidate per lane:
//   - per (tile, cluster): distance of each query to the representative,
//     the cluster lower bound d(q,b) - r_b > theta_q skips the query, and the
//     point-level bound |d(q,b) - d(s,b)| <= theta_q (the reference's
//     break/continue, knnjoin.cu:796-800) is applied per 64-row chunk
//     through the ascending member distances;
//   - candidate keys (squared distance, row) below the query's threshold are
//     appended to a per-query LDS buffer (ballot + mbcnt, no atomics); a full
//     buffer is bitonic-sorted and rank-merged into the query's sorted
//     KCAP-key list, and the threshold drops to the list's k-th key.
// Integer inputs (exact mode): binary32 squared distances are exact, the key
// row is the original row.  There the buffer is an unsorted pool of
// pool_cap(KCAP) keys: when it fills, a select by counting (the k-th
// smallest distance found bit by bit, one ballot per register) drops every
// key that already has k smaller ones and lowers the threshold, with no
// sort; the pool is sorted once, at the end, and its first k keys are the
// answer.  (Sorted merges every 64 appends took 60 % of a tile at d = 3,
// k = 200.)  Real-valued inputs: the
// binary32 key is a filter with relative error <= delta; every candidate
// within fac = (1+delta)/(1-delta) of the k-th filter key is kept, then
// re-scored with the oracle's binary64 sequence and sorted by (binary64
// squared distance, original row).  A list whose last key is still within
// that window may have dropped a needed candidate: the query is reported as
// overflowed and re-run at a larger capacity by the host.
#include <cstdio>
#include <cstdlib>

#include "knnj_common.h"
#include "knnj_kernels.h"
#include "wave_sort.h"
#include "join_common.h"

// Waves per workgroup, each on its own tile.  One: tile lengths vary
// several-fold, and a workgroup of several waves holds its slot until its
// slowest wave ends (measured at d = 3, k = 200: 4 waves per workgroup ran
// 30 % fewer waves at once than single waves).  The kernel's workgroup
// barriers order one wave's LDS accesses and rely on this being 1.
#define VALU_WPB 1

// Exact mode keeps each query's candidates in an unsorted pool of
// pool_cap(KCAP) keys instead of a sorted list: at least 64 keys above the
// largest k it serves plus one chunk of 64 appends (KCAP >= max(k, 64)).
__host__ __device__ constexpr int pool_cap(int kcap) { return 2 * kcap > kcap + 256 ? 2 * kcap : kcap + 256; }

// LDS words (u64) of one wave
__host__ __device__ constexpr size_t valu_wave_words(int tq, int kcap, bool exact) {
  return exact ? (size_t)tq * pool_cap(kcap) : (size_t)tq * kcap + (size_t)tq * KNNJ_BCAP + (size_t)kcap;
}

// number of this wave's values dv[0..EP) (all lanes) that are <= t
template <int EP>
__device__ __forceinline__ int wave_count_le(const u32 (&dv)[EP], u32 t) {
  int c = 0;
#pragma unroll
  for (int u = 0; u < EP; ++u) c += __popcll(wave_ballot(dv[u] <= t));
  return c;
}

__device__ __forceinline__ u32 wave_max_u32(u32 v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = max(v, xor_lanes(v, m));
  return v;
}

// Exact-mode pool shrink.  pool[0..cnt) holds cnt >= k distinct keys whose
// high words are exact integer squared distances (binary32).  D*, the k-th
// smallest distance, is found bit by bit from the top (a count of the
// pool's distances <= t per bit: one ballot per register).  Every key of
// distance > D* has at least k smaller keys, so it can never be among the k
// nearest; the pool keeps the keys of distance <= D* and the pass threshold
// becomes (D*, any row).  When those are more than keep_max keys (ties at
// D*), the (k - #{distance < D*})-th smallest row among the ties is found the
// same way and exactly the k smallest keys stay, the threshold then being
// just below the k-th key.  Returns the threshold (a key passes when <= it)
// and sets cnt.  rowbits: width of the row ids.
template <int P>
__device__ __forceinline__ u64 pool_shrink(u64 *pool, int &cnt, int k, int rowbits, int keep_max) {
  constexpr int EP = P / 64;
  const int lane = threadIdx.x & 63;
  u64 v[EP];
  u32 dv[EP];
  u32 dmax = 0;
#pragma unroll
  for (int u = 0; u < EP; ++u) {
    const int e = lane + 64 * u;
    v[u] = e < cnt ? pool[e] : KEY_MAX;
    dv[u] = e < cnt ? (u32)key_dist(v[u]) : 0xFFFFFFFFu;
    if (e < cnt) dmax = max(dmax, dv[u]);
  }
  dmax = wave_max_u32(dmax);
  u32 r = 0;
  for (int b = 31 - __builtin_clz(dmax | 1u); b >= 0; --b) {
    const u32 t = r | ((1u << b) - 1u);
    if (wave_count_le(dv, t) < k) r |= 1u << b;
  }
  const int cle = wave_count_le(dv, r);
  u64 tau, thr;
  if (cle <= keep_max) {
    tau = ((u64)__float_as_uint((float)r) << 32) | 0xFFFFFFFFull;
    thr = tau;
  } else {
    const int kk = k - (r > 0 ? wave_count_le(dv, r - 1u) : 0);
    u32 rr = 0;
    for (int b = rowbits - 1; b >= 0; --b) {
      const u32 t = rr | ((1u << b) - 1u);
      int c = 0;
#pragma unroll
      for (int u = 0; u < EP; ++u) c += __popcll(wave_ballot(dv[u] == r && key_row(v[u]) <= t));
      if (c < kk) rr |= 1u << b;
    }
    tau = ((u64)__float_as_uint((float)r) << 32) | rr;
    thr = tau - 1;
  }
  wave_sync();
  int base = 0;
#pragma unroll
  for (int u = 0; u < EP; ++u) {
    const bool keep = v[u] <= tau;
    const u64 m = wave_ballot(keep);
    if (keep) pool[base + (int)lanes_below(m)] = v[u];
    base += __popcll(m);
  }
  wave_sync();
  cnt = base;
  return thr;
}

// Exact mode: row qo of the answer from a query's pool (cnt keys): the k
// smallest keys, sorted, are the answer (ties by original row).  The pool is
// first cut to at most KCAP keys (the ties at the k-th distance kept when
// they fit, so the row search runs only when they do not).
template <int KCAP, int P>
__device__ __forceinline__ void pool_emit(const JoinArgs &a, u64 *pool, int cnt, int64_t qo, int rowbits) {
  const int lane = threadIdx.x & 63;
  const int k = a.k;
  if (cnt > KCAP) pool_shrink<P>(pool, cnt, k, rowbits, KCAP);
  constexpr int E = KCAP / 64;
  u64 w[E];
#pragma unroll
  for (int u = 0; u < E; ++u) w[u] = lane + 64 * u < cnt ? pool[lane + 64 * u] : KEY_MAX;
  wave_sort_u64<E>(w);
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int j = lane + 64 * u;
    if (j < k) {
      const float s = key_dist(w[u]);
      const int32_t row = (int32_t)key_row(w[u]);
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
}

template <int TQ, int KCAP, bool EXACT>
__device__ __forceinline__ void valu_tile(const JoinArgs &a, const Tile tile, const int slot) {
  extern __shared__ u64 smem[];
  constexpr int P = pool_cap(KCAP);
  u64 *wsm = smem + (threadIdx.x >> 6) * valu_wave_words(TQ, KCAP, EXACT);
  u64 *lists = wsm;                        // real mode: [TQ][KCAP] sorted lists
  u64 *bufs = wsm + TQ * KCAP;             // real mode: [TQ][BCAP] buffers
  u64 *tmp = bufs + TQ * KNNJ_BCAP;        // real mode: [KCAP]
  u64 *pools = wsm;                        // exact mode: [TQ][P] pools
  const int lane = threadIdx.x & 63;
  const u64 clk0 = a.tile_prof ? wall_clock64() : 0;
  u64 nvisit = 0, nsweepb = 0;  // diagnostics: visits, screened blocks
  u64 nchunks = 0, mticks = 0;   // diagnostics: 64-row chunks computed, ticks in merges
  const int A = tile.cluster;
  const int nq = tile.count;
  const int d = a.d, dpR = a.R.dpad, dpQ = a.Q.dpad, k = a.k;
  const float eps = dist_relerr(d);

  if constexpr (!EXACT) {
#pragma unroll
    for (int i = 0; i < TQ; ++i)
#pragma unroll
      for (int u = 0; u < KCAP / 64; ++u) lists[i * KCAP + lane + 64 * u] = KEY_MAX;
    __syncthreads();
  }
  const int rowbits = 32 - __builtin_clz((u32)max(a.R.n - 1, (int64_t)1));

  const double th = a.theta[A];
  const float T0 = (float)(th * th * (1.0 + a.delta) * (1.0 + 1e-6));
  u64 thr[TQ];
  int cnt[TQ];
  const float *qrow[TQ];
#pragma unroll
  for (int i = 0; i < TQ; ++i) {
    thr[i] = ((u64)__float_as_uint(T0) << 32) | 0xFFFFFFFFull;
    cnt[i] = 0;
    qrow[i] = a.Q.px + (int64_t)(tile.start + (i < nq ? i : 0)) * dpQ;
  }
  u64 evals = 0, nmerge = 0, npass = 0;
  // the first 16 bytes of each query row, held for the whole tile
  float4 qv0[TQ];
#pragma unroll
  for (int i = 0; i < TQ; ++i) qv0[i] = *(const float4 *)qrow[i];

  // the candidates appended for query i since its last merge (real mode:
  // buffer into the sorted list) or shrink (exact mode: pool)
  auto merge = [&](int i) {
    ++nmerge;
    if constexpr (EXACT) {
      // ties kept while a chunk of room remains above them
      const u64 t = pool_shrink<P>(pools + i * P, cnt[i], k, rowbits, P - 128);
      if (t < thr[i]) thr[i] = t;
    } else {
      wave_merge<KCAP>(lists + i * KCAP, bufs + i * KNNJ_BCAP, cnt[i], tmp);
      cnt[i] = 0;
      const u64 kth = lists[i * KCAP + k - 1];
      if (kth != KEY_MAX) {
        const float T = (float)((double)key_dist(kth) * a.fac);
        const u64 t = ((u64)__float_as_uint(T) << 32) | 0xFFFFFFFFull;
        if (t < thr[i]) thr[i] = t;
      }
    }
  };
  // appends from one more chunk would not fit
  auto full = [&](int i) { return EXACT ? cnt[i] > P - 64 : cnt[i] >= 64; };

  const int np = a.pcnt[A];
  const int32_t *pl = a.plist + (int64_t)A * a.R.m;
  const float *pll = a.plb + (int64_t)A * a.R.m;
  // d(q, a) of each query: with the list ascending by lb_b <= d(a, s), every
  // s of a later cluster has d(q, s) >= lb_b - d(q, a)
  float dqa[TQ];
#pragma unroll
  for (int i = 0; i < TQ; ++i) dqa[i] = i < nq ? a.Q.pdist[tile.start + i] * (1.f + 2.f * eps) : 0.f;
  for (int pb = 0; pb < np; pb += 64) {
    ++nsweepb;
    {
      float reach = 0.f;
#pragma unroll
      for (int i = 0; i < TQ; ++i)
        if (i < nq) reach = fmaxf(reach, sqrtf(key_dist(thr[i])) * (1.f + eps) + dqa[i]);
      if (pll[pb] * (1.f - 2.f * eps) > reach) break;  // this and every later cluster is out of reach
    }
    // Screening: lane j takes listed cluster pb + j, computes the distance of
    // every query of the tile to its representative, and keeps the cluster
    // if the lower bound d(q,b) - r_b does not exceed some query's threshold.
    const int pj = pb + lane;
    const bool lv = pj < np;
    const int bj = lv ? pl[pj] : pl[pb];
    const float rbj = (float)a.R.rad[bj] * (1.f + 2.4e-7f);
    const float *repj = a.R.reps + (int64_t)bj * dpR;
    // member range of each listed cluster, loaded beside its representative
    const int s0j = a.R.off[bj], s1j = a.R.off[bj + 1];
    float dql[TQ];
#pragma unroll
    for (int i = 0; i < TQ; ++i) dql[i] = 0.f;
    for (int t = 0; t < dpR; t += 4) {
      const float4 c4 = *(const float4 *)(repj + t);
#pragma unroll
      for (int i = 0; i < TQ; ++i) {
        if (i < nq) {
          const float4 q4 = *(const float4 *)(qrow[i] + t);
          float df = q4.x - c4.x;
          dql[i] = __builtin_fmaf(df, df, dql[i]);
          df = q4.y - c4.y;
          dql[i] = __builtin_fmaf(df, df, dql[i]);
          df = q4.z - c4.z;
          dql[i] = __builtin_fmaf(df, df, dql[i]);
          df = q4.w - c4.w;
          dql[i] = __builtin_fmaf(df, df, dql[i]);
        }
      }
    }
    bool needj = false;
#pragma unroll
    for (int i = 0; i < TQ; ++i) {
      dql[i] = sqrtf(dql[i]);
      if (i < nq) {
        const float thi = sqrtf(key_dist(thr[i])) * (1.f + eps);
        needj |= !(dql[i] * (1.f - 2.f * eps) - rbj > thi);
      }
    }
    u64 pending = wave_ballot(lv && needj);
    // The member distances at the chunk starts of the next pending cluster
    // (lane c: chunk c; lane 63 also chunk 64) are loaded one visit ahead, so
    // that a visit's window test does not wait on loads of its own.
    float wc = INFINITY, wn = INFINITY;
    auto window_load = [&](const int jl) {
      const int s0 = __builtin_amdgcn_readlane(s0j, jl), s1 = __builtin_amdgcn_readlane(s1j, jl);
      const int nch = (s1 - s0 + 63) >> 6;
      wc = lane < nch ? a.R.pdist[s0 + 64 * lane] : INFINITY;
      wn = (lane == 63 && 64 < nch) ? a.R.pdist[s0 + 64 * 64] : INFINITY;
    };
    if (pending) window_load(__builtin_ctzll(pending));
    while (pending) {
    const int jl = __builtin_ctzll(pending);
    pending &= pending - 1;
    const float wc0 = wc, wn0 = wn;
    if (pending) window_load(__builtin_ctzll(pending));
    const int s0 = __builtin_amdgcn_readlane(s0j, jl), s1 = __builtin_amdgcn_readlane(s1j, jl);
    const float rb = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, rbj), jl));
    bool use[TQ];
    bool any = false;
    float lo = INFINITY, hi = -INFINITY;
#pragma unroll
    for (int i = 0; i < TQ; ++i) {
      use[i] = false;
      if (i < nq) {
        const float dqb = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, dql[i]), jl));
        const float thi = sqrtf(key_dist(thr[i])) * (1.f + eps);
        const float lb = dqb * (1.f - 2.f * eps) - rb;
        if (!(lb > thi)) {
          use[i] = true;
          any = true;
          lo = fminf(lo, dqb * (1.f - 2.f * eps) - thi);
          hi = fmaxf(hi, dqb * (1.f + 2.f * eps) + thi);
        }
      }
    }
    if (!any) continue;
    ++nvisit;
    lo = lo > 0.f ? lo * (1.f - eps) : -INFINITY;
    hi = hi * (1.f + eps);

    const int nch = (s1 - s0 + 63) >> 6;
    for (int cg = 0; cg < nch; cg += 64) {
      const int c = cg + lane;
      float cmin, nmin;
      if (cg == 0) {
        cmin = wc0;
        nmin = __shfl_down(wc0, 1);
        if (lane == 63) nmin = wn0;
      } else {
        cmin = c < nch ? a.R.pdist[s0 + 64 * c] : INFINITY;
        nmin = __shfl_down(cmin, 1);
        if (lane == 63) nmin = (c + 1 < nch) ? a.R.pdist[s0 + 64 * (c + 1)] : INFINITY;
      }
      const bool need = (c < nch) && (cmin <= hi) && (nmin >= lo);
      const u64 m = wave_ballot(need);
      if (m == 0) {
        if (__shfl(cmin, 0) > hi) break;
        continue;
      }
      const int cf = cg + __builtin_ctzll(m);
      const int cl = cg + 63 - __builtin_clzll(m);
      // each lane's first 16 row bytes and row id are loaded one chunk
      // ahead, so a chunk's distances do not wait on its own loads (rows of
      // d <= 4 are just those 16 bytes)
      auto chunk_pos = [&](int c) {
        const int q = s0 + 64 * c + lane;
        return q < s1 ? q : s0;
      };
      int ppn = chunk_pos(cf);
      float4 r0n = *(const float4 *)(a.R.px + (int64_t)ppn * dpR);
      u32 rown = EXACT ? (u32)a.R.perm[ppn] : (u32)ppn;
      for (int c2 = cf; c2 <= cl; ++c2) {
        const int p = s0 + 64 * c2 + lane;
        const bool valid = p < s1;
        const int pp = ppn;
        const float4 r0 = r0n;
        const u32 row = rown;
        if (c2 < cl) {
          ppn = chunk_pos(c2 + 1);
          r0n = *(const float4 *)(a.R.px + (int64_t)ppn * dpR);
          rown = EXACT ? (u32)a.R.perm[ppn] : (u32)ppn;
        }
        const float *rp = a.R.px + (int64_t)pp * dpR;
        float acc[TQ];
        auto step4 = [&](const float4 r4, int t) {
#pragma unroll
          for (int i = 0; i < TQ; ++i) {
            if (use[i]) {
              const float4 q4 = t == 0 ? qv0[i] : *(const float4 *)(qrow[i] + t);
              float df = q4.x - r4.x;
              acc[i] = __builtin_fmaf(df, df, acc[i]);
              df = q4.y - r4.y;
              acc[i] = __builtin_fmaf(df, df, acc[i]);
              df = q4.z - r4.z;
              acc[i] = __builtin_fmaf(df, df, acc[i]);
              df = q4.w - r4.w;
              acc[i] = __builtin_fmaf(df, df, acc[i]);
            }
          }
        };
#pragma unroll
        for (int i = 0; i < TQ; ++i) acc[i] = 0.f;
        step4(r0, 0);
        for (int t = 4; t < dpR; t += 4) step4(*(const float4 *)(rp + t), t);
        const int nvalid = __popcll(wave_ballot(valid));
#pragma unroll
        for (int i = 0; i < TQ; ++i) {
          if (use[i]) {
            const u64 key = valid ? make_key(acc[i], row) : KEY_MAX;
            const bool pass = key <= thr[i];
            const u64 pm = wave_ballot(pass);
            if (pm) {
              if (pass) (EXACT ? pools + i * P : bufs + i * KNNJ_BCAP)[cnt[i] + (int)lanes_below(pm)] = key;
              cnt[i] += __popcll(pm);
              npass += __popcll(pm);
            }
            evals += nvalid;
          }
        }
        ++nchunks;
        bool need_merge = false;
#pragma unroll
        for (int i = 0; i < TQ; ++i) need_merge |= full(i);
        if (need_merge) {
          const u64 cm0 = a.tile_prof ? wall_clock64() : 0;
          __syncthreads();
#pragma unroll
          for (int i = 0; i < TQ; ++i)
            if (full(i)) merge(i);
          if (a.tile_prof) mticks += wall_clock64() - cm0;
        }
      }
    }
  }
  }
  const u64 clk_sweep = a.tile_prof ? wall_clock64() : 0;
  __syncthreads();
  if constexpr (!EXACT) {
#pragma unroll
    for (int i = 0; i < TQ; ++i)
      if (cnt[i] > 0) merge(i);
    __syncthreads();
  }
  const u64 clk_merge = a.tile_prof ? wall_clock64() : 0;

  // ---- results
  if constexpr (EXACT) {
#pragma unroll
    for (int i = 0; i < TQ; ++i)
      if (i < nq) pool_emit<KCAP, P>(a, pools + i * P, cnt[i], a.Q.perm[tile.start + i], rowbits);
  } else {
    for (int i = 0; i < nq; ++i) {
      // the threshold of query i (register array, selected without dynamic indexing)
      u64 ti = thr[0];
#pragma unroll
      for (int j = 1; j < TQ; ++j)
        if (j == i) ti = thr[j];
      emit_query<KCAP, EXACT>(a, lists + i * KCAP, tile.start + i, A, key_dist(ti),
                              a.Q.px + (int64_t)(tile.start + i) * dpQ);
    }
  }
  if (lane == 0 && a.tile_prof) {
    u64 *tp = a.tile_prof + 8 * (size_t)slot;
    tp[0] = clk0;
    tp[1] = wall_clock64();
    // phase clocks as the MFMA kernel's record: no seed (4 = 7 = start),
    // sweep end, final merges end; B groups = 64-row chunks
    tp[2] = (nchunks << 32) | min(mticks, (u64)0xFFFFFFFFull);
    tp[4] = tp[7] = clk0;
    tp[5] = clk_sweep;
    tp[6] = clk_merge;
    tp[3] = (min(nvisit, (u64)0xFFFF) << 48) | (min(nsweepb, (u64)0xFFFF) << 16) | (u64)min(np, 0xFFFF);
  }
  if (lane == 0) {
    atomicAdd(a.n_evals, evals);
    atomicAdd(a.n_evals + 1, nmerge);
    atomicAdd(a.n_evals + 2, npass);
  }
}

// Wave w of workgroup b runs tile slot VALU_WPB b + w.
// At most 128 registers (4 waves per SIMD) for tiles of up to 2 queries and
// capacity 256: the sweep waits on memory, so waves in flight matter more
// than registers (larger tiles, and the many-join kernel, would spill).
#define VALU_WAVES_PER_EU(TQ, KCAP) __attribute__((amdgpu_waves_per_eu((TQ) <= 2 && (KCAP) <= 256 ? 4 : 1)))

template <int TQ, int KCAP, bool EXACT>
__global__ __launch_bounds__(64 * VALU_WPB) VALU_WAVES_PER_EU(TQ, KCAP) void k_join_valu(JoinArgs a) {
  const int slot = (int)blockIdx.x * VALU_WPB + (int)(threadIdx.x >> 6);
  if (slot >= *a.ntiles) return;
  const Tile tile = a.tiles[slot];
  if (tile.count <= 0) return;
  valu_tile<TQ, KCAP, EXACT>(a, tile, slot);
}

// Many joins in one launch (see k_join_mfma_multi).
template <int TQ, int KCAP, bool EXACT>
__global__ __launch_bounds__(64 * VALU_WPB) void k_join_valu_multi(const JoinArgs *__restrict__ argv,
                                                                   const Tile *__restrict__ tiles, int nslots) {
  const int slot = (int)blockIdx.x * VALU_WPB + (int)(threadIdx.x >> 6);
  if (slot >= nslots) return;
  const Tile tile = tiles[slot];
  if (tile.count <= 0) return;
  valu_tile<TQ, KCAP, EXACT>(argv[tile.pad], tile, slot);
}

// The persistent form of the many-joins launch: a grid of as many
// single-wave workgroups as the device holds at once, each taking tile slots
// from the eight device counters queue[0..8) (zero at launch; see
// queue_next) until none are left.  Every wave leaves the loop once the
// counters have passed nslots, and no wave waits on another (one wave per
// workgroup: the tile's barriers are its own).
template <int TQ, int KCAP, bool EXACT>
__global__ __launch_bounds__(64 * VALU_WPB) VALU_WAVES_PER_EU(TQ, KCAP) void k_join_valu_queue(
    const JoinArgs *__restrict__ argv, const Tile *__restrict__ tiles, int nslots, int32_t *queue) {
  static_assert(VALU_WPB == 1, "queue slots are taken per wave");
  const int home = queue_home();
  int turn = 0;
  for (;;) {
    const int slot = queue_next(queue, nslots, home, turn);
    if (slot < 0) break;
    const Tile tile = tiles[slot];
    if (tile.count > 0) valu_tile<TQ, KCAP, EXACT>(argv[tile.pad], tile, slot);
  }
}

// workgroups of a persistent launch: twice what the occupancy query says
// the device holds at once (at least two per compute unit), at most one per
// slot.  Workgroups beyond what fits start as others end and find the queue
// drained, so asking for more costs nothing, while asking for fewer would
// leave the device part idle.  KNNJ_QUEUE_DEBUG=1 prints the figures.
int64_t queue_grid(const void *kernel, int block, size_t lds, int64_t nslots) {
  int per_cu = 0, dev = 0, ncu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds) != hipSuccess) per_cu = 1;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    ncu = 256;
  const int64_t g = std::min<int64_t>(nslots, 2 * (int64_t)std::max(per_cu, 1) * std::max(ncu, 1));
  if (std::getenv("KNNJ_QUEUE_DEBUG"))
    fprintf(stderr, "[knnj queue] slots=%lld per_cu=%d cus=%d lds=%zu grid=%lld\n", (long long)nslots, per_cu, ncu,
            lds, (long long)g);
  return g;
}

int join_tile_q(int kcap) {
  // measured on MI355X: 4 queries per wave at capacity 64 (d = 16, k = 16),
  // 2 at 256 (d = 3, k = 200: 4.9 ms vs 6.7 ms at 4)
  if (kcap <= 64) return 4;
  if (kcap <= 512) return 2;
  return 1;
}

static size_t join_lds_bytes(int tq, int kcap, bool exact) {
  return sizeof(u64) * VALU_WPB * valu_wave_words(tq, kcap, exact);
}

template <int TQ, int KCAP, bool EX>
static hipError_t launch_tq_kcap_ex(const JoinArgs &a, int64_t grid, hipStream_t s, const JoinArgs *argv,
                                    const Tile *tiles, int32_t *queue) {
  const size_t lds = join_lds_bytes(TQ, KCAP, EX);
  const unsigned blocks = (unsigned)((grid + VALU_WPB - 1) / VALU_WPB);
  if (argv && queue) {
    const void *kq = (const void *)k_join_valu_queue<TQ, KCAP, EX>;
    if (lds > 65536) {
      const hipError_t e = hipFuncSetAttribute(kq, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
    }
    const int64_t g = queue_grid(kq, 64 * VALU_WPB, lds, grid);
    hipLaunchKernelGGL((k_join_valu_queue<TQ, KCAP, EX>), dim3((unsigned)g), dim3(64 * VALU_WPB), lds, s, argv, tiles,
                       (int)grid, queue);
  } else if (argv) {
    if (lds > 65536) {
      const hipError_t e = hipFuncSetAttribute((const void *)k_join_valu_multi<TQ, KCAP, EX>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_join_valu_multi<TQ, KCAP, EX>), dim3(blocks), dim3(64 * VALU_WPB), lds, s, argv, tiles,
                       (int)grid);
  } else {
    if (lds > 65536) {
      const hipError_t e = hipFuncSetAttribute((const void *)k_join_valu<TQ, KCAP, EX>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_join_valu<TQ, KCAP, EX>), dim3(blocks), dim3(64 * VALU_WPB), lds, s, a);
  }
  return hipGetLastError();
}

template <int TQ, int KCAP>
static hipError_t launch_tq_kcap(const JoinArgs &a, int64_t grid, hipStream_t s, const JoinArgs *argv,
                                 const Tile *tiles, int32_t *queue) {
  return a.exact ? launch_tq_kcap_ex<TQ, KCAP, true>(a, grid, s, argv, tiles, queue)
                 : launch_tq_kcap_ex<TQ, KCAP, false>(a, grid, s, argv, tiles, queue);
}

template <int KCAP>
static hipError_t launch_kcap(const JoinArgs &a, int tq, int64_t grid, hipStream_t s, const JoinArgs *argv,
                              const Tile *tiles, int32_t *queue) {
  switch (tq) {
    case 8:
      if constexpr (KCAP <= 256) return launch_tq_kcap<8, KCAP>(a, grid, s, argv, tiles, queue);
      break;
    case 4:
      if constexpr (KCAP <= 256) return launch_tq_kcap<4, KCAP>(a, grid, s, argv, tiles, queue);
      break;
    case 2:
      if constexpr (KCAP <= 512) return launch_tq_kcap<2, KCAP>(a, grid, s, argv, tiles, queue);
      break;
    case 1:
      return launch_tq_kcap<1, KCAP>(a, grid, s, argv, tiles, queue);
  }
  return hipErrorInvalidValue;
}

static hipError_t launch_join_any(const JoinArgs &a, int tq, int64_t grid, hipStream_t s, const JoinArgs *argv,
                                  const Tile *tiles, int32_t *queue) {
  if (grid <= 0) return hipSuccess;
  switch (a.kcap) {
    case 64: return launch_kcap<64>(a, tq, grid, s, argv, tiles, queue);
    case 128: return launch_kcap<128>(a, tq, grid, s, argv, tiles, queue);
    case 256: return launch_kcap<256>(a, tq, grid, s, argv, tiles, queue);
    case 512: return launch_kcap<512>(a, tq, grid, s, argv, tiles, queue);
    case 1024: return launch_kcap<1024>(a, tq, grid, s, argv, tiles, queue);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_join(const JoinArgs &a, int tq, int64_t max_tiles, hipStream_t s) {
  return launch_join_any(a, tq, max_tiles, s, nullptr, nullptr, nullptr);
}

hipError_t launch_join_multi(const JoinArgs &proto, int tq, const JoinArgs *argv, const Tile *tiles, int nslots,
                             hipStream_t s, int32_t *queue) {
  return launch_join_any(proto, tq, nslots, s, argv, tiles, queue);
}
