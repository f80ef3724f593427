// k_dedup.hip -- exact duplicate queries are answered once.
//
// The reference runs one thread per query whatever the data
// (KNNQuery_base, src/knnjoin.cu:762-858).  Integer data repeats points
// (the Skin file holds 51,433 distinct B,G,R triples in 245,057 rows), and
// two queries with bit-identical coordinates have the same answer under the
// (squared distance, row) order.  After the query side's cluster sort, such
// queries sit in one cluster and in one run of equal distance to its
// representative (the sort key is (distance, row)), so:
//   k_dup_mark:    per query position, its leader = the first position of its
//                  run with bit-identical coordinates; per cluster, the
//                  number of leaders;
//   k_dup_compact: the leaders' rows, original ids and distances, cluster by
//                  cluster (a query side the join kernels read as is);
//   k_dup_expand:  after the join has written the leaders' rows, every other
//                  query's output row is copied from its leader's.
#include <algorithm>

#include "knnj_common.h"
#include "knnj_kernels.h"

namespace {

constexpr int DD_THREADS = 256;

// Inclusive block scan (sum or max) of one value per thread, 256 threads.
template <bool MAX>
__device__ int block_scan(int v, int *sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int w = 1; w < DD_THREADS; w <<= 1) {
    const int o = t >= w ? sh[t - w] : (MAX ? -1 : 0);
    __syncthreads();
    sh[t] = MAX ? max(sh[t], o) : sh[t] + o;
    __syncthreads();
  }
  const int r = sh[t];
  __syncthreads();
  return r;
}

__device__ __forceinline__ bool rows_equal(const float *a, const float *b, int d) {
  for (int j = 0; j < d; ++j)
    if (__float_as_uint(a[j]) != __float_as_uint(b[j])) return false;
  return true;
}

}  // namespace

// One workgroup per cluster b, its positions in chunks of 256: the start of
// each position's run of equal pdist is a running max of run-start
// positions (carried across chunks), then a forward search from the run
// start finds the first bit-identical row.
__global__ __launch_bounds__(DD_THREADS) void k_dup_mark(SideIndex Q, int d, int32_t *__restrict__ lead,
                                                        int32_t *__restrict__ ucnt) {
  __shared__ int sh[DD_THREADS];
  const int b = blockIdx.x;
  const int c0 = Q.off[b], c1 = Q.off[b + 1];
  int carry = c0, leaders = 0;
  for (int base = c0; base < c1; base += DD_THREADS) {
    const int p = base + (int)threadIdx.x;
    const bool in = p < c1;
    const bool start = in && (p == c0 || Q.pdist[p] != Q.pdist[p - 1]);
    const int s = max(carry, block_scan<true>(start ? p : -1, sh));
    if (in) {
      const float *xp = Q.px + (int64_t)p * Q.dpad;
      int L = p;
      for (int q = s; q < p; ++q)
        if (rows_equal(Q.px + (int64_t)q * Q.dpad, xp, d)) {
          L = q;
          break;
        }
      lead[p] = L;
      leaders += L == p;
    }
    // the last position's run start carries into the next chunk
    const int last = min(DD_THREADS, c1 - base) - 1;
    if ((int)threadIdx.x == last) sh[0] = s;
    __syncthreads();
    carry = sh[0];
    __syncthreads();
  }
  const int tot = block_scan<false>(leaders, sh);
  if (threadIdx.x == DD_THREADS - 1) ucnt[b] = tot;
}

// Leaders of cluster b, in position order, to uoff[b] + their rank.
__global__ __launch_bounds__(DD_THREADS) void k_dup_compact(SideIndex Q, const int32_t *__restrict__ lead,
                                                           const int32_t *__restrict__ uoff, float *__restrict__ upx,
                                                           int32_t *__restrict__ uperm, float *__restrict__ updist) {
  __shared__ int sh[DD_THREADS];
  const int b = blockIdx.x;
  const int c0 = Q.off[b], c1 = Q.off[b + 1];
  int out = uoff[b];
  const int q4 = Q.dpad / 4;
  for (int base = c0; base < c1; base += DD_THREADS) {
    const int p = base + (int)threadIdx.x;
    const bool u = p < c1 && lead[p] == p;
    const int incl = block_scan<false>(u ? 1 : 0, sh);
    if (u) {
      const int o = out + incl - 1;
      const float4 *src = (const float4 *)(Q.px + (int64_t)p * Q.dpad);
      float4 *dst = (float4 *)(upx + (int64_t)o * Q.dpad);
      for (int j = 0; j < q4; ++j) dst[j] = src[j];
      uperm[o] = Q.perm[p];
      updist[o] = Q.pdist[p];
    }
    if (threadIdx.x == DD_THREADS - 1) sh[0] = incl;
    __syncthreads();
    out += sh[0];
    __syncthreads();
  }
}

// One wave per query position: a non-leader's output row is its leader's.
__global__ __launch_bounds__(DD_THREADS) void k_dup_expand(int64_t n, const int32_t *__restrict__ perm,
                                                          const int32_t *__restrict__ lead, int k, int aos,
                                                          int32_t *__restrict__ out_idx,
                                                          float *__restrict__ out_dist) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = (int64_t)gridDim.x * (DD_THREADS / 64);
  for (int64_t p = blockIdx.x * (int64_t)(DD_THREADS / 64) + (threadIdx.x >> 6); p < n; p += waves) {
    const int L = lead[p];
    if (L == p) continue;  // wave-uniform
    const int64_t dst = perm[p], src = perm[L];
    if (aos) {
      const int2 *s = (const int2 *)out_idx + src * k;
      int2 *t = (int2 *)out_idx + dst * k;
      for (int j = lane; j < k; j += 64) t[j] = s[j];
    } else {
      for (int j = lane; j < k; j += 64) {
        out_idx[dst * k + j] = out_idx[src * k + j];
        out_dist[dst * k + j] = out_dist[src * k + j];
      }
    }
  }
}

void launch_dedup(const SideIndex &Q, int d, int32_t *lead, int32_t *ucnt, int32_t *uoff, float *upx,
                  int32_t *uperm, float *updist, hipStream_t s) {
  hipLaunchKernelGGL(k_dup_mark, dim3(Q.m), dim3(DD_THREADS), 0, s, Q, d, lead, ucnt);
  launch_scan_counts(ucnt, Q.m, uoff, s);
  hipLaunchKernelGGL(k_dup_compact, dim3(Q.m), dim3(DD_THREADS), 0, s, Q, lead, uoff, upx, uperm, updist);
}

void launch_dedup_expand(int64_t n, const int32_t *perm, const int32_t *lead, int k, int aos, int32_t *out_idx,
                         float *out_dist, hipStream_t s) {
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>((n + 3) / 4, 8192);
  hipLaunchKernelGGL(k_dup_expand, dim3((unsigned)blocks), dim3(DD_THREADS), 0, s, n, perm, lead, k, aos, out_idx,
                     out_dist);
}
