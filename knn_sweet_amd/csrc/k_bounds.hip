// k_bounds.hip -- per-query-cluster K-th-distance bound and cluster-pair lists.
//
// Replaces RepsUpperBound (reference src/knnjoin.cu:459-504), FilterReps
// (knnjoin.cu:506-525) and the host OpenMP sort of every replist
// (knnjoin.cu:1076-1094).  One workgroup per query cluster A:
//  1. d(a,b) to every reference representative, in binary64;
//  2. theta_A = k-th smallest upper bound r_A + d(a,b) + d(s,b) over distinct
//     reference points s, gathered from the clusters with the smallest
//     r_A + d(a,b) + min_s d(s,b) until they hold 2k candidates (any k
//     distinct points bound the k-th nearest distance, so a subset gives a
//     valid bound; the reference keeps a K-array that drops tied entries);
//  3. keep b unless d(a,b) - r_A - r_b > theta_A (both radii; the reference
//     never computes r_b), list sorted by d(a,b) ascending.
// All bound arithmetic is binary64; reference-member distances stored as
// binary32 are scaled up by (1 + 2^-22) before entering an upper bound.
#include "knnj_common.h"
#include "knnj_kernels.h"

#define BD_THREADS 256
#define BD_CAND 4096  // >= 3 * KNNJ_MAX_K rounded to a power of two

__device__ __forceinline__ void block_sort_u64(u64 *a, int p2) {
  for (int size = 2; size <= p2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const int mask = (stride == (size >> 1)) ? (size - 1) : stride;
      for (int i = threadIdx.x; i < p2; i += blockDim.x) {
        const int j = i ^ mask;
        if (j > i) {
          const u64 x = a[i], y = a[j];
          if (y < x) {
            a[i] = y;
            a[j] = x;
          }
        }
      }
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(BD_THREADS) void k_bounds(SideIndex Qi, SideIndex Ri, int d, int k,
                                                       double *theta, int32_t *plist, int32_t *pcnt) {
  __shared__ double dab[KNNJ_MAX_REPS];
  __shared__ u64 keys[KNNJ_MAX_REPS];
  __shared__ int32_t pre[KNNJ_MAX_REPS];
  __shared__ u64 cand[BD_CAND];
  __shared__ int s_T, s_cnt;

  const int A = blockIdx.x;
  const int mR = Ri.m;
  const int na = Qi.off[A + 1] - Qi.off[A];
  if (na == 0) {
    if (threadIdx.x == 0) {
      pcnt[A] = 0;
      theta[A] = 0.0;
    }
    return;
  }
  const double rA = Qi.rad[A];
  const float *ra = Qi.reps + (int64_t)A * Qi.dpad;
  int P = 1;
  while (P < mR) P <<= 1;

  for (int b = threadIdx.x; b < P; b += blockDim.x) {
    u64 key = KEY_MAX;
    if (b < mR) {
      const int o = Ri.off[b], nb = Ri.off[b + 1] - o;
      if (nb > 0) {
        const double dd = sqrt(sq64_pair(ra, Ri.reps + (int64_t)b * Ri.dpad, d));
        dab[b] = dd;
        const double lbub = rA + dd + (double)Ri.pdist[o];
        key = ((u64)__float_as_uint((float)lbub) << 32) | (u32)b;
      } else {
        dab[b] = INFINITY;
      }
    }
    keys[b] = key;
  }
  __syncthreads();
  block_sort_u64(keys, P);

  // inclusive prefix of min(k, |b_t|) over the sorted order
  const int per = P / BD_THREADS > 0 ? P / BD_THREADS : 1;
  const int t0 = threadIdx.x * per;
  int local = 0;
  for (int j = 0; j < per; ++j) {
    const int t = t0 + j;
    if (t < P) {
      const u64 key = keys[t];
      int c = 0;
      if (key != KEY_MAX) {
        const int b = (int)(u32)key;
        c = min(k, Ri.off[b + 1] - Ri.off[b]);
      }
      local += c;
      pre[t] = local;
    }
  }
  __shared__ int32_t part[BD_THREADS];
  part[threadIdx.x] = (t0 < P) ? local : 0;
  __syncthreads();
  for (int w = 1; w < BD_THREADS; w <<= 1) {
    const int32_t v = threadIdx.x >= (unsigned)w ? part[threadIdx.x - w] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  const int32_t base = part[threadIdx.x] - ((t0 < P) ? local : 0);
  for (int j = 0; j < per; ++j)
    if (t0 + j < P) pre[t0 + j] += base;
  __syncthreads();
  const int total = pre[P - 1];
  const int target = min(2 * k, total);
  for (int t = threadIdx.x; t < P; t += blockDim.x)
    if (pre[t] >= target && (t == 0 || pre[t - 1] < target)) s_T = t;
  __syncthreads();
  const int T = s_T;
  const int C = pre[T];
  int P2 = 1;
  while (P2 < C) P2 <<= 1;

  for (int f = threadIdx.x; f < P2; f += blockDim.x) {
    u64 v = KEY_MAX;
    if (f < C) {
      // cluster slot t: smallest t with pre[t] > f
      int lo = 0, hi = T;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] > f)
          hi = mid;
        else
          lo = mid + 1;
      }
      const int b = (int)(u32)keys[lo];
      const int j = f - (lo > 0 ? pre[lo - 1] : 0);
      const double ub = rA + dab[b] + (double)Ri.pdist[Ri.off[b] + j] * (1.0 + 0x1p-22);
      v = (u64)__double_as_longlong(ub);
    }
    cand[f] = v;
  }
  __syncthreads();
  block_sort_u64(cand, P2);
  const double th = __longlong_as_double((long long)cand[k - 1]) * (1.0 + 1e-9);

  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  int mine = 0;
  for (int b = threadIdx.x; b < P; b += blockDim.x) {
    u64 key = KEY_MAX;
    if (b < mR && Ri.off[b + 1] > Ri.off[b]) {
      const double dd = dab[b], rb = Ri.rad[b];
      const double lb = dd - rA - rb;
      if (!(lb - th > 1e-12 * (dd + rA + rb + th))) {
        key = ((u64)__float_as_uint((float)dd) << 32) | (u32)b;
        ++mine;
      }
    }
    keys[b] = key;
  }
  atomicAdd(&s_cnt, mine);
  __syncthreads();
  block_sort_u64(keys, P);
  const int cnt = s_cnt;
  for (int j = threadIdx.x; j < cnt; j += blockDim.x) plist[(int64_t)A * mR + j] = (int32_t)(u32)keys[j];
  if (threadIdx.x == 0) {
    pcnt[A] = cnt;
    theta[A] = th;
  }
}

void launch_bounds(const SideIndex &Qi, const SideIndex &Ri, int d, int k, double *theta, int32_t *plist,
                   int32_t *pcnt, hipStream_t s) {
  hipLaunchKernelGGL(k_bounds, dim3(Qi.m), dim3(BD_THREADS), 0, s, Qi, Ri, d, k, theta, plist, pcnt);
}
