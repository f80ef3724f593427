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
//     never computes r_b).  The list is sorted by lb_b = max(d(a,b) - r_b, 0),
//     a lower bound of d(a, s) for every s in b, stored beside it (plb,
//     binary32 rounded down): a query q of A has d(q, s) >= lb_b - d(q, a), so
//     a join that walks the list in order stops at the first lb_b that exceeds
//     every query's d(q, a) + search radius.  (Sorted by d(a,b) instead, the
//     walk could not stop before the list's end: measured 2.4x more screened
//     clusters on the d=128 mixture.)
// All bound arithmetic is binary64; reference-member distances stored as
// binary32 are scaled up by (1 + 2^-22) before entering an upper bound.
//
// k_bounds (64 KiB of LDS, two workgroups per CU) computes the distances with
// 4 lanes per representative and 16-byte loads, takes the clusters of step 2
// one at a time by a workgroup min-reduction (a few clusters usually hold 2k
// points), and sorts only the surviving list.  When BD_SEL clusters still
// hold fewer than k points it marks the query cluster (pcnt = -1) and
// k_bounds_full, which sorts every representative, redoes it.
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

// lb_b = max(d(a,b) - r_b, 0) as binary32, rounded below the binary64 value
__device__ __forceinline__ float list_lb(double dd, double rb) {
  const double lb = dd - rb;
  return lb > 0.0 ? (float)(lb * (1.0 - 0x1p-20)) : 0.f;
}

// Writes the sorted list (keys[j] = (lb_b bits, b), j < cnt) of query cluster A.
__device__ __forceinline__ void write_list(const u64 *keys, int cnt, int mR, int A, int32_t *plist, float *plb) {
  for (int j = threadIdx.x; j < cnt; j += blockDim.x) {
    plist[(int64_t)A * mR + j] = (int32_t)(u32)keys[j];
    plb[(int64_t)A * mR + j] = __uint_as_float((u32)(keys[j] >> 32));
  }
}

__global__ __launch_bounds__(BD_THREADS) void k_bounds_full(SideIndex Qi, SideIndex Ri, int d, int k,
                                                            double *theta, int32_t *plist, float *plb,
                                                            int32_t *pcnt) {
  __shared__ double dab[KNNJ_MAX_REPS];
  __shared__ u64 keys[KNNJ_MAX_REPS];
  __shared__ int32_t pre[KNNJ_MAX_REPS];
  __shared__ u64 cand[BD_CAND];
  __shared__ int s_T, s_cnt;

  const int A = blockIdx.x;
  if (pcnt[A] >= 0) return;  // done by k_bounds
  const int mR = Ri.m;
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
        key = ((u64)__float_as_uint(list_lb(dd, rb)) << 32) | (u32)b;
        ++mine;
      }
    }
    keys[b] = key;
  }
  atomicAdd(&s_cnt, mine);
  __syncthreads();
  block_sort_u64(keys, P);
  const int cnt = s_cnt;
  write_list(keys, cnt, mR, A, plist, plb);
  if (threadIdx.x == 0) {
    pcnt[A] = cnt;
    theta[A] = th;
  }
}

#define BD_SEL 256  // clusters gathered one at a time before the fallback
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float bd_f32x4;

__device__ __forceinline__ u64 wave_min_u64(u64 v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const u64 o = xor_lanes_u64(v, off);
    v = o < v ? o : v;
  }
  return v;
}

// Binary64 distances of every (query representative, reference
// representative) pair, dab[a][b]: tiles of 64 x 64 pairs whose 16-dimension
// slices are staged through LDS, 4 x 4 pairs per thread.  Each reference
// representative is read from L2 once per 64 query clusters (k_bounds read
// all of them once per query cluster: 2 MB per workgroup at m_R = 4,096,
// d = 128).  (double)x - (double)y is exact; the squares are summed with
// fused multiply-adds in ascending dimension order.
#define RD_T 64
__global__ __launch_bounds__(256) void k_rep_dist(const float *__restrict__ qr, int mq, int qdp,
                                                  const float *__restrict__ rr, int mr, int rdp, int d,
                                                  double *__restrict__ dab) {
  __shared__ float sa[RD_T][17], sb[RD_T][17];  // 16 dimensions + 1 against bank conflicts
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int a0 = blockIdx.y * RD_T, b0 = blockIdx.x * RD_T;
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
  for (int k0 = 0; k0 < d; k0 += 16) {
    for (int e = threadIdx.x; e < RD_T * 16; e += 256) {
      const int r = e >> 4, c = e & 15, dim = k0 + c;
      sa[r][c] = (a0 + r < mq && dim < d) ? qr[(int64_t)(a0 + r) * qdp + dim] : 0.f;
      sb[r][c] = (b0 + r < mr && dim < d) ? rr[(int64_t)(b0 + r) * rdp + dim] : 0.f;
    }
    __syncthreads();
#pragma unroll 4
    for (int c = 0; c < 16; ++c) {
      double av[4], bv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        av[i] = (double)sa[ty + 16 * i][c];
        bv[i] = (double)sb[tx + 16 * i][c];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const double t = av[i] - bv[j];
          acc[i][j] = __builtin_fma(t, t, acc[i][j]);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int a = a0 + ty + 16 * i, b = b0 + tx + 16 * j;
      if (a < mq && b < mr) dab[(int64_t)a * mr + b] = sqrt(acc[i][j]);
    }
}

__global__ __launch_bounds__(BD_THREADS, 2) void k_bounds(SideIndex Qi, SideIndex Ri, int d, int k, double *theta,
                                                          int32_t *plist, float *plb, int32_t *pcnt,
                                                          const double *__restrict__ dab_all) {
  __shared__ double dab[KNNJ_MAX_REPS];
  __shared__ u64 keys[KNNJ_MAX_REPS];  // cluster keys, then candidate bounds, then the surviving list
  __shared__ u64 wmin[BD_THREADS / 64];
  __shared__ int32_t sel[BD_SEL], spre[BD_SEL];
  __shared__ int s_cnt;
  const int A = blockIdx.x, mR = Ri.m, tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int na = Qi.off[A + 1] - Qi.off[A];
  if (na == 0) {
    if (tid == 0) {
      pcnt[A] = 0;
      theta[A] = 0.0;
    }
    return;
  }
  const double rA = Qi.rad[A];

  // 1. d(a, b) in binary64 (row A of k_rep_dist's matrix)
  for (int b = tid; b < mR; b += BD_THREADS) dab[b] = dab_all[(int64_t)A * mR + b];
  __syncthreads();
  // the clusters' keys r_A + d(a,b) + min_s d(s,b) (empty clusters: never taken)
  for (int b = tid; b < mR; b += BD_THREADS) {
    const int o = Ri.off[b], nb = Ri.off[b + 1] - o;
    const double dd = dab[b];
    if (nb > 0) {
      keys[b] = ((u64)__float_as_uint((float)(rA + dd + (double)Ri.pdist[o])) << 32) | (u32)b;
    } else {
      dab[b] = INFINITY;
      keys[b] = KEY_MAX;
    }
  }
  __syncthreads();

  // 2. clusters by smallest r_A + d(a,b) + min_s d(s,b), one per round (thread
  // b mod BD_THREADS owns b), until they hold 2k points
  u64 lmin = KEY_MAX;
  for (int b = tid; b < mR; b += BD_THREADS) lmin = lmin < keys[b] ? lmin : keys[b];
  int T = 0, C = 0;
  for (;;) {
    const u64 w = wave_min_u64(lmin);
    if (lane == 0) wmin[wid] = w;
    __syncthreads();
    u64 best = wmin[0];
#pragma unroll
    for (int x = 1; x < BD_THREADS / 64; ++x) best = wmin[x] < best ? wmin[x] : best;
    __syncthreads();
    if (best == KEY_MAX) break;  // every cluster taken
    const int b = (int)(u32)best;
    C += min(k, Ri.off[b + 1] - Ri.off[b]);
    if (tid == 0) {
      sel[T] = b;
      spre[T] = C;
    }
    ++T;
    if ((b & (BD_THREADS - 1)) == tid) {
      keys[b] = KEY_MAX;
      lmin = KEY_MAX;
      for (int bb = tid; bb < mR; bb += BD_THREADS) lmin = lmin < keys[bb] ? lmin : keys[bb];
    }
    if (C >= 2 * k || T == BD_SEL) break;
  }
  if (C < k) {  // too many tiny clusters: k_bounds_full takes this one
    if (tid == 0) pcnt[A] = -1;
    return;
  }
  __syncthreads();

  // 3. the k-th smallest of the C candidate upper bounds
  int P2 = 1;
  while (P2 < C) P2 <<= 1;
  for (int f = tid; f < P2; f += BD_THREADS) {
    u64 v = KEY_MAX;
    if (f < C) {
      int t = 0;
      while (spre[t] <= f) ++t;  // T <= BD_SEL entries
      const int b = sel[t];
      const int j = f - (t > 0 ? spre[t - 1] : 0);
      const double ub = rA + dab[b] + (double)Ri.pdist[Ri.off[b] + j] * (1.0 + 0x1p-22);
      v = (u64)__double_as_longlong(ub);
    }
    keys[f] = v;
  }
  __syncthreads();
  block_sort_u64(keys, P2);
  const double th = __longlong_as_double((long long)keys[k - 1]) * (1.0 + 1e-9);
  if (tid == 0) s_cnt = 0;
  __syncthreads();

  // 4. the surviving clusters, compacted, then sorted by lb_b
  for (int b = tid; b < mR; b += BD_THREADS) {
    if (Ri.off[b + 1] > Ri.off[b]) {
      const double dd = dab[b], rb = Ri.rad[b];
      const double lb = dd - rA - rb;
      if (!(lb - th > 1e-12 * (dd + rA + rb + th))) {
        const int pos = atomicAdd(&s_cnt, 1);
        keys[pos] = ((u64)__float_as_uint(list_lb(dd, rb)) << 32) | (u32)b;
      }
    }
  }
  __syncthreads();
  const int cnt = s_cnt;
  int P3 = 1;
  while (P3 < cnt) P3 <<= 1;
  for (int f = cnt + tid; f < P3; f += BD_THREADS) keys[f] = KEY_MAX;
  __syncthreads();
  block_sort_u64(keys, P3);
  write_list(keys, cnt, mR, A, plist, plb);
  if (tid == 0) {
    pcnt[A] = cnt;
    theta[A] = th;
  }
}

void launch_bounds(const SideIndex &Qi, const SideIndex &Ri, int d, int k, double *theta, int32_t *plist,
                   float *plb, int32_t *pcnt, double *dab, hipStream_t s) {
  hipLaunchKernelGGL(k_rep_dist, dim3((unsigned)((Ri.m + RD_T - 1) / RD_T), (unsigned)((Qi.m + RD_T - 1) / RD_T)),
                     dim3(256), 0, s, Qi.reps, Qi.m, Qi.dpad, Ri.reps, Ri.m, Ri.dpad, d, dab);
  hipLaunchKernelGGL(k_bounds, dim3(Qi.m), dim3(BD_THREADS), 0, s, Qi, Ri, d, k, theta, plist, plb, pcnt, dab);
  hipLaunchKernelGGL(k_bounds_full, dim3(Qi.m), dim3(BD_THREADS), 0, s, Qi, Ri, d, k, theta, plist, plb, pcnt);
}
