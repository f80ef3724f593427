// k_index.hip -- cluster index build on gfx950.
//
// Replaces clusterReps (reference src/knnjoin.cu:166-411): representative
// sampling (selectReps_cuda/_max/_copy, knnjoin.cu:125-154), the
// cuBLAS + AddAll + findQCluster/findTCluster assignment (knnjoin.cu:43-97,
// 270-342), the atomic member bucketing (fillQMembers/fillTMembers/
// reorderMembers, knnjoin.cu:98-124) and the host OpenMP sort of every
// reference cluster by distance to its representative (knnjoin.cu:380-400).
// Here: no N x m distance matrix is materialised (argmin is fused into the
// distance loop), radii are binary64 maxima for both sides, bucketing is a
// counting sort, the per-cluster sort runs on the device, and the rows are
// physically permuted into cluster order so the join streams contiguous
// memory.
#include <algorithm>

#include "knnj_common.h"
#include "knnj_kernels.h"
#include "wave_sort.h"

// ---------------------------------------------------------------------------
// Zero fill (a kernel, not a memset node, so the enqueue path can be
// captured into a HIP graph and replayed).
__global__ void k_zero32(u32 *p, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

void launch_zero(void *p, size_t bytes, hipStream_t s) {
  const int64_t n = (int64_t)(bytes / 4);
  if (n > 0) hipLaunchKernelGGL(k_zero32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (u32 *)p, n);
}

// Up to three buffers in one launch (an index build clears several small
// per-cluster arrays at once; each launch costs the GPU a few microseconds
// whatever its size, which many small joins pay many times).
struct ZeroSet {
  u32 *p[3];
  int64_t n[3];
};
__global__ void k_zero32_set(ZeroSet z) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
#pragma unroll
  for (int j = 0; j < 3; ++j)
    if (i < z.n[j]) z.p[j][i] = 0u;
}

void launch_zeros(hipStream_t s, void *p0, size_t b0, void *p1, size_t b1, void *p2, size_t b2) {
  ZeroSet z{{(u32 *)p0, (u32 *)p1, (u32 *)p2},
            {p0 ? (int64_t)(b0 / 4) : 0, p1 ? (int64_t)(b1 / 4) : 0, p2 ? (int64_t)(b2 / 4) : 0}};
  const int64_t n = std::max(z.n[0], std::max(z.n[1], z.n[2]));
  if (n > 0) hipLaunchKernelGGL(k_zero32_set, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, z);
}

// Representative draws: element i is SplitMix64's output number i (the
// state after i + 1 increments, mixed) modulo n -- the same sequence a host
// loop of splitmix64() calls gives, computed independently per element.
__global__ void k_draw(int32_t *cand, int count, uint64_t seed, int64_t n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  uint64_t z = seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  cand[i] = (int32_t)(z % (uint64_t)n);
}

void launch_draw(int32_t *cand, int count, uint64_t seed, int64_t n, hipStream_t s) {
  if (count > 0) hipLaunchKernelGGL(k_draw, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, cand, count, seed, n);
}

// ---------------------------------------------------------------------------
// Input scan: non-finite values, non-integers, max |x|.
// flags[0]: any non-finite; flags[1]: float bits of max |x|; flags[2]: any non-integer.
__global__ void k_scan_input(const float *__restrict__ X, int64_t n, int d, int64_t ld, u32 *flags) {
  u32 bad = 0, nonint = 0;
  float mx = 0.f;
  const int64_t total = n * (int64_t)d;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / d;
    const int c = (int)(e - r * d);
    const float x = X[r * ld + c];
    if (!isfinite(x)) {
      bad = 1;
    } else {
      mx = fmaxf(mx, fabsf(x));
      if (x != rintf(x)) nonint = 1;
    }
  }
  bad = __ballot(bad) != 0;
  nonint = __ballot(nonint) != 0;
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0) {
    if (bad) atomicOr(&flags[0], 1u);
    atomicMax(&flags[1], __float_as_uint(mx));
    if (nonint) atomicOr(&flags[2], 1u);
  }
}

// Packed, 16-byte aligned rows: one flat float4 stream.
__global__ void k_scan_input4(const float4 *__restrict__ X4, int64_t n4, u32 *flags) {
  u32 bad = 0, nonint = 0;
  float mx = 0.f;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n4; e += (int64_t)gridDim.x * blockDim.x) {
    const float4 v = X4[e];
    const float x[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!isfinite(x[j])) {
        bad = 1;
      } else {
        mx = fmaxf(mx, fabsf(x[j]));
        if (x[j] != rintf(x[j])) nonint = 1;
      }
    }
  }
  // one set of atomics per workgroup (thousands on one address serialise)
  __shared__ u32 sh[3][4];
  bad = __ballot(bad) != 0;
  nonint = __ballot(nonint) != 0;
  mx = wave_max(mx);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sh[0][w] = bad;
    sh[1][w] = __float_as_uint(mx);
    sh[2][w] = nonint;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    u32 b = 0, x = 0, ni = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
      b |= sh[0][i];
      x = max(x, sh[1][i]);
      ni |= sh[2][i];
    }
    if (b) atomicOr(&flags[0], 1u);
    atomicMax(&flags[1], x);
    if (ni) atomicOr(&flags[2], 1u);
  }
}

void launch_scan_input(const float *X, int64_t n, int d, int64_t ld, u32 *flags, hipStream_t s) {
  const int64_t total = n * (int64_t)d;
  if (ld == d && total % 4 == 0 && ((uintptr_t)X & 15) == 0 && total > 0) {
    const int64_t n4 = total / 4;
    int64_t blocks = std::min<int64_t>((n4 + 255) / 256, 2048);
    hipLaunchKernelGGL(k_scan_input4, dim3((unsigned)blocks), dim3(256), 0, s, (const float4 *)X, n4, flags);
    return;
  }
  int64_t blocks = (total + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_scan_input, dim3((unsigned)blocks), dim3(256), 0, s, X, n, d, ld, flags);
}

// Many inputs in one launch (a batch of joins): workgroup row y scans input
// y, its workgroups striding over that input's values as above.
__global__ __launch_bounds__(256) void k_scan_inputs(const ScanDesc *__restrict__ desc) {
  const ScanDesc D = desc[blockIdx.y];
  const int64_t total = D.n * (int64_t)D.d;
  const bool packed = D.ld == D.d && (total & 3) == 0 && ((uintptr_t)D.X & 15) == 0;
  const int64_t items = packed ? total / 4 : total;
  if ((int64_t)blockIdx.x * blockDim.x >= items) return;  // workgroup-uniform
  u32 bad = 0, nonint = 0;
  float mx = 0.f;
  auto take = [&](const float x) {
    if (!isfinite(x)) {
      bad = 1;
    } else {
      mx = fmaxf(mx, fabsf(x));
      if (x != rintf(x)) nonint = 1;
    }
  };
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  if (packed) {
    const float4 *X4 = (const float4 *)D.X;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < items; e += stride) {
      const float4 v = X4[e];
      take(v.x);
      take(v.y);
      take(v.z);
      take(v.w);
    }
  } else {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < items; e += stride) {
      const int64_t r = e / D.d;
      take(D.X[r * D.ld + (e - r * D.d)]);
    }
  }
  __shared__ u32 sh[3][4];
  bad = __ballot(bad) != 0;
  nonint = __ballot(nonint) != 0;
  mx = wave_max(mx);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sh[0][w] = bad;
    sh[1][w] = __float_as_uint(mx);
    sh[2][w] = nonint;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    u32 b = 0, x = 0, ni = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
      b |= sh[0][i];
      x = max(x, sh[1][i]);
      ni |= sh[2][i];
    }
    if (b) atomicOr(&D.flags[0], 1u);
    atomicMax(&D.flags[1], x);
    if (ni) atomicOr(&D.flags[2], 1u);
  }
}

void launch_scan_inputs(const ScanDesc *desc, int nd, int64_t max_values, hipStream_t s) {
  const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((max_values + 255) / 256, 256));
  for (int i = 0; i < nd; i += 65535)  // the grid's y extent
    hipLaunchKernelGGL(k_scan_inputs, dim3((unsigned)blocks, (unsigned)std::min(65535, nd - i)), dim3(256), 0, s,
                       desc + i);
}

// ---------------------------------------------------------------------------
// Representative sampling score: for trial t, the sum over unordered pairs
// i < j of the FP32 distance between candidate rows (the reference sums
// int(distance) over ordered pairs, knnjoin.cu:125-133; the argmax is the
// same up to that truncation).  The m x m pair matrix of a trial is cut into
// 64 x 64 blocks (upper triangle); a workgroup stages the two blocks' rows in
// LDS 32 dimensions at a time.  The distances take the GEMM form
// sqrt(max(|x_i|^2 + |x_j|^2 - 2 x_i.x_j, 0)) with the dots on
// v_mfma_f32_16x16x4_f32: wave w owns rows 16 w .. 16 w + 15 of the block
// against its 64 columns (four 16 x 16 accumulators); the score only ranks
// the trials, so the FP32 GEMM form is precise enough.  Block partial sums
// are reduced in a fixed order (k_rep_pick), so the choice is deterministic.
#define RS_B 64
#define RS_DC 32
template <bool VEC>
__global__ __launch_bounds__(256) void k_rep_score(const float *__restrict__ X, int64_t ld, int d,
                                                   const int32_t *__restrict__ cand, int m, double *partial) {
  __shared__ float si[RS_DC][RS_B + 1], sj[RS_DC][RS_B + 1];
  __shared__ float ni[RS_B], nj[RS_B];
  __shared__ double red[256];
  const int t = blockIdx.y;
  const int nb = (m + RS_B - 1) / RS_B;
  // block index -> (bi, bj), bi <= bj
  int rem = blockIdx.x, bi = 0;
  while (rem >= nb - bi) {
    rem -= nb - bi;
    ++bi;
  }
  const int bj = bi + rem;
  const int32_t *c = cand + (int64_t)t * m;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, g = lane >> 4, c16 = lane & 15;
  typedef __attribute__((__vector_size__(4 * sizeof(float)))) float rs_f32x4;
  rs_f32x4 acc[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) acc[b] = rs_f32x4{0.f, 0.f, 0.f, 0.f};
  float nacc = 0.f;  // threads 0..63: |x_i|^2 of row tid; 64..127: |x_j|^2 of row tid - 64
  for (int t0 = 0; t0 < d; t0 += RS_DC) {
    __syncthreads();
    if constexpr (VEC) {
      // 16-byte loads (rows 16-byte aligned, d % 4 == 0): thread e stages
      // dimensions 4 q4 .. 4 q4 + 3 of row r
      for (int e = tid; e < RS_B * RS_DC / 4; e += blockDim.x) {
        const int r = e / (RS_DC / 4), q4 = e % (RS_DC / 4), q = 4 * q4;
        const int gi = bi * RS_B + r, gj = bj * RS_B + r;
        // guarded loads (a select between a load and a constant can become a
        // load through a selected pointer, with the constant in scratch memory)
        float4 vi = make_float4(0.f, 0.f, 0.f, 0.f), vj = vi;
        if (gi < m && t0 + q < d) vi = *(const float4 *)(X + (int64_t)c[gi] * ld + t0 + q);
        if (gj < m && t0 + q < d) vj = *(const float4 *)(X + (int64_t)c[gj] * ld + t0 + q);
        si[q][r] = vi.x;
        si[q + 1][r] = vi.y;
        si[q + 2][r] = vi.z;
        si[q + 3][r] = vi.w;
        sj[q][r] = vj.x;
        sj[q + 1][r] = vj.y;
        sj[q + 2][r] = vj.z;
        sj[q + 3][r] = vj.w;
      }
    } else {
      for (int e = tid; e < RS_B * RS_DC; e += blockDim.x) {
        const int r = e / RS_DC, q = e % RS_DC;
        const int gi = bi * RS_B + r, gj = bj * RS_B + r;
        si[q][r] = (gi < m && t0 + q < d) ? X[(int64_t)c[gi] * ld + t0 + q] : 0.f;
        sj[q][r] = (gj < m && t0 + q < d) ? X[(int64_t)c[gj] * ld + t0 + q] : 0.f;
      }
    }
    __syncthreads();
    if (tid < 2 * RS_B) {
      const int r = tid & (RS_B - 1);
#pragma unroll 8
      for (int q = 0; q < RS_DC; ++q) {
        const float v = tid < RS_B ? si[q][r] : sj[q][r];
        nacc = __builtin_fmaf(v, v, nacc);
      }
    }
    // k-step q4: dimensions 4 q4 .. 4 q4 + 3; lane (c16, g) supplies A[c16][g]
    // = x_i(row 16 w + c16, dim 4 q4 + g) and B[g][c16] = x_j(col 16 b + c16, same dim)
#pragma unroll 4
    for (int q4 = 0; q4 < RS_DC / 4; ++q4) {
      const float av = si[4 * q4 + g][16 * w + c16];
#pragma unroll
      for (int b = 0; b < 4; ++b)
        acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, sj[4 * q4 + g][16 * b + c16], acc[b], 0, 0, 0);
    }
  }
  if (tid < RS_B)
    ni[tid] = nacc;
  else if (tid < 2 * RS_B)
    nj[tid - RS_B] = nacc;
  __syncthreads();
  // acc[b][i]: pair (row 16 w + 4 g + i, column 16 b + c16)
  double sum = 0.0;
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int ri = 16 * w + 4 * g + i, cj = 16 * b + c16;
      const int gi = bi * RS_B + ri, gj = bj * RS_B + cj;
      if (gi < m && gj < m && gi < gj) {
        const float sq = (ni[ri] + nj[cj]) - 2.f * acc[b][i];
        sum += (double)sqrtf(fmaxf(sq, 0.f));
      }
    }
  red[tid] = sum;
  __syncthreads();
  for (int s2 = 128; s2 > 0; s2 >>= 1) {
    if (tid < s2) red[tid] += red[tid + s2];
    __syncthreads();
  }
  if (tid == 0) partial[(int64_t)t * gridDim.x + blockIdx.x] = red[0];
}

// The trial with the largest score (first maximum, as the reference's strict
// compare, knnjoin.cu:135-148); its index goes to partial[trials * nblk] (one
// slot past the partial sums).
__global__ void k_rep_pick(double *__restrict__ partial, int nblk, int trials) {
  __shared__ double tot[64];
  // one wave per trial slot: lane-strided sums, then a fixed shuffle tree
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int t = w; t < trials; t += blockDim.x / 64) {
    double s = 0.0;
    for (int b = lane; b < nblk; b += 64) s += partial[(int64_t)t * nblk + b];
    s = wave_sum_d(s);
    if (lane == 0) tot[t] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double bs = -1.0;
    int bt = 0;
    for (int t = 0; t < trials; ++t)
      if (tot[t] > bs) {
        bs = tot[t];
        bt = t;
      }
    *(int32_t *)(partial + (int64_t)trials * nblk) = bt;
  }
}

// Gathers the winning trial's rows into reps[m][dpad] (zero padded).
__global__ void k_rep_gather(const double *__restrict__ partial, int nblk, int trials,
                             const int32_t *__restrict__ cand, int m, const float *__restrict__ X, int64_t ld,
                             int d, int dpad, float *reps, int32_t *rep_rows) {
  const int best = *(const int32_t *)(partial + (int64_t)trials * nblk);
  const int32_t *c = cand + (int64_t)best * m;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < (int64_t)m * dpad;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e / dpad), q = (int)(e - (int64_t)r * dpad);
    reps[e] = q < d ? X[(int64_t)c[r] * ld + q] : 0.f;
    if (q == 0 && rep_rows) rep_rows[r] = c[r];
  }
}

int rep_score_blocks(int m) {
  const int nb = (m + RS_B - 1) / RS_B;
  return nb * (nb + 1) / 2;
}

void launch_rep_select(const float *X, int64_t ld, int d, int dpad, const int32_t *cand, int m,
                       int trials, double *partial, int nblk, float *reps, int32_t *rep_rows,
                       hipStream_t s) {
  if (d % 4 == 0 && ld % 4 == 0 && ((uintptr_t)X & 15) == 0)
    hipLaunchKernelGGL(k_rep_score<true>, dim3(nblk, trials), dim3(256), 0, s, X, ld, d, cand, m, partial);
  else
    hipLaunchKernelGGL(k_rep_score<false>, dim3(nblk, trials), dim3(256), 0, s, X, ld, d, cand, m, partial);
  hipLaunchKernelGGL(k_rep_pick, dim3(1), dim3(1024), 0, s, partial, nblk, trials);
  const int64_t tot = (int64_t)m * dpad;
  const unsigned gb = (unsigned)std::min<int64_t>(1024, (tot + 255) / 256);
  hipLaunchKernelGGL(k_rep_gather, dim3(gb), dim3(256), 0, s, partial, nblk, trials, cand, m, X, ld, d, dpad, reps,
                     rep_rows);
}

// ---------------------------------------------------------------------------
// Assignment: nearest representative (FP32 direct form, lowest index on
// ties, as findQCluster's strict compare, knnjoin.cu:64-80) for small d.
// Representatives are staged in LDS, 32 at a time, up to 32 dimensions at a
// time.
#define AS_RB 32
template <int AS_DC>
__global__ __launch_bounds__(256) void k_assign(const float *__restrict__ X, int64_t n, int64_t ld, int d,
                                                int dpad, const float *__restrict__ reps, int m, int32_t *cl) {
  __shared__ float sr[AS_RB][AS_DC];
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const bool live = i < n;
  const float *x = X + (live ? i : 0) * ld;
  float best = INFINITY;
  int bi = 0;
  for (int r0 = 0; r0 < m; r0 += AS_RB) {
    const int rn = min(AS_RB, m - r0);
    float acc[AS_RB];
#pragma unroll
    for (int r = 0; r < AS_RB; ++r) acc[r] = 0.f;
    for (int t0 = 0; t0 < dpad; t0 += AS_DC) {
      const int tn = min(AS_DC, dpad - t0);
      __syncthreads();
      for (int e = threadIdx.x; e < AS_RB * AS_DC; e += blockDim.x) {
        const int r = e / AS_DC, t = e % AS_DC;
        sr[r][t] = (r < rn && t < tn) ? reps[(int64_t)(r0 + r) * dpad + t0 + t] : 0.f;
      }
      __syncthreads();
      float xv[AS_DC];
#pragma unroll
      for (int t = 0; t < AS_DC; ++t) xv[t] = (live && t < tn && t0 + t < d) ? x[t0 + t] : 0.f;
#pragma unroll
      for (int r = 0; r < AS_RB; ++r) {
        float a = acc[r];
#pragma unroll
        for (int t = 0; t < AS_DC; ++t) {
          const float df = xv[t] - sr[r][t];
          a = __builtin_fmaf(df, df, a);
        }
        acc[r] = a;
      }
    }
#pragma unroll
    for (int r = 0; r < AS_RB; ++r)
      if (r < rn && acc[r] < best) {
        best = acc[r];
        bi = r0 + r;
      }
  }
  if (live) cl[i] = bi;
}

// Binary64 distance of every row to its assigned representative, the cluster
// radius (atomic max on the bit pattern of a non-negative double) and the
// cluster histogram.
__global__ void k_assign_finish(const float *__restrict__ X, int64_t n, int64_t ld, int d, int dpad,
                                const float *__restrict__ reps, const int32_t *__restrict__ cl, float *d2r,
                                u64 *radbits, int32_t *cnt) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = cl[i];
  const double dd = sqrt(sq64_pair(X + i * ld, reps + (int64_t)b * dpad, d));
  d2r[i] = (float)dd;
  atomicMax(&radbits[b], (u64)__double_as_longlong(dd));
  atomicAdd(&cnt[b], 1);
}

__global__ void k_count(const int32_t *__restrict__ cl, int64_t n, int32_t *cnt) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&cnt[cl[i]], 1);
}

// Nearest representative on the matrix cores.  Rows x and representatives
// c are shifted by the first representative c0 (so an offset common to the
// data costs no bf16 precision) and rounded to bf16: x~, c~.  The MFMA
// accumulators start at -||x~||^2 / 2, so D = ||c~||^2 - 2 acc is the FP32
// squared distance ||x~ - c~||^2.  The argmin only has to be good, not exact:
// every later bound uses the binary64 distance to whichever representative
// was chosen.
//
// DK (a multiple of 32, >= dk) is the compiled width: columns dk .. DK-1 are
// zero in both operands.  A workgroup is 4 waves x (16 AB) rows, A operands in registers; the
// representatives stream through LDS in chunks of 64 (double-buffered, one
// barrier per chunk, rows padded by 16 bytes against bank conflicts), so
// every chunk is read from L2 once per workgroup.  Per (row, 16-rep block)
// the running minimum is one signed 32-bit min over a key: D's bits with the
// low `nb` bits replaced by the block number (D >= 0 up to rounding, and a
// negative D -- a distance that rounded below zero -- is smaller than every
// positive one).  Equal keys go to the lower representative.
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 as_bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float as_f32x4;

template <int DK>
__global__ __launch_bounds__(256, 2) void k_assign_mfma(const float *__restrict__ X, int64_t n, int64_t ld, int d,
                                                        const float *__restrict__ c0, const __bf16 *__restrict__ rb,
                                                        const float *__restrict__ rn, int m, int dk, int32_t *cl) {
  // CH representatives per LDS chunk: 128 for DK <= 128, so a chunk's MFMA
  // work covers the global loads of the next chunk (64 measured stalls)
  constexpr int NS = DK / 32, AB = DK <= 128 ? 4 : 2, CH = DK <= 128 ? 128 : 64, RS = DK + 8, PPR = DK / 8;
  constexpr int PT = CH * PPR / 256;
  __shared__ __attribute__((aligned(16))) __bf16 sb[2][CH * RS];
  __shared__ float sn[2][CH];
  const int tid = threadIdx.x, lane = tid & 63, g = lane >> 4, c16 = lane & 15;
  const int64_t base = ((int64_t)blockIdx.x * 4 + (tid >> 6)) * (16 * AB);
  const bool live = base < n;  // wave-uniform; every wave takes part in the staging
  const int nblk = (m + 15) >> 4;
  const int nb = nblk > 1 ? 32 - __builtin_clz((unsigned)(nblk - 1)) : 0;
  const int lowmask = (1 << nb) - 1, himask = ~lowmask;

  as_bf16x8 a[AB][NS];
  as_f32x4 cinit[AB];
  if (live) {
    float cv[NS][8];
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int dim = 32 * s + 8 * g + j;
        cv[s][j] = dim < d ? c0[dim] : 0.f;
      }
#pragma unroll
    for (int ab = 0; ab < AB; ++ab) {
      const int64_t p = base + 16 * ab + c16;
      float part = 0.f;
#pragma unroll
      for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int dim = 32 * s + 8 * g + j;
          const __bf16 v = (__bf16)((dim < d && p < n) ? X[p * ld + dim] - cv[s][j] : 0.f);
          a[ab][s][j] = v;
          const float f = (float)v;
          part = __builtin_fmaf(f, f, part);
        }
      // ||x~||^2 of row c16 in all four lanes holding it, then to the
      // accumulator layout (lane g holds rows 4g .. 4g+3)
      part += xor_lanes_f(part, 16);
      part += xor_lanes_f(part, 32);
#pragma unroll
      for (int i = 0; i < 4; ++i) cinit[ab][i] = -0.5f * __shfl(part, 4 * g + i);
    }
  }
  int best[AB][4];
#pragma unroll
  for (int ab = 0; ab < AB; ++ab)
#pragma unroll
    for (int i = 0; i < 4; ++i) best[ab][i] = 0x7fffffff;

  as_bf16x8 stg[PT];
  float stn = 0.f;
  auto gload = [&](int r0) {
#pragma unroll
    for (int q = 0; q < PT; ++q) {
      const int e = tid + 256 * q, r = e / PPR, t = e % PPR;
      const int rep = r0 + r;
      if (rep < m && 8 * t < dk)
        stg[q] = *(const as_bf16x8 *)(rb + (int64_t)rep * dk + 8 * t);
      else
        stg[q] = as_bf16x8{};
    }
    if (tid < CH) stn = r0 + tid < m ? rn[r0 + tid] : INFINITY;
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int q = 0; q < PT; ++q) {
      const int e = tid + 256 * q, r = e / PPR, t = e % PPR;
      *(as_bf16x8 *)&sb[buf][r * RS + 8 * t] = stg[q];
    }
    if (tid < CH) sn[buf][tid] = stn;
  };
  gload(0);
  lstore(0);
  __syncthreads();
  const int nch = (m + CH - 1) / CH;
  for (int ch = 0; ch < nch; ++ch) {
    const int buf = ch & 1;
    if (ch + 1 < nch) gload((ch + 1) * CH);
    if (live) {
      // the chunk's four 16-representative blocks, unrolled: the next block's
      // LDS reads are issued before this block's MFMAs, and the AB
      // accumulation chains are interleaved step by step
      as_bf16x8 bv[2][NS];
      float nrc[2];
      auto lds_b = [&](int blk, int slot) {
        const int rl = blk * 16 + c16;
#pragma unroll
        for (int s = 0; s < NS; ++s) bv[slot][s] = *(const as_bf16x8 *)&sb[buf][rl * RS + 32 * s + 8 * g];
        nrc[slot] = sn[buf][rl];
      };
      lds_b(0, 0);
#pragma unroll
      for (int blk = 0; blk < CH / 16; ++blk) {
        const int cur = blk & 1;
        if (blk + 1 < CH / 16) lds_b(blk + 1, cur ^ 1);
        const int bno = __builtin_amdgcn_readfirstlane(ch * (CH / 16) + blk);
        as_f32x4 c[AB];
#pragma unroll
        for (int ab = 0; ab < AB; ++ab) c[ab] = cinit[ab];
#pragma unroll
        for (int s = 0; s < NS; ++s)
#pragma unroll
          for (int ab = 0; ab < AB; ++ab)
            c[ab] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ab][s], bv[cur][s], c[ab], 0, 0, 0);
#pragma unroll
        for (int ab = 0; ab < AB; ++ab)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const float dd = __builtin_fmaf(-2.f, c[ab][i], nrc[cur]);
            int key;  // (bits(dd) & himask) | bno in one instruction
            asm("v_and_or_b32 %0, %1, %2, %3" : "=v"(key) : "v"(dd), "v"(himask), "s"(bno));
            best[ab][i] = min(best[ab][i], key);
          }
      }
    }
    if (ch + 1 < nch) lstore(buf ^ 1);
    __syncthreads();
  }
  if (!live) return;
#pragma unroll
  for (int ab = 0; ab < AB; ++ab)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int kb = best[ab][i];
      int bi = ((kb & lowmask) << 4) + c16;
#pragma unroll
      for (int msk = 1; msk < 16; msk <<= 1) {
        const int ok = xor_lanes_i(kb, msk);
        const int oi = xor_lanes_i(bi, msk);
        if (ok < kb || (ok == kb && oi < bi)) {
          kb = ok;
          bi = oi;
        }
      }
      const int64_t row = base + 16 * ab + 4 * g + i;
      if (c16 == 0 && row < n) cl[row] = bi;
    }
}

// bf16 copy of the representatives shifted by the first one (row width dk)
// and the FP32 squared norms of the rounded rows.
__global__ void k_reps_bf16(const float *__restrict__ reps, int m, int d, int dpad, int dk, __bf16 *rb,
                            float *rn) {
  const int r = blockIdx.x;
  if (r >= m) return;
  const int lane = threadIdx.x;
  float part = 0.f;
  for (int t = lane; t < dk; t += 64) {
    const __bf16 v = (__bf16)(t < d ? reps[(int64_t)r * dpad + t] - reps[t] : 0.f);
    rb[(int64_t)r * dk + t] = v;
    const float f = (float)v;
    part = __builtin_fmaf(f, f, part);
  }
  part = wave_sum(part);
  if (lane == 0) rn[r] = part;
}

// Mean of each cluster's members (binary64 sums over members in ascending
// row order, then rounded to FP32; to the nearest integer for integer
// inputs, which keeps the GEMM form exact there).  Empty clusters keep their
// representative.  This is one Lloyd step on the sampled representatives:
// in high dimension a member of a Gaussian cluster lies about sqrt(2) times
// farther from another member than from the centre, so centroids shrink the
// cluster radii that every triangle-inequality bound adds.
__global__ __launch_bounds__(256) void k_centroid(const float *__restrict__ X, int64_t ld, int d, int dpad,
                                                  const int32_t *__restrict__ off,
                                                  const int32_t *__restrict__ perm, int integral, float *reps) {
  // wave w sums members w, w+4, ... for 256 dimensions at a time (lane t
  // holds t, t+64, t+128, t+192); the four partial sums are added in wave
  // order, so the result does not depend on scheduling
  __shared__ double part[4][256];
  const int b = blockIdx.x;
  const int s0 = off[b], s1 = off[b + 1];
  if (s1 <= s0) return;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int t0 = 0; t0 < d; t0 += 256) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    // members p, p+4, p+8, p+12 per step: independent loads in flight,
    // still summed in ascending member order
    for (int p = s0 + w; p < s1; p += 16) {
      int rows[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) rows[q] = p + 4 * q < s1 ? perm[p + 4 * q] : -1;
      float v[4][4];
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int u = 0; u < 4; ++u)
          v[q][u] = (rows[q] >= 0 && t0 + lane + 64 * u < d) ? X[(int64_t)rows[q] * ld + t0 + lane + 64 * u] : 0.f;
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (rows[q] >= 0) {
#pragma unroll
          for (int u = 0; u < 4; ++u) acc[u] += (double)v[q][u];
        }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) part[w][lane + 64 * u] = acc[u];
    __syncthreads();
    const int t = t0 + threadIdx.x;
    if (t < d) {
      const double sum = ((part[0][threadIdx.x] + part[1][threadIdx.x]) + part[2][threadIdx.x]) + part[3][threadIdx.x];
      float v = (float)(sum / (double)(s1 - s0));
      if (integral) v = rintf(v);
      reps[(int64_t)b * dpad + t] = v;
    }
    __syncthreads();
  }
}

void launch_assign(const float *X, int64_t n, int64_t ld, int d, int dpad, const float *reps, int m,
                   int32_t *cl, __bf16 *rb, float *rn, hipStream_t s) {
  if (n <= 0) return;
  const int dk = (d + 31) & ~31;
  if (d >= 16 && dk <= 256 && rb && rn) {
    hipLaunchKernelGGL(k_reps_bf16, dim3(m), dim3(64), 0, s, reps, m, d, dpad, dk, rb, rn);
    // 4 waves of 64 rows (DK <= 128) or of 32 rows
    const dim3 g64((unsigned)((n + 255) / 256)), g32((unsigned)((n + 127) / 128));
    if (dk <= 32)
      hipLaunchKernelGGL(k_assign_mfma<32>, g64, dim3(256), 0, s, X, n, ld, d, reps, rb, rn, m, dk, cl);
    else if (dk <= 64)
      hipLaunchKernelGGL(k_assign_mfma<64>, g64, dim3(256), 0, s, X, n, ld, d, reps, rb, rn, m, dk, cl);
    else if (dk <= 96)
      hipLaunchKernelGGL(k_assign_mfma<96>, g64, dim3(256), 0, s, X, n, ld, d, reps, rb, rn, m, dk, cl);
    else if (dk <= 128)
      hipLaunchKernelGGL(k_assign_mfma<128>, g64, dim3(256), 0, s, X, n, ld, d, reps, rb, rn, m, dk, cl);
    else if (dk <= 192)
      hipLaunchKernelGGL(k_assign_mfma<192>, g32, dim3(256), 0, s, X, n, ld, d, reps, rb, rn, m, dk, cl);
    else
      hipLaunchKernelGGL(k_assign_mfma<256>, g32, dim3(256), 0, s, X, n, ld, d, reps, rb, rn, m, dk, cl);
    return;
  }
  const dim3 g((unsigned)((n + 255) / 256));
  if (dpad <= 4)
    hipLaunchKernelGGL(k_assign<4>, g, dim3(256), 0, s, X, n, ld, d, dpad, reps, m, cl);
  else if (dpad <= 16)
    hipLaunchKernelGGL(k_assign<16>, g, dim3(256), 0, s, X, n, ld, d, dpad, reps, m, cl);
  else
    hipLaunchKernelGGL(k_assign<32>, g, dim3(256), 0, s, X, n, ld, d, dpad, reps, m, cl);
}

// The same for 16-byte aligned rows with ld % 4 == 0: a group of L = dpad/4
// lanes (a power of two, <= 64) reads one row with float4 loads and reduces
// its binary64 partial sums across the group; 64/L rows per wave step.
template <int L>
__global__ __launch_bounds__(256) void k_assign_finish4(const float *__restrict__ X, int64_t n, int64_t ld, int d,
                                                        int dpad, const float *__restrict__ reps,
                                                        const int32_t *__restrict__ cl, float *d2r, u64 *radbits,
                                                        int32_t *cnt) {
  // a wave step takes 2 RPW rows: lane group sub holds rows i and i + RPW,
  // both rows' loads in flight before the binary64 sums
  constexpr int RPW = 64 / L;  // rows per lane-group pass
  const int lane = threadIdx.x & 63, sub = lane / L, j = lane % L;
  const int64_t wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  int64_t ir[2];
  bool ok[2];
  int b[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    ir[h] = wave * 2 * RPW + h * RPW + sub;
    ok[h] = ir[h] < n;
    b[h] = ok[h] ? cl[ir[h]] : 0;
  }
  double acc[2] = {0.0, 0.0};
  for (int c = j; 4 * c < d; c += L) {
    float4 x[2], r[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      x[h] = ok[h] ? *(const float4 *)(X + ir[h] * ld + 4 * c) : make_float4(0.f, 0.f, 0.f, 0.f);
      r[h] = *(const float4 *)(reps + (int64_t)b[h] * dpad + 4 * c);
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const float xa[4] = {x[h].x, x[h].y, x[h].z, x[h].w}, ra[4] = {r[h].x, r[h].y, r[h].z, r[h].w};
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (4 * c + u < d) {
          const double t = (double)xa[u] - (double)ra[u];
          acc[h] += t * t;
        }
    }
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
#pragma unroll
    for (int msk = 1; msk < L; msk <<= 1)
      acc[h] += __builtin_bit_cast(double, xor_lanes_u64(__builtin_bit_cast(u64, acc[h]), msk));
    if (ok[h] && j == 0) {
      const double dd = sqrt(acc[h]);
      d2r[ir[h]] = (float)dd;
      atomicMax(&radbits[b[h]], (u64)__double_as_longlong(dd));
      atomicAdd(&cnt[b[h]], 1);
    }
  }
}

void launch_assign_finish(const float *X, int64_t n, int64_t ld, int d, int dpad, const float *reps,
                          const int32_t *cl, float *d2r, u64 *radbits, int32_t *cnt, hipStream_t s) {
  if (n <= 0) return;
  const bool vec = ld % 4 == 0 && ((uintptr_t)X & 15) == 0 && dpad <= 256;
  if (vec) {
    int L = 1;
    while (L < dpad / 4) L <<= 1;
    const int64_t waves = (n + 2 * (64 / L) - 1) / (2 * (64 / L));  // 2 x 64/L rows per wave
    const dim3 g((unsigned)((waves + 3) / 4));
#define KJ_FIN(LL) hipLaunchKernelGGL(k_assign_finish4<LL>, g, dim3(256), 0, s, X, n, ld, d, dpad, reps, cl, d2r, radbits, cnt)
    switch (L) {
      case 1: KJ_FIN(1); return;
      case 2: KJ_FIN(2); return;
      case 4: KJ_FIN(4); return;
      case 8: KJ_FIN(8); return;
      case 16: KJ_FIN(16); return;
      case 32: KJ_FIN(32); return;
      case 64: KJ_FIN(64); return;
    }
#undef KJ_FIN
  }
  hipLaunchKernelGGL(k_assign_finish, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, X, n, ld, d, dpad,
                     reps, cl, d2r, radbits, cnt);
}

// ---------------------------------------------------------------------------
// Split round of the representative selection.  Uniform sampling leaves
// some dense regions without a representative (with m draws over g equally
// weighted components, a fraction ~exp(-m/g) of them); their points then
// join far-away representatives, whose clusters get radii several times the
// typical one and survive every triangle-inequality test.  After a first
// assignment to m0 sampled representatives, the `extra` clusters of largest
// radius give up their farthest member as a new representative (Gonzalez's
// farthest-point rule, one point per cluster).

// far[b] = max over members of (distance bits << 32 | row).
__global__ void k_far_member(const int32_t *__restrict__ cl, const float *__restrict__ d2r, int64_t n, u64 *far) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) atomicMax(&far[cl[i]], ((u64)__float_as_uint(d2r[i]) << 32) | (u64)(u32)i);
}

// Representatives m0 .. m0 + extra - 1 = farthest members of the `extra`
// clusters with the largest farthest-member distance (ties: lower cluster).
// One workgroup of 1024.
__global__ __launch_bounds__(1024) void k_split_far(const float *__restrict__ X, int64_t ld, int d, int dpad,
                                                    int m0, int extra, const u64 *__restrict__ far, float *reps) {
  __shared__ u64 ord[KNNJ_MAX_REPS];
  int P = 1;
  while (P < m0) P <<= 1;
  for (int c = threadIdx.x; c < P; c += blockDim.x)
    ord[c] = c < m0 ? (((u64)(0xFFFFFFFFu - (u32)(far[c] >> 32)) << 32) | (u32)c) : KEY_MAX;
  __syncthreads();
  for (int size = 2; size <= P; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const int mask = (stride == (size >> 1)) ? (size - 1) : stride;
      for (int i = threadIdx.x; i < P; i += blockDim.x) {
        const int j = i ^ mask;
        if (j > i) {
          const u64 x = ord[i], y = ord[j];
          if (y < x) {
            ord[i] = y;
            ord[j] = x;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int e = threadIdx.x; e < extra * dpad; e += blockDim.x) {
    const int j = e / dpad, t = e % dpad;
    const int c = (int)(u32)ord[j];
    const int64_t row = (int64_t)(u32)far[c];
    reps[(int64_t)(m0 + j) * dpad + t] = t < d ? X[row * ld + t] : 0.f;
  }
}

void launch_split_reps(const float *X, int64_t n, int64_t ld, int d, int dpad, int m0, int extra,
                       const int32_t *cl, const float *d2r, u64 *far, float *reps, hipStream_t s) {
  launch_zero(far, sizeof(u64) * m0, s);
  hipLaunchKernelGGL(k_far_member, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, cl, d2r, n, far);
  hipLaunchKernelGGL(k_split_far, dim3(1), dim3(1024), 0, s, X, ld, d, dpad, m0, extra, (const u64 *)far, reps);
}

void launch_count(const int32_t *cl, int64_t n, int32_t *cnt, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_count, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, cl, n, cnt);
}

// The same mean for 16-byte rows (d % 4 == 0, ld % 4 == 0, X 16-byte
// aligned, d <= 256): thread (r, t) of RP = 256 / (d/4) row slots sums
// float4 t of members r, r + RP, ... (four rows in flight per step) in
// binary64; the row slots' partial sums are then added in slot order, so
// the result does not depend on scheduling.
__global__ __launch_bounds__(256) void k_centroid4(const float4 *__restrict__ X4, int64_t ld4, int d, int dpad,
                                                   const int32_t *__restrict__ off,
                                                   const int32_t *__restrict__ perm, int integral, float *reps) {
  __shared__ double part[256][4];
  const int b = blockIdx.x;
  const int s0 = off[b], s1 = off[b + 1];
  if (s1 <= s0) return;
  const int q4 = d >> 2, RP = 256 / q4;
  const int tid = threadIdx.x, r = tid / q4, t = tid - r * q4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (r < RP) {
    for (int p = s0 + r; p < s1; p += 4 * RP) {
      float4 v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int pu = p + u * RP;
        v[u] = pu < s1 ? X4[(int64_t)perm[pu] * ld4 + t] : make_float4(0.f, 0.f, 0.f, 0.f);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        acc[0] += (double)v[u].x;
        acc[1] += (double)v[u].y;
        acc[2] += (double)v[u].z;
        acc[3] += (double)v[u].w;
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) part[tid][c] = acc[c];
  __syncthreads();
  if (tid < q4) {
    double sum[4] = {0.0, 0.0, 0.0, 0.0};
    for (int rr = 0; rr < RP; ++rr)
#pragma unroll
      for (int c = 0; c < 4; ++c) sum[c] += part[rr * q4 + tid][c];
    const double cntd = (double)(s1 - s0);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      float v = (float)(sum[c] / cntd);
      if (integral) v = rintf(v);
      reps[(int64_t)b * dpad + 4 * tid + c] = v;
    }
  }
}

void launch_centroid(const float *X, int64_t ld, int d, int dpad, int m, const int32_t *off, const int32_t *perm,
                     int integral, float *reps, hipStream_t s) {
  if (d % 4 == 0 && ld % 4 == 0 && ((uintptr_t)X & 15) == 0 && d <= 256)
    hipLaunchKernelGGL(k_centroid4, dim3(m), dim3(256), 0, s, (const float4 *)X, ld / 4, d, dpad, off, perm, integral,
                       reps);
  else
    hipLaunchKernelGGL(k_centroid, dim3(m), dim3(256), 0, s, X, ld, d, dpad, off, perm, integral, reps);
}

// ---------------------------------------------------------------------------
// Exclusive scan of m counts (one workgroup of 1024); off[m] = total.
__global__ __launch_bounds__(1024) void k_scan_counts(const int32_t *__restrict__ cnt, int m, int32_t *off) {
  __shared__ int32_t part[1024];
  const int per = (m + 1023) / 1024;
  const int b = threadIdx.x * per;
  int32_t s = 0;
  for (int j = 0; j < per; ++j)
    if (b + j < m) s += cnt[b + j];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 1; w < 1024; w <<= 1) {
    const int32_t v = threadIdx.x >= (unsigned)w ? part[threadIdx.x - w] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  int32_t run = part[threadIdx.x] - s;
  for (int j = 0; j < per; ++j)
    if (b + j < m) {
      off[b + j] = run;
      run += cnt[b + j];
    }
  if (threadIdx.x == 1023) off[m] = part[1023];
}

void launch_scan_counts(const int32_t *cnt, int m, int32_t *off, hipStream_t s) {
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, s, cnt, m, off);
}

// Places each row in its cluster's segment (slot order is arbitrary here;
// the segment sort below makes it deterministic).
__global__ void k_scatter(const int32_t *__restrict__ cl, int64_t n, const int32_t *__restrict__ off,
                          int32_t *fill, int32_t *perm) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cl[i];
  perm[off[c] + atomicAdd(&fill[c], 1)] = (int32_t)i;
}

// Sorts every cluster segment ascending by (distance to rep, original row)
// and writes pdist; with d2r == NULL, by original row only.  Segments up to SEG_LDS keys sort in LDS; longer ones
// sort in place in `gkeys` (global, same indexing as perm).
#define SEG_LDS 8192
__device__ __forceinline__ void cex(u64 *a, int i, int j) {
  const u64 x = a[i], y = a[j];
  if (y < x) {
    a[i] = y;
    a[j] = x;
  }
}

// Clusters of at most SEG_WAVE rows: one wave sorts the keys in registers
// (wave_sort_u64, no workgroup barriers); k_segsort (LDS bitonic network,
// global memory above SEG_LDS keys) takes the larger ones.
#define SEG_WAVE 512
template <int E>
__device__ __forceinline__ void seg_wave_sort(int32_t *perm, const float *__restrict__ d2r, float *pdist, int s0,
                                              int n) {
  const int lane = threadIdx.x & 63;
  u64 v[E];
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int j = lane + 64 * u;
    if (j < n) {
      const int row = perm[s0 + j];
      v[u] = ((u64)(d2r ? __float_as_uint(d2r[row]) : 0u) << 32) | (u32)row;
    } else {
      v[u] = KEY_MAX;
    }
  }
  wave_sort_u64<E>(v);
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int j = lane + 64 * u;
    if (j < n) {
      perm[s0 + j] = (int32_t)(u32)v[u];
      if (pdist) pdist[s0 + j] = __uint_as_float((u32)(v[u] >> 32));
    }
  }
}

__global__ __launch_bounds__(64) void k_segsort_wave(const int32_t *__restrict__ off, int32_t *perm,
                                                     const float *__restrict__ d2r, float *pdist) {
  const int c = blockIdx.x;
  const int s0 = off[c], n = off[c + 1] - s0;
  if (n <= 1 || n > SEG_WAVE) return;
  if (n <= 64)
    seg_wave_sort<1>(perm, d2r, pdist, s0, n);
  else if (n <= 128)
    seg_wave_sort<2>(perm, d2r, pdist, s0, n);
  else if (n <= 256)
    seg_wave_sort<4>(perm, d2r, pdist, s0, n);
  else
    seg_wave_sort<8>(perm, d2r, pdist, s0, n);
}

__global__ __launch_bounds__(1024) void k_segsort(const int32_t *__restrict__ off, int32_t *perm,
                                                  const float *__restrict__ d2r, float *pdist,
                                                  u64 *gkeys) {
  __shared__ u64 sk[SEG_LDS];
  const int c = blockIdx.x;
  const int s0 = off[c], n = off[c + 1] - s0;
  if (n <= SEG_WAVE) {
    // sorted by k_segsort_wave; a single row still gets its distance
    if (n == 1 && threadIdx.x == 0 && pdist) pdist[s0] = d2r ? d2r[perm[s0]] : 0.f;
    return;
  }
  int p2 = 1;
  while (p2 < n) p2 <<= 1;
  const bool in_lds = p2 <= SEG_LDS;
  u64 *a = in_lds ? sk : gkeys + s0;
  for (int j = threadIdx.x; j < (in_lds ? p2 : n); j += blockDim.x) {
    if (j < n) {
      const int row = perm[s0 + j];
      a[j] = ((u64)(d2r ? __float_as_uint(d2r[row]) : 0u) << 32) | (u32)row;
    } else {
      a[j] = KEY_MAX;
    }
  }
  __syncthreads();
  for (int size = 2; size <= p2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const int mask = (stride == (size >> 1)) ? (size - 1) : stride;
      for (int i = threadIdx.x; i < p2; i += blockDim.x) {
        const int j = i ^ mask;
        if (j > i && j < (in_lds ? p2 : n)) cex(a, i, j);
      }
      __syncthreads();
    }
  }
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    const u64 key = a[j];
    perm[s0 + j] = (int32_t)(u32)key;
    if (pdist) pdist[s0 + j] = __uint_as_float((u32)(key >> 32));
  }
}

// Permuted, zero-padded copy of the rows: px[p][0..dpad).
__global__ void k_permute(const float *__restrict__ X, int64_t ld, int d, int dpad,
                          const int32_t *__restrict__ perm, int64_t n, float *px) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n * dpad) return;
  const int64_t p = e / dpad;
  const int t = (int)(e - p * dpad);
  px[e] = t < d ? X[(int64_t)perm[p] * ld + t] : 0.f;
}

// The permuted copy for 16-byte rows (d % 4 == 0, ld % 4 == 0, X 16-byte
// aligned; then dpad == d): one float4 per thread, 32-bit index arithmetic.
__global__ void k_permute4(const float4 *__restrict__ X4, int64_t ld4, int q4, const int32_t *__restrict__ perm,
                           uint32_t total, float4 *px4) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const uint32_t p = e / (uint32_t)q4, t = e - p * (uint32_t)q4;
  px4[e] = X4[(int64_t)perm[p] * ld4 + t];
}

void launch_bucket_sort_permute(const float *X, int64_t n, int64_t ld, int d, int dpad, int m,
                                const int32_t *cl, const int32_t *cnt, int32_t *off, int32_t *fill,
                                int32_t *perm, const float *d2r, float *pdist, u64 *gkeys, float *px,
                                hipStream_t s) {
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, s, cnt, m, off);
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scatter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, cl, n, off, fill,
                     perm);
  hipLaunchKernelGGL(k_segsort_wave, dim3(m), dim3(64), 0, s, off, perm, d2r, pdist);
  hipLaunchKernelGGL(k_segsort, dim3(m), dim3(1024), 0, s, off, perm, d2r, pdist, gkeys);
  if (px) launch_permute(X, n, ld, d, dpad, perm, px, s);
}

void launch_permute(const float *X, int64_t n, int64_t ld, int d, int dpad, const int32_t *perm, float *px,
                    hipStream_t s) {
  if (n <= 0) return;
  const int64_t tot = n * dpad;
  if (d % 4 == 0 && ld % 4 == 0 && ((uintptr_t)X & 15) == 0 && tot / 4 < ((int64_t)1 << 32)) {
    const int64_t t4 = tot / 4;
    hipLaunchKernelGGL(k_permute4, dim3((unsigned)((t4 + 255) / 256)), dim3(256), 0, s, (const float4 *)X, ld / 4,
                       d / 4, perm, (uint32_t)t4, (float4 *)px);
  } else {
    hipLaunchKernelGGL(k_permute, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, X, ld, d, dpad,
                       perm, n, px);
  }
}

// ---------------------------------------------------------------------------
// Query tiles: cluster A of the query side is cut into ceil(|A| / tq) tiles
// of consecutive permuted queries, clusters taken in descending order of
// their k-th neighbour bound theta_A (the work estimate: a query's search
// radius, hence the number of reference rows it must visit, grows with it)
// so the longest tiles start first.  With nxcd = 8 the placement is XCD-aware: workgroup b runs on XCD
// b % 8, so tile slot 8 i + x is the i-th tile of XCD x; clusters are dealt
// to the XCDs round-robin and all tiles of one cluster go to the same XCD,
// where they share the reference rows in that XCD's L2 (at the cost of
// balance when a few clusters are heavy).  nxcd = 1 is the plain order.
// Every slot not holding a tile, up to max_slots, holds an empty tile
// (count 0); *nslots receives nxcd x the longest sequence.  Every tile
// carries `inst` (the join it belongs to in a batched launch) in its pad
// field.  One workgroup of 1024.
__global__ __launch_bounds__(1024) void k_make_tiles(const int32_t *__restrict__ off, int m,
                                                     const double *__restrict__ theta, int tq, int inst,
                                                     Tile *tiles, int32_t *nslots, int max_slots, int nxcd) {
  __shared__ u64 ord[KNNJ_MAX_REPS];
  __shared__ int32_t part[1024];
  __shared__ int32_t seqlen[8];
  const int NXCD = nxcd;
  int P = 1;
  while (P < m) P <<= 1;
  for (int c = threadIdx.x; c < P; c += blockDim.x)
    ord[c] = c < m ? (((u64)(0xFFFFFFFFu - __float_as_uint((float)theta[c])) << 32) | (u32)c) : KEY_MAX;
  __syncthreads();
  for (int size = 2; size <= P; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const int mask = (stride == (size >> 1)) ? (size - 1) : stride;
      for (int i = threadIdx.x; i < P; i += blockDim.x) {
        const int j = i ^ mask;
        if (j > i) {
          const u64 x = ord[i], y = ord[j];
          if (y < x) {
            ord[i] = y;
            ord[j] = x;
          }
        }
      }
      __syncthreads();
    }
  }
  // regroup: position q of the XCD-major order is sorted cluster
  // j = x + NXCD * t with x = XCD, t = rank within the XCD
  const int per_x = (m + NXCD - 1) / NXCD;
  auto sorted_at = [&](int q) -> int {  // -1 if q maps past m
    const int x = q / per_x, t = q % per_x;
    const int j = x + NXCD * t;
    return j < m ? (int)(u32)ord[j] : -1;
  };
  const int Mq = per_x * NXCD;
  const int per = (Mq + 1023) / 1024;
  const int b = threadIdx.x * per;
  int32_t s = 0;
  for (int jq = 0; jq < per; ++jq)
    if (b + jq < Mq) {
      const int c = sorted_at(b + jq);
      if (c >= 0) s += (off[c + 1] - off[c] + tq - 1) / tq;
    }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 1; w < 1024; w <<= 1) {
    const int32_t v = threadIdx.x >= (unsigned)w ? part[threadIdx.x - w] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  // exclusive prefix of the whole XCD-major sequence at q = x * per_x, via a
  // second pass: a thread that owns the first position of XCD x records it
  int32_t run = part[threadIdx.x] - s;
  for (int jq = 0; jq < per; ++jq) {
    const int q = b + jq;
    if (q >= Mq) break;
    if (q % per_x == 0) seqlen[q / per_x] = run;  // start of XCD q / per_x
    const int c = sorted_at(q);
    if (c >= 0) run += (off[c + 1] - off[c] + tq - 1) / tq;
  }
  __syncthreads();
  const int total = part[1023];
  __shared__ int32_t seqstart[9];
  if (threadIdx.x == 0) {
    for (int x = 0; x < NXCD; ++x) seqstart[x] = seqlen[x];
    seqstart[NXCD] = total;
    int L = 0;
    for (int x = 0; x < NXCD; ++x) L = max(L, seqstart[x + 1] - seqstart[x]);
    *nslots = min(NXCD * L, max_slots);
  }
  __syncthreads();
  int Lmax = 0;
  for (int x = 0; x < NXCD; ++x) Lmax = max(Lmax, seqstart[x + 1] - seqstart[x]);
  // empty slots first, then the tiles
  for (int e = threadIdx.x; e < max_slots; e += blockDim.x) {
    Tile t;
    t.start = 0;
    t.count = 0;
    t.cluster = 0;
    t.pad = inst;
    tiles[e] = t;
  }
  __syncthreads();
  run = part[threadIdx.x] - s;
  for (int jq = 0; jq < per; ++jq) {
    const int q = b + jq;
    if (q >= Mq) break;
    const int c = sorted_at(q);
    if (c < 0) continue;
    const int x = q / per_x;
    const int a0 = off[c], na = off[c + 1] - a0;
    for (int qq = 0; qq < na; qq += tq) {
      const int i = run - seqstart[x];  // rank within XCD x
      const int slot = NXCD * i + x;
      if (slot < max_slots) {
        Tile t;
        t.start = a0 + qq;
        t.count = min(tq, na - qq);
        t.cluster = c;
        t.pad = inst;
        tiles[slot] = t;
      }
      ++run;
    }
  }
}

void launch_make_tiles(const int32_t *off, int m, const double *theta, int tq, int inst, Tile *tiles,
                       int32_t *nslots, int max_slots, int nxcd, hipStream_t s) {
  hipLaunchKernelGGL(k_make_tiles, dim3(1), dim3(1024), 0, s, off, m, theta, tq, inst, tiles, nslots, max_slots,
                     nxcd);
}

// Members of every cluster in ascending row order (counts already in cnt).
void launch_bucket_rows(const int32_t *cl, int64_t n, int m, const int32_t *cnt, int32_t *off, int32_t *fill,
                        int32_t *perm, u64 *gkeys, hipStream_t s) {
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, s, cnt, m, off);
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scatter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, cl, n, off, fill, perm);
  hipLaunchKernelGGL(k_segsort_wave, dim3(m), dim3(64), 0, s, off, perm, (const float *)nullptr, (float *)nullptr);
  hipLaunchKernelGGL(k_segsort, dim3(m), dim3(1024), 0, s, off, perm, (const float *)nullptr, (float *)nullptr,
                     gkeys);
}
