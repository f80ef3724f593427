// k_center.hip -- reference rows centred on their cluster representative, for
// the MFMA distance tiles.
//
// Row p of cluster b becomes r' = fl32(r - c_b), stored as fp32, as bf16
// (round to nearest even), or as two bf16 planes hi = bf16(r'),
// lo = bf16(r' - hi) (row layout [hi[dk] lo[dk]]) for the compensated
// three-product filter, in rows of dk >= d elements (zero padded), together
// with ||r'||^2 accumulated in binary32 from the fp32 differences.  Centring
// keeps the operands of the GEMM form ||q-c||^2 + ||r-c||^2 - 2 (q-c).(r-c)
// small, which is what makes its rounding error small relative to the
// neighbour distances (the error bound is in k_join_mfma.h).
#include "knnj_common.h"
#include "knnj_kernels.h"

__global__ __launch_bounds__(256) void k_center(SideIndex Ri, int d, int dk, int bf16, void *rc, float *nr,
                                                float *reps32, float *rn32) {
  // bf16: 0 fp32 rows, 1 one bf16 plane, 2 hi/lo bf16 planes
  const int b = blockIdx.x;
  const int s0 = Ri.off[b], s1 = Ri.off[b + 1];
  const float *rep = Ri.reps + (int64_t)b * Ri.dpad;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (wave == 0) {
    // the representative itself, padded to dk, and its squared norm
    float part = 0.f;
    for (int t = lane; t < dk; t += 64) {
      const float v = t < d ? rep[t] : 0.f;
      reps32[(int64_t)b * dk + t] = v;
      part = __builtin_fmaf(v, v, part);
    }
    part = wave_sum(part);
    if (lane == 0) rn32[b] = part;
  }
  for (int p = s0 + wave; p < s1; p += 4) {
    const float *x = Ri.px + (int64_t)p * Ri.dpad;
    float part = 0.f;
    for (int t = lane; t < dk; t += 64) {
      const float v = t < d ? x[t] - rep[t] : 0.f;
      part = __builtin_fmaf(v, v, part);
      if (bf16 == 2) {
        const __bf16 h = (__bf16)v;
        ((__bf16 *)rc)[(int64_t)p * 2 * dk + t] = h;
        ((__bf16 *)rc)[(int64_t)p * 2 * dk + dk + t] = (__bf16)(v - (float)h);
      } else if (bf16 == 1) {
        ((__bf16 *)rc)[(int64_t)p * dk + t] = (__bf16)v;
      } else {
        ((float *)rc)[(int64_t)p * dk + t] = v;
      }
    }
    part = wave_sum(part);
    if (lane == 0) nr[p] = part;
  }
}

void launch_center(const SideIndex &Ri, int d, int dk, int planes, void *rc, float *nr, float *reps32, float *rn32,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_center, dim3(Ri.m), dim3(256), 0, s, Ri, d, dk, planes, rc, nr, reps32, rn32);
}
