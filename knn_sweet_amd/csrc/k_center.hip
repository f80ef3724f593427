// k_center.hip -- reference rows centred on their cluster representative, for
// the MFMA distance tiles.
//
// Row p of cluster b becomes r' = fl32(r - c_b), stored as fp32, as bf16
// (round to nearest even), as two bf16 planes hi = bf16(r'),
// lo = bf16(r' - hi) (row layout [hi[dk] lo[dk]]) for the compensated
// three-product filter, or as fp8 (e4m3) of r' scaled by a power of two
// 2^e that brings its largest magnitude into [64, 128) (below the largest
// finite value of both e4m3 encodings), with 2^-e kept per row in rinv; rows
// of dk >= d elements (zero padded), together with ||r'||^2 accumulated in
// binary32 from the fp32 differences.  Centring
// keeps the operands of the GEMM form ||q-c||^2 + ||r-c||^2 - 2 (q-c).(r-c)
// small, which is what makes its rounding error small relative to the
// neighbour distances (the error bound is in k_join_mfma.h).
#include "knnj_common.h"
#include "knnj_kernels.h"

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 kc_bf16x8;

// Workgroup per cluster.  Wave 0 first writes the representative (padded to
// dk) and its squared norm.  Rows: a group of G lanes (G = the power of two
// >= dk/8) takes one row, lane c dimensions 8c .. 8c+7 with 16-byte loads
// and stores; 2 x 64/G rows per wave step.
template <int G>
__global__ __launch_bounds__(256) void k_center(SideIndex Ri, int d, int dk, int planes, void *rc, float *nr,
                                                float *reps32, float *rn32, float *rinv) {
  // planes: 0 fp32 rows, 1 one bf16 plane, 2 hi/lo bf16 planes, 3 scaled fp8
  const int b = blockIdx.x;
  const int s0 = Ri.off[b], s1 = Ri.off[b + 1];
  const float *rep = Ri.reps + (int64_t)b * Ri.dpad;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (wave == 0) {
    float part = 0.f;
    for (int t = lane; t < dk; t += 64) {
      const float v = t < d ? rep[t] : 0.f;
      reps32[(int64_t)b * dk + t] = v;
      part = __builtin_fmaf(v, v, part);
    }
    part = wave_sum(part);
    if (lane == 0) rn32[b] = part;
  }
  constexpr int RPW = 64 / G;
  const int sub = lane / G, c = lane % G;
  const bool act = 8 * c < dk;
  float rv[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) rv[u] = (8 * c + u < d) ? rep[8 * c + u] : 0.f;
  // two rows per lane group per step (a workgroup step covers 8 RPW rows):
  // both rows' loads are in flight before either is centred
  auto load_row = [&](int p, float (&xv)[8]) {
    const float *x = Ri.px + (int64_t)p * Ri.dpad + 8 * c;
    if (8 * c < Ri.dpad) {
      const float4 x0 = *(const float4 *)x;
      xv[0] = x0.x; xv[1] = x0.y; xv[2] = x0.z; xv[3] = x0.w;
    } else {
      xv[0] = xv[1] = xv[2] = xv[3] = 0.f;
    }
    if (8 * c + 4 < Ri.dpad) {
      const float4 x1 = *(const float4 *)(x + 4);
      xv[4] = x1.x; xv[5] = x1.y; xv[6] = x1.z; xv[7] = x1.w;
    } else {
      xv[4] = xv[5] = xv[6] = xv[7] = 0.f;
    }
  };
  auto store_row = [&](int p, const float (&v)[8], float part, float vmax) {
    if (planes == 3) {
      // 2^e with vmax 2^e in [64, 128): exact scaling, then round to nearest
      const int e = vmax > 0.f ? 6 - ilogbf(vmax) : 0;
      float w[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) w[u] = ldexpf(v[u], e);
      int lo = __builtin_amdgcn_cvt_pk_fp8_f32(w[0], w[1], 0, false);
      lo = __builtin_amdgcn_cvt_pk_fp8_f32(w[2], w[3], lo, true);
      int hi = __builtin_amdgcn_cvt_pk_fp8_f32(w[4], w[5], 0, false);
      hi = __builtin_amdgcn_cvt_pk_fp8_f32(w[6], w[7], hi, true);
      *(uint2 *)((unsigned char *)rc + (int64_t)p * dk + 8 * c) = make_uint2((unsigned)lo, (unsigned)hi);
      if (c == 0) rinv[p] = ldexpf(1.f, -e);
    } else if (planes == 2) {
      kc_bf16x8 h, l;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        h[u] = (__bf16)v[u];
        l[u] = (__bf16)(v[u] - (float)h[u]);
      }
      *(kc_bf16x8 *)((__bf16 *)rc + (int64_t)p * 2 * dk + 8 * c) = h;
      *(kc_bf16x8 *)((__bf16 *)rc + (int64_t)p * 2 * dk + dk + 8 * c) = l;
    } else if (planes == 1) {
      kc_bf16x8 h;
#pragma unroll
      for (int u = 0; u < 8; ++u) h[u] = (__bf16)v[u];
      *(kc_bf16x8 *)((__bf16 *)rc + (int64_t)p * dk + 8 * c) = h;
    } else {
      float *o = (float *)rc + (int64_t)p * dk + 8 * c;
      *(float4 *)o = make_float4(v[0], v[1], v[2], v[3]);
      *(float4 *)(o + 4) = make_float4(v[4], v[5], v[6], v[7]);
    }
    if (c == 0) nr[p] = part;
  };
  for (int p0 = s0 + wave * 2 * RPW; p0 < s1; p0 += 8 * RPW) {
    const int pa = p0 + sub, pb = p0 + RPW + sub;
    const bool oka = pa < s1 && act, okb = pb < s1 && act;
    float xa[8], xb[8];
    if (oka) load_row(pa, xa);
    if (okb) load_row(pb, xb);
    float va[8], vb[8];
    float parta = 0.f, partb = 0.f, ma = 0.f, mb = 0.f;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      va[u] = (oka && 8 * c + u < d) ? xa[u] - rv[u] : 0.f;
      vb[u] = (okb && 8 * c + u < d) ? xb[u] - rv[u] : 0.f;
      parta = __builtin_fmaf(va[u], va[u], parta);
      partb = __builtin_fmaf(vb[u], vb[u], partb);
      ma = fmaxf(ma, fabsf(va[u]));
      mb = fmaxf(mb, fabsf(vb[u]));
    }
#pragma unroll
    for (int msk = 1; msk < G; msk <<= 1) {
      parta += xor_lanes_f(parta, msk);
      partb += xor_lanes_f(partb, msk);
      ma = fmaxf(ma, xor_lanes_f(ma, msk));
      mb = fmaxf(mb, xor_lanes_f(mb, msk));
    }
    if (oka) store_row(pa, va, parta, ma);
    if (okb) store_row(pb, vb, partb, mb);
  }
}

void launch_center(const SideIndex &Ri, int d, int dk, int planes, void *rc, float *nr, float *reps32, float *rn32,
                   float *rinv, hipStream_t s) {
  int G = 1;
  while (G < dk / 8) G <<= 1;
#define KJ_CENTER(GG) \
  hipLaunchKernelGGL(k_center<GG>, dim3(Ri.m), dim3(256), 0, s, Ri, d, dk, planes, rc, nr, reps32, rn32, rinv)
  switch (G) {
    case 4: KJ_CENTER(4); break;
    case 8: KJ_CENTER(8); break;
    case 16: KJ_CENTER(16); break;
    default: KJ_CENTER(32); break;
  }
#undef KJ_CENTER
}
