// wave_sort.h -- wave64 bitonic sorting and merging of u64 keys.
//
// A wave holds E*64 elements, element e = lane + 64*u in register v[u].
// The network is the direction-free bitonic sort: within each merge of size
// `size`, the first step pairs e with e ^ (size-1) (the mirror element) and
// later steps pair e with e ^ stride; every compare-exchange puts the
// smaller key at the smaller index, so padding with KEY_MAX at the end of the
// array never moves a real key into the padding.
#pragma once

#include "knnj_common.h"

template <int E>
__device__ __forceinline__ void wave_sort_u64(u64 (&v)[E]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int size = 2; size <= 64 * E; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const int mask = (stride == (size >> 1)) ? (size - 1) : stride;
      const int lm = mask & 63, um = mask >> 6;
      u64 nv[E];
#pragma unroll
      for (int u = 0; u < E; ++u) {
        u64 other = v[u ^ um];
        if (lm) other = xor_lanes_u64(other, lm);
        const int e = lane + 64 * u;
        const bool lower = e < (e ^ mask);
        const u64 mn = v[u] < other ? v[u] : other;
        const u64 mx = v[u] < other ? other : v[u];
        nv[u] = lower ? mn : mx;
      }
#pragma unroll
      for (int u = 0; u < E; ++u) v[u] = nv[u];
    }
  }
}

// Same network on (hi, lo) pairs compared lexicographically.
template <int E>
__device__ __forceinline__ void wave_sort_pair(u64 (&hi)[E], u32 (&lo)[E]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int size = 2; size <= 64 * E; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const int mask = (stride == (size >> 1)) ? (size - 1) : stride;
      const int lm = mask & 63, um = mask >> 6;
      u64 nh[E];
      u32 nl[E];
#pragma unroll
      for (int u = 0; u < E; ++u) {
        u64 oh = hi[u ^ um];
        u32 ol = lo[u ^ um];
        if (lm) {
          oh = xor_lanes_u64(oh, lm);
          ol = xor_lanes(ol, lm);
        }
        const int e = lane + 64 * u;
        const bool lower = e < (e ^ mask);
        const bool mine_less = (hi[u] < oh) || (hi[u] == oh && lo[u] < ol);
        const bool take_mine = (lower == mine_less);
        nh[u] = take_mine ? hi[u] : oh;
        nl[u] = take_mine ? lo[u] : ol;
      }
#pragma unroll
      for (int u = 0; u < E; ++u) {
        hi[u] = nh[u];
        lo[u] = nl[u];
      }
    }
  }
}

// number of elements of sorted a[0..n) that are < x (LT) or <= x (!LT);
// p2 is a power of two >= n.
template <bool LT>
__device__ __forceinline__ int count_below(const u64 *a, int n, int p2, u64 x) {
  int c = 0;
  for (int s = p2; s > 0; s >>= 1) {
    const int t = c + s;
    if (t <= n) {
      const u64 y = a[t - 1];
      if (LT ? (y < x) : (y <= x)) c = t;
    }
  }
  return c;
}

struct NoDrop {
  static constexpr bool tracks = false;
  __device__ float operator()(u64) const { return INFINITY; }
};

// Merges the unsorted candidate buffer buf[0..cnt) (cnt <= BCAP) into the
// sorted list[0..KCAP), keeping the KCAP smallest keys.  Keys in list and
// buf are distinct except for KEY_MAX padding.  tmp: KCAP scratch keys, all
// three arrays private to the calling wave (other waves of its workgroup may
// be elsewhere: only wave-level ordering is used).  Returns the wave-wide minimum of
// drop(key) over the real keys that fell off the end (INFINITY if none).
//
// A buffer of up to 128 keys into a list of up to 256 is merged by counting: every lane reads the
// buffer keys as LDS broadcasts and counts, for its own buffer and list keys,
// the buffer keys below them; a buffer key's list rank is a binary search.
// Each key's new position is then its index among the smaller keys of both
// sequences (keys are distinct), with no cross-lane dependency chain.  Larger
// buffers sort the buffer with the bitonic network and rank-merge.
template <int KCAP, int BCAP = KNNJ_BCAP, class Drop = NoDrop>
__device__ __forceinline__ float wave_merge(u64 *list, u64 *buf, int cnt, u64 *tmp, Drop drop = Drop()) {
  constexpr int EB = BCAP / 64;
  const int lane = threadIdx.x & 63;
  if constexpr (BCAP <= 64 && KCAP <= 256) {
    constexpr int EL = KCAP / 64;
    u64 v[EB], y[EL];
    int rv[EB], ry[EL];
#pragma unroll
    for (int u = 0; u < EB; ++u) {
      v[u] = lane + 64 * u < cnt ? buf[lane + 64 * u] : KEY_MAX;
      rv[u] = 0;
    }
#pragma unroll
    for (int u = 0; u < EL; ++u) {
      y[u] = list[lane + 64 * u];
      ry[u] = 0;
    }
    // pad the buffer to a multiple of 16 keys with KEY_MAX (below no key),
    // then count in batches of 16 broadcast reads
#pragma unroll
    for (int u = 0; u < EB; ++u)
      if (lane + 64 * u >= cnt) buf[lane + 64 * u] = KEY_MAX;
    wave_sync();
    for (int i = 0; i < cnt; i += 16) {
      u64 b[16];
#pragma unroll
      for (int t = 0; t < 16; ++t) b[t] = buf[i + t];
#pragma unroll
      for (int t = 0; t < 16; ++t) {
#pragma unroll
        for (int u = 0; u < EB; ++u) rv[u] += b[t] < v[u] ? 1 : 0;
#pragma unroll
        for (int u = 0; u < EL; ++u) ry[u] += b[t] < y[u] ? 1 : 0;
      }
    }
    float dmin = INFINITY;
#pragma unroll
    for (int u = 0; u < EL; ++u) {
      const int pos = lane + 64 * u + ry[u];
      if (pos < KCAP)
        tmp[pos] = y[u];
      else if (y[u] != KEY_MAX)
        dmin = fminf(dmin, drop(y[u]));
    }
#pragma unroll
    for (int u = 0; u < EB; ++u) {
      const int pos = rv[u] + count_below<false>(list, KCAP, KCAP, v[u]);
      if (pos < KCAP)
        tmp[pos] = v[u];
      else if (v[u] != KEY_MAX)
        dmin = fminf(dmin, drop(v[u]));
    }
    wave_sync();
#pragma unroll
    for (int u = 0; u < EL; ++u) list[lane + 64 * u] = tmp[lane + 64 * u];
    wave_sync();
    if constexpr (Drop::tracks)
      return wave_min(dmin);
    else
      return INFINITY;
  }
  u64 v[EB];
#pragma unroll
  for (int u = 0; u < EB; ++u) v[u] = lane + 64 * u < cnt ? buf[lane + 64 * u] : KEY_MAX;
  wave_sort_u64<EB>(v);
#pragma unroll
  for (int u = 0; u < EB; ++u) buf[lane + 64 * u] = v[u];
  wave_sync();
  float dmin = INFINITY;
#pragma unroll
  for (int u = 0; u < KCAP / 64; ++u) {
    const int j = lane + 64 * u;
    const u64 x = list[j];
    const int pos = j + count_below<true>(buf, BCAP, BCAP, x);
    if (pos < KCAP)
      tmp[pos] = x;
    else if (x != KEY_MAX)
      dmin = fminf(dmin, drop(x));
  }
#pragma unroll
  for (int u = 0; u < EB; ++u) {
    const int j = lane + 64 * u;
    const u64 x = v[u];
    const int pos = j + count_below<false>(list, KCAP, KCAP, x);
    if (pos < KCAP)
      tmp[pos] = x;
    else if (x != KEY_MAX)
      dmin = fminf(dmin, drop(x));
  }
  wave_sync();
#pragma unroll
  for (int u = 0; u < KCAP / 64; ++u) list[lane + 64 * u] = tmp[lane + 64 * u];
  wave_sync();
  if constexpr (Drop::tracks)
    return wave_min(dmin);
  else
    return INFINITY;
}
