/*
 * knnj_oracle.c -- CPU oracle for the exact kNN join.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this file: it is used by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg, as the checker.
 *
 * What it restates.  The reference (gychen1991/knn_sweet) computes, for every
 * query point q of Q, the K nearest points of the source set R
 * (src/knnjoin.cu:762-858 KNNQuery_base, output element IndexDist
 * include/common8.h:37-40, Euclidean distance include/common8.h:153-175 and
 * 220-227).  The reference has no defined tie-break among equal distances
 * (its order depends on atomics and an unstable std::sort, knnjoin.cu:102,
 * 110, 390, 1087), so the oracle fixes one:
 *
 *   key(q, r) = (s64(q, r), index of r)     compared lexicographically,
 *   s64(q, r) = (((0 + t_0) + t_1) + ...) + t_{d-1}  in IEEE binary64,
 *   t_j       = ((double)q_j - (double)r_j) * ((double)q_j - (double)r_j),
 *
 * every operation rounded to nearest, no fused multiply-add, dimensions
 * summed in ascending order.  Row q of the answer holds the k smallest keys
 * in ascending order; the reported distance is (float)sqrt(s64).
 *
 * Because s64 is a fixed sequence of correctly rounded IEEE operations, any
 * implementation that performs the same sequence (the GPU rerank does)
 * obtains bit-identical keys.  For integer-valued coordinates with
 * d*(2*max|x|)^2 < 2^24 every operation is exact, in binary32 as well.
 *
 * knnj_oracle_certify() re-evaluates adjacent pairs of an answer in
 * __float128 (113-bit significand: exact for differences and squares of
 * binary32 values of moderate range) and counts pairs whose exact order
 * differs from the binary64 key order, which says whether the binary64
 * ordering is also the exact real-number ordering for that input.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  double s;
  int32_t i;
} okey;

/* (s, i) lexicographic less-than */
static inline int key_lt(double s1, int32_t i1, double s2, int32_t i2) {
  return (s1 < s2) || (s1 == s2 && i1 < i2);
}

/* max-heap on (s, i): h[0] is the largest key */
static void heap_sift_down(okey *h, int n, int p) {
  for (;;) {
    int l = 2 * p + 1, r = l + 1, m = p;
    if (l < n && key_lt(h[m].s, h[m].i, h[l].s, h[l].i)) m = l;
    if (r < n && key_lt(h[m].s, h[m].i, h[r].s, h[r].i)) m = r;
    if (m == p) return;
    okey t = h[p];
    h[p] = h[m];
    h[m] = t;
    p = m;
  }
}

static void heap_sift_up(okey *h, int c) {
  while (c > 0) {
    int p = (c - 1) / 2;
    if (!key_lt(h[p].s, h[p].i, h[c].s, h[c].i)) return;
    okey t = h[p];
    h[p] = h[c];
    h[c] = t;
    c = p;
  }
}

static int okey_cmp(const void *a, const void *b) {
  const okey *x = (const okey *)a, *y = (const okey *)b;
  if (key_lt(x->s, x->i, y->s, y->i)) return -1;
  if (key_lt(y->s, y->i, x->s, x->i)) return 1;
  return 0;
}

/* The defining scalar sequence for one pair. */
double knnj_oracle_sq64(const float *q, const float *r, int32_t d) {
  double s = 0.0;
  for (int32_t j = 0; j < d; ++j) {
    double t = (double)q[j] - (double)r[j];
    s = s + t * t;
  }
  return s;
}

/* s64 for an explicit list of (query row, reference row) pairs. */
void knnj_oracle_pair_sq64(const float *R, int64_t ldR, const float *Q, int64_t ldQ,
                           int32_t d, const int32_t *qi, const int32_t *ri, int64_t n,
                           double *out) {
  for (int64_t p = 0; p < n; ++p)
    out[p] = knnj_oracle_sq64(Q + (int64_t)qi[p] * ldQ, R + (int64_t)ri[p] * ldR, d);
}

#define RB 512 /* reference rows per block */
#define QB 32  /* queries per block */

/*
 * FP64 brute-force exact kNN join (the north-star oracle).
 * R: nR rows of d floats with row stride ldR; Q likewise.  k <= nR.
 * Outputs are row-major [nQ][k]; out_sq may be NULL.
 * Returns 0, or -1 on bad arguments / allocation failure.
 */
int knnj_oracle_bruteforce(const float *R, int64_t nR, int64_t ldR, const float *Q, int64_t nQ,
                           int64_t ldQ, int32_t d, int32_t k, int32_t *out_idx, float *out_dist,
                           double *out_sq, int32_t nthreads) {
  if (!R || !Q || d < 1 || k < 1 || k > nR || nQ < 0 || ldR < d || ldQ < d) return -1;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  int64_t nqb = (nQ + QB - 1) / QB;
  int failed = 0;
#pragma omp parallel
  {
    double *rt = (double *)malloc(sizeof(double) * (size_t)d * RB); /* transposed block, [j][r] */
    double *s = (double *)malloc(sizeof(double) * RB);
    okey *heaps = (okey *)malloc(sizeof(okey) * (size_t)k * QB);
    int *hn = (int *)malloc(sizeof(int) * QB);
    double *qd = (double *)malloc(sizeof(double) * (size_t)d * QB);
    if (!rt || !s || !heaps || !hn || !qd) {
#pragma omp atomic write
      failed = 1;
    } else {
#pragma omp for schedule(dynamic, 1)
      for (int64_t b = 0; b < nqb; ++b) {
        int64_t q0 = b * QB;
        int nq = (int)((nQ - q0) < QB ? (nQ - q0) : QB);
        for (int a = 0; a < nq; ++a) {
          hn[a] = 0;
          for (int32_t j = 0; j < d; ++j) qd[(size_t)a * d + j] = (double)Q[(q0 + a) * ldQ + j];
        }
        for (int64_t r0 = 0; r0 < nR; r0 += RB) {
          int nr = (int)((nR - r0) < RB ? (nR - r0) : RB);
          for (int rr = 0; rr < nr; ++rr)
            for (int32_t j = 0; j < d; ++j)
              rt[(size_t)j * RB + rr] = (double)R[(r0 + rr) * ldR + j];
          for (int a = 0; a < nq; ++a) {
            const double *qa = qd + (size_t)a * d;
            /* same operation sequence as knnj_oracle_sq64, vectorised over rows */
            for (int rr = 0; rr < nr; ++rr) {
              double t = qa[0] - rt[rr];
              s[rr] = t * t;
            }
            for (int32_t j = 1; j < d; ++j) {
              const double qj = qa[j];
              const double *rj = rt + (size_t)j * RB;
              for (int rr = 0; rr < nr; ++rr) {
                double t = qj - rj[rr];
                s[rr] = s[rr] + t * t;
              }
            }
            okey *h = heaps + (size_t)a * k;
            int n = hn[a];
            for (int rr = 0; rr < nr; ++rr) {
              int32_t idx = (int32_t)(r0 + rr);
              if (n < k) {
                h[n].s = s[rr];
                h[n].i = idx;
                heap_sift_up(h, n);
                ++n;
              } else if (key_lt(s[rr], idx, h[0].s, h[0].i)) {
                h[0].s = s[rr];
                h[0].i = idx;
                heap_sift_down(h, k, 0);
              }
            }
            hn[a] = n;
          }
        }
        for (int a = 0; a < nq; ++a) {
          okey *h = heaps + (size_t)a * k;
          qsort(h, (size_t)k, sizeof(okey), okey_cmp);
          int64_t q = q0 + a;
          for (int32_t j = 0; j < k; ++j) {
            out_idx[q * k + j] = h[j].i;
            if (out_dist) out_dist[q * k + j] = (float)sqrt(h[j].s);
            if (out_sq) out_sq[q * k + j] = h[j].s;
          }
        }
      }
    }
    free(rt);
    free(s);
    free(heaps);
    free(hn);
    free(qd);
  }
  return failed ? -1 : 0;
}

/* exact squared distance in binary128 */
static __float128 sq128(const float *q, const float *r, int32_t d) {
  __float128 s = 0;
  for (int32_t j = 0; j < d; ++j) {
    __float128 t = (__float128)q[j] - (__float128)r[j];
    s += t * t;
  }
  return s;
}

/*
 * Counts adjacent pairs (j, j+1) of each answer row whose order by exact
 * (binary128) squared distance, index as tie-break, differs from the row's
 * order.  idx is [nQ][kk].  Returns the count (0 means the binary64 key
 * order is the exact order on this answer).
 */
int64_t knnj_oracle_certify(const float *R, int64_t ldR, const float *Q, int64_t nQ, int64_t ldQ,
                            int32_t d, int32_t kk, const int32_t *idx) {
  int64_t bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
  for (int64_t q = 0; q < nQ; ++q) {
    const float *qp = Q + q * ldQ;
    for (int32_t j = 0; j + 1 < kk; ++j) {
      int32_t a = idx[q * kk + j], b = idx[q * kk + j + 1];
      __float128 sa = sq128(qp, R + (int64_t)a * ldR, d);
      __float128 sb = sq128(qp, R + (int64_t)b * ldR, d);
      if (sb < sa || (sb == sa && b < a)) ++bad;
    }
  }
  return bad;
}
