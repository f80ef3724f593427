/*
 * ti_join.c -- CPU triangle-inequality kNN join (the CPU baseline).
 *
 * TEST INFRASTRUCTURE ONLY (timed beside the GPU engine by bench.py's
 * cpu_baseline leg and checked against knnj_oracle_bruteforce by tests/).
 *
 * A restatement of the reference algorithm with its exactness defects fixed:
 *   1. representative sampling: `trials` draws of m indices with
 *      replacement; the draw with the largest sum of pairwise distances wins
 *      (src/knnjoin.cu:195-239, kernels 125-154).  Indices come from
 *      SplitMix64 instead of rand(); the score is a binary64 sum instead of a
 *      truncated int sum.
 *   2. assignment of every point to its nearest representative, lowest
 *      representative index on ties (knnjoin.cu:64-97), with the radius of
 *      BOTH the query and the reference clusters (the reference never
 *      computes the reference-cluster radius, knnjoin.cu:95, 391-392).
 *   3. reference-cluster members sorted ascending by distance to their
 *      representative (knnjoin.cu:380-400).
 *   4. theta_A = k-th smallest of the multiset
 *      { r_A + d(a,b) + d(s,b) : b, s among the min(k,|b|) members of b
 *      nearest to b }, kept as a true multiset (knnjoin.cu:459-504 replaces
 *      every entry equal to the maximum, which loses ties).
 *   5. cluster-pair pruning d(a,b) - r_A - r_b > theta_A with both radii
 *      (knnjoin.cu:506-525), surviving pairs sorted by d(a,b)
 *      (knnjoin.cu:1076-1094).
 *   6. per-query scan of surviving clusters nearest first, members from the
 *      farthest to the nearest with the point-level bound
 *      |d(q,b) - d(s,b)| <= theta (knnjoin.cu:762-858), the running theta
 *      being the current k-th distance.
 * Distances used for ordering are the oracle's binary64 keys (s64, index),
 * so the answer equals knnj_oracle_bruteforce() exactly.  Every pruning test
 * "x > y" is evaluated as x - y > EPS * (|x| + |y|) with EPS = 1e-12, which
 * covers the binary64 rounding of the distances and sums that form x and y
 * (each is accurate to a few units of 2^-53 relative to its magnitude).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

double knnj_oracle_sq64(const float *q, const float *r, int32_t d);

#define EPS 1e-12
/* x > y beyond the rounding error of quantities of magnitude mag */
#define SURELY_GT(x, y, mag) ((x) - (y) > EPS * (mag))

static uint64_t splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

typedef struct {
  double v;
  int32_t i;
} dv;

static int dv_cmp(const void *a, const void *b) {
  const dv *x = (const dv *)a, *y = (const dv *)b;
  if (x->v < y->v) return -1;
  if (x->v > y->v) return 1;
  return (x->i > y->i) - (x->i < y->i);
}

/* pick m representative rows of X (n rows); writes their row indices */
static void pick_reps(const float *X, int64_t n, int64_t ld, int32_t d, int32_t m, int32_t trials,
                      uint64_t seed, int32_t *rep) {
  int32_t *cand = (int32_t *)malloc(sizeof(int32_t) * (size_t)m * trials);
  double *score = (double *)calloc((size_t)trials, sizeof(double));
  uint64_t st = seed;
  for (int64_t i = 0; i < (int64_t)m * trials; ++i) cand[i] = (int32_t)(splitmix64(&st) % (uint64_t)n);
  for (int32_t t = 0; t < trials; ++t) {
    const int32_t *c = cand + (size_t)t * m;
    double sc = 0.0;
#pragma omp parallel for reduction(+ : sc) schedule(static)
    for (int32_t i = 0; i < m; ++i)
      for (int32_t j = 0; j < m; ++j)
        if (i != j) sc += sqrt(knnj_oracle_sq64(X + (int64_t)c[i] * ld, X + (int64_t)c[j] * ld, d));
    score[t] = sc;
  }
  int32_t best = 0;
  for (int32_t t = 1; t < trials; ++t)
    if (score[t] > score[best]) best = t;
  memcpy(rep, cand + (size_t)best * m, sizeof(int32_t) * m);
  free(cand);
  free(score);
}

/* Nearest representative for every row; returns cluster ids and distances.
 * Like the reference (cublasSgemm of -2 X C^T, then AddAll and the argmin,
 * knnjoin.cu:270-342, 64-97) the choice is made on the binary32 GEMM form
 * ||c||^2 - 2 x.c, lowest index on ties; it only has to be good: every
 * bound then uses the binary64 distance to the chosen representative.  The
 * dots run with the representatives transposed, [d][m], so the inner loop
 * over representatives is independent sums the compiler vectorises. */
static void assign(const float *X, int64_t n, int64_t ld, int32_t d, const float *reps, int32_t m,
                   int32_t *cl, double *dist) {
  float *rt = (float *)malloc(sizeof(float) * (size_t)d * m);
  float *rn = (float *)malloc(sizeof(float) * (size_t)m);
  for (int32_t c = 0; c < m; ++c) {
    float s = 0.f;
    for (int32_t j = 0; j < d; ++j) {
      const float v = reps[(int64_t)c * d + j];
      rt[(int64_t)j * m + c] = v;
      s += v * v;
    }
    rn[c] = s;
  }
#pragma omp parallel
  {
    float *acc = (float *)malloc(sizeof(float) * (size_t)m);
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      const float *x = X + i * ld;
      for (int32_t c = 0; c < m; ++c) acc[c] = 0.f;
      for (int32_t j = 0; j < d; ++j) {
        const float xj = x[j];
        const float *col = rt + (int64_t)j * m;
        for (int32_t c = 0; c < m; ++c) acc[c] += xj * col[c];
      }
      float best = INFINITY;
      int32_t bi = 0;
      for (int32_t c = 0; c < m; ++c) {
        const float s = rn[c] - 2.f * acc[c];
        if (s < best) {
          best = s;
          bi = c;
        }
      }
      cl[i] = bi;
      dist[i] = sqrt(knnj_oracle_sq64(x, reps + (int64_t)bi * d, d));
    }
    free(acc);
  }
  free(rt);
  free(rn);
}

/* counting sort of rows by cluster: off[m+1], mem[n] (row ids, ascending within cluster) */
static void bucket(const int32_t *cl, int64_t n, int32_t m, int64_t *off, int32_t *mem) {
  memset(off, 0, sizeof(int64_t) * (size_t)(m + 1));
  for (int64_t i = 0; i < n; ++i) off[cl[i] + 1]++;
  for (int32_t c = 0; c < m; ++c) off[c + 1] += off[c];
  int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * (size_t)m);
  memcpy(pos, off, sizeof(int64_t) * (size_t)m);
  for (int64_t i = 0; i < n; ++i) mem[pos[cl[i]]++] = (int32_t)i;
  free(pos);
}

typedef struct {
  double s;
  int32_t i;
} hk;

static inline int hk_lt(double s1, int32_t i1, double s2, int32_t i2) {
  return (s1 < s2) || (s1 == s2 && i1 < i2);
}

static void hk_down(hk *h, int n, int p) {
  for (;;) {
    int l = 2 * p + 1, r = l + 1, m = p;
    if (l < n && hk_lt(h[m].s, h[m].i, h[l].s, h[l].i)) m = l;
    if (r < n && hk_lt(h[m].s, h[m].i, h[r].s, h[r].i)) m = r;
    if (m == p) return;
    hk t = h[p];
    h[p] = h[m];
    h[m] = t;
    p = m;
  }
}

static void hk_up(hk *h, int c) {
  while (c > 0) {
    int p = (c - 1) / 2;
    if (!hk_lt(h[p].s, h[p].i, h[c].s, h[c].i)) return;
    hk t = h[p];
    h[p] = h[c];
    h[c] = t;
    c = p;
  }
}

static int hk_cmp(const void *a, const void *b) {
  const hk *x = (const hk *)a, *y = (const hk *)b;
  if (hk_lt(x->s, x->i, y->s, y->i)) return -1;
  if (hk_lt(y->s, y->i, x->s, x->i)) return 1;
  return 0;
}

/* max-heap of plain doubles (theta computation) */
static void d_down(double *h, int n, int p) {
  for (;;) {
    int l = 2 * p + 1, r = l + 1, m = p;
    if (l < n && h[l] > h[m]) m = l;
    if (r < n && h[r] > h[m]) m = r;
    if (m == p) return;
    double t = h[p];
    h[p] = h[m];
    h[m] = t;
    p = m;
  }
}

static void d_up(double *h, int c) {
  while (c > 0) {
    int p = (c - 1) / 2;
    if (h[p] >= h[c]) return;
    double t = h[p];
    h[p] = h[c];
    h[c] = t;
    c = p;
  }
}

/*
 * Exact kNN join by clustering + triangle-inequality pruning.
 * m_q, m_r: representative counts (clamped to the set sizes); trials: draws
 * per side (the reference uses 10); seed: SplitMix64 seed for both sides.
 * n_evals (nullable) receives the number of exact distance evaluations;
 * t_phases (nullable) the wall seconds of [0] the index build and bounds of
 * both sides and [1] the per-query scan.
 * Returns 0, or -1 on bad arguments.
 */
int knnj_oracle_ti_join(const float *R, int64_t nR, int64_t ldR, const float *Q, int64_t nQ,
                        int64_t ldQ, int32_t d, int32_t k, int32_t m_q, int32_t m_r, int32_t trials,
                        uint64_t seed, int32_t *out_idx, float *out_dist, double *out_sq,
                        int64_t *n_evals, int32_t nthreads, double *t_phases) {
  if (!R || !Q || d < 1 || k < 1 || k > nR || nQ < 0 || ldR < d || ldQ < d || m_q < 1 || m_r < 1 ||
      trials < 1)
    return -1;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  if (nQ == 0) {
    if (n_evals) *n_evals = 0;
    return 0;
  }
  if (m_q > nQ) m_q = (int32_t)nQ;
  if (m_r > nR) m_r = (int32_t)nR;
  const double t0 = omp_get_wtime();

  int32_t *rq = (int32_t *)malloc(sizeof(int32_t) * m_q), *rr = (int32_t *)malloc(sizeof(int32_t) * m_r);
  pick_reps(Q, nQ, ldQ, d, m_q, trials, seed, rq);
  pick_reps(R, nR, ldR, d, m_r, trials, seed, rr);
  float *repq = (float *)malloc(sizeof(float) * (size_t)m_q * d);
  float *repr = (float *)malloc(sizeof(float) * (size_t)m_r * d);
  for (int32_t c = 0; c < m_q; ++c) memcpy(repq + (size_t)c * d, Q + (int64_t)rq[c] * ldQ, sizeof(float) * d);
  for (int32_t c = 0; c < m_r; ++c) memcpy(repr + (size_t)c * d, R + (int64_t)rr[c] * ldR, sizeof(float) * d);

  int32_t *clq = (int32_t *)malloc(sizeof(int32_t) * nQ), *clr = (int32_t *)malloc(sizeof(int32_t) * nR);
  double *dq = (double *)malloc(sizeof(double) * nQ), *dr = (double *)malloc(sizeof(double) * nR);
  assign(Q, nQ, ldQ, d, repq, m_q, clq, dq);
  assign(R, nR, ldR, d, repr, m_r, clr, dr);

  int64_t *offq = (int64_t *)malloc(sizeof(int64_t) * (m_q + 1));
  int64_t *offr = (int64_t *)malloc(sizeof(int64_t) * (m_r + 1));
  int32_t *memq = (int32_t *)malloc(sizeof(int32_t) * nQ), *memr = (int32_t *)malloc(sizeof(int32_t) * nR);
  bucket(clq, nQ, m_q, offq, memq);
  bucket(clr, nR, m_r, offr, memr);

  /* radii, and R members sorted ascending by (distance to rep, row) */
  double *radq = (double *)calloc((size_t)m_q, sizeof(double));
  double *radr = (double *)calloc((size_t)m_r, sizeof(double));
  for (int64_t i = 0; i < nQ; ++i)
    if (dq[i] > radq[clq[i]]) radq[clq[i]] = dq[i];
  for (int64_t i = 0; i < nR; ++i)
    if (dr[i] > radr[clr[i]]) radr[clr[i]] = dr[i];
  dv *sm = (dv *)malloc(sizeof(dv) * nR);
#pragma omp parallel for schedule(dynamic, 4)
  for (int32_t c = 0; c < m_r; ++c) {
    for (int64_t p = offr[c]; p < offr[c + 1]; ++p) {
      sm[p].v = dr[memr[p]];
      sm[p].i = memr[p];
    }
    qsort(sm + offr[c], (size_t)(offr[c + 1] - offr[c]), sizeof(dv), dv_cmp);
  }

  /* representative-to-representative distances */
  double *rrd = (double *)malloc(sizeof(double) * (size_t)m_q * m_r);
#pragma omp parallel for schedule(static)
  for (int32_t a = 0; a < m_q; ++a)
    for (int32_t b = 0; b < m_r; ++b)
      rrd[(size_t)a * m_r + b] = sqrt(knnj_oracle_sq64(repq + (size_t)a * d, repr + (size_t)b * d, d));

  /* theta_A and the surviving, sorted pair lists */
  double *theta = (double *)malloc(sizeof(double) * m_q);
  dv *plist = (dv *)malloc(sizeof(dv) * (size_t)m_q * m_r);
  int32_t *pcnt = (int32_t *)malloc(sizeof(int32_t) * m_q);
#pragma omp parallel
  {
    dv *ord = (dv *)malloc(sizeof(dv) * m_r);
    double *h = (double *)malloc(sizeof(double) * k);
#pragma omp for schedule(dynamic, 1)
    for (int32_t a = 0; a < m_q; ++a) {
      if (offq[a + 1] == offq[a]) {
        pcnt[a] = 0;
        theta[a] = 0;
        continue;
      }
      int32_t no = 0;
      for (int32_t b = 0; b < m_r; ++b) {
        if (offr[b + 1] == offr[b]) continue;
        ord[no].v = radq[a] + rrd[(size_t)a * m_r + b] + sm[offr[b]].v;
        ord[no].i = b;
        ++no;
      }
      qsort(ord, (size_t)no, sizeof(dv), dv_cmp);
      int n = 0;
      for (int32_t t = 0; t < no; ++t) {
        if (n == k && ord[t].v >= h[0]) break;
        int32_t b = ord[t].i;
        double base = radq[a] + rrd[(size_t)a * m_r + b];
        int64_t cnt = offr[b + 1] - offr[b];
        if (cnt > k) cnt = k;
        for (int64_t j = 0; j < cnt; ++j) {
          double ub = base + sm[offr[b] + j].v;
          if (n < k) {
            h[n] = ub;
            d_up(h, n);
            ++n;
          } else if (ub < h[0]) {
            h[0] = ub;
            d_down(h, k, 0);
          } else {
            break;
          }
        }
      }
      theta[a] = h[0] * (1.0 + EPS);
      int32_t np = 0;
      for (int32_t b = 0; b < m_r; ++b) {
        if (offr[b + 1] == offr[b]) continue;
        double dab = rrd[(size_t)a * m_r + b];
        if (!SURELY_GT(dab - radq[a] - radr[b], theta[a], dab + radq[a] + radr[b] + theta[a])) {
          plist[(size_t)a * m_r + np].v = dab;
          plist[(size_t)a * m_r + np].i = b;
          ++np;
        }
      }
      qsort(plist + (size_t)a * m_r, (size_t)np, sizeof(dv), dv_cmp);
      pcnt[a] = np;
    }
    free(ord);
    free(h);
  }

  /* the per-query scan */
  const double t1 = omp_get_wtime();
  int64_t evals = 0;
#pragma omp parallel reduction(+ : evals)
  {
    hk *h = (hk *)malloc(sizeof(hk) * k);
#pragma omp for schedule(dynamic, 64)
    for (int64_t q = 0; q < nQ; ++q) {
      const float *qp = Q + q * ldQ;
      int32_t a = clq[q];
      double th = theta[a];
      int n = 0;
      for (int32_t t = 0; t < pcnt[a]; ++t) {
        int32_t b = plist[(size_t)a * m_r + t].i;
        double dqb = sqrt(knnj_oracle_sq64(qp, repr + (size_t)b * d, d));
        for (int64_t p = offr[b + 1] - 1; p >= offr[b]; --p) {
          double lb = dqb - sm[p].v, mag = dqb + sm[p].v + th;
          if (SURELY_GT(lb, th, mag)) break;
          if (SURELY_GT(-lb, th, mag)) continue;
          int32_t idx = sm[p].i;
          double s = knnj_oracle_sq64(qp, R + (int64_t)idx * ldR, d);
          ++evals;
          if (n < k) {
            h[n].s = s;
            h[n].i = idx;
            hk_up(h, n);
            ++n;
            if (n == k && sqrt(h[0].s) < th) th = sqrt(h[0].s);
          } else if (hk_lt(s, idx, h[0].s, h[0].i)) {
            h[0].s = s;
            h[0].i = idx;
            hk_down(h, k, 0);
            th = sqrt(h[0].s);
          }
        }
      }
      qsort(h, (size_t)n, sizeof(hk), hk_cmp);
      for (int32_t j = 0; j < k; ++j) {
        int ok = j < n;
        out_idx[q * k + j] = ok ? h[j].i : -1;
        if (out_dist) out_dist[q * k + j] = ok ? (float)sqrt(h[j].s) : INFINITY;
        if (out_sq) out_sq[q * k + j] = ok ? h[j].s : INFINITY;
      }
    }
    free(h);
  }
  if (n_evals) *n_evals = evals;
  if (t_phases) {
    t_phases[0] = t1 - t0;
    t_phases[1] = omp_get_wtime() - t1;
  }

  free(rq); free(rr); free(repq); free(repr); free(clq); free(clr); free(dq); free(dr);
  free(offq); free(offr); free(memq); free(memr); free(radq); free(radr); free(sm);
  free(rrd); free(theta); free(plist); free(pcnt);
  return 0;
}
