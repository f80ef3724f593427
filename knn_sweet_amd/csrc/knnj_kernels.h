// knnj_kernels.h -- host-side launchers of the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "knnj_common.h"

// k_index.hip
void launch_zero(void *p, size_t bytes, hipStream_t s);  // bytes: a multiple of 4
// zero up to three buffers (null pointers skipped) in one launch
void launch_zeros(hipStream_t s, void *p0, size_t b0, void *p1 = nullptr, size_t b1 = 0, void *p2 = nullptr,
                  size_t b2 = 0);
// cand[i] = SplitMix64 output i (state seed, i + 1 increments) mod n
void launch_draw(int32_t *cand, int count, uint64_t seed, int64_t n, hipStream_t s);
void launch_scan_input(const float *X, int64_t n, int d, int64_t ld, u32 *flags, hipStream_t s);
// the same scan of nd inputs in one launch (desc: device array; max_values:
// the largest n d among them)
struct ScanDesc {
  const float *X;
  int64_t n, ld;
  int d, pad;
  u32 *flags;
};
void launch_scan_inputs(const ScanDesc *desc, int nd, int64_t max_values, hipStream_t s);
int rep_score_blocks(int m);  // partial sums per trial of launch_rep_select
// partial: trials * nblk partial sums + one slot for the winning trial
void launch_rep_select(const float *X, int64_t ld, int d, int dpad, const int32_t *cand, int m, int trials,
                       double *partial, int nblk, float *reps, int32_t *rep_rows, hipStream_t s);
// nearest representative (MFMA bf16 for 16 <= d <= 256 when rb/rn are given, VALU otherwise)
void launch_assign(const float *X, int64_t n, int64_t ld, int d, int dpad, const float *reps, int m,
                   int32_t *cl, __bf16 *rb, float *rn, hipStream_t s);
void launch_assign_finish(const float *X, int64_t n, int64_t ld, int d, int dpad, const float *reps,
                          const int32_t *cl, float *d2r, u64 *radbits, int32_t *cnt, hipStream_t s);
void launch_count(const int32_t *cl, int64_t n, int32_t *cnt, hipStream_t s);
// split round: representatives m0 .. m0+extra-1 = farthest members of the
// extra widest clusters (cl, d2r from an assignment to the first m0; far: m0 scratch)
void launch_split_reps(const float *X, int64_t n, int64_t ld, int d, int dpad, int m0, int extra,
                       const int32_t *cl, const float *d2r, u64 *far, float *reps, hipStream_t s);
void launch_centroid(const float *X, int64_t ld, int d, int dpad, int m, const int32_t *off, const int32_t *perm,
                     int integral, float *reps, hipStream_t s);
void launch_bucket_rows(const int32_t *cl, int64_t n, int m, const int32_t *cnt, int32_t *off, int32_t *fill,
                        int32_t *perm, u64 *gkeys, hipStream_t s);
// (px null: the permuted copy is left to launch_permute)
void launch_bucket_sort_permute(const float *X, int64_t n, int64_t ld, int d, int dpad, int m,
                                const int32_t *cl, const int32_t *cnt, int32_t *off, int32_t *fill,
                                int32_t *perm, const float *d2r, float *pdist, u64 *gkeys, float *px,
                                hipStream_t s);
// px[p] = X[perm[p]] (rows zero padded to dpad)
void launch_permute(const float *X, int64_t n, int64_t ld, int d, int dpad, const int32_t *perm, float *px,
                    hipStream_t s);
// XCD-aware tile slots; *nslots = slots used (empty tiles have count 0);
// every slot's pad field = inst
void launch_make_tiles(const int32_t *off, int m, const double *theta, int tq, int inst, Tile *tiles,
                       int32_t *nslots, int max_slots, int nxcd, hipStream_t s);

// exclusive scan of m counts (m <= KNNJ_MAX_REPS): off[0..m], off[m] = total
void launch_scan_counts(const int32_t *cnt, int m, int32_t *off, hipStream_t s);

// k_dedup.hip: exact duplicate queries answered once.  launch_dedup marks
// each position's leader (first bit-identical row of its cluster) in lead,
// counts leaders per cluster (ucnt, scanned into uoff) and writes the
// leaders' side (rows, original ids, distances) into upx/uperm/updist.
void launch_dedup(const SideIndex &Q, int d, int32_t *lead, int32_t *ucnt, int32_t *uoff, float *upx,
                  int32_t *uperm, float *updist, hipStream_t s);
// The leaders' side as the join reads it: Q's representatives and radii
// with the buffers launch_dedup fills (a host-side view, known before any
// launch, so kernel arguments can be filled ahead of the work).
inline SideIndex dedup_side(const SideIndex &Q, const int32_t *uoff, const float *upx, const int32_t *uperm,
                            const float *updist) {
  SideIndex u = Q;
  u.off = uoff;
  u.px = upx;
  u.perm = uperm;
  u.pdist = updist;
  return u;
}
// copies every non-leader's output row from its leader's (after the join)
void launch_dedup_expand(int64_t n, const int32_t *perm, const int32_t *lead, int k, int aos, int32_t *out_idx,
                         float *out_dist, hipStream_t s);

// k_bounds.hip
// plist[A][j]: surviving reference clusters of query cluster A, ascending by
// plb[A][j] = max(d(a,b) - r_b, 0) rounded down (a lower bound of d(a, s), s in b);
// dab: [Qi.m][Ri.m] binary64 scratch for the representative distances
void launch_bounds(const SideIndex &Qi, const SideIndex &Ri, int d, int k, double *theta, int32_t *plist,
                   float *plb, int32_t *pcnt, double *dab, hipStream_t s);

// k_join.hip
struct JoinArgs {
  SideIndex Q, R;
  int d;
  int k;
  int kcap;
  int exact;             // 1: integer mode (FP32 keys exact, key row = original row)
  double fac;            // real mode: threshold factor (1+delta)/(1-delta), with margin
  double delta;          // real mode: relative error bound of the FP32 squared distance
  const Tile *tiles;
  const int32_t *ntiles;
  const double *theta;   // [mQ]
  const int32_t *plist;  // [mQ][mR]
  const float *plb;      // [mQ][mR] list order key: lower bound of d(a, s), s in the cluster
  const int32_t *pcnt;   // [mQ]
  int32_t *out_idx;      // [nQ][k] or IndexDist [nQ][k] when out_mode == 2
  float *out_dist;
  int out_mode;          // 0 Euclidean, 1 squared, 2 AoS IndexDist (Euclidean)
  u64 *n_evals;          // [0] distance evaluations, [1] merges, [2] candidates appended
  Tile *ovf_tiles;       // overflowed queries, as one-query tiles
  int32_t *ovf_count;
  u32 *max_err_bits;     // max |key - s64| seen in rerank (float bits)
  u32 *ferr_bits;        // [3] rerank: max |s_f - s64|, max E, max |s_f - s64| / E (float bits)
  int relbound;          // filter keys are the FP32 direct-form values, |s_f - s| <= delta s
  // MFMA paths: reference rows centred on their cluster representative
  const void *rc;        // [nR][dk] bf16 or fp32
  const float *nr;       // [nR] squared norm of the centred row (binary32)
  int dk;                // padded width of rc and reps32: roundup(d, 32)
  const float *reps32;   // [mR][dk] reference representatives, FP32, zero padded
  const float *rn32;     // [mR] their squared norms
  const float *rinv;     // [nR] fp8 filter: inverse power-of-two scale of each centred row
  float c1, c2;          // error-bound coefficients (see k_join_mfma.h)
  // optional per-tile record {start clock, end clock, clusters listed, clusters visited}
  u64 *tile_prof;
  float *ovf_diag;       // optional: first 64 overflow records {cutoff, k-th key, last key, first key}
  int ecode;             // MFMA filter keys carry per-entry error codes (rows < 2^26)
  ECode ec;              // their layout: base, steps per octave, row field width
  int waves;             // MFMA join: > 1 asks for the multi-wave workgroup (tiles of 16 x waves queries)
  int lockstep;          // multi-wave workgroup: 1 = waves walk the union of their cluster lists together
  int split;             // multi-wave workgroup: > 1 = its waves split one 16-query tile's cluster list
  int seed_blocks;       // MFMA join: the seed searches the first seed_blocks x 16 listed clusters (0: all)
  // MFMA join, optional: [original query row] its nearest reference
  // representative (the index build's assignment of the query rows); the
  // seed then screens the tile's 16 of these instead of searching the list
  const int32_t *qseed;
};

// Picks the wave tile size for a candidate capacity; returns the dynamic LDS bytes.
int join_tile_q(int kcap);
size_t join_lds_bytes(int tq, int kcap);
hipError_t launch_join(const JoinArgs &a, int tq, int64_t max_tiles, hipStream_t s);

// k_center.hip
// planes: 0 fp32 rows, 1 bf16 rows, 2 bf16 hi/lo rows, 3 fp8 rows scaled per
// row by a power of two (rinv[nR]: the inverse scales; unused otherwise)
void launch_center(const SideIndex &Ri, int d, int dk, int planes, void *rc, float *nr, float *reps32, float *rn32,
                   float *rinv, hipStream_t s);

// k_join_mfma.hip (kernels per operand width in k_join_mfma_d*.hip)
hipError_t launch_join_mfma_f32(const JoinArgs &a, int64_t max_tiles, hipStream_t s);
// terms: 1 single bf16 product, 3 compensated hi/lo products
hipError_t launch_join_mfma_bf16(const JoinArgs &a, int terms, int64_t max_tiles, hipStream_t s);
// fp8 (e4m3) filter, rows scaled per row; operand widths 64, 96, 128
hipError_t launch_join_mfma_fp8(const JoinArgs &a, int64_t max_tiles, hipStream_t s);
hipError_t launch_join_mfma_fp8_multi(const JoinArgs &proto, const JoinArgs *argv, const Tile *tiles, int nslots,
                                      hipStream_t s, int32_t *queue = nullptr);
// queries per workgroup tile of the MFMA join (16 per wave; want_waves > 1
// asks for the multi-wave workgroup)
int join_mfma_tile_q(int kcap, int dk, int want_waves);
// waves of the MFMA join's split mode for (kcap, dk): 1 if it does not fit
int join_mfma_split_waves(int kcap, int dk);
// Operand width of the MFMA join for dimension d: the planes (centred rows,
// representatives, query rows) are zero padded to one of the compiled widths.
inline int mfma_dk(int d) { return d <= 32 ? 32 : d <= 64 ? 64 : d <= 96 ? 96 : d <= 128 ? 128 : 256; }

// Fused launches over the tiles of many instances: slot b runs tile tiles[b]
// of instance tiles[b].pad with arguments argv[tiles[b].pad]; proto carries
// the template selectors (kcap, exact, dk) shared by all those instances.
// queue (eight device ints, zero at launch): the persistent form -- as many
// single-wave workgroups as the device holds take slots from the per-XCD
// counters until none are left (multi-wave MFMA workgroups keep one
// workgroup per slot).
hipError_t launch_join_multi(const JoinArgs &proto, int tq, const JoinArgs *argv, const Tile *tiles, int nslots,
                             hipStream_t s, int32_t *queue = nullptr);
hipError_t launch_join_mfma_f32_multi(const JoinArgs &proto, const JoinArgs *argv, const Tile *tiles, int nslots,
                                      hipStream_t s, int32_t *queue = nullptr);
hipError_t launch_join_mfma_bf16_multi(const JoinArgs &proto, int terms, const JoinArgs *argv, const Tile *tiles,
                                       int nslots, hipStream_t s, int32_t *queue = nullptr);
// workgroups of a persistent launch of `kernel`: what the device holds at
// once (at least one per compute unit), at most nslots
int64_t queue_grid(const void *kernel, int block, size_t lds, int64_t nslots);
