/*
 * knnj.h -- C ABI of the MI355X exact kNN-join engine (libknnj.so).
 *
 * The reference (gychen1991/knn_sweet) has no library API: its join is the
 * program `knnjoin` (src/knnjoin.cu:983-1172), which reads
 *   query_nb source_nb dim qrep_nb srep_nb K qfile sfile
 * from argv (knnjoin.cu:990-1001), builds a clustered index of both sets
 * (clusterReps, knnjoin.cu:166-411), bounds and prunes cluster pairs
 * (RepsUpperBound / FilterReps, knnjoin.cu:459-525), scans the survivors
 * (KNNQuery_base / KNNQuery + final, knnjoin.cu:762-979) and leaves
 * IndexDist knearest[query_nb][K] (include/common8.h:37-40) on the device.
 * The entry points below replace that program's compute path; each cites the
 * part it replaces.
 *
 * Conventions
 *  - Matrices are row-major FP32, row stride `ld` floats (ld >= d).
 *  - Output row q belongs to the q-th row of Q.  Each row lists the k
 *    reference rows with the smallest (squared distance, row index) keys,
 *    ascending; the squared distance is the binary64 value
 *    sum_j ((double)q_j - (double)r_j)^2 accumulated in ascending j without
 *    fused multiply-add, so ties and near-ties are ordered reproducibly.
 *    out_dist holds (float)sqrt of that value, or the value itself rounded
 *    to float when knnj_opts.out_squared is set.
 *  - Output pointers are device pointers when Q.on_device is set, host
 *    pointers otherwise.  Inputs may be host or device independently.
 *  - Every call is synchronous and returns a status; knnj_last_error() gives
 *    the message of the calling thread's last failure.  No exception crosses
 *    the ABI.  A knnj_ctx / knnj_index is used by one host thread at a time;
 *    several contexts may share a device.
 */
#ifndef KNNJ_H
#define KNNJ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  KNNJ_OK = 0,
  KNNJ_EINVAL = 1,       /* bad argument: k < 1, k > |R|, d < 1, null pointer, non-finite input */
  KNNJ_ENOMEM = 2,       /* device memory or the context's memory budget exhausted */
  KNNJ_EDEVICE = 3,      /* HIP runtime error */
  KNNJ_EUNSUPPORTED = 4, /* valid request outside what this build implements (e.g. k > 1024) */
  KNNJ_EOVERFLOW = 5     /* candidate list overflowed even at the largest capacity */
} knnj_status;

typedef enum {
  KNNJ_PREC_AUTO = 0, /* BF16 filter for d >= 64 real-valued data, FP32 otherwise */
  KNNJ_PREC_FP32 = 1, /* plain FP32 distance tiles */
  KNNJ_PREC_BF16 = 2, /* BF16 MFMA filter, compensated: operands split into bf16 hi + lo,
                         three products; + exact rerank (implies the MFMA kernel) */
  KNNJ_PREC_FP8 = 3,  /* FP8 (e4m3) MFMA filter, 32 < d <= 128: centred rows scaled per row
                         by a power of two, one fp8 product per term, with its rigorous
                         (wide) error bound + exact rerank (implies the MFMA kernel) */
  KNNJ_PREC_BF16X1 = 4 /* BF16 MFMA filter with a single product per term (wider error
                          band, larger candidate lists) + exact rerank */
} knnj_precision;

typedef enum {
  KNNJ_KERNEL_AUTO = 0, /* MFMA for 32 <= d <= 256, VALU otherwise */
  KNNJ_KERNEL_VALU = 1, /* direct-difference FP32 tiles on the vector ALUs (any d) */
  KNNJ_KERNEL_MFMA = 2  /* GEMM-form tiles on the matrix cores (d <= 256) */
} knnj_kernel;

typedef struct knnj_ctx knnj_ctx;
typedef struct knnj_index knnj_index;

typedef struct {
  const float *data;
  int64_t n;         /* rows */
  int32_t d;         /* dimension */
  int64_t ld;        /* row stride in floats, >= d (0 means d) */
  int32_t on_device; /* 1: `data` is a device pointer on the context's device */
} knnj_matrix;

/* Options; a NULL pointer or zero fields select tuned defaults. */
typedef struct {
  int32_t m_q;         /* query representatives (reference: qrep_nb, knnjoin.cu:993) */
  int32_t m_r;         /* reference representatives (srep_nb, knnjoin.cu:994) */
  int32_t trials;      /* representative draws per side (reference: 10, knnjoin.cu:195) */
  int32_t precision;   /* knnj_precision */
  uint64_t seed;       /* representative sampling seed (reference: 2015, knnjoin.cu:199) */
  int32_t out_squared; /* 1: out_dist holds squared distances */
  int32_t out_aos;     /* 1: out_idx is IndexDist{int32 index; float dist}[nQ][k]
                          (include/common8.h:37-40); out_dist is then ignored */
  int32_t brute_force; /* 1: one cluster per side (no pruning) */
  int32_t tile_q;      /* queries per wave tile of the VALU kernel, 0 = auto */
  int32_t kernel;      /* knnj_kernel: distance-tile implementation, 0 = auto */
  int32_t refine_iters;/* Lloyd steps moving each representative to its members' mean
                          (0 = default: 1 per join, none per knnj_join_batch instance, 2 for
                          knnj_index_build; negative = keep the sampled points) */
  int32_t xcd_tiles;   /* 1: place all tiles of a query cluster on one XCD (shared L2) */
  int32_t tile_waves;  /* MFMA kernel: waves per workgroup tile, 0 = auto, 1 = one 16-query
                          wave, > 1 = the most waves whose LDS fits (2 or 4), walking the
                          reference clusters in lockstep */
  int32_t tile_split;  /* MFMA kernel: 0 = auto (joins with few tiles, e.g. small query
                          batches), 1 = off, > 1 = on: the waves of a workgroup share one
                          16-query tile, each taking every W-th block of its cluster list */
  int32_t split_rounds;/* representative selection: rounds that replace sampled representatives
                          by the farthest members of the widest clusters (0 = default: none
                          per join, 16 for knnj_index_build; negative = none) */
  int32_t autotune;    /* knnj_join: 1 = on the first call of a problem shape (d, k, log2 of
                          both sizes, self-join) try the representative count x Lloyd step x
                          VALU tile candidates the other fields leave open, keep the fastest
                          for this context and use it for later calls of that shape */
  int32_t graph;       /* knnj_join_batch: 1 = capture the index builds and fused joins as a
                          HIP graph and replay it while the plan (every pointer, size and
                          decision) is unchanged; the input scan and overflow re-runs stay
                          outside it */
  int32_t kcap;        /* candidate capacity per query of the first pass: 0 = auto, else a
                          power of two in [64, 1024], >= k (queries whose candidate lists
                          overflow are re-run at larger capacities, so any value gives the
                          same answer) */
  int32_t kcap_mfma_max; /* largest capacity of the MFMA kernel's overflow re-runs before the
                          VALU kernel at 1024 takes the rest: 0 = 512, else a power of two in
                          [64, 512] */
  int32_t dedup;       /* queries with bit-identical coordinates share one answer: 0 = auto
                          (integer-mode joins on the vector ALUs), 1 = off, 2 = on: the join
                          runs once per distinct query and copies the rows to the others */
  int32_t filter_stats;/* 1: measure the filter's error against binary64 on every re-scored
                          candidate (knnj_stats.filter_err_max / filter_bound_max /
                          filter_err_ratio; a diagnostic: it re-scores each list once more) */
  int32_t work_queue;  /* knnj_join_batch: 1 = each fused launch of single-wave workgroups
                          runs persistent -- as many workgroups as the device holds take tile
                          slots from eight per-XCD device counters until none are left --
                          instead of one workgroup per slot (0); the batch autotuner tries it
                          when this is 0 */
  int32_t rep_scale;   /* representatives per side where m_r / m_q are 0: rep_scale (1..8) times
                          the default 2 sqrt(n), at most 4096 (0 = 1; knnj_join_batch: per
                          instance, and its autotuner tries 2 when this is 0) */
  int32_t reserved[8];
} knnj_opts;

/* Device-side timings (hipEvents on the context's stream) and counters. */
typedef struct {
  double t_total_ms;        /* whole call, including host<->device copies */
  double t_index_ms;        /* representative sampling, assignment, sorting, permuted copies */
  double t_bounds_ms;       /* theta_A and surviving cluster-pair lists */
  double t_join_ms;         /* distance tiles, k-select and rerank (the join proper) */
  double t_fallback_ms;     /* overflow re-run, if any */
  int64_t n_dist_evals;     /* executed (query, reference) distance evaluations */
  double pruning_ratio;     /* 1 - n_dist_evals / (|Q| |R|) */
  int64_t n_cluster_pairs;  /* surviving (query cluster, reference cluster) pairs */
  int64_t n_overflow;       /* queries re-run at a larger candidate capacity */
  int32_t exact_mode;       /* 1: integer inputs, FP32 keys are exact */
  int32_t precision_used;   /* knnj_precision actually run */
  int32_t m_q, m_r;         /* representative counts used */
  int32_t tile_q;           /* queries per wave tile used */
  int32_t kcap;             /* candidate capacity per query used */
  int64_t n_tiles;          /* query tiles launched */
  double filter_max_abs_err;/* largest |list key - binary64 distance^2| over re-scored
                               candidates (VALU: the FP32 filter error; MFMA: the key is the
                               upper bound s_f + E, so this is the filter error plus the
                               bound's width; filter_err_max below is the error alone) */
  int32_t kernel_used;      /* knnj_kernel actually run */
  int32_t n_launches;       /* join kernel launches (knnj_join_batch: fused launches) */
  int64_t n_merges;         /* k-select buffer merges */
  int64_t n_appended;       /* candidates that passed a query's threshold */
  int32_t tuned;            /* autotune: 0 off, else 1 + the candidate the context chose */
  int32_t graph;            /* knnj_join_batch: 0 eager, 1 captured this call, 2 replayed */
  int32_t rerun_stages;     /* overflow re-run stages executed (0: none; each stage doubles the
                               capacity on the MFMA kernel, the last one is the VALU kernel at 1024) */
  int32_t queue_launches;   /* knnj_join_batch: fused launches run in the persistent work-queue form */
  double filter_err_max;    /* with opts.filter_stats: the filter value's own error, largest
                               |s_f - s64| over re-scored candidates, s_f the filter's squared
                               distance and s64 the binary64 one (0 in exact mode) */
  double filter_bound_max;  /* largest rigorous error bound E attached to a re-scored candidate */
  double filter_err_ratio;  /* largest |s_f - s64| / E over them: <= 1 when the bound holds */
  int64_t n_distinct_queries; /* queries the join ran for (|Q| unless duplicates were folded;
                                 0 from knnj_join_batch) */
} knnj_stats;

/* One independent join of a batch (knnj_join_batch). */
typedef struct {
  knnj_matrix R, Q;
  int32_t k;
  int32_t *out_idx;
  float *out_dist;
} knnj_instance;

/* Library version string. */
const char *knnj_version(void);

/* Thread-local message of the last failed call on this thread ("" if none). */
const char *knnj_last_error(void);

/* Context on HIP device `device`: stream, events, grow-only workspace.
 * Replaces the reference's implicit CUDA context and its warm-up thread
 * (knnjoin.cu:980-986). */
knnj_status knnj_ctx_create(int device, knnj_ctx **out);
void knnj_ctx_destroy(knnj_ctx *ctx);

/* Caps the context's workspace in bytes (0 = no cap).  A call whose plan
 * needs more returns KNNJ_ENOMEM before touching the device. */
knnj_status knnj_ctx_set_budget(knnj_ctx *ctx, int64_t bytes);

/* Exact kNN join of Q against R: the whole reference program
 * (knnjoin.cu:1043-1141: clusterReps, RepsUpperBound, FilterReps, the
 * replist sort and KNNQuery_base / KNNQuery+final), results in original
 * query order.  out_idx/out_dist: [Q.n][k]. */
knnj_status knnj_join(knnj_ctx *ctx, knnj_matrix R, knnj_matrix Q, int32_t k,
                      const knnj_opts *opts, int32_t *out_idx, float *out_dist,
                      knnj_stats *stats);

/* Many independent joins (sizes, d and k may differ; one opts for all) in
 * one call.  The reference has no batched form: each instance runs the
 * whole program of knnj_join (knnjoin.cu:983-1172), with every instance's
 * index build queued back to back on the context's stream and the join
 * tiles of all instances that share a kernel instantiation (kernel,
 * precision, candidate capacity, exact mode, width class) executed by one
 * launch, each tile carrying its instance number.  Results are those of
 * knnj_join on each instance alone.  stats: the plan of the first
 * non-empty instance, counters summed over all, t_index_ms covering index
 * builds, bounds and tile lists. */
knnj_status knnj_join_batch(knnj_ctx *ctx, int32_t n_inst, const knnj_instance *inst,
                            const knnj_opts *opts, knnj_stats *stats);

/* ---- Introspection of the last join (tests and diagnostics) -------------
 * After knnj_join or knnj_index_query returns, the context still holds that
 * join's cluster index and bounds until its next call: the parts the
 * reference keeps in R2all_static / R2all_dyn (include/common8.h:50-69) and
 * in its replists (knnjoin.cu:506-525, 1076-1094).  They let a caller check
 * the pruning invariants the join's exactness rests on (radius = largest
 * member distance, theta_A >= every member query's true k-th distance, no
 * true neighbour in a pruned cluster pair). */
typedef struct {
  int64_t n_r, n_q;    /* rows of each side */
  int32_t d, dpad;     /* dimension; row stride of the representative rows */
  int32_t m_r, m_q;    /* clusters per side */
  int32_t k;
  int32_t self_join;   /* 1: the query side is the reference side's index */
  int32_t valid;       /* 0: the context has run no single join yet */
  int32_t reserved[5];
} knnj_plan_desc;

typedef enum {
  KNNJ_PART_R_REPS = 0,    /* float [m_r][dpad]: representative rows (zero padded) */
  KNNJ_PART_R_RADIUS = 1,  /* double [m_r]: largest member distance to the representative */
  KNNJ_PART_R_OFFSETS = 2, /* int32 [m_r + 1]: cluster b holds positions off[b] .. off[b+1]-1 */
  KNNJ_PART_R_PERM = 3,    /* int32 [n_r]: position -> original row */
  KNNJ_PART_R_PDIST = 4,   /* float [n_r]: distance to the representative, ascending per cluster */
  KNNJ_PART_Q_REPS = 5,    /* the same five parts of the query side */
  KNNJ_PART_Q_RADIUS = 6,
  KNNJ_PART_Q_OFFSETS = 7,
  KNNJ_PART_Q_PERM = 8,
  KNNJ_PART_Q_PDIST = 9,
  KNNJ_PART_THETA = 10,     /* double [m_q]: theta_A, an upper bound of the k-th neighbour
                               distance of every query of cluster A */
  KNNJ_PART_PAIR_COUNT = 11,/* int32 [m_q]: surviving reference clusters of A */
  KNNJ_PART_PAIR_LIST = 12, /* int32 [m_q][m_r]: their ids (first count entries of row A) */
  KNNJ_PART_PAIR_LB = 13    /* float [m_q][m_r]: lower bound of d(a, s) over the members s of
                               each listed cluster, ascending along the list */
} knnj_plan_part;

knnj_status knnj_ctx_last_plan(knnj_ctx *ctx, knnj_plan_desc *out);
/* Copies part `part` of the last join's plan to host memory `dst` of
 * `bytes` bytes (exactly the part's size, see knnj_plan_part). */
knnj_status knnj_ctx_copy_plan_part(knnj_ctx *ctx, int32_t part, void *dst, int64_t bytes);

/* ---- Resident reference index (the serving path) ----------------------
 * The reference builds R's clusters inside every run (knnjoin.cu:1043-1068:
 * clusterReps and the sorts that follow).  A knnj_index keeps that half of
 * the program -- representatives, radii, the cluster-sorted rows and, for
 * the MFMA kernel, the centred operand rows -- resident in HBM, so that
 * query batches pay only for their own side (knnjoin.cu:1069-1141). */
typedef struct knnj_index knnj_index;

typedef struct {
  int64_t n;          /* reference rows */
  int32_t d;
  int32_t m_r;        /* reference clusters */
  int32_t kernel;     /* knnj_kernel fixed at build */
  int32_t precision;  /* knnj_precision of the operand planes fixed at build */
  int32_t device;
  int32_t reserved0;
  int64_t bytes;      /* device bytes held */
  double t_build_ms;
} knnj_index_desc;

/* Builds the index of R on ctx's device (opts: m_r, trials, seed,
 * refine_iters, kernel, precision).  R may be freed afterwards. */
knnj_status knnj_index_build(knnj_ctx *ctx, knnj_matrix R, const knnj_opts *opts, knnj_index **out);
void knnj_index_destroy(knnj_index *index);
knnj_status knnj_index_info(const knnj_index *index, knnj_index_desc *info);

/* Exact kNN of Q against the indexed R: the result of knnj_join(R, Q, k)
 * with the index's build options.  Any number of contexts on the index's
 * device may query one index concurrently (it is read-only). */
knnj_status knnj_index_query(knnj_ctx *ctx, const knnj_index *index, knnj_matrix Q, int32_t k,
                             const knnj_opts *opts, int32_t *out_idx, float *out_dist,
                             knnj_stats *stats);

/* Allocation-free variant of knnj_index_query, with the same arguments and
 * results.  It has no counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Every device buffer of the call
 * -- staged host rows, the query side's cluster index, bounds, tile lists,
 * counters, staged host outputs -- is carved from a fixed scratch array of
 * KNNJ_NOALLOC_SCRATCH_BYTES per device, which the library holds from load to
 * process end; the call makes no hipMalloc / hipFree.  A query whose plan
 * needs more returns KNNJ_ENOMEM before any index work (a 4,096-row batch at
 * d = 96, k = 50 needs about 10 MB).  All such calls on one device share that
 * array and the last call's contents stay in it: calls that overlap on one
 * GPU race, so serialize them. */
#define KNNJ_NOALLOC_SCRATCH_BYTES (64ll << 20)
knnj_status knnj_index_query_noalloc(knnj_ctx *ctx, const knnj_index *index, knnj_matrix Q, int32_t k,
                                     const knnj_opts *opts, int32_t *out_idx, float *out_dist,
                                     knnj_stats *stats);

/* ---- Serving: continuous batching over a resident index -----------------
 * The reference answers one query set per process (knnjoin.cu:983-1172;
 * its small-batch path knnjoin.cu:1103-1127).  A knnj_server owns `slots`
 * worker threads, each with its own context (HIP stream and workspace).
 * Requests (host rows, 1..max_batch_rows queries) are queued on arrival; a
 * free worker fuses the oldest waiting requests while their rows fit
 * max_batch_rows (the slot's workspace is sized for that many rows: the
 * memory-footprint admission rule) and runs them as one knnj_index_query,
 * so batches of different slots overlap on the device.  Every query's
 * answer is its exact k-NN, the same alone and in any mix. */
typedef struct knnj_server knnj_server;

typedef struct {
  int32_t k;
  int32_t slots;           /* concurrent batches (worker threads / streams), 1..16 */
  int64_t max_batch_rows;  /* rows fused into one batch (and the largest request) */
  int64_t slot_budget;     /* device workspace bytes per slot, 0 = no cap */
  knnj_opts opts;          /* query options (kernel and precision come from the index) */
} knnj_server_config;

typedef struct {
  int32_t status;          /* knnj_status of the request */
  int32_t batch_rows;      /* rows of the fused batch it ran in */
  int32_t batch_requests;  /* requests in that batch */
  int32_t slot;
  double queue_ms;         /* submit -> its batch started */
  double latency_ms;       /* submit -> results written */
  double t_submit_ms;      /* submit, batch start and batch end on the server's clock */
  double t_start_ms;       /*   (milliseconds since knnj_server_create), so callers can */
  double t_end_ms;         /*   see which batches of different slots overlapped */
  char message[160];       /* error text when status != KNNJ_OK */
} knnj_request_info;

knnj_status knnj_server_create(const knnj_index *index, const knnj_server_config *cfg, knnj_server **out);
/* Queues Q (host, n rows of stride ld floats, 0 = d); out_idx/out_dist:
 * host [n][k], written before the request completes.  Thread-safe. */
knnj_status knnj_server_submit(knnj_server *srv, const float *Q, int64_t n, int64_t ld, int32_t *out_idx,
                               float *out_dist, uint64_t *ticket);
/* Blocks until the request completes; returns its status.  Each ticket is
 * waited for once: a second wait on the same ticket (also one that races
 * the first from another thread) returns KNNJ_EINVAL. */
knnj_status knnj_server_wait(knnj_server *srv, uint64_t ticket, knnj_request_info *info);
/* Fused batches run and rows served so far. */
knnj_status knnj_server_counters(knnj_server *srv, int64_t *batches, int64_t *rows);
/* Serves every queued request, then stops the workers. */
void knnj_server_destroy(knnj_server *srv);

#ifdef __cplusplus
}
#endif

#endif /* KNNJ_H */
