// api.hip -- the C ABI (include/knnj.h) and the host-side planner.
//
// The planner replaces the reference driver main() (src/knnjoin.cu:983-1172):
// argument checks instead of print-and-continue (knnjoin.cu:38-41), one
// grow-only device workspace per context instead of 3*m_Q + m_R cudaMallocs
// that are never freed (knnjoin.cu:305-309, 347-349), one stream with hipEvent
// stage timers instead of CLOCK_REALTIME around cudaDeviceSynchronize
// (common8.h:83-90), and device-side sorts instead of the OpenMP
// cudaMemcpy round trips (knnjoin.cu:380-400, 1076-1094).
//
// A join is planned as an `Inst`: decide() picks the kernel, precision and
// capacities from the input scan; alloc_inst() lays out its device buffers
// through an Alloc (named context workspace for knnj_join, one carved arena
// for knnj_join_batch); enqueue_index() queues the index build, centring,
// bounds and tile list; fill_args() fills the join kernel arguments.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/knnj.h"
#include "knnj_common.h"
#include "knnj_kernels.h"

namespace {

thread_local std::string g_err;

void set_err(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

// Prefixes the current error message.
void prefix_err(const char *fmt, int v) {
  char buf[64];
  snprintf(buf, sizeof(buf), fmt, v);
  g_err = std::string(buf) + g_err;
}

struct DevBuf {
  void *p = nullptr;
  size_t n = 0;
};

}  // namespace

struct knnj_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::map<std::string, DevBuf> ws;
  int64_t budget = 0;
  hipEvent_t ev[10] = {};
  // knnj_join_batch: streams the instances' index builds are spread over
  // (default priority, created on first use) and their completion events
  static constexpr int kAux = 32;  // created on first use; batches spread over aux_streams() of them
  hipStream_t aux[kAux] = {};
  hipEvent_t aux_ev[kAux + 1] = {};
  // autotune: problem shape -> the chosen candidate: its index in the first
  // phase's list, in the second phase's variants of it (-1: none), and its
  // number (1 + index over both lists).  The candidates are rebuilt from
  // each call's own options, so per-call fields (output form, statistics)
  // are the caller's.
  struct Tuned {
    int b1, b2, id;
  };
  std::map<std::vector<int64_t>, Tuned> tuned;
  // knnj_join_batch with opts.graph: the captured index builds + fused joins
  // of the last plan, and that plan's signature
  hipGraphExec_t batch_exec = nullptr;
  std::vector<unsigned char> batch_sig;
  // this device's address of the allocation-free entry point's scratch array
  void *scratch = nullptr;
  // the last single join's index and bounds (knnj_ctx_last_plan): device
  // pointers into this context's workspace or a resident index
  struct Plan {
    bool valid = false;
    int64_t n_r = 0, n_q = 0;
    int d = 0, dpad = 0, m_r = 0, m_q = 0, k = 0, self = 0;
    SideIndex Ri{}, Qi{};
    const double *theta = nullptr;
    const int32_t *plist = nullptr, *pcnt = nullptr;
    const float *plb = nullptr;
  } plan;
};

#define HIPCHK(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) {                                                               \
      set_err("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
      return KNNJ_EDEVICE;                                                                \
    }                                                                                     \
  } while (0)

#define TRY(expr)                 \
  do {                            \
    knnj_status s_ = (expr);      \
    if (s_ != KNNJ_OK) return s_; \
  } while (0)

namespace {

size_t ws_total(const knnj_ctx *c) {
  size_t t = 0;
  for (auto &kv : c->ws) t += kv.second.n;
  return t;
}

// Device buffer of at least `bytes` under `name`; grows, never shrinks.
knnj_status ws_get(knnj_ctx *c, const std::string &name, size_t bytes, void **out) {
  bytes = std::max<size_t>(bytes, 256);
  DevBuf &b = c->ws[name];
  if (b.n >= bytes) {
    *out = b.p;
    return KNNJ_OK;
  }
  const size_t cur = ws_total(c) - b.n;
  if (c->budget > 0 && (int64_t)(cur + bytes) > c->budget) {
    set_err("workspace budget exceeded: %s needs %zu bytes, %zu in use, budget %lld", name.c_str(), bytes, cur,
            (long long)c->budget);
    return KNNJ_ENOMEM;
  }
  if (b.p) {
    hipFree(b.p);
    b.p = nullptr;
    b.n = 0;
  }
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    set_err("hipMalloc(%zu) for %s failed: %s", bytes, name.c_str(), hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? KNNJ_ENOMEM : KNNJ_EDEVICE;
  }
  b.n = bytes;
  *out = b.p;
  return KNNJ_OK;
}

template <class T>
knnj_status ws(knnj_ctx *c, const std::string &name, size_t count, T **out) {
  void *p = nullptr;
  knnj_status s = ws_get(c, name, count * sizeof(T), &p);
  *out = (T *)p;
  return s;
}

// Buffer source: the named context workspace, or consecutive 256-byte
// aligned carves of one arena.  An arena Alloc with a null base is a sizing
// pass: it hands out null pointers and only advances `off`.  cap > 0 bounds
// an arena (the allocation-free entry point's fixed scratch): a carve past
// it fails before anything is handed out.
struct Alloc {
  knnj_ctx *c = nullptr;
  bool arena = false;
  char *base = nullptr;
  size_t off = 0;
  size_t cap = 0;
  bool no_alloc = false;  // the caller cannot allocate: no diagnostic buffers either
  std::vector<void *> *owned = nullptr;  // set: a hipMalloc per buffer, recorded here
  template <class T>
  knnj_status get(const std::string &name, size_t count, T **out) {
    if (owned) {
      void *p = nullptr;
      const hipError_t e = hipMalloc(&p, std::max<size_t>(count * sizeof(T), 256));
      if (e != hipSuccess) {
        set_err("hipMalloc(%zu) for %s failed: %s", count * sizeof(T), name.c_str(), hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? KNNJ_ENOMEM : KNNJ_EDEVICE;
      }
      owned->push_back(p);
      off += count * sizeof(T);
      *out = (T *)p;
      return KNNJ_OK;
    }
    if (!arena) return ws(c, name, count, out);
    const size_t bytes = (std::max<size_t>(count * sizeof(T), 1) + 255) & ~(size_t)255;
    if (cap > 0 && off + bytes > cap) {
      set_err("the call needs more than the %zu-byte scratch of the allocation-free entry point (%s: %zu bytes "
              "at offset %zu)", cap, name.c_str(), bytes, off);
      return KNNJ_ENOMEM;
    }
    *out = base ? (T *)(base + off) : nullptr;
    off += bytes;
    return KNNJ_OK;
  }
};

int pow2_at_least(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

int default_reps(int64_t n) {
  int64_t m = (int64_t)std::llround(2.0 * std::sqrt((double)n));
  return (int)std::max<int64_t>(1, std::min<int64_t>({m, (int64_t)KNNJ_MAX_REPS, n}));
}

knnj_status check_matrix(const knnj_matrix &M, const char *name) {
  if (M.n < 0) {
    set_err("%s.n < 0", name);
    return KNNJ_EINVAL;
  }
  if (M.d < 1) {
    set_err("%s.d < 1", name);
    return KNNJ_EINVAL;
  }
  if (M.ld != 0 && M.ld < M.d) {
    set_err("%s.ld < %s.d", name, name);
    return KNNJ_EINVAL;
  }
  if (M.n > 0 && !M.data) {
    set_err("%s.data is NULL", name);
    return KNNJ_EINVAL;
  }
  if (M.n > (int64_t)INT32_MAX - 1) {
    set_err("%s.n exceeds 2^31-2 rows", name);
    return KNNJ_EUNSUPPORTED;
  }
  return KNNJ_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
  return ms;
}

knnj_status check_join_args(const knnj_matrix &R, const knnj_matrix &Q, int32_t k, const knnj_opts &o,
                            const void *out_idx, const void *out_dist) {
  TRY(check_matrix(R, "R"));
  TRY(check_matrix(Q, "Q"));
  if (R.d != Q.d) {
    set_err("R.d (%d) != Q.d (%d)", R.d, Q.d);
    return KNNJ_EINVAL;
  }
  if (k < 1 || k > R.n) {
    set_err("k = %d outside [1, |R| = %lld]", k, (long long)R.n);
    return KNNJ_EINVAL;
  }
  if (k > KNNJ_MAX_K) {
    set_err("k = %d above the supported maximum %d", k, KNNJ_MAX_K);
    return KNNJ_EUNSUPPORTED;
  }
  if (Q.n > 0 && (!out_idx || (!o.out_aos && !out_dist))) {
    set_err("output pointer is NULL");
    return KNNJ_EINVAL;
  }
  if (o.precision < 0 || o.precision > KNNJ_PREC_BF16X1) {
    set_err("unknown precision %d", o.precision);
    return KNNJ_EINVAL;
  }
  if (o.precision == KNNJ_PREC_FP8 && (R.d <= 32 || R.d > 128)) {
    set_err("the FP8 filter supports 32 < d <= 128 (d = %d)", R.d);
    return KNNJ_EUNSUPPORTED;
  }
  if (o.tile_q != 0 && o.tile_q != 1 && o.tile_q != 2 && o.tile_q != 4 && o.tile_q != 8) {
    set_err("tile_q must be 0, 1, 2, 4 or 8");
    return KNNJ_EINVAL;
  }
  if (o.rep_scale < 0 || o.rep_scale > 8) {
    set_err("rep_scale must be in [0, 8]");
    return KNNJ_EINVAL;
  }
  if (o.tile_split < 0 || o.tile_split > 16) {
    set_err("tile_split must be in [0, 16]");
    return KNNJ_EINVAL;
  }
  if (o.tile_waves < 0 || o.tile_waves > 16) {
    set_err("tile_waves must be in [0, 16]");
    return KNNJ_EINVAL;
  }
  if (o.kernel < 0 || o.kernel > KNNJ_KERNEL_MFMA) {
    set_err("unknown kernel %d", o.kernel);
    return KNNJ_EINVAL;
  }
  if ((o.precision == KNNJ_PREC_BF16 || o.precision == KNNJ_PREC_BF16X1 || o.precision == KNNJ_PREC_FP8) &&
      o.kernel == KNNJ_KERNEL_VALU) {
    set_err("the BF16 and FP8 filters run on the MFMA kernel only");
    return KNNJ_EINVAL;
  }
  const auto pow2 = [](int v) { return v > 0 && (v & (v - 1)) == 0; };
  if (o.kcap != 0 && (!pow2(o.kcap) || o.kcap < 64 || o.kcap > KNNJ_MAX_K || o.kcap < k)) {
    set_err("kcap must be 0 or a power of two in [64, %d] and >= k", KNNJ_MAX_K);
    return KNNJ_EINVAL;
  }
  if (o.dedup < 0 || o.dedup > 2) {
    set_err("dedup must be 0 (auto), 1 (off) or 2 (on)");
    return KNNJ_EINVAL;
  }
  if (o.kcap_mfma_max != 0 && (!pow2(o.kcap_mfma_max) || o.kcap_mfma_max < 64 || o.kcap_mfma_max > 512)) {
    set_err("kcap_mfma_max must be 0 or a power of two in [64, 512]");
    return KNNJ_EINVAL;
  }
  return KNNJ_OK;
}

// Buffers of one side's cluster index.
struct SideBufs {
  float *reps = nullptr, *d2r = nullptr, *pdist = nullptr, *px = nullptr, *rn = nullptr;
  u64 *radbits = nullptr, *gkeys = nullptr;
  int32_t *off = nullptr, *perm = nullptr, *cl = nullptr, *cnt = nullptr, *fill = nullptr, *cand = nullptr,
          *rep_rows = nullptr;
  double *partial = nullptr;
  __bf16 *rb = nullptr;
  int nblk = 0;
};

}  // namespace

// A resident reference-side index (knnj_index_build): the cluster index of
// R, its permuted rows and, for the MFMA kernel, the centred rows.
struct knnj_index {
  int device = 0;
  int d = 0;
  int64_t n = 0;
  knnj_opts o{};     // build options
  u32 flags[3] = {};  // input scan of R: non-finite, max |x| bits, non-integral
  int kind = 0, dpad = 0, dk = 0, m_r = 0;
  bool mfma = false, bf16 = false, bf16x1 = false, fp8 = false;
  SideIndex Ri{};
  void *rc = nullptr;
  float *nr = nullptr, *reps32 = nullptr, *rn32 = nullptr, *rinv = nullptr;
  std::vector<void *> owned;
  size_t bytes = 0;
  double t_build_ms = 0.0;
};

namespace {

// One planned join.
struct Inst {
  int id = 0;  // position in a batch (the tiles' pad field)
  const knnj_index *pre = nullptr;  // reference side prebuilt (knnj_index_query)
  // inputs: device rows, or host rows still to be copied
  const float *dR = nullptr, *dQ = nullptr;
  const float *hR = nullptr, *hQ = nullptr;
  int64_t nR = 0, nQ = 0, ldR = 0, ldQ = 0;
  int d = 0, k = 0;
  bool self = false;
  knnj_opts o{};
  // outputs: the caller's pointers, and device staging when they are host
  int32_t *user_idx = nullptr;
  float *user_dist = nullptr;
  int32_t *out_idx = nullptr;
  float *out_dist = nullptr;
  bool out_host = false, aos = false;
  // decisions
  int kind = 0, kcap = 0, tq = 0, dpad = 0, dk = 0, m_r = 0, m_q = 0, trials = 10, refine = 1, waves = 1,
      split = 1;
  bool mfma = false, bf16 = false, bf16x1 = false, fp8 = false, exact = false, exact_valu = false,
       integral = false;
  bool use_mfma = false;
  int split_rounds = 0;  // split rounds in the representative selection
  bool dedup = false;     // duplicate queries answered once (Qu: the leaders' side)
  bool resident = false;  // building a resident index (knnj_index_build): full Lloyd means
  double span = 0.0;
  uint64_t seed = 2015;
  int64_t max_slots = 0;
  // device state
  SideBufs rb, qb;
  SideIndex Ri{}, Qi{};
  void *rc = nullptr;
  float *nr = nullptr, *reps32 = nullptr, *rn32 = nullptr, *rinv = nullptr;
  double *theta = nullptr;
  double *dab = nullptr;  // [m_q][m_r] binary64 representative distances (bounds scratch)
  int32_t *qseed = nullptr;  // [nQ] MFMA join: each query row's nearest reference representative
  __bf16 *qs_rb = nullptr;   // the query rows' assignment's own bf16 representatives
  float *qs_rn = nullptr;
  int32_t *plist = nullptr, *pcnt = nullptr;
  float *plb = nullptr;
  int32_t *lead = nullptr, *ucnt = nullptr, *uoff = nullptr, *uperm = nullptr;
  float *upx = nullptr, *updist = nullptr;
  Tile *tiles = nullptr, *ovf_tiles = nullptr, *ovf2_tiles = nullptr;
  int32_t *counters = nullptr;  // [0] slots, [1] overflow, [2] overflow 2, [3] max err bits, [4] overflow 3
  u64 *evals = nullptr;         // [0] evaluations, [1] merges, [2] appended
  bool batch_counters = false;  // counters and evals are slices of a batch's arrays, cleared by the batch
  JoinArgs args{};

  // exact flag of the kernel that runs the tiles
  int kernel_exact() const { return use_mfma ? (exact ? 1 : 0) : ((mfma ? exact_valu : exact) ? 1 : 0); }
  // kernel instantiation: joins with equal keys can share one launch
  std::tuple<int, int, int, int, int, int> launch_key() const {
    return std::make_tuple(use_mfma ? (fp8 ? 4 : bf16 ? (bf16x1 ? 3 : 2) : 1) : 0, kcap, kernel_exact(), tq,
                           use_mfma ? dk : 0, split);
  }
};

// Input-side fields of a join (no device work).
void describe(Inst &I, const knnj_matrix &R, const knnj_matrix &Q, int32_t k, const knnj_opts &o,
              int32_t *out_idx, float *out_dist) {
  I.d = R.d;
  I.k = k;
  I.o = o;
  I.nR = R.n;
  I.nQ = Q.n;
  const int64_t ldR = R.ld ? R.ld : R.d, ldQ = Q.ld ? Q.ld : Q.d;
  I.self = R.data == Q.data && R.n == Q.n && ldR == ldQ && R.on_device == Q.on_device;
  (R.on_device ? I.dR : I.hR) = R.data;
  I.ldR = ldR;
  (Q.on_device ? I.dQ : I.hQ) = Q.data;
  I.ldQ = ldQ;
  I.aos = o.out_aos != 0;
  I.out_host = !Q.on_device;
  I.user_idx = out_idx;
  I.user_dist = out_dist;
}

// Host rows (stride ld) to packed device rows (stride d): one linear copy
// when the rows are already packed (a strided copy of many short rows from
// pageable memory is far slower), a 2-D copy otherwise.
knnj_status copy_rows_h2d(float *dst, const float *src, int64_t n, int d, int64_t ld, hipStream_t s) {
  if (n <= 0) return KNNJ_OK;
  if (ld == d)
    HIPCHK(hipMemcpyAsync(dst, src, sizeof(float) * (size_t)n * d, hipMemcpyHostToDevice, s));
  else
    HIPCHK(hipMemcpy2DAsync(dst, sizeof(float) * d, src, sizeof(float) * ld, sizeof(float) * d, n,
                            hipMemcpyHostToDevice, s));
  return KNNJ_OK;
}

// Device copies of host inputs (rows packed to ld = d), then the H2D copies.
knnj_status stage_inputs(Alloc &A, Inst &I, hipStream_t s) {
  if (I.hR) {
    float *b;
    TRY(A.get("in_r", (size_t)std::max<int64_t>(I.nR, 1) * I.d, &b));
    if (b) TRY(copy_rows_h2d(b, I.hR, I.nR, I.d, I.ldR, s));
    I.dR = b;
    I.ldR = I.d;
  }
  if (I.self) {
    I.dQ = I.dR;
    I.ldQ = I.ldR;
  } else if (I.hQ) {
    float *b;
    TRY(A.get("in_q", (size_t)std::max<int64_t>(I.nQ, 1) * I.d, &b));
    if (b) TRY(copy_rows_h2d(b, I.hQ, I.nQ, I.d, I.ldQ, s));
    I.dQ = b;
    I.ldQ = I.d;
  }
  return KNNJ_OK;
}

// Kernel, precision and capacity choices from the input scan flags
// {non-finite, max |x| bits, non-integral}.
knnj_status decide(Inst &I, const u32 *flags) {
  if (flags[0]) {
    set_err("non-finite value in R or Q");
    return KNNJ_EINVAL;
  }
  const int d = I.d, k = I.k;
  const knnj_opts &o = I.o;
  float maxabs;
  std::memcpy(&maxabs, &flags[1], 4);
  I.span = 2.0 * (double)maxabs;
  I.integral = !flags[2];
  I.exact_valu = I.integral && (double)d * I.span * I.span < 16777216.0;
  const bool exact_mfma = I.integral && 4.0 * (double)d * I.span * I.span < 16777216.0;
  int kind = o.kernel;
  if (o.precision == KNNJ_PREC_BF16 || o.precision == KNNJ_PREC_BF16X1 || o.precision == KNNJ_PREC_FP8)
    kind = KNNJ_KERNEL_MFMA;
  if (kind == KNNJ_KERNEL_AUTO) kind = (d >= 32 && d <= 256) ? KNNJ_KERNEL_MFMA : KNNJ_KERNEL_VALU;
  if (kind == KNNJ_KERNEL_MFMA && d > 256) {
    set_err("the MFMA kernel supports d <= 256 (d = %d)", d);
    return KNNJ_EUNSUPPORTED;
  }
  I.kind = kind;
  I.mfma = kind == KNNJ_KERNEL_MFMA;
  I.bf16 = I.mfma && (o.precision == KNNJ_PREC_BF16 || o.precision == KNNJ_PREC_BF16X1 ||
                      (o.precision == KNNJ_PREC_AUTO && d >= 64 && !exact_mfma));
  I.bf16x1 = I.bf16 && o.precision == KNNJ_PREC_BF16X1;
  I.fp8 = I.mfma && o.precision == KNNJ_PREC_FP8;
  I.exact = I.mfma ? (!I.bf16 && !I.fp8 && exact_mfma) : I.exact_valu;
  I.dpad = (d + 3) & ~3;
  I.dk = I.mfma ? mfma_dk(d) : 0;
  if (I.pre) {
    // the index fixed the kernel and the operand planes of the reference rows
    const knnj_index &X = *I.pre;
    I.kind = X.kind;
    I.mfma = X.mfma;
    I.bf16 = X.bf16;
    I.bf16x1 = X.bf16x1;
    I.fp8 = X.fp8;
    I.exact = I.mfma ? (!I.bf16 && !I.fp8 && exact_mfma) : I.exact_valu;
    I.dk = X.dk;
  }
  const bool brute = o.brute_force != 0;
  // the default count scaled by rep_scale, within [1, min(n, KNNJ_MAX_REPS)]
  const int rs = std::max(1, o.rep_scale);
  auto scaled = [&](int64_t n) { return (int)std::min<int64_t>(n, (int64_t)rs * default_reps(n)); };
  int m_r = brute ? 1 : (o.m_r > 0 ? (int)std::min<int64_t>(o.m_r, I.nR) : scaled(I.nR));
  int m_q = brute ? 1 : (o.m_q > 0 ? (int)std::min<int64_t>(o.m_q, I.nQ) : scaled(I.nQ));
  I.m_r = std::max(1, std::min(m_r, KNNJ_MAX_REPS));
  I.m_q = std::max(1, std::min(m_q, KNNJ_MAX_REPS));
  if (I.self) I.m_q = I.m_r;
  if (I.pre) I.m_r = I.pre->m_r;
  I.trials = o.trials > 0 ? std::min(o.trials, 64) : 10;
  I.seed = o.seed ? o.seed : 2015;
  I.refine = o.refine_iters < 0 ? 0 : (o.refine_iters > 0 ? o.refine_iters : 1);
  I.split_rounds = o.split_rounds > 0 ? std::min(o.split_rounds, 16) : 0;  // joins: none by default
  int kcap = I.exact    ? pow2_at_least(std::max(k, 64))
             : I.fp8    ? pow2_at_least(std::max(4 * k, 64))
             : I.bf16x1 ? pow2_at_least(std::max(2 * k, 64))
                        : pow2_at_least(std::max(k + std::max(8, k / 8), 64));
  if (o.kcap > 0) kcap = o.kcap;  // the caller's first-pass capacity (>= k, checked)
  kcap = std::min(kcap, I.mfma ? 512 : KNNJ_MAX_K);
  if (kcap < k) kcap = KNNJ_MAX_K;  // only the VALU kernel holds k > 512
  I.kcap = kcap;
  I.use_mfma = I.mfma && kcap <= 512;
  // integer data repeats points: there the VALU join runs once per distinct query
  I.dedup = o.dedup == 2 || (o.dedup == 0 && I.exact && !I.use_mfma);
  I.waves = o.tile_waves > 1 ? o.tile_waves : 1;  // auto: one wave per tile
  I.split = 1;
  if (I.use_mfma && I.waves == 1 && o.tile_split != 1) {
    // few tiles (small query sets): W waves share each tile's cluster list
    const int sw = join_mfma_split_waves(kcap, I.dk);
    const bool want = o.tile_split > 1 || (o.tile_split == 0 && I.nQ / 16 + I.m_q < 1024);
    if (sw > 1 && want) I.split = I.waves = sw;
  }
  I.tq = I.use_mfma ? (I.split > 1 ? 16 : join_mfma_tile_q(kcap, I.dk, I.waves))
                    : (o.tile_q ? std::min(o.tile_q, join_tile_q(kcap)) : join_tile_q(kcap));
  // XCD-aware slots, empty ones included
  I.max_slots = (int64_t)(o.xcd_tiles ? 8 : 1) * (I.nQ / I.tq + I.m_q + 1);
  return KNNJ_OK;
}

knnj_status alloc_side(Alloc &A, const std::string &p, int64_t n, int m, int dpad, int d, int trials,
                       SideBufs &b) {
  const int64_t nn = std::max<int64_t>(n, 1);
  TRY(A.get(p + "reps", (size_t)m * dpad, &b.reps));
  TRY(A.get(p + "rad", (size_t)m, &b.radbits));
  TRY(A.get(p + "off", (size_t)m + 1, &b.off));
  TRY(A.get(p + "cnt", (size_t)m, &b.cnt));
  TRY(A.get(p + "fill", (size_t)m, &b.fill));
  TRY(A.get(p + "perm", (size_t)nn, &b.perm));
  TRY(A.get(p + "cl", (size_t)nn, &b.cl));
  TRY(A.get(p + "d2r", (size_t)nn, &b.d2r));
  TRY(A.get(p + "pdist", (size_t)nn, &b.pdist));
  TRY(A.get(p + "gkeys", (size_t)nn, &b.gkeys));
  TRY(A.get(p + "px", (size_t)nn * dpad, &b.px));
  TRY(A.get(p + "cand", (size_t)m * trials, &b.cand));
  TRY(A.get(p + "reprows", (size_t)m, &b.rep_rows));
  b.nblk = rep_score_blocks(m);
  TRY(A.get(p + "partial", (size_t)b.nblk * trials + 1, &b.partial));  // + the winning trial's slot
  const int dk = (d + 31) & ~31;
  if (d >= 16 && dk <= 256) {  // the MFMA assignment's bf16 representatives
    TRY(A.get(p + "repsbf16", (size_t)m * dk, &b.rb));
    TRY(A.get(p + "repsnrm", (size_t)m, &b.rn));
  }
  return KNNJ_OK;
}

// The MFMA join's seeds from an assignment of the query rows to the
// reference representatives (a join that builds both sides): each query row
// then starts at its nearest representative without searching its cluster
// list (measured on the d = 128 mixture, 2M x 200k, 4,096 representatives:
// the list search was a seventh of a tile).  Small reference sides keep the
// search, which is cheap there.  A self-join's rows already carry their
// assignment (the reference side's).  A query of a resident index without
// allocation keeps the search (its scratch array holds no assignment).
bool seed_by_assignment(const Inst &I) { return I.mfma && I.m_r >= 512 && !std::getenv("KNNJ_SEED_LIST"); }
bool use_qseed(const Inst &I, const Alloc &A) { return seed_by_assignment(I) && !I.self && !(I.pre && A.no_alloc); }

// Every device buffer of one join after decide(); `tiles` given: the join's
// slot range of a batch's shared tile array.
knnj_status alloc_inst(Alloc &A, Inst &I, Tile *tiles) {
  if (I.pre) {
    I.Ri = I.pre->Ri;
    I.rc = I.pre->rc;
    I.nr = I.pre->nr;
    I.reps32 = I.pre->reps32;
    I.rn32 = I.pre->rn32;
    I.rinv = I.pre->rinv;
  } else {
    TRY(alloc_side(A, "r_", I.nR, I.m_r, I.dpad, I.d, I.trials, I.rb));
  }
  if (!I.self) TRY(alloc_side(A, "q_", I.nQ, I.m_q, I.dpad, I.d, I.trials, I.qb));
  if (I.mfma && !I.pre) {
    char *rc;
    TRY(A.get("r_rc", (size_t)I.nR * I.dk * 4, &rc));
    I.rc = rc;
    // fp8 rows take nR dk bytes of it, their inverse scales the next 4 nR
    I.rinv = rc ? (float *)(rc + (size_t)I.nR * I.dk) : nullptr;
    TRY(A.get("r_nr", (size_t)I.nR, &I.nr));
    TRY(A.get("r_reps32", (size_t)I.m_r * I.dk, &I.reps32));
    TRY(A.get("r_rn32", (size_t)I.m_r, &I.rn32));
  }
  if (I.dedup) {
    const size_t nq = (size_t)std::max<int64_t>(I.nQ, 1);
    TRY(A.get("dd_lead", nq, &I.lead));
    TRY(A.get("dd_ucnt", (size_t)I.m_q, &I.ucnt));
    TRY(A.get("dd_uoff", (size_t)I.m_q + 1, &I.uoff));
    TRY(A.get("dd_perm", nq, &I.uperm));
    TRY(A.get("dd_pdist", nq, &I.updist));
    TRY(A.get("dd_px", nq * I.dpad, &I.upx));
  }
  TRY(A.get("theta", (size_t)I.m_q, &I.theta));
  TRY(A.get("dab", (size_t)I.m_q * I.m_r, &I.dab));
  if (use_qseed(I, A)) {
    TRY(A.get("qseed", (size_t)std::max<int64_t>(I.nQ, 1), &I.qseed));
    const int dk = (I.d + 31) & ~31;
    if (I.d >= 16 && dk <= 256) {  // the assignment's own bf16 representatives
      TRY(A.get("qs_rb", (size_t)I.m_r * dk, &I.qs_rb));
      TRY(A.get("qs_rn", (size_t)I.m_r, &I.qs_rn));
    }
  }
  TRY(A.get("plist", (size_t)I.m_q * I.m_r, &I.plist));
  TRY(A.get("plb", (size_t)I.m_q * I.m_r, &I.plb));
  TRY(A.get("pcnt", (size_t)I.m_q, &I.pcnt));
  if (tiles)
    I.tiles = tiles;
  else
    TRY(A.get("tiles", (size_t)I.max_slots, &I.tiles));
  TRY(A.get("ovf_tiles", (size_t)I.nQ, &I.ovf_tiles));
  TRY(A.get("ovf2_tiles", (size_t)I.nQ, &I.ovf2_tiles));
  if (!I.batch_counters) {
    TRY(A.get("counters", 8, &I.counters));
    TRY(A.get("evals", 4, &I.evals));
  }
  const size_t nout = (size_t)I.nQ * I.k;
  if (I.out_host) {
    TRY(A.get("out_idx", nout * (I.aos ? 2 : 1), &I.out_idx));
    if (!I.aos) TRY(A.get("out_dist", nout, &I.out_dist));
  } else {
    I.out_idx = I.user_idx;
    I.out_dist = I.user_dist;
  }
  return KNNJ_OK;
}

// The SideIndex view of one side's buffers (host-side only).
void fill_side_index(Inst &I, bool ref_side) {
  SideBufs &b = ref_side ? I.rb : I.qb;
  SideIndex &out = ref_side ? I.Ri : I.Qi;
  out.n = ref_side ? I.nR : I.nQ;
  out.m = ref_side ? I.m_r : I.m_q;
  out.dpad = I.dpad;
  out.reps = b.reps;
  out.rad = (const double *)b.radbits;
  out.off = b.off;
  out.perm = b.perm;
  out.pdist = b.pdist;
  out.px = b.px;
}

// Auxiliary streams a batch spreads its index builds over: 32, or
// KNNJ_AUX_STREAMS (1..32).  The phase is some 40 short dependent launches
// per instance, bound by the wait between dependent launches on a stream,
// so more streams overlap more of them -- also beyond the 4 hardware queues
// HIP gives a process by default (measured at 512 joins: index phase 103 ms
// on 4 streams, 71 ms on 16, 67 ms on 24, 64 ms on 32; 76 ms on 16 streams
// with 16 hardware queues).
int aux_streams() {
  static const int n = [] {
    const char *e = std::getenv("KNNJ_AUX_STREAMS");
    const int v = e ? std::atoi(e) : 32;
    return std::min(std::max(v, 1), (int)knnj_ctx::kAux);
  }();
  return n;
}

// The context's first n auxiliary streams (default priority) and their
// events, created on first use (a single join needs three: the query side's
// stream, the events of the reference representatives, aux_ev[1], and of the
// centred reference rows, aux_ev[2]; a batch aux_streams()), and the fork
// event aux_ev[kAux].
knnj_status ensure_aux(knnj_ctx *c, int n) {
  n = std::min(n, (int)knnj_ctx::kAux);
  if (!c->aux_ev[knnj_ctx::kAux]) HIPCHK(hipEventCreateWithFlags(&c->aux_ev[knnj_ctx::kAux], hipEventDisableTiming));
  for (int j = 0; j < n; ++j)
    if (!c->aux[j]) {
      HIPCHK(hipStreamCreate(&c->aux[j]));
      HIPCHK(hipEventCreateWithFlags(&c->aux_ev[j], hipEventDisableTiming));
    }
  return KNNJ_OK;
}

// One side's cluster index: sampled representatives, Lloyd refinement,
// assignment, radii, and the cluster-sorted permuted copy.
knnj_status enqueue_side(knnj_ctx *c, Inst &I, bool ref_side, hipStream_t s, hipEvent_t reps_final = nullptr,
                         bool defer_permute = false) {
  SideBufs &b = ref_side ? I.rb : I.qb;
  const float *X = ref_side ? I.dR : I.dQ;
  const int64_t n = ref_side ? I.nR : I.nQ, ld = ref_side ? I.ldR : I.ldQ;
  const int m = ref_side ? I.m_r : I.m_q, d = I.d, dpad = I.dpad;
  launch_draw(b.cand, m * I.trials, I.seed, n, s);
  // m0 sampled representatives, then `rounds` split rounds, each adding the
  // farthest members of the `step` widest clusters of the current set
  const int rounds = (m >= 64 && n >= 4 * (int64_t)m) ? I.split_rounds : 0;
  const int step = rounds > 0 ? std::max(1, m / (8 * rounds)) : 0;
  int mc = m - rounds * step;
  launch_rep_select(X, ld, d, dpad, b.cand, mc, I.trials, b.partial, rep_score_blocks(mc), b.reps, b.rep_rows, s);
  for (int r = 0; r < rounds; ++r) {
    launch_assign(X, n, ld, d, dpad, b.reps, mc, b.cl, b.rb, b.rn, s);
    launch_zeros(s, b.radbits, sizeof(u64) * mc, b.cnt, sizeof(int32_t) * mc);
    launch_assign_finish(X, n, ld, d, dpad, b.reps, b.cl, b.d2r, b.radbits, b.cnt, s);
    launch_split_reps(X, n, ld, d, dpad, mc, step, b.cl, b.d2r, b.gkeys, b.reps, s);
    mc += step;
  }
  // Lloyd refinement: members of each sampled representative (ascending row
  // order), their mean becomes the representative.  In a join the means are
  // taken over every ls-th row (a view with ls times the row stride), ls a
  // power of two up to 16 that leaves 16 or more rows per representative:
  // an estimate of the mean is all the step needs, and the assignment it
  // runs costs ls times less (measured on the d = 128 mixture, 2M rows, 4,096
  // representatives: strides 4, 8, 16 give the same 3.907e8 evaluated pairs
  // and index times of 6.04, 5.33, 5.01 ms).  A resident index, built once
  // and queried many times, uses every row.
  int ls = 1;
  if (!I.resident)
    while (ls < 16 && n >= 32 * (int64_t)ls * m) ls *= 2;
  if (const char *e = std::getenv("KNNJ_LLOYD_STRIDE")) ls = std::max(1, std::atoi(e));  // diagnostic override
  const int64_t ns = (n + ls - 1) / ls;
  for (int it = 0; it < I.refine; ++it) {
    launch_assign(X, ns, ld * ls, d, dpad, b.reps, m, b.cl, b.rb, b.rn, s);
    launch_zeros(s, b.cnt, sizeof(int32_t) * m, b.fill, sizeof(int32_t) * m);
    launch_count(b.cl, ns, b.cnt, s);
    launch_bucket_rows(b.cl, ns, m, b.cnt, b.off, b.fill, b.perm, b.gkeys, s);
    launch_centroid(X, ld * ls, d, dpad, m, b.off, b.perm, I.integral ? 1 : 0, b.reps, s);
  }
  if (reps_final) HIPCHK(hipEventRecord(reps_final, s));  // the representatives change no more
  launch_assign(X, n, ld, d, dpad, b.reps, m, b.cl, b.rb, b.rn, s);
  launch_zeros(s, b.radbits, sizeof(u64) * m, b.cnt, sizeof(int32_t) * m, b.fill, sizeof(int32_t) * m);
  launch_assign_finish(X, n, ld, d, dpad, b.reps, b.cl, b.d2r, b.radbits, b.cnt, s);
  launch_bucket_sort_permute(X, n, ld, d, dpad, m, b.cl, b.cnt, b.off, b.fill, b.perm, b.d2r, b.pdist, b.gkeys,
                             defer_permute ? nullptr : b.px, s);
  HIPCHK(hipGetLastError());
  fill_side_index(I, ref_side);
  return KNNJ_OK;
}

// The operand planes of the centred reference rows (launch_center).
int center_planes(const Inst &I) { return I.fp8 ? 3 : I.bf16 ? (I.bf16x1 ? 1 : 2) : 0; }

// Index build, centring, bounds and the tile list of one join.  Optional
// events are recorded after the index (with centring) and after the bounds.
// fork: when set (a single join with two sides to index), the query side is
// built on that auxiliary stream while the reference side is built on s;
// the two sides share no buffer, and s waits for the query side before the
// bounds.
knnj_status enqueue_index(knnj_ctx *c, Inst &I, hipStream_t s, hipEvent_t ev_index, hipEvent_t ev_bounds,
                          hipStream_t fork = nullptr) {
  const bool two = !I.pre && !I.self && fork;
  if (two) {
    HIPCHK(hipEventRecord(c->aux_ev[knnj_ctx::kAux], s));
    HIPCHK(hipStreamWaitEvent(fork, c->aux_ev[knnj_ctx::kAux], 0));
    TRY(enqueue_side(c, I, false, fork));
  }
  // with the query side on the auxiliary stream, the query rows'
  // assignment runs there too, once the reference representatives are final
  // (beside the reference side's own assignment, into buffers of its own)
  const bool seed_fork = two && I.qseed && I.qs_rb;
  // with two streams the permuted copy and the centring of the reference
  // rows (read by the join only) run on the auxiliary stream beside the
  // bounds, once the reference side is indexed; s waits for them after the
  // tile list
  const bool center = I.mfma && !I.pre;
  const bool center_fork = two && center;
  if (!I.pre) TRY(enqueue_side(c, I, true, s, seed_fork ? c->aux_ev[1] : nullptr, center_fork));
  if (I.qseed) {
    if (I.pre) {
      launch_assign(I.dQ, I.nQ, I.ldQ, I.d, I.dpad, I.pre->Ri.reps, I.m_r, I.qseed, I.qs_rb, I.qs_rn, s);
    } else if (seed_fork) {
      HIPCHK(hipStreamWaitEvent(fork, c->aux_ev[1], 0));
      launch_assign(I.dQ, I.nQ, I.ldQ, I.d, I.dpad, I.rb.reps, I.m_r, I.qseed, I.qs_rb, I.qs_rn, fork);
    } else {
      launch_assign(I.dQ, I.nQ, I.ldQ, I.d, I.dpad, I.rb.reps, I.m_r, I.qseed, I.rb.rb, I.rb.rn, s);
    }
    HIPCHK(hipGetLastError());
  }
  if (I.self)
    I.Qi = I.Ri;
  else if (!two)
    TRY(enqueue_side(c, I, false, s));
  if (two) {
    HIPCHK(hipEventRecord(c->aux_ev[0], fork));  // query side indexed
    if (center_fork) {
      HIPCHK(hipEventRecord(c->aux_ev[knnj_ctx::kAux], s));  // reference side indexed
      HIPCHK(hipStreamWaitEvent(fork, c->aux_ev[knnj_ctx::kAux], 0));
      launch_permute(I.dR, I.nR, I.ldR, I.d, I.dpad, I.rb.perm, I.rb.px, fork);
      launch_center(I.Ri, I.d, I.dk, center_planes(I), I.rc, I.nr, I.reps32, I.rn32, I.rinv, fork);
      HIPCHK(hipGetLastError());
      HIPCHK(hipEventRecord(c->aux_ev[2], fork));
    }
    HIPCHK(hipStreamWaitEvent(s, c->aux_ev[0], 0));
  }
  if (center && !center_fork) {
    launch_center(I.Ri, I.d, I.dk, center_planes(I), I.rc, I.nr, I.reps32, I.rn32, I.rinv, s);
    HIPCHK(hipGetLastError());
  }
  if (ev_index) HIPCHK(hipEventRecord(ev_index, s));
  launch_bounds(I.Qi, I.Ri, I.d, I.k, I.theta, I.plist, I.plb, I.pcnt, I.dab, s);
  HIPCHK(hipGetLastError());
  if (ev_bounds) HIPCHK(hipEventRecord(ev_bounds, s));
  if (I.dedup) launch_dedup(I.Qi, I.d, I.lead, I.ucnt, I.uoff, I.upx, I.uperm, I.updist, s);
  if (!I.batch_counters) launch_zeros(s, I.counters, 8 * sizeof(int32_t), I.evals, 4 * sizeof(u64));
  launch_make_tiles(I.dedup ? I.uoff : I.Qi.off, I.m_q, I.theta, I.tq, I.id, I.tiles, I.counters,
                    (int)I.max_slots, I.o.xcd_tiles ? 8 : 1, s);
  HIPCHK(hipGetLastError());
  if (center_fork) HIPCHK(hipStreamWaitEvent(s, c->aux_ev[2], 0));  // the centred rows
  return KNNJ_OK;
}

void fill_args(Inst &I) {
  JoinArgs &a = I.args;
  a = JoinArgs{};
  const int d = I.d;
  // the join reads the distinct queries only (the leaders' side, whose
  // buffers are fixed by the allocation: fill_args may run before the index
  // and the dedup kernels are queued, as knnj_join_batch does)
  a.Q = I.dedup ? dedup_side(I.Qi, I.uoff, I.upx, I.uperm, I.updist) : I.Qi;
  a.R = I.Ri;
  a.d = d;
  a.k = I.k;
  a.kcap = I.kcap;
  a.waves = I.waves;
  a.split = I.split;
  a.lockstep = std::getenv("KNNJ_NOLOCK") ? 0 : 1;
  a.seed_blocks = 0;
  if (const char *sb = std::getenv("KNNJ_SEED_BLOCKS")) a.seed_blocks = std::atoi(sb);
  a.qseed = I.qseed ? I.qseed : (I.self && !I.pre && seed_by_assignment(I)) ? I.rb.cl : nullptr;
  a.exact = I.kernel_exact();
  const double delta = I.exact ? 0.0 : (double)(d + 4) * 0x1p-23 + 0x1p-40;
  a.delta = delta;
  a.fac = (1.0 + delta) / (1.0 - delta) * (1.0 + 0x1p-20);
  a.tiles = I.tiles;
  a.ntiles = I.counters;
  a.theta = I.theta;
  a.plist = I.plist;
  a.plb = I.plb;
  a.pcnt = I.pcnt;
  a.out_idx = I.out_idx;
  a.out_dist = I.out_dist;
  a.out_mode = I.aos ? 2 : (I.o.out_squared ? 1 : 0);
  a.n_evals = I.evals;
  a.ovf_tiles = I.ovf_tiles;
  a.ovf_count = I.counters + 1;
  a.max_err_bits = (u32 *)(I.counters + 3);
  a.ferr_bits = I.o.filter_stats ? (u32 *)(I.counters + 5) : nullptr;
  a.relbound = (!I.use_mfma && !a.exact) ? 1 : 0;
  a.rc = I.rc;
  a.nr = I.nr;
  a.reps32 = I.reps32;
  a.rn32 = I.rn32;
  a.rinv = I.rinv;
  a.dk = I.dk;
  // per-entry error codes in the MFMA filter keys: rows must fit 26 bits;
  // codes span 63 binades below the largest possible squared distance
  // the finest code that leaves the rows their field: 2^lgs steps per octave
  int lgs = -1;
  for (int t = 3; t >= 0 && lgs < 0; --t)
    if (I.nR <= ((int64_t)1 << (26 - t))) lgs = t;
  a.ecode = (I.use_mfma && !I.exact && lgs >= 0) ? 1 : 0;
  const double smax = std::max(1e-30, (double)d * I.span * I.span);
  a.ec.lgs = std::max(lgs, 0);
  a.ec.rowbits = 26 - a.ec.lgs;
  a.ec.ebase = std::ilogb(smax) + 1 - 62;
  const double u = 0x1p-24, ub = 0x1p-8;
  auto gam = [&](double n) { return n * u / (1.0 - n * u); };
  // |computed dot - (q-c).(r-c)| <= cdot ||q-c|| ||r-c||
  double cdot = gam(d);  // fp32 operands
  if (I.fp8) {
    // e4m3 rows scaled by powers of two (k_join_mfma.h): relative rounding
    // 2^-4, subnormal parts <= 2^-16 ||x|| per element
    const double dl = 0x1p-4 + std::sqrt((double)d) * 0x1p-16;
    cdot = 2.0 * dl + dl * dl + gam(d) * (1.0 + dl) * (1.0 + dl);
  } else if (I.bf16x1)
    cdot = 2.0 * ub + ub * ub + gam(d) * (1.0 + ub) * (1.0 + ub);  // one bf16 product
  else if (I.bf16)
    cdot = 3.1 * ub * ub + 1.02 * gam(3.0 * d);  // hi/lo, three products
  a.c1 = (float)((2.0 * cdot + 4.0 * u) * 1.01);
  a.c2 = (float)((gam(d + 3) + 8.0 * u) * 1.01);
}

// The join kernel's buffers, checked on the host before it is launched: a
// plan whose arguments point nowhere is a bug in this file, reported as an
// error rather than left to fault the device.
knnj_status check_args(const JoinArgs &a, int64_t nq) {
  if (nq <= 0) return KNNJ_OK;
  const void *need[] = {a.Q.px,    a.Q.perm,  a.Q.pdist, a.Q.off,   a.Q.reps, a.R.px,   a.R.perm,
                        a.R.pdist, a.R.off,   a.R.reps,  a.R.rad,   a.tiles,  a.ntiles, a.theta,
                        a.plist,   a.plb,     a.pcnt,    a.out_idx, a.n_evals, a.ovf_tiles, a.ovf_count};
  for (size_t i = 0; i < sizeof(need) / sizeof(need[0]); ++i)
    if (!need[i]) {
      set_err("internal error: join argument %zu is unset", i);
      return KNNJ_EDEVICE;
    }
  if (a.out_mode != 2 && !a.out_dist) {
    set_err("internal error: join distance output is unset");
    return KNNJ_EDEVICE;
  }
  return KNNJ_OK;
}

hipError_t launch_main(const Inst &I, int64_t grid, hipStream_t s) {
  if (I.use_mfma) {
    if (I.fp8) return launch_join_mfma_fp8(I.args, grid, s);
    if (I.bf16) return launch_join_mfma_bf16(I.args, I.bf16x1 ? 1 : 3, grid, s);
    return launch_join_mfma_f32(I.args, grid, s);
  }
  return launch_join(I.args, I.tq, grid, s);
}

// Overflowed queries of one join, one per tile, re-run at larger
// capacities: on the MFMA kernel at twice the capacity (at most 512; where
// that capacity's LDS still allows several waves, they split each query's
// cluster list), then on the VALU kernel at 1024; KNNJ_EOVERFLOW only if
// that overflows too.  Each stage's overflows feed the next.
knnj_status rerun_overflow(Inst &I, int32_t novf, hipStream_t s, int *stages) {
  Tile *in = I.ovf_tiles, *out = I.ovf2_tiles;
  int32_t *in_cnt = I.counters + 1, *out_cnt = I.counters + 2, *spare = I.counters + 4;
  int kc = I.kcap;
  bool mfma = I.use_mfma;
  const int mfma_max = I.o.kcap_mfma_max > 0 ? I.o.kcap_mfma_max : 512;
  int32_t n = novf;
  while (n > 0) {
    ++*stages;
    JoinArgs b = I.args;
    b.tiles = in;
    b.ntiles = in_cnt;
    b.ovf_tiles = out;
    b.ovf_count = out_cnt;
    b.tile_prof = nullptr;
    b.ovf_diag = nullptr;
    b.waves = 1;  // one-query tiles
    b.split = 1;
    HIPCHK(hipMemsetAsync(out_cnt, 0, sizeof(int32_t), s));
    if (mfma && kc < mfma_max) {
      kc = std::min(mfma_max, 2 * kc);
      b.kcap = kc;
      const int sw = join_mfma_split_waves(kc, I.dk);
      if (sw > 1) b.split = b.waves = sw;
      if (I.fp8)
        HIPCHK(launch_join_mfma_fp8(b, n, s));
      else if (I.bf16)
        HIPCHK(launch_join_mfma_bf16(b, I.bf16x1 ? 1 : 3, n, s));
      else
        HIPCHK(launch_join_mfma_f32(b, n, s));
    } else if (kc < KNNJ_MAX_K) {
      mfma = false;
      kc = KNNJ_MAX_K;
      b.kcap = kc;
      b.exact = I.exact_valu ? 1 : 0;
      b.ecode = 0;
      b.relbound = b.exact ? 0 : 1;
      HIPCHK(launch_join(b, 1, n, s));
    } else {
      set_err("%d queries overflowed the largest candidate capacity (%d)", n, KNNJ_MAX_K);
      return KNNJ_EOVERFLOW;
    }
    int32_t h = 0;
    HIPCHK(hipMemcpyAsync(&h, out_cnt, sizeof(h), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    n = h;
    std::swap(in, out);
    // the next stage reads this stage's count and writes into the third slot
    int32_t *next_out = (in_cnt == I.counters + 1) ? spare : in_cnt;
    in_cnt = out_cnt;
    out_cnt = next_out;
  }
  return KNNJ_OK;
}

// Every duplicate query's output row from its leader's (after the join and
// any overflow re-runs have written the leaders' rows).
void expand_duplicates(const Inst &I, hipStream_t s) {
  if (I.dedup && I.nQ > 0)
    launch_dedup_expand(I.nQ, I.Qi.perm, I.lead, I.k, I.aos ? 1 : 0, I.out_idx, I.out_dist, s);
}

knnj_status copy_outputs(const Inst &I, hipStream_t s) {
  if (!I.out_host || I.nQ == 0) return KNNJ_OK;
  const size_t nout = (size_t)I.nQ * I.k;
  HIPCHK(hipMemcpyAsync(I.user_idx, I.out_idx, nout * sizeof(int32_t) * (I.aos ? 2 : 1), hipMemcpyDeviceToHost,
                        s));
  if (!I.aos) HIPCHK(hipMemcpyAsync(I.user_dist, I.out_dist, nout * sizeof(float), hipMemcpyDeviceToHost, s));
  return KNNJ_OK;
}

void plan_stats(const Inst &I, knnj_stats &st) {
  st.exact_mode = I.exact ? 1 : 0;
  st.precision_used = I.fp8 ? KNNJ_PREC_FP8 : I.bf16 ? (I.bf16x1 ? KNNJ_PREC_BF16X1 : KNNJ_PREC_BF16) : KNNJ_PREC_FP32;
  st.kernel_used = I.kind;
  st.m_q = I.m_q;
  st.m_r = I.m_r;
  st.tile_q = I.tq;
  st.kcap = I.kcap;
}

// Diagnostic print of the per-tile record (KNNJ_TILE_PROFILE set).
knnj_status print_tile_profile(const Inst &I, int32_t nslots, int32_t novf) {
  if (nslots <= 0) return KNNJ_OK;
  std::vector<u64> tp((size_t)nslots * 8);
  HIPCHK(hipMemcpy(tp.data(), I.args.tile_prof, tp.size() * sizeof(u64), hipMemcpyDeviceToHost));
  std::vector<double> dur;
  double listed = 0, visited = 0, seedb = 0, sweepb = 0;
  double ph[7] = {0, 0, 0, 0, 0, 0, 0};  // ticks: seed, sweep, final merges, emit, seed search; groups; merge ticks
  u64 t0 = ~0ull, t1 = 0;
  for (int i = 0; i < nslots; ++i) {
    if (tp[8 * i + 1] == 0) continue;  // empty slot
    dur.push_back((double)(tp[8 * i + 1] - tp[8 * i]));
    listed += (double)(tp[8 * i + 3] & 0xFFFF);
    sweepb += (double)((tp[8 * i + 3] >> 16) & 0xFFFF);
    seedb += (double)((tp[8 * i + 3] >> 32) & 0xFFFF);
    visited += (double)(tp[8 * i + 3] >> 48);
    ph[0] += tp[8 * i + 4] ? (double)(tp[8 * i + 4] - tp[8 * i]) : 0.0;
    ph[4] += tp[8 * i + 7] ? (double)(tp[8 * i + 7] - tp[8 * i]) : 0.0;
    ph[5] += (double)(tp[8 * i + 2] >> 32);
    ph[6] += (double)(tp[8 * i + 2] & 0xFFFFFFFFull);
    ph[1] += tp[8 * i + 5] ? (double)(tp[8 * i + 5] - (tp[8 * i + 4] ? tp[8 * i + 4] : tp[8 * i])) : 0.0;
    ph[2] += tp[8 * i + 6] ? (double)(tp[8 * i + 6] - tp[8 * i + 5]) : 0.0;
    ph[3] += tp[8 * i + 6] ? (double)(tp[8 * i + 1] - tp[8 * i + 6]) : 0.0;
    t0 = std::min(t0, tp[8 * i]);
    t1 = std::max(t1, tp[8 * i + 1]);
  }
  if (const char *dump = std::getenv("KNNJ_TILE_DUMP")) {
    if (FILE *f = fopen(dump, "wb")) {
      fwrite(tp.data(), sizeof(u64), tp.size(), f);
      fclose(f);
    }
  }
  std::sort(dur.begin(), dur.end());
  double sum = 0;
  for (double v : dur) sum += v;
  const int nt = (int)dur.size();
  if (nt > 0)
    fprintf(stderr,
            "[knnj tile profile] tiles=%d span=%.0f ticks sum=%.0f mean=%.1f p50=%.0f p99=%.0f max=%.0f "
            "listed/tile=%.1f seed-blocks/tile=%.1f sweep-blocks/tile=%.1f visited/tile=%.1f\n",
            nt, (double)(t1 - t0), sum, sum / nt, dur[nt / 2], dur[(size_t)(0.99 * (nt - 1))], dur.back(),
            listed / nt, seedb / nt, sweepb / nt, visited / nt);
  if (nt > 0 && ph[0] + ph[1] > 0)
    fprintf(stderr,
            "[knnj tile profile] phase ticks per tile: seed=%.0f (search %.0f) sweep=%.0f merges=%.0f emit=%.0f; "
            "B groups %.1f, ticks in visit merges %.0f\n",
            ph[0] / nt, ph[4] / nt, ph[1] / nt, ph[2] / nt, ph[3] / nt, ph[5] / nt, ph[6] / nt);
  if (novf > 0 && I.args.ovf_diag) {
    float dg[256];
    HIPCHK(hipMemcpy(dg, I.args.ovf_diag, sizeof(dg), hipMemcpyDeviceToHost));
    for (int i = 0; i < std::min(novf, 8); ++i)
      fprintf(stderr, "[knnj overflow] cutoff=%.6g kth=%.6g last=%.6g first=%.6g\n", dg[4 * i], dg[4 * i + 1],
              dg[4 * i + 2], dg[4 * i + 3]);
  }
  return KNNJ_OK;
}

}  // namespace

namespace {

// The planned join after its input scan: index (both sides, or the query
// side against a prebuilt index), bounds, tiles, the join, overflow re-runs,
// outputs and stats.  ev[0] was recorded at the call's start.
knnj_status run_single(knnj_ctx *c, Inst &I, Alloc &A, knnj_stats *stats) {
  hipStream_t s = c->stream;
  hipEvent_t *ev = c->ev;
  c->plan.valid = false;  // set again once this join's plan is complete
  knnj_stats st{};
  plan_stats(I, st);
  if (I.nQ == 0) {
    if (stats) *stats = st;
    return KNNJ_OK;
  }
  TRY(alloc_inst(A, I, nullptr));

  HIPCHK(hipEventRecord(ev[1], s));
  // the query side's index on an auxiliary stream beside the reference
  // side's (the allocation-free path creates nothing and stays on one stream)
  hipStream_t fork = nullptr;
  if (!A.no_alloc && !I.pre && !I.self) {
    TRY(ensure_aux(c, 3));  // aux_ev[1]: representatives final, aux_ev[2]: rows centred
    fork = c->aux[0];
  }
  TRY(enqueue_index(c, I, s, ev[2], ev[3], fork));
  int32_t nslots = 0;
  HIPCHK(hipMemcpyAsync(&nslots, I.counters, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));

  fill_args(I);
  const bool tile_prof = !A.no_alloc && std::getenv("KNNJ_TILE_PROFILE") != nullptr;
  if (tile_prof) {
    u64 *tp;
    float *dg;
    TRY(ws(c, "tile_prof", (size_t)I.max_slots * 8, &tp));
    TRY(ws(c, "ovf_diag", 256, &dg));
    HIPCHK(hipMemsetAsync(tp, 0, sizeof(u64) * 8 * I.max_slots, s));
    HIPCHK(hipMemsetAsync(dg, 0, sizeof(float) * 256, s));
    I.args.tile_prof = tp;
    I.args.ovf_diag = dg;
  }
  TRY(check_args(I.args, I.nQ));
  HIPCHK(hipEventRecord(ev[8], s));
  HIPCHK(launch_main(I, nslots, s));
  HIPCHK(hipEventRecord(ev[4], s));

  int32_t hc[4];
  HIPCHK(hipMemcpyAsync(hc, I.counters, sizeof(hc), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  st.n_tiles = hc[0];
  st.n_overflow = hc[1];
  st.n_launches = 1;
  if (tile_prof) TRY(print_tile_profile(I, nslots, hc[1]));
  HIPCHK(hipEventRecord(ev[5], s));
  int stages = 0;
  if (hc[1] > 0) TRY(rerun_overflow(I, hc[1], s, &stages));
  st.rerun_stages = stages;
  HIPCHK(hipEventRecord(ev[6], s));
  expand_duplicates(I, s);
  HIPCHK(hipEventRecord(ev[9], s));
  int32_t ndist = (int32_t)I.nQ;
  if (I.dedup) HIPCHK(hipMemcpyAsync(&ndist, I.uoff + I.m_q, sizeof(ndist), hipMemcpyDeviceToHost, s));
  TRY(copy_outputs(I, s));
  u64 hcnt[4] = {0, 0, 0, 0};
  u32 herr = 0, hferr[3] = {0, 0, 0};
  HIPCHK(hipMemcpyAsync(hcnt, I.evals, sizeof(hcnt), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(&herr, I.counters + 3, sizeof(herr), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(hferr, I.counters + 5, sizeof(hferr), hipMemcpyDeviceToHost, s));
  HIPCHK(hipEventRecord(ev[7], s));
  HIPCHK(hipStreamSynchronize(s));

  st.t_total_ms = elapsed(ev[0], ev[7]);
  st.t_index_ms = elapsed(ev[1], ev[2]);
  st.t_bounds_ms = elapsed(ev[2], ev[3]);
  // the join proper: the tiles' launch, and the copies to duplicate queries
  st.t_join_ms = elapsed(ev[8], ev[4]) + (I.dedup ? elapsed(ev[6], ev[9]) : 0.0);
  st.t_fallback_ms = hc[1] > 0 ? elapsed(ev[5], ev[6]) : 0.0;
  st.n_distinct_queries = ndist;
  st.n_dist_evals = (int64_t)hcnt[0];
  st.n_merges = (int64_t)hcnt[1];
  st.n_appended = (int64_t)hcnt[2];
  const double tot = (double)I.nR * (double)I.nQ;
  st.pruning_ratio = tot > 0 ? 1.0 - (double)hcnt[0] / tot : 0.0;
  float ferr, fe3[3];
  std::memcpy(&ferr, &herr, 4);
  std::memcpy(fe3, hferr, sizeof(fe3));
  st.filter_max_abs_err = ferr;
  st.filter_err_max = fe3[0];
  st.filter_bound_max = fe3[1];
  st.filter_err_ratio = fe3[2];
  {
    knnj_ctx::Plan &P = c->plan;
    P.valid = true;
    P.n_r = I.nR;
    P.n_q = I.nQ;
    P.d = I.d;
    P.dpad = I.dpad;
    P.m_r = I.m_r;
    P.m_q = I.m_q;
    P.k = I.k;
    P.self = I.self ? 1 : 0;
    P.Ri = I.Ri;
    P.Qi = I.Qi;
    P.theta = I.theta;
    P.plist = I.plist;
    P.pcnt = I.pcnt;
    P.plb = I.plb;
  }
  {
    int32_t hp[KNNJ_MAX_REPS];  // m_q <= KNNJ_MAX_REPS; no host allocation either
    HIPCHK(hipMemcpy(hp, I.pcnt, sizeof(int32_t) * I.m_q, hipMemcpyDeviceToHost));
    int64_t t = 0;
    for (int i = 0; i < I.m_q; ++i) t += hp[i];
    st.n_cluster_pairs = t;
  }
  if (stats) *stats = st;
  return KNNJ_OK;
}


}  // namespace

// Scratch of the allocation-free entry point (knnj_index_query_noalloc):
// one fixed array per device, held from library load to process end.  Its
// name is a host-side stand-in; host code uses only the device address
// hipGetSymbolAddress returns for it, once per device with that device
// current (knnj_ctx_create), kept in g_scratch_addr and in each context.
static __device__ __align__(256) unsigned char scratch[KNNJ_NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
std::map<int, void *> g_scratch_addr;

// The current device's scratch address (device `device` must be current).
knnj_status scratch_address(int device, void **out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_addr.find(device);
  if (it == g_scratch_addr.end()) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_err("hipGetSymbolAddress(scratch) failed on device %d: %s", device, hipGetErrorString(e));
      return KNNJ_EDEVICE;
    }
    it = g_scratch_addr.emplace(device, p).first;
  }
  *out = it->second;
  return KNNJ_OK;
}

}  // namespace

extern "C" {

const char *knnj_version(void) { return "knnj 0.3.0 (gfx950)"; }

const char *knnj_last_error(void) { return g_err.c_str(); }

knnj_status knnj_ctx_create(int device, knnj_ctx **out) {
  g_err.clear();
  if (!out) {
    set_err("out is NULL");
    return KNNJ_EINVAL;
  }
  *out = nullptr;
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) {
    set_err("device %d out of range (%d devices)", device, n);
    return KNNJ_EINVAL;
  }
  HIPCHK(hipSetDevice(device));
  knnj_ctx *c = new knnj_ctx();
  c->device = device;
  if (hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    set_err("hipStreamCreate failed");
    return KNNJ_EDEVICE;
  }
  for (auto &e : c->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      set_err("hipEventCreate failed");
      knnj_ctx_destroy(c);
      return KNNJ_EDEVICE;
    }
  }
  {
    const knnj_status e = scratch_address(device, &c->scratch);
    if (e != KNNJ_OK) {
      knnj_ctx_destroy(c);
      return e;
    }
  }
  *out = c;
  return KNNJ_OK;
}

void knnj_ctx_destroy(knnj_ctx *c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  for (auto &kv : c->ws)
    if (kv.second.p) hipFree(kv.second.p);
  for (auto &e : c->ev)
    if (e) hipEventDestroy(e);
  if (c->batch_exec) hipGraphExecDestroy(c->batch_exec);
  for (auto &e : c->aux_ev)
    if (e) hipEventDestroy(e);
  for (auto &x : c->aux)
    if (x) hipStreamDestroy(x);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

knnj_status knnj_ctx_set_budget(knnj_ctx *c, int64_t bytes) {
  if (!c || bytes < 0) {
    set_err("bad context or budget");
    return KNNJ_EINVAL;
  }
  c->budget = bytes;
  return KNNJ_OK;
}

}  // extern "C"

namespace {

knnj_status join_once(knnj_ctx *c, const knnj_matrix &R, const knnj_matrix &Q, int32_t k, const knnj_opts &o,
                      int32_t *out_idx, float *out_dist, knnj_stats *stats) {
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  hipEvent_t *ev = c->ev;
  Inst I;
  describe(I, R, Q, k, o, out_idx, out_dist);
  HIPCHK(hipEventRecord(ev[0], s));

  // device-resident inputs, then the input scan: finiteness, integrality, magnitude
  Alloc A;
  A.c = c;
  TRY(stage_inputs(A, I, s));
  u32 *flags;
  TRY(ws(c, "flags", 8, &flags));
  HIPCHK(hipMemsetAsync(flags, 0, 8 * sizeof(u32), s));
  launch_scan_input(I.dR, I.nR, I.d, I.ldR, flags, s);
  if (!I.self && I.nQ > 0) launch_scan_input(I.dQ, I.nQ, I.d, I.ldQ, flags, s);
  u32 hflags[3];
  HIPCHK(hipMemcpyAsync(hflags, flags, sizeof(hflags), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  TRY(decide(I, hflags));
  return run_single(c, I, A, stats);
}

// The BF16 hi/lo filter as an autotune candidate: dimensions whose default
// is plain FP32 (16 <= d < 64), precision and kernel left to the library.
bool low_d_filter_candidate(const knnj_opts &o, int d) {
  return o.precision == KNNJ_PREC_AUTO && o.kernel == KNNJ_KERNEL_AUTO && d >= 16 && d < 64;
}

// Autotune candidates for the fields the caller left open (0): Lloyd step
// on/off, representatives 1x / 2x the default count, and for the VALU
// kernel 2 or 4 queries per tile; then the BF16 filter where it is not the
// default.  Every candidate computes the same exact answer; only the time
// differs.
std::vector<knnj_opts> tune_candidates(const knnj_opts &o, const knnj_matrix &R, const knnj_matrix &Q, bool self) {
  std::vector<knnj_opts> v;
  const int d = R.d;
  const bool valu = o.kernel == KNNJ_KERNEL_VALU ||
                    (o.kernel == KNNJ_KERNEL_AUTO && !(d >= 32 && d <= 256) && o.precision != KNNJ_PREC_BF16 &&
                     o.precision != KNNJ_PREC_BF16X1);
  // Lloyd steps: 1, none, or 2 (cheap with the strided means; C3 takes 2)
  const int refines[3] = {o.refine_iters ? o.refine_iters : 1, -1, 2};
  const int nref = o.refine_iters ? 1 : 3;
  const int scales[2] = {1, 2};
  const int nsc = (o.m_r || o.m_q || o.brute_force) ? 1 : 2;
  const int tqs[2] = {o.tile_q ? o.tile_q : 2, 4};
  const int ntq = (valu && !o.tile_q) ? 2 : 1;
  for (int a = 0; a < nref; ++a)
    for (int b = 0; b < nsc; ++b)
      for (int t = 0; t < ntq; ++t) {
        knnj_opts x = o;
        x.autotune = 0;
        x.refine_iters = refines[a];
        if (scales[b] > 1) {
          x.m_r = std::min(KNNJ_MAX_REPS, (int)std::min<int64_t>(R.n, 2 * default_reps(R.n)));
          x.m_q = self ? x.m_r : std::min(KNNJ_MAX_REPS, (int)std::min<int64_t>(Q.n, 2 * default_reps(Q.n)));
        }
        if (valu) x.tile_q = tqs[t];
        v.push_back(x);
      }
  // and the BF16 hi/lo filter on the matrix cores where plain FP32 is the
  // default (16 <= d < 64): the first candidate's filtered form
  if (low_d_filter_candidate(o, d)) {
    knnj_opts x = v[0];
    x.precision = KNNJ_PREC_BF16;
    x.tile_q = 0;
    v.push_back(x);
  }
  return v;
}

// Second-phase autotune variants of the first phase's winner, for the fields
// the caller left open: the k-select capacity doubled (kmax: the largest k
// served; a larger pool or list trades LDS, hence waves in flight, for fewer
// merges and overflow re-runs) and, where the matrix cores may run, the
// 4-wave lockstep tile shape (reference rows shared by 64 queries through
// L2, one barrier per visited cluster).
std::vector<knnj_opts> tune_variants(const knnj_opts &best, const knnj_opts &o, int kmax, bool mfma) {
  std::vector<knnj_opts> v;
  if (!o.kcap) {
    knnj_opts x = best;
    x.kcap = pow2_at_least(std::max(2 * kmax, 64));
    if (x.kcap <= KNNJ_MAX_K) v.push_back(x);
  }
  if (!o.tile_waves && mfma) {
    knnj_opts x = best;
    x.tile_waves = 4;
    v.push_back(x);
  }
  return v;
}

bool tune_may_use_mfma(const knnj_opts &o, int d) {
  if (o.kernel == KNNJ_KERNEL_VALU) return false;
  return o.kernel == KNNJ_KERNEL_MFMA || (d >= 32 && d <= 256) || o.precision == KNNJ_PREC_BF16 ||
         o.precision == KNNJ_PREC_BF16X1 || o.precision == KNNJ_PREC_FP8;
}

// Times run(opts) for each candidate (each writes the same exact answer) and
// returns the index of the fastest; *st gets the fastest run's statistics.
template <class Run>
knnj_status time_candidates(knnj_ctx *c, const std::vector<knnj_opts> &cand, Run run, int *best, double *best_ms,
                            knnj_stats *st_best) {
  knnj_stats st{};
  for (size_t i = 0; i < cand.size(); ++i) {
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    const auto t0 = std::chrono::steady_clock::now();
    TRY(run(cand[i], &st));
    HIPCHK(hipStreamSynchronize(c->stream));
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (*best < 0 || ms < *best_ms) {
      *best = (int)i;
      *best_ms = ms;
      if (st_best) *st_best = st;
    }
  }
  return KNNJ_OK;
}

// The two-phase search: the first-phase list, then the winner's variants;
// the chosen candidate is cached under sig for later calls of that shape,
// which rebuild it from their own options (cand, o) and run it.
template <class Run>
knnj_status autotune_call(knnj_ctx *c, const std::vector<int64_t> &sig, const std::vector<knnj_opts> &cand,
                          const knnj_opts &o, int kmax, bool mfma, Run run, knnj_stats *stats) {
  auto it = c->tuned.find(sig);
  if (it != c->tuned.end()) {
    const knnj_ctx::Tuned t = it->second;
    knnj_opts x = cand[t.b1];
    if (t.b2 >= 0) x = tune_variants(x, o, kmax, mfma)[t.b2];
    const knnj_status e = run(x, stats);
    if (stats) stats->tuned = t.id;
    return e;
  }
  int b1 = -1;
  double ms1 = 0.0;
  knnj_stats st1{};
  TRY(time_candidates(c, cand, run, &b1, &ms1, &st1));
  knnj_ctx::Tuned t{b1, -1, 1 + b1};
  const std::vector<knnj_opts> var = tune_variants(cand[b1], o, kmax, mfma);
  int b2 = -1;
  double ms2 = 0.0;
  knnj_stats st2{};
  TRY(time_candidates(c, var, run, &b2, &ms2, &st2));
  if (b2 >= 0 && ms2 < ms1) {
    t = knnj_ctx::Tuned{b1, b2, 1 + (int)cand.size() + b2};
    st1 = st2;
  }
  c->tuned[sig] = t;
  if (stats) {
    *stats = st1;
    stats->tuned = t.id;
  }
  return KNNJ_OK;
}

}  // namespace

extern "C" {

knnj_status knnj_join(knnj_ctx *c, knnj_matrix R, knnj_matrix Q, int32_t k, const knnj_opts *opts,
                      int32_t *out_idx, float *out_dist, knnj_stats *stats) {
  g_err.clear();
  if (!c) {
    set_err("context is NULL");
    return KNNJ_EINVAL;
  }
  knnj_opts o{};
  if (opts) o = *opts;
  TRY(check_join_args(R, Q, k, o, out_idx, out_dist));
  if (!o.autotune || Q.n == 0) return join_once(c, R, Q, k, o, out_idx, out_dist, stats);
  const int64_t ldR = R.ld ? R.ld : R.d, ldQ = Q.ld ? Q.ld : Q.d;
  const bool self = R.data == Q.data && R.n == Q.n && ldR == ldQ && R.on_device == Q.on_device;
  auto lg = [](int64_t n) { return (int64_t)std::ceil(std::log2((double)std::max<int64_t>(n, 1))); };
  std::vector<int64_t> sig = {R.d, k, lg(R.n), lg(Q.n), self, o.precision, o.kernel, o.m_r, o.m_q,
                              o.refine_iters, o.tile_q, o.brute_force};
  sig.push_back(o.kcap);
  sig.push_back(o.tile_waves);
  const std::vector<knnj_opts> cand = tune_candidates(o, R, Q, self);
  return autotune_call(
      c, sig, cand, o, k, tune_may_use_mfma(o, R.d) || low_d_filter_candidate(o, R.d),
      [&](const knnj_opts &x, knnj_stats *st) { return join_once(c, R, Q, k, x, out_idx, out_dist, st); }, stats);
}

void knnj_index_destroy(knnj_index *x) {
  if (!x) return;
  hipSetDevice(x->device);
  hipDeviceSynchronize();
  for (void *p : x->owned) hipFree(p);
  delete x;
}

knnj_status knnj_index_build(knnj_ctx *c, knnj_matrix R, const knnj_opts *opts, knnj_index **out) {
  g_err.clear();
  if (!c || !out) {
    set_err("context or output pointer is NULL");
    return KNNJ_EINVAL;
  }
  *out = nullptr;
  knnj_opts o{};
  if (opts) o = *opts;
  TRY(check_matrix(R, "R"));
  if (R.n < 1) {
    set_err("R is empty");
    return KNNJ_EINVAL;
  }
  // the option checks of a join (a one-row dummy query side)
  {
    knnj_matrix Q1 = R;
    Q1.n = 0;
    TRY(check_join_args(R, Q1, 1, o, nullptr, nullptr));
  }
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  HIPCHK(hipEventRecord(c->ev[0], s));
  Inst I;
  I.d = R.d;
  I.k = 1;
  // a resident index is built once: more split rounds and Lloyd steps by
  // default, its Lloyd means over every row
  if (o.split_rounds == 0) o.split_rounds = 16;
  if (o.refine_iters == 0) o.refine_iters = 2;
  I.o = o;
  I.resident = true;
  I.nR = R.n;
  I.ldR = R.ld ? R.ld : R.d;
  (R.on_device ? I.dR : I.hR) = R.data;
  Alloc A;
  A.c = c;
  TRY(stage_inputs(A, I, s));
  u32 *flags;
  TRY(ws(c, "flags", 8, &flags));
  HIPCHK(hipMemsetAsync(flags, 0, 8 * sizeof(u32), s));
  launch_scan_input(I.dR, I.nR, I.d, I.ldR, flags, s);
  u32 hflags[3];
  HIPCHK(hipMemcpyAsync(hflags, flags, sizeof(hflags), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  TRY(decide(I, hflags));

  knnj_index *x = new knnj_index();
  x->device = c->device;
  x->d = I.d;
  x->n = I.nR;
  x->o = o;
  std::memcpy(x->flags, hflags, sizeof(hflags));
  x->kind = I.kind;
  x->dpad = I.dpad;
  x->dk = I.dk;
  x->m_r = I.m_r;
  x->mfma = I.mfma;
  x->bf16 = I.bf16;
  x->bf16x1 = I.bf16x1;
  x->fp8 = I.fp8;
  // persistent buffers owned by the index; build scratch from the context
  Alloc P;
  P.c = c;
  P.owned = &x->owned;
  SideBufs &b = I.rb;
  const int64_t n = I.nR;
  const int m = I.m_r;
  auto fail = [&](knnj_status e) {
    knnj_index_destroy(x);
    return e;
  };
#define OWN(expr)                            \
  do {                                       \
    const knnj_status e_ = (expr);           \
    if (e_ != KNNJ_OK) return fail(e_);      \
  } while (0)
  OWN(P.get("reps", (size_t)m * I.dpad, &b.reps));
  OWN(P.get("rad", (size_t)m, &b.radbits));
  OWN(P.get("off", (size_t)m + 1, &b.off));
  OWN(P.get("perm", (size_t)n, &b.perm));
  OWN(P.get("pdist", (size_t)n, &b.pdist));
  OWN(P.get("px", (size_t)n * I.dpad, &b.px));
  if (I.mfma) {
    char *rc;
    OWN(P.get("rc", (size_t)n * I.dk * 4, &rc));
    I.rc = rc;
    I.rinv = (float *)(rc + (size_t)n * I.dk);
    OWN(P.get("nr", (size_t)n, &I.nr));
    OWN(P.get("reps32", (size_t)m * I.dk, &I.reps32));
    OWN(P.get("rn32", (size_t)m, &I.rn32));
  }
  OWN(A.get("ix_cnt", (size_t)m, &b.cnt));
  OWN(A.get("ix_fill", (size_t)m, &b.fill));
  OWN(A.get("ix_cl", (size_t)n, &b.cl));
  OWN(A.get("ix_d2r", (size_t)n, &b.d2r));
  OWN(A.get("ix_gkeys", (size_t)n, &b.gkeys));
  OWN(A.get("ix_cand", (size_t)m * I.trials, &b.cand));
  OWN(A.get("ix_reprows", (size_t)m, &b.rep_rows));
  b.nblk = rep_score_blocks(m);
  OWN(A.get("ix_partial", (size_t)b.nblk * I.trials + 1, &b.partial));  // + the winning trial's slot
  if (I.d >= 16 && ((I.d + 31) & ~31) <= 256) {
    OWN(A.get("ix_repsbf16", (size_t)m * ((I.d + 31) & ~31), &b.rb));
    OWN(A.get("ix_repsnrm", (size_t)m, &b.rn));
  }
  OWN(enqueue_side(c, I, true, s));
  if (I.mfma) {
    launch_center(I.Ri, I.d, I.dk, center_planes(I), I.rc, I.nr, I.reps32, I.rn32, I.rinv, s);
    if (hipGetLastError() != hipSuccess) {
      set_err("centring launch failed");
      return fail(KNNJ_EDEVICE);
    }
  }
  if (hipEventRecord(c->ev[1], s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
    set_err("index build failed on the device");
    return fail(KNNJ_EDEVICE);
  }
#undef OWN
  x->Ri = I.Ri;
  x->rc = I.rc;
  x->nr = I.nr;
  x->reps32 = I.reps32;
  x->rn32 = I.rn32;
  x->rinv = I.rinv;
  x->bytes = P.off;
  x->t_build_ms = elapsed(c->ev[0], c->ev[1]);
  *out = x;
  return KNNJ_OK;
}

knnj_status knnj_index_info(const knnj_index *x, knnj_index_desc *info) {
  if (!x || !info) {
    set_err("index or output pointer is NULL");
    return KNNJ_EINVAL;
  }
  knnj_index_desc v{};
  v.n = x->n;
  v.d = x->d;
  v.m_r = x->m_r;
  v.kernel = x->kind;
  v.precision = x->fp8 ? KNNJ_PREC_FP8 : x->bf16 ? (x->bf16x1 ? KNNJ_PREC_BF16X1 : KNNJ_PREC_BF16) : KNNJ_PREC_FP32;
  v.device = x->device;
  v.bytes = (int64_t)x->bytes;
  v.t_build_ms = x->t_build_ms;
  *info = v;
  return KNNJ_OK;
}

}  // extern "C"

namespace {

// knnj_index_query with its buffers from the context workspace, or (no_alloc)
// carved from the device's fixed scratch array.
knnj_status index_query_impl(knnj_ctx *c, const knnj_index *x, knnj_matrix Q, int32_t k, const knnj_opts *opts,
                             int32_t *out_idx, float *out_dist, knnj_stats *stats, bool no_alloc) {
  g_err.clear();
  if (!c || !x) {
    set_err("context or index is NULL");
    return KNNJ_EINVAL;
  }
  if (c->device != x->device) {
    set_err("index lives on device %d, context on device %d", x->device, c->device);
    return KNNJ_EINVAL;
  }
  knnj_opts o{};
  if (opts) o = *opts;
  // the index fixed these: a query may not ask for others
  o.kernel = x->kind;
  o.precision = x->fp8 ? KNNJ_PREC_FP8 : x->bf16 ? (x->bf16x1 ? KNNJ_PREC_BF16X1 : KNNJ_PREC_BF16) : KNNJ_PREC_FP32;
  o.brute_force = 0;
  knnj_matrix R{};
  R.data = nullptr;
  R.n = x->n;
  R.d = x->d;
  R.ld = x->d;
  R.on_device = 1;
  {
    // argument checks of a join with R of the index's shape (R.data unused)
    knnj_matrix R1 = R;
    R1.data = (const float *)1;
    TRY(check_join_args(R1, Q, k, o, out_idx, out_dist));
  }
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  Inst I;
  I.pre = x;
  I.d = Q.d;
  I.k = k;
  I.o = o;
  I.nR = x->n;
  I.nQ = Q.n;
  I.ldQ = Q.ld ? Q.ld : Q.d;
  (Q.on_device ? I.dQ : I.hQ) = Q.data;
  I.aos = o.out_aos != 0;
  I.out_host = !Q.on_device;
  I.user_idx = out_idx;
  I.user_dist = out_dist;
  Alloc A;
  A.c = c;
  if (no_alloc) {
    if (!c->scratch) {
      set_err("the context has no scratch array");
      return KNNJ_EDEVICE;
    }
    A.arena = true;
    A.no_alloc = true;
    A.base = (char *)c->scratch;
    A.cap = (size_t)KNNJ_NOALLOC_SCRATCH_BYTES;
  }
  HIPCHK(hipEventRecord(c->ev[0], s));
  TRY(stage_inputs(A, I, s));
  u32 *flags;
  TRY(A.get("flags", 8, &flags));
  HIPCHK(hipMemsetAsync(flags, 0, 8 * sizeof(u32), s));
  if (I.nQ > 0) launch_scan_input(I.dQ, I.nQ, I.d, I.ldQ, flags, s);
  u32 hflags[3];
  HIPCHK(hipMemcpyAsync(hflags, flags, sizeof(hflags), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  // combined flags of the index's R and this Q
  float mq, mr;
  std::memcpy(&mq, &hflags[1], 4);
  std::memcpy(&mr, &x->flags[1], 4);
  const float mx = std::max(mq, mr);
  std::memcpy(&hflags[1], &mx, 4);
  hflags[0] |= x->flags[0];
  hflags[2] |= x->flags[2];
  TRY(decide(I, hflags));
  return run_single(c, I, A, stats);
}

}  // namespace

extern "C" {

knnj_status knnj_index_query(knnj_ctx *c, const knnj_index *x, knnj_matrix Q, int32_t k, const knnj_opts *opts,
                             int32_t *out_idx, float *out_dist, knnj_stats *stats) {
  return index_query_impl(c, x, Q, k, opts, out_idx, out_dist, stats, false);
}

knnj_status knnj_index_query_noalloc(knnj_ctx *c, const knnj_index *x, knnj_matrix Q, int32_t k,
                                     const knnj_opts *opts, int32_t *out_idx, float *out_dist, knnj_stats *stats) {
  return index_query_impl(c, x, Q, k, opts, out_idx, out_dist, stats, true);
}

// One batched call with fixed options (arguments already checked).
static knnj_status join_batch_once(knnj_ctx *c, int32_t n_inst, const knnj_instance *inst, const knnj_opts &o,
                                   knnj_stats *stats) {
  knnj_stats st{};
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  hipEvent_t *ev = c->ev;
  HIPCHK(hipEventRecord(ev[0], s));
  std::vector<Inst> I(n_inst);
  for (int i = 0; i < n_inst; ++i) {
    describe(I[i], inst[i].R, inst[i].Q, inst[i].k, o, inst[i].out_idx, inst[i].out_dist);
    I[i].id = i;
    // the fused launches have tiles enough: no automatic split mode
    if (I[i].o.tile_split == 0) I[i].o.tile_split = 1;
    // batched instances are small: the Lloyd step costs more index time than
    // it saves join time (measured on the 512-instance workload: 14.5 M vs
    // 11.7 M queries/s), so it is off unless asked for
    if (I[i].o.refine_iters == 0) I[i].o.refine_iters = -1;
  }

  // inputs: one arena for the host copies, one flag quadruple per join, one sync
  {
    Alloc sz;
    sz.c = c;
    sz.arena = true;
    for (Inst x : I) TRY(stage_inputs(sz, x, s));  // sizing pass on copies, no device work
    char *base = nullptr;
    if (sz.off > 0) TRY(ws(c, "batch_inputs", sz.off, &base));
    Alloc A;
    A.c = c;
    A.arena = true;
    A.base = base;
    for (Inst &x : I) TRY(stage_inputs(A, x, s));
  }
  u32 *flags;
  TRY(ws(c, "batch_flags", (size_t)4 * n_inst, &flags));
  HIPCHK(hipMemsetAsync(flags, 0, sizeof(u32) * 4 * n_inst, s));
  // every input scanned in one launch (the descriptors stay alive until the
  // synchronization below)
  std::vector<ScanDesc> sd;
  {
    int64_t mv = 0;
    for (const Inst &x : I) {
      if (x.nR > 0) sd.push_back(ScanDesc{x.dR, x.nR, x.ldR, x.d, 0, flags + 4 * x.id});
      if (!x.self && x.nQ > 0) sd.push_back(ScanDesc{x.dQ, x.nQ, x.ldQ, x.d, 0, flags + 4 * x.id});
      mv = std::max(mv, std::max(x.nR, x.self ? 0 : x.nQ) * (int64_t)x.d);
    }
    if (!sd.empty()) {
      ScanDesc *dsd;
      TRY(ws(c, "batch_scan", sd.size(), &dsd));
      HIPCHK(hipMemcpyAsync(dsd, sd.data(), sd.size() * sizeof(ScanDesc), hipMemcpyHostToDevice, s));
      launch_scan_inputs(dsd, (int)sd.size(), mv, s);
    }
  }
  std::vector<u32> hflags((size_t)4 * n_inst);
  HIPCHK(hipMemcpyAsync(hflags.data(), flags, sizeof(u32) * hflags.size(), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (Inst &x : I) {
    const knnj_status e = decide(x, &hflags[4 * x.id]);
    if (e != KNNJ_OK) {
      prefix_err("instance %d: ", x.id);
      return e;
    }
  }

  // launch groups: joins sharing a kernel instantiation run in one launch,
  // their tile slots contiguous in one array
  std::vector<int> order;
  for (const Inst &x : I)
    if (x.nQ > 0) order.push_back(x.id);
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return I[a].launch_key() < I[b].launch_key(); });
  int64_t total_slots = 0;
  std::vector<int64_t> slot0(n_inst, 0);
  for (int id : order) {
    slot0[id] = total_slots;
    total_slots += I[id].max_slots;
  }
  if (total_slots > (int64_t)INT32_MAX) {
    set_err("batch needs %lld tile slots, above 2^31-1", (long long)total_slots);
    return KNNJ_EUNSUPPORTED;
  }
  Tile *all_tiles;
  TRY(ws(c, "batch_tiles", (size_t)std::max<int64_t>(total_slots, 1), &all_tiles));
  // every join's counters in two arrays: cleared in one launch and read back
  // in one copy each (per-join launches and copies cost as much as the
  // small joins' kernels)
  int32_t *all_counters;
  u64 *all_evals;
  TRY(ws(c, "batch_counters", (size_t)8 * n_inst, &all_counters));
  TRY(ws(c, "batch_evals", (size_t)4 * n_inst, &all_evals));
  // eight tile-slot counters per fused launch (the persistent launches' work queues)
  int32_t *queues;
  TRY(ws(c, "batch_queues", (size_t)8 * (n_inst + 1), &queues));
  for (Inst &x : I) {
    x.counters = all_counters + 8 * (size_t)x.id;
    x.evals = all_evals + 4 * (size_t)x.id;
    x.batch_counters = true;
  }
  {
    // one arena for every join's buffers: a sizing pass, then the carve
    Alloc sz;
    sz.c = c;
    sz.arena = true;
    for (int id : order) {
      Inst tmp = I[id];
      TRY(alloc_inst(sz, tmp, all_tiles + slot0[id]));
    }
    char *arena = nullptr;
    TRY(ws(c, "batch_arena", sz.off, &arena));
    Alloc A;
    A.c = c;
    A.arena = true;
    A.base = arena;
    for (int id : order) TRY(alloc_inst(A, I[id], all_tiles + slot0[id]));
  }

  // kernel arguments: fixed by the plan (buffers and decisions), so they are
  // filled and copied before any work is queued
  for (int id : order) {
    Inst &x = I[id];
    fill_side_index(x, true);
    if (x.self)
      x.Qi = x.Ri;
    else
      fill_side_index(x, false);
    fill_args(x);
  }
  JoinArgs *dargs;
  TRY(ws(c, "batch_args", (size_t)n_inst, &dargs));
  std::vector<JoinArgs> hargs(n_inst);
  for (int id : order) {
    TRY(check_args(I[id].args, I[id].nQ));
    hargs[id] = I[id].args;
  }
  HIPCHK(hipMemcpyAsync(dargs, hargs.data(), sizeof(JoinArgs) * n_inst, hipMemcpyHostToDevice, s));
  const int naux = aux_streams();
  TRY(ensure_aux(c, naux));

  int n_launch = 0;
  const bool use_queue = o.work_queue != 0;
  // fused launches run persistent (multi-wave MFMA workgroups keep one workgroup per slot)
  int n_queued = 0;
  if (use_queue)
    for (size_t g = 0; g < order.size(); ++g)
      if (g == 0 || !(I[order[g]].launch_key() == I[order[g - 1]].launch_key())) {
        const Inst &P = I[order[g]];
        if (!P.use_mfma || join_mfma_tile_q(P.args.kcap, P.args.dk, P.args.waves) == 16) ++n_queued;
      }
  // index builds, bounds and tile lists: independent per instance, so they
  // are spread over the auxiliary streams and overlap on the device (most
  // are many small launches); the fused joins wait for all of them
  auto enqueue_all = [&](bool timed) -> knnj_status {
    launch_zeros(s, all_counters, sizeof(int32_t) * 8 * (size_t)n_inst, all_evals, sizeof(u64) * 4 * (size_t)n_inst,
                 queues, sizeof(int32_t) * 8 * ((size_t)n_inst + 1));
    HIPCHK(hipEventRecord(c->aux_ev[knnj_ctx::kAux], s));
    for (int j = 0; j < naux; ++j) HIPCHK(hipStreamWaitEvent(c->aux[j], c->aux_ev[knnj_ctx::kAux], 0));
    {
      // largest instances first, each to the stream with the least queued rows
      std::vector<int> by_size(order);
      std::stable_sort(by_size.begin(), by_size.end(),
                       [&](int a, int b) { return I[a].nR + I[a].nQ > I[b].nR + I[b].nQ; });
      int64_t load[knnj_ctx::kAux] = {};
      for (int id : by_size) {
        const int j = (int)(std::min_element(load, load + naux) - load);
        load[j] += I[id].nR + I[id].nQ;
        TRY(enqueue_index(c, I[id], c->aux[j], nullptr, nullptr));
      }
    }
    for (int j = 0; j < naux; ++j) {
      HIPCHK(hipEventRecord(c->aux_ev[j], c->aux[j]));
      HIPCHK(hipStreamWaitEvent(s, c->aux_ev[j], 0));
    }
    if (timed) {
      HIPCHK(hipEventRecord(ev[2], s));
      HIPCHK(hipEventRecord(ev[8], s));
    }
    n_launch = 0;
    for (size_t g = 0; g < order.size();) {
      size_t h = g;
      while (h < order.size() && I[order[h]].launch_key() == I[order[g]].launch_key()) ++h;
      const Inst &P = I[order[g]];
      const int64_t first = slot0[order[g]];
      const int64_t last = slot0[order[h - 1]] + I[order[h - 1]].max_slots;
      const int nslots = (int)(last - first);
      const Tile *t0 = all_tiles + first;
      // the persistent form takes the launch's slots from its own eight
      // counters (multi-wave MFMA workgroups keep one workgroup per slot)
      int32_t *q = use_queue ? queues + 8 * n_launch : nullptr;
      if (P.use_mfma) {
        if (P.fp8)
          HIPCHK(launch_join_mfma_fp8_multi(P.args, dargs, t0, nslots, s, q));
        else if (P.bf16)
          HIPCHK(launch_join_mfma_bf16_multi(P.args, P.bf16x1 ? 1 : 3, dargs, t0, nslots, s, q));
        else
          HIPCHK(launch_join_mfma_f32_multi(P.args, dargs, t0, nslots, s, q));
      } else {
        HIPCHK(launch_join_multi(P.args, P.tq, dargs, t0, nslots, s, q));
      }
      ++n_launch;
      g = h;
    }
    return KNNJ_OK;
  };

  HIPCHK(hipEventRecord(ev[1], s));
  bool graphed = false;
  int graph_state = 0;
  // fused launches of this plan (one per kernel instantiation)
  int n_groups = 0;
  for (size_t g = 0; g < order.size(); ++g)
    if (g == 0 || !(I[order[g]].launch_key() == I[order[g - 1]].launch_key())) ++n_groups;
  if (o.graph) {
    // the plan's signature: every kernel argument and every host-side
    // parameter of the queued work
    std::vector<unsigned char> sig(reinterpret_cast<const unsigned char *>(hargs.data()),
                                   reinterpret_cast<const unsigned char *>(hargs.data() + n_inst));
    auto put = [&](const void *p, size_t n) {
      sig.insert(sig.end(), (const unsigned char *)p, (const unsigned char *)p + n);
    };
    for (int id : order) {
      const Inst &x = I[id];
      const int64_t v[] = {x.id, (int64_t)(uintptr_t)x.dR, (int64_t)(uintptr_t)x.dQ, x.ldR, x.ldQ, x.nR, x.nQ,
                           x.m_r, x.m_q, x.trials, (int64_t)x.seed, x.refine, x.split_rounds, x.integral,
                           x.bf16, x.bf16x1, x.mfma, x.tq, x.max_slots, slot0[id], x.self, x.fp8};
      put(v, sizeof(v));
    }
    put(&all_tiles, sizeof(all_tiles));
    put(&dargs, sizeof(dargs));
    put(&all_counters, sizeof(all_counters));
    put(&all_evals, sizeof(all_evals));
    put(&queues, sizeof(queues));
    put(&use_queue, sizeof(use_queue));
    if (c->batch_exec && sig == c->batch_sig) {
      HIPCHK(hipGraphLaunch(c->batch_exec, s));
      graphed = true;
      graph_state = 2;
    } else {
      if (c->batch_exec) {
        hipGraphExecDestroy(c->batch_exec);
        c->batch_exec = nullptr;
        c->batch_sig.clear();
      }
      HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
      const knnj_status e = enqueue_all(false);
      hipGraph_t graph = nullptr;
      const hipError_t ee = hipStreamEndCapture(s, &graph);
      if (e != KNNJ_OK || ee != hipSuccess) {
        if (graph) hipGraphDestroy(graph);
        if (e != KNNJ_OK) return e;
        set_err("graph capture failed: %s", hipGetErrorString(ee));
        return KNNJ_EDEVICE;
      }
      const hipError_t ie = hipGraphInstantiate(&c->batch_exec, graph, nullptr, nullptr, 0);
      hipGraphDestroy(graph);
      if (ie != hipSuccess) {
        c->batch_exec = nullptr;
        set_err("graph instantiation failed: %s", hipGetErrorString(ie));
        return KNNJ_EDEVICE;
      }
      c->batch_sig = std::move(sig);
      HIPCHK(hipGraphLaunch(c->batch_exec, s));
      graphed = true;
      graph_state = 1;
    }
  } else {
    TRY(enqueue_all(true));
  }
  HIPCHK(hipEventRecord(ev[4], s));
  // counters of every join: one copy per array
  std::vector<int32_t> hc((size_t)8 * n_inst, 0);
  std::vector<u64> hcnt((size_t)4 * n_inst, 0);
  HIPCHK(hipMemcpyAsync(hc.data(), all_counters, hc.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(hcnt.data(), all_evals, hcnt.size() * sizeof(u64), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));

  HIPCHK(hipEventRecord(ev[5], s));
  int64_t novf = 0, evals = 0, merges = 0, appended = 0;
  double tot = 0.0, ferr = 0.0;
  for (int id : order) {
    Inst &x = I[id];
    if (hc[8 * id + 1] > 0) {
      int stages = 0;
      const knnj_status e = rerun_overflow(x, hc[8 * id + 1], s, &stages);
      st.rerun_stages = std::max(st.rerun_stages, stages);
      if (e != KNNJ_OK) {
        prefix_err("instance %d: ", id);
        return e;
      }
    }
    novf += hc[8 * id + 1];
    expand_duplicates(x, s);
    TRY(copy_outputs(x, s));
    evals += (int64_t)hcnt[4 * id];
    merges += (int64_t)hcnt[4 * id + 1];
    appended += (int64_t)hcnt[4 * id + 2];
    tot += (double)x.nR * (double)x.nQ;
    float fe;
    std::memcpy(&fe, &hc[8 * id + 3], 4);
    ferr = std::max(ferr, (double)fe);
    float f3[3];
    std::memcpy(f3, &hc[8 * id + 5], sizeof(f3));
    st.filter_err_max = std::max(st.filter_err_max, (double)f3[0]);
    st.filter_bound_max = std::max(st.filter_bound_max, (double)f3[1]);
    st.filter_err_ratio = std::max(st.filter_err_ratio, (double)f3[2]);
  }
  HIPCHK(hipEventRecord(ev[6], s));
  HIPCHK(hipStreamSynchronize(s));
  st.t_total_ms = elapsed(ev[0], ev[6]);
  // index builds, centring, bounds and tile lists of all joins; a replayed
  // graph reports its whole time as join time
  st.t_index_ms = graphed ? 0.0 : elapsed(ev[1], ev[2]);
  st.t_join_ms = graphed ? elapsed(ev[1], ev[4]) : elapsed(ev[8], ev[4]);
  st.t_fallback_ms = novf > 0 ? elapsed(ev[5], ev[6]) : 0.0;
  st.n_dist_evals = evals;
  st.n_merges = merges;
  st.n_appended = appended;
  st.n_overflow = novf;
  st.n_tiles = total_slots;
  st.pruning_ratio = tot > 0 ? 1.0 - (double)evals / tot : 0.0;
  st.filter_max_abs_err = ferr;
  if (!order.empty()) plan_stats(I[order[0]], st);
  st.n_launches = n_groups;
  st.queue_launches = n_queued;
  st.graph = graph_state;
  (void)n_launch;
  if (stats) *stats = st;
  return KNNJ_OK;
}

knnj_status knnj_join_batch(knnj_ctx *c, int32_t n_inst, const knnj_instance *inst, const knnj_opts *opts,
                            knnj_stats *stats) {
  g_err.clear();
  if (!c) {
    set_err("context is NULL");
    return KNNJ_EINVAL;
  }
  if (n_inst < 0 || (n_inst > 0 && !inst)) {
    set_err("bad instance array");
    return KNNJ_EINVAL;
  }
  knnj_opts o{};
  if (opts) o = *opts;
  if (n_inst == 0) {
    if (stats) *stats = knnj_stats{};
    return KNNJ_OK;
  }
  for (int i = 0; i < n_inst; ++i) {
    const knnj_status e = check_join_args(inst[i].R, inst[i].Q, inst[i].k, o, inst[i].out_idx, inst[i].out_dist);
    if (e != KNNJ_OK) {
      prefix_err("instance %d: ", i);
      return e;
    }
  }
  if (!o.autotune) return join_batch_once(c, n_inst, inst, o, stats);
  // Batch autotune: the batch's shape (every instance's sizes, d, k and
  // self-join flag, and the options the caller fixed) keys the choice.
  // First phase: the Lloyd step off (the batch default) or on, and where an
  // instance may run on the vector ALUs, their 2 or 4 queries per tile;
  // second phase: the doubled k-select capacity and the 4-wave lockstep tile.
  std::vector<int64_t> sig = {-1, n_inst, o.precision, o.kernel, o.m_r, o.m_q, o.refine_iters, o.tile_q,
                              o.tile_waves, o.kcap, o.brute_force, o.graph, o.split_rounds, o.work_queue,
                              o.rep_scale};
  int kmax = 1;
  bool any_valu = false, any_mfma = false, any_low_d = false;
  int min_d = 1 << 30;
  for (int i = 0; i < n_inst; ++i) {
    const knnj_matrix &R = inst[i].R, &Q = inst[i].Q;
    const int64_t ldR = R.ld ? R.ld : R.d, ldQ = Q.ld ? Q.ld : Q.d;
    const bool self = R.data == Q.data && R.n == Q.n && ldR == ldQ && R.on_device == Q.on_device;
    sig.insert(sig.end(), {R.n, Q.n, (int64_t)R.d, (int64_t)inst[i].k, (int64_t)self});
    kmax = std::max(kmax, (int)inst[i].k);
    const bool mf = tune_may_use_mfma(o, R.d);
    any_mfma |= mf;
    any_valu |= !mf;
    any_low_d |= R.d >= 16 && R.d < 64;
    min_d = std::min(min_d, (int)R.d);
  }
  std::vector<knnj_opts> cand;
  const int refs[2] = {o.refine_iters, 1};
  const int nref = o.refine_iters ? 1 : 2;
  const int tqs[3] = {o.tile_q, 2, 4};
  const int ntq = (any_valu && !o.tile_q) ? 3 : 1;
  for (int r = 0; r < nref; ++r)
    for (int t = 0; t < ntq; ++t) {
      knnj_opts x = o;
      x.autotune = 0;
      x.refine_iters = refs[r];
      x.tile_q = tqs[t];
      cand.push_back(x);
    }
  if (!o.work_queue) {  // and the first candidate's persistent form
    knnj_opts x = cand[0];
    x.work_queue = 1;
    cand.push_back(x);
  }
  if (!o.rep_scale && !o.m_r && !o.m_q && !o.brute_force) {  // and twice the representatives
    knnj_opts x = cand[0];
    x.rep_scale = 2;
    cand.push_back(x);
  }
  // and the BF16 hi/lo filter on the matrix cores for every instance, when
  // all have d >= 16: by default the instances of 16 <= d < 64 run plain
  // FP32 (on the vector ALUs below d = 32); at 512 uneven joins of d 16, 32,
  // 96 the filter's join phase took 82 against 109 ms
  if (any_low_d && min_d >= 16 && low_d_filter_candidate(o, 16)) {
    knnj_opts x = cand[0];
    x.precision = KNNJ_PREC_BF16;
    cand.push_back(x);
    any_mfma = true;
  }
  return autotune_call(
      c, sig, cand, o, kmax, any_mfma,
      [&](const knnj_opts &x, knnj_stats *st) { return join_batch_once(c, n_inst, inst, x, st); }, stats);
}


}  // extern "C"

extern "C" {

knnj_status knnj_ctx_last_plan(knnj_ctx *c, knnj_plan_desc *out) {
  g_err.clear();
  if (!c || !out) {
    set_err("context or output pointer is NULL");
    return KNNJ_EINVAL;
  }
  const knnj_ctx::Plan &P = c->plan;
  knnj_plan_desc v{};
  v.valid = P.valid ? 1 : 0;
  v.n_r = P.n_r;
  v.n_q = P.n_q;
  v.d = P.d;
  v.dpad = P.dpad;
  v.m_r = P.m_r;
  v.m_q = P.m_q;
  v.k = P.k;
  v.self_join = P.self;
  *out = v;
  return KNNJ_OK;
}

knnj_status knnj_ctx_copy_plan_part(knnj_ctx *c, int32_t part, void *dst, int64_t bytes) {
  g_err.clear();
  if (!c || !dst) {
    set_err("context or destination is NULL");
    return KNNJ_EINVAL;
  }
  const knnj_ctx::Plan &P = c->plan;
  if (!P.valid) {
    set_err("the context has run no single join");
    return KNNJ_EINVAL;
  }
  const void *src = nullptr;
  size_t n = 0;
  const bool q = part >= KNNJ_PART_Q_REPS && part <= KNNJ_PART_Q_PDIST;
  const SideIndex &S = q ? P.Qi : P.Ri;
  const int side = q ? part - KNNJ_PART_Q_REPS : part;
  if (part >= KNNJ_PART_R_REPS && part <= KNNJ_PART_Q_PDIST) {
    switch (side) {
      case 0: src = S.reps, n = sizeof(float) * (size_t)S.m * S.dpad; break;
      case 1: src = S.rad, n = sizeof(double) * (size_t)S.m; break;
      case 2: src = S.off, n = sizeof(int32_t) * ((size_t)S.m + 1); break;
      case 3: src = S.perm, n = sizeof(int32_t) * (size_t)S.n; break;
      case 4: src = S.pdist, n = sizeof(float) * (size_t)S.n; break;
    }
  } else if (part == KNNJ_PART_THETA) {
    src = P.theta, n = sizeof(double) * (size_t)P.m_q;
  } else if (part == KNNJ_PART_PAIR_COUNT) {
    src = P.pcnt, n = sizeof(int32_t) * (size_t)P.m_q;
  } else if (part == KNNJ_PART_PAIR_LIST) {
    src = P.plist, n = sizeof(int32_t) * (size_t)P.m_q * P.m_r;
  } else if (part == KNNJ_PART_PAIR_LB) {
    src = P.plb, n = sizeof(float) * (size_t)P.m_q * P.m_r;
  } else {
    set_err("unknown plan part %d", part);
    return KNNJ_EINVAL;
  }
  if ((int64_t)n != bytes) {
    set_err("plan part %d holds %zu bytes, %lld asked for", part, n, (long long)bytes);
    return KNNJ_EINVAL;
  }
  if (n == 0) return KNNJ_OK;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(dst, src, n, hipMemcpyDeviceToHost));
  return KNNJ_OK;
}

}  // extern "C"
