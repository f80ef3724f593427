// api.hip -- the C ABI (include/knnj.h) and the host-side planner.
//
// The planner replaces the reference driver main() (src/knnjoin.cu:983-1172):
// argument checks instead of print-and-continue (knnjoin.cu:38-41), one
// grow-only device workspace per context instead of 3*m_Q + m_R cudaMallocs
// that are never freed (knnjoin.cu:305-309, 347-349), one stream with hipEvent
// stage timers instead of CLOCK_REALTIME around cudaDeviceSynchronize
// (common8.h:83-90), and device-side sorts instead of the OpenMP
// cudaMemcpy round trips (knnjoin.cu:380-400, 1076-1094).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
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

struct DevBuf {
  void *p = nullptr;
  size_t n = 0;
};

uint64_t splitmix64(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

struct Status {
  knnj_status code = KNNJ_OK;
  bool ok() const { return code == KNNJ_OK; }
};

}  // namespace

struct knnj_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::map<std::string, DevBuf> ws;
  int64_t budget = 0;
  std::vector<int32_t> host_cand[2];
  hipEvent_t ev[8] = {};
};

#define HIPCHK(expr)                                                             \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess) {                                                      \
      set_err("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
      return KNNJ_EDEVICE;                                                       \
    }                                                                            \
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
    set_err("workspace budget exceeded: %s needs %zu bytes, %zu in use, budget %lld", name.c_str(), bytes,
            cur, (long long)c->budget);
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

#define TRY(expr)                      \
  do {                                 \
    knnj_status s_ = (expr);           \
    if (s_ != KNNJ_OK) return s_;      \
  } while (0)

int pow2_at_least(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

int default_reps(int64_t n) {
  int64_t m = (int64_t)std::llround(2.0 * std::sqrt((double)n));
  return (int)std::max<int64_t>(1, std::min<int64_t>({m, (int64_t)KNNJ_MAX_REPS, n}));
}

// Builds one side's cluster index from device rows X.
knnj_status build_side(knnj_ctx *c, int side, const float *X, int64_t n, int64_t ld, int d, int dpad, int m,
                       int trials, uint64_t seed, int refine, bool integral, SideIndex *out) {
  const std::string p = side == 0 ? "r_" : "q_";
  hipStream_t s = c->stream;
  float *reps, *d2r, *pdist, *px;
  u64 *radbits, *gkeys;
  int32_t *off, *perm, *cl, *cnt, *fill, *cand, *rep_rows;
  double *partial;
  const int64_t nn = std::max<int64_t>(n, 1);
  TRY(ws(c, p + "reps", (size_t)m * dpad, &reps));
  TRY(ws(c, p + "rad", (size_t)m, &radbits));
  TRY(ws(c, p + "off", (size_t)m + 1, &off));
  TRY(ws(c, p + "cnt", (size_t)m, &cnt));
  TRY(ws(c, p + "fill", (size_t)m, &fill));
  TRY(ws(c, p + "perm", (size_t)nn, &perm));
  TRY(ws(c, p + "cl", (size_t)nn, &cl));
  TRY(ws(c, p + "d2r", (size_t)nn, &d2r));
  TRY(ws(c, p + "pdist", (size_t)nn, &pdist));
  TRY(ws(c, p + "gkeys", (size_t)nn, &gkeys));
  TRY(ws(c, p + "px", (size_t)nn * dpad, &px));
  TRY(ws(c, p + "cand", (size_t)m * trials, &cand));
  TRY(ws(c, p + "reprows", (size_t)m, &rep_rows));
  const int nblk = rep_score_blocks(m);
  TRY(ws(c, p + "partial", (size_t)nblk * trials, &partial));
  __bf16 *rb = nullptr;
  float *rn = nullptr;
  const int dk = (d + 31) & ~31;
  if (d >= 16 && dk <= 256) {
    TRY(ws(c, p + "repsbf16", (size_t)m * dk, &rb));
    TRY(ws(c, p + "repsnrm", (size_t)m, &rn));
  }

  std::vector<int32_t> &hc = c->host_cand[side];
  hc.resize((size_t)m * trials);
  uint64_t st = seed;
  for (auto &v : hc) v = (int32_t)(splitmix64(st) % (uint64_t)n);
  HIPCHK(hipMemcpyAsync(cand, hc.data(), hc.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
  launch_rep_select(X, ld, d, dpad, cand, m, trials, partial, nblk, reps, rep_rows, s);
  // Lloyd refinement: members of each sampled representative (ascending row
  // order), their mean becomes the representative
  for (int it = 0; it < refine; ++it) {
    launch_assign(X, n, ld, d, dpad, reps, m, cl, rb, rn, s);
    HIPCHK(hipMemsetAsync(cnt, 0, sizeof(int32_t) * m, s));
    HIPCHK(hipMemsetAsync(fill, 0, sizeof(int32_t) * m, s));
    launch_count(cl, n, cnt, s);
    launch_bucket_rows(cl, n, m, cnt, off, fill, perm, gkeys, s);
    launch_centroid(X, ld, d, dpad, m, off, perm, integral ? 1 : 0, reps, s);
  }
  launch_assign(X, n, ld, d, dpad, reps, m, cl, rb, rn, s);
  HIPCHK(hipMemsetAsync(radbits, 0, sizeof(u64) * m, s));
  HIPCHK(hipMemsetAsync(cnt, 0, sizeof(int32_t) * m, s));
  HIPCHK(hipMemsetAsync(fill, 0, sizeof(int32_t) * m, s));
  launch_assign_finish(X, n, ld, d, dpad, reps, cl, d2r, radbits, cnt, s);
  launch_bucket_sort_permute(X, n, ld, d, dpad, m, cl, cnt, off, fill, perm, d2r, pdist, gkeys, px, s);
  HIPCHK(hipGetLastError());
  out->n = n;
  out->m = m;
  out->dpad = dpad;
  out->reps = reps;
  out->rad = (const double *)radbits;
  out->off = off;
  out->perm = perm;
  out->pdist = pdist;
  out->px = px;
  return KNNJ_OK;
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

}  // namespace

extern "C" {

const char *knnj_version(void) { return "knnj 0.1.0 (gfx950)"; }

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

knnj_status knnj_join(knnj_ctx *c, knnj_matrix R, knnj_matrix Q, int32_t k, const knnj_opts *opts,
                      int32_t *out_idx, float *out_dist, knnj_stats *stats) {
  g_err.clear();
  if (!c) {
    set_err("context is NULL");
    return KNNJ_EINVAL;
  }
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
  knnj_opts o{};
  if (opts) o = *opts;
  const bool aos = o.out_aos != 0;
  if (Q.n > 0 && (!out_idx || (!aos && !out_dist))) {
    set_err("output pointer is NULL");
    return KNNJ_EINVAL;
  }
  if (o.precision == KNNJ_PREC_FP8) {
    set_err("FP8 filter is not built");
    return KNNJ_EUNSUPPORTED;
  }
  if (o.tile_q != 0 && o.tile_q != 1 && o.tile_q != 2 && o.tile_q != 4 && o.tile_q != 8) {
    set_err("tile_q must be 0, 1, 2, 4 or 8");
    return KNNJ_EINVAL;
  }
  knnj_stats st{};
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const int d = R.d;
  const int64_t ldR = R.ld ? R.ld : d, ldQ = Q.ld ? Q.ld : d;
  const bool self = R.data == Q.data && R.n == Q.n && ldR == ldQ && R.on_device == Q.on_device;
  HIPCHK(hipEventRecord(c->ev[0], s));

  // device-resident inputs
  const float *dR = R.data, *dQ = Q.data;
  int64_t dldR = ldR, dldQ = ldQ;
  if (!R.on_device) {
    float *b;
    TRY(ws(c, "in_r", (size_t)std::max<int64_t>(R.n, 1) * d, &b));
    HIPCHK(hipMemcpy2DAsync(b, sizeof(float) * d, R.data, sizeof(float) * ldR, sizeof(float) * d, R.n,
                            hipMemcpyHostToDevice, s));
    dR = b;
    dldR = d;
  }
  if (self) {
    dQ = dR;
    dldQ = dldR;
  } else if (!Q.on_device && Q.n > 0) {
    float *b;
    TRY(ws(c, "in_q", (size_t)Q.n * d, &b));
    HIPCHK(hipMemcpy2DAsync(b, sizeof(float) * d, Q.data, sizeof(float) * ldQ, sizeof(float) * d, Q.n,
                            hipMemcpyHostToDevice, s));
    dQ = b;
    dldQ = d;
  }

  // input scan: finiteness, integrality, magnitude
  u32 *flags;
  TRY(ws(c, "flags", 8, &flags));
  HIPCHK(hipMemsetAsync(flags, 0, 8 * sizeof(u32), s));
  launch_scan_input(dR, R.n, d, dldR, flags, s);
  if (!self && Q.n > 0) launch_scan_input(dQ, Q.n, d, dldQ, flags, s);
  u32 hflags[3];
  HIPCHK(hipMemcpyAsync(hflags, flags, sizeof(hflags), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (hflags[0]) {
    set_err("non-finite value in R or Q");
    return KNNJ_EINVAL;
  }
  float maxabs;
  std::memcpy(&maxabs, &hflags[1], 4);
  const double span = 2.0 * (double)maxabs;
  const bool integral = !hflags[2];
  const bool exact_valu = integral && (double)d * span * span < 16777216.0;
  const bool exact_mfma = integral && 4.0 * (double)d * span * span < 16777216.0;
  int kind = o.kernel;
  if (o.precision < 0 || o.precision > KNNJ_PREC_BF16X1) {
    set_err("unknown precision %d", o.precision);
    return KNNJ_EINVAL;
  }
  if (o.precision == KNNJ_PREC_BF16 || o.precision == KNNJ_PREC_BF16X1) {
    if (kind == KNNJ_KERNEL_VALU) {
      set_err("the BF16 filter runs on the MFMA kernel only");
      return KNNJ_EINVAL;
    }
    kind = KNNJ_KERNEL_MFMA;
  }
  if (kind == KNNJ_KERNEL_AUTO) kind = (d >= 32 && d <= 256) ? KNNJ_KERNEL_MFMA : KNNJ_KERNEL_VALU;
  if (kind != KNNJ_KERNEL_VALU && kind != KNNJ_KERNEL_MFMA) {
    set_err("unknown kernel %d", o.kernel);
    return KNNJ_EINVAL;
  }
  if (kind == KNNJ_KERNEL_MFMA && d > 256) {
    set_err("the MFMA kernel supports d <= 256 (d = %d)", d);
    return KNNJ_EUNSUPPORTED;
  }
  const bool mfma = kind == KNNJ_KERNEL_MFMA;
  const bool bf16 = mfma && (o.precision == KNNJ_PREC_BF16 || o.precision == KNNJ_PREC_BF16X1 ||
                             (o.precision == KNNJ_PREC_AUTO && d >= 64 && !exact_mfma));
  const bool bf16x1 = bf16 && o.precision == KNNJ_PREC_BF16X1;
  const bool exact = mfma ? (!bf16 && exact_mfma) : exact_valu;
  st.exact_mode = exact ? 1 : 0;
  st.precision_used = bf16 ? (bf16x1 ? KNNJ_PREC_BF16X1 : KNNJ_PREC_BF16) : KNNJ_PREC_FP32;
  st.kernel_used = kind;
  if (Q.n == 0) {
    if (stats) *stats = st;
    return KNNJ_OK;
  }

  const int dpad = (d + 3) & ~3;
  const bool brute = o.brute_force != 0;
  int m_r = brute ? 1 : (o.m_r > 0 ? std::min<int64_t>(o.m_r, R.n) : default_reps(R.n));
  int m_q = brute ? 1 : (o.m_q > 0 ? std::min<int64_t>(o.m_q, Q.n) : default_reps(Q.n));
  m_r = std::min(m_r, KNNJ_MAX_REPS);
  m_q = std::min(m_q, KNNJ_MAX_REPS);
  if (self) m_q = m_r;
  const int trials = o.trials > 0 ? std::min(o.trials, 64) : 10;
  const uint64_t seed = o.seed ? o.seed : 2015;
  st.m_q = m_q;
  st.m_r = m_r;

  HIPCHK(hipEventRecord(c->ev[1], s));
  SideIndex Ri{}, Qi{};
  const int refine = o.refine_iters < 0 ? 0 : (o.refine_iters > 0 ? o.refine_iters : 1);
  TRY(build_side(c, 0, dR, R.n, dldR, d, dpad, m_r, trials, seed, refine, integral, &Ri));
  if (self)
    Qi = Ri;
  else
    TRY(build_side(c, 1, dQ, Q.n, dldQ, d, dpad, m_q, trials, seed, refine, integral, &Qi));
  const int dk = mfma ? (d + 31) & ~31 : 0;
  void *rc = nullptr;
  float *nr = nullptr, *reps32 = nullptr, *rn32 = nullptr;
  if (mfma) {
    TRY(ws_get(c, "r_rc", (size_t)R.n * dk * 4, &rc));
    TRY(ws(c, "r_nr", (size_t)R.n, &nr));
    TRY(ws(c, "r_reps32", (size_t)m_r * dk, &reps32));
    TRY(ws(c, "r_rn32", (size_t)m_r, &rn32));
    launch_center(Ri, d, dk, bf16 ? (bf16x1 ? 1 : 2) : 0, rc, nr, reps32, rn32, s);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipEventRecord(c->ev[2], s));

  double *theta;
  int32_t *plist, *pcnt;
  TRY(ws(c, "theta", (size_t)m_q, &theta));
  TRY(ws(c, "plist", (size_t)m_q * m_r, &plist));
  TRY(ws(c, "pcnt", (size_t)m_q, &pcnt));
  launch_bounds(Qi, Ri, d, k, theta, plist, pcnt, s);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c->ev[3], s));

  // capacity, tile size, tiles
  int kcap = exact ? pow2_at_least(std::max(k, 64))
             : bf16x1 ? pow2_at_least(std::max(2 * k, 64))
                    : pow2_at_least(std::max(k + std::max(8, k / 8), 64));
  kcap = std::min(kcap, mfma ? 512 : KNNJ_MAX_K);
  if (kcap < k) kcap = KNNJ_MAX_K;  // only the VALU kernel holds k > 512
  const bool use_mfma = mfma && kcap <= 512;
  const int tq = use_mfma ? 16 : (o.tile_q ? std::min(o.tile_q, join_tile_q(kcap)) : join_tile_q(kcap));
  const int64_t max_tiles = 8 * (Q.n / tq + m_q + 1);  // XCD-aware slots, empty ones included
  Tile *tiles, *ovf_tiles, *ovf2_tiles;
  int32_t *ntiles, *counters;
  u64 *n_evals;
  TRY(ws(c, "tiles", (size_t)max_tiles, &tiles));
  TRY(ws(c, "ovf_tiles", (size_t)Q.n, &ovf_tiles));
  TRY(ws(c, "ovf2_tiles", (size_t)Q.n, &ovf2_tiles));
  TRY(ws(c, "counters", 8, &counters));  // [0] ntiles, [1] ovf, [2] ovf2, [3] max_err bits, [4] ovf3
  TRY(ws(c, "evals", 4, &n_evals));
  ntiles = counters;
  HIPCHK(hipMemsetAsync(counters, 0, 8 * sizeof(int32_t), s));
  HIPCHK(hipMemsetAsync(n_evals, 0, 4 * sizeof(u64), s));
  launch_make_tiles(Qi.off, m_q, pcnt, tq, tiles, ntiles, (int)max_tiles, o.xcd_tiles ? 8 : 1, s);
  int32_t nslots = 0;
  HIPCHK(hipMemcpyAsync(&nslots, ntiles, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));

  // outputs
  int32_t *oidx = out_idx;
  float *odist = out_dist;
  const size_t nout = (size_t)Q.n * k;
  if (!Q.on_device) {
    TRY(ws(c, "out_idx", nout * (aos ? 2 : 1), &oidx));
    if (!aos) TRY(ws(c, "out_dist", nout, &odist));
  }

  JoinArgs a{};
  a.Q = Qi;
  a.R = Ri;
  a.d = d;
  a.k = k;
  a.kcap = kcap;
  a.exact = exact ? 1 : 0;
  const double delta = exact ? 0.0 : (double)(d + 4) * 0x1p-23 + 0x1p-40;
  a.delta = delta;
  a.fac = (1.0 + delta) / (1.0 - delta) * (1.0 + 0x1p-20);
  a.tiles = tiles;
  a.ntiles = ntiles;
  a.theta = theta;
  a.plist = plist;
  a.pcnt = pcnt;
  a.out_idx = oidx;
  a.out_dist = odist;
  a.out_mode = aos ? 2 : (o.out_squared ? 1 : 0);
  a.n_evals = n_evals;
  a.ovf_tiles = ovf_tiles;
  a.ovf_count = counters + 1;
  a.max_err_bits = (u32 *)(counters + 3);
  a.rc = rc;
  a.nr = nr;
  // per-entry error codes in the MFMA filter keys: rows must fit 26 bits;
  // codes span 63 binades below the largest possible squared distance
  a.ecode = (use_mfma && !exact && R.n < (int64_t)(1 << 26)) ? 1 : 0;
  {
    const double smax = std::max(1e-30, (double)d * span * span);
    a.ebase = std::ilogb(smax) + 1 - 62;
  }
  a.reps32 = reps32;
  a.rn32 = rn32;
  const bool tile_prof = std::getenv("KNNJ_TILE_PROFILE") != nullptr;
  if (tile_prof) {
    u64 *tp;
    TRY(ws(c, "tile_prof", (size_t)max_tiles * 4, &tp));
    HIPCHK(hipMemsetAsync(tp, 0, sizeof(u64) * 4 * max_tiles, s));
    a.tile_prof = tp;
    float *dg;
    TRY(ws(c, "ovf_diag", 256, &dg));
    HIPCHK(hipMemsetAsync(dg, 0, sizeof(float) * 256, s));
    a.ovf_diag = dg;
  }
  a.dk = dk;
  {
    const double u = 0x1p-24, ub = 0x1p-8;
    auto gam = [&](double n) { return n * u / (1.0 - n * u); };
    // |computed dot - (q-c).(r-c)| <= cdot ||q-c|| ||r-c||
    double cdot = gam(d);                                                     // fp32 operands
    if (bf16x1) cdot = 2.0 * ub + ub * ub + gam(d) * (1.0 + ub) * (1.0 + ub);  // one bf16 product
    else if (bf16) cdot = 3.1 * ub * ub + 1.02 * gam(3.0 * d);              // hi/lo, three products
    a.c1 = (float)((2.0 * cdot + 4.0 * u) * 1.01);
    a.c2 = (float)((gam(d + 3) + 8.0 * u) * 1.01);
  }
  if (use_mfma) {
    a.exact = exact ? 1 : 0;
    if (bf16)
      HIPCHK(launch_join_mfma_bf16(a, bf16x1 ? 1 : 3, nslots, s));
    else
      HIPCHK(launch_join_mfma_f32(a, nslots, s));
  } else {
    a.exact = (mfma ? exact_valu : exact) ? 1 : 0;
    HIPCHK(launch_join(a, tq, nslots, s));
  }
  HIPCHK(hipEventRecord(c->ev[4], s));

  int32_t hc[4];
  HIPCHK(hipMemcpyAsync(hc, counters, sizeof(hc), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  st.n_tiles = hc[0];
  st.n_overflow = hc[1];
  if (tile_prof && hc[0] > 0) {
    // diagnostic: distribution of tile durations (wall clock ticks) and cluster counts
    std::vector<u64> tp((size_t)hc[0] * 4);
    HIPCHK(hipMemcpy(tp.data(), a.tile_prof, tp.size() * sizeof(u64), hipMemcpyDeviceToHost));
    std::vector<double> dur;
    double listed = 0, visited = 0;
    u64 t0 = ~0ull, t1 = 0;
    for (int i = 0; i < hc[0]; ++i) {
      if (tp[4 * i + 1] == 0) continue;  // empty slot
      dur.push_back((double)(tp[4 * i + 1] - tp[4 * i]));
      listed += (double)(tp[4 * i + 3] & 0xFFFFFFFFull);
      visited += (double)(tp[4 * i + 3] >> 32);
      t0 = std::min(t0, tp[4 * i]);
      t1 = std::max(t1, tp[4 * i + 1]);
    }
    if (const char *dump = std::getenv("KNNJ_TILE_DUMP")) {
      if (FILE *f = fopen(dump, "wb")) {
        fwrite(tp.data(), sizeof(u64), tp.size(), f);
        fclose(f);
      }
    }
    std::vector<double> sd = dur;
    std::sort(sd.begin(), sd.end());
    double sum = 0;
    for (double v : sd) sum += v;
    const int nt = (int)sd.size();
    if (nt > 0)
    fprintf(stderr,
            "[knnj tile profile] tiles=%d span=%.0f ticks sum=%.0f mean=%.1f p50=%.0f p99=%.0f max=%.0f "
            "listed/tile=%.1f visited/tile=%.1f\n",
            nt, (double)(t1 - t0), sum, sum / nt, sd[nt / 2], sd[(size_t)(0.99 * (nt - 1))], sd.back(),
            listed / nt, visited / nt);
    if (hc[1] > 0) {
      float dg[256];
      HIPCHK(hipMemcpy(dg, a.ovf_diag, sizeof(dg), hipMemcpyDeviceToHost));
      for (int i = 0; i < std::min(hc[1], 8); ++i)
        fprintf(stderr, "[knnj overflow] cutoff=%.6g kth=%.6g last=%.6g first=%.6g\n", dg[4 * i], dg[4 * i + 1],
                dg[4 * i + 2], dg[4 * i + 3]);
    }
  }
  HIPCHK(hipEventRecord(c->ev[5], s));
  if (hc[1] > 0) {
    // overflowed queries, one per tile, at a larger capacity: the MFMA kernel
    // at 512 if it ran below that, else the VALU kernel at 1024
    JoinArgs b = a;
    b.tiles = ovf_tiles;
    b.ntiles = counters + 1;
    b.ovf_tiles = ovf2_tiles;
    b.ovf_count = counters + 2;
    b.tile_prof = nullptr;
    b.ovf_diag = nullptr;
    b.ecode = (b.kcap <= 512 && use_mfma && kcap < 512) ? a.ecode : 0;
    if (use_mfma && kcap < 512) {
      b.kcap = 512;
      if (bf16)
        HIPCHK(launch_join_mfma_bf16(b, bf16x1 ? 1 : 3, hc[1], s));
      else
        HIPCHK(launch_join_mfma_f32(b, hc[1], s));
    } else {
      b.kcap = KNNJ_MAX_K;
      b.exact = exact_valu ? 1 : 0;
      HIPCHK(launch_join(b, 1, hc[1], s));
    }
    int32_t h2 = 0;
    HIPCHK(hipMemcpyAsync(&h2, counters + 2, sizeof(h2), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (h2 > 0) {
      // last resort: the VALU kernel at the largest capacity
      JoinArgs c2 = b;
      c2.tiles = ovf2_tiles;
      c2.ntiles = counters + 2;
      c2.ovf_tiles = ovf_tiles;
      c2.ovf_count = counters + 4;
      c2.kcap = KNNJ_MAX_K;
      c2.exact = exact_valu ? 1 : 0;
      c2.ecode = 0;
      if (!(use_mfma && kcap < 512)) {
        set_err("%d queries overflowed the largest candidate capacity (%d)", h2, KNNJ_MAX_K);
        return KNNJ_EOVERFLOW;
      }
      HIPCHK(launch_join(c2, 1, h2, s));
      int32_t h3 = 0;
      HIPCHK(hipMemcpyAsync(&h3, counters + 4, sizeof(h3), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      if (h3 > 0) {
        set_err("%d queries overflowed the largest candidate capacity (%d)", h3, KNNJ_MAX_K);
        return KNNJ_EOVERFLOW;
      }
    }
  }
  HIPCHK(hipEventRecord(c->ev[6], s));
  if (!Q.on_device) {
    HIPCHK(hipMemcpyAsync(out_idx, oidx, nout * sizeof(int32_t) * (aos ? 2 : 1), hipMemcpyDeviceToHost, s));
    if (!aos) HIPCHK(hipMemcpyAsync(out_dist, odist, nout * sizeof(float), hipMemcpyDeviceToHost, s));
  }
  u64 hcnt[4] = {0, 0, 0, 0};
  u32 herr = 0;
  HIPCHK(hipMemcpyAsync(hcnt, n_evals, sizeof(hcnt), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(&herr, counters + 3, sizeof(herr), hipMemcpyDeviceToHost, s));
  HIPCHK(hipEventRecord(c->ev[7], s));
  HIPCHK(hipStreamSynchronize(s));

  st.t_total_ms = elapsed(c->ev[0], c->ev[7]);
  st.t_index_ms = elapsed(c->ev[1], c->ev[2]);
  st.t_bounds_ms = elapsed(c->ev[2], c->ev[3]);
  st.t_join_ms = elapsed(c->ev[3], c->ev[4]);
  st.t_fallback_ms = hc[1] > 0 ? elapsed(c->ev[5], c->ev[6]) : 0.0;
  const u64 hev = hcnt[0];
  st.n_dist_evals = (int64_t)hev;
  st.n_merges = (int64_t)hcnt[1];
  st.n_appended = (int64_t)hcnt[2];
  const double tot = (double)R.n * (double)Q.n;
  st.pruning_ratio = tot > 0 ? 1.0 - (double)hev / tot : 0.0;
  st.tile_q = tq;
  st.kcap = kcap;
  float ferr;
  std::memcpy(&ferr, &herr, 4);
  st.filter_max_abs_err = ferr;
  // surviving cluster pairs
  {
    std::vector<int32_t> hp(m_q);
    HIPCHK(hipMemcpy(hp.data(), pcnt, sizeof(int32_t) * m_q, hipMemcpyDeviceToHost));
    int64_t t = 0;
    for (int32_t v : hp) t += v;
    st.n_cluster_pairs = t;
  }
  if (stats) *stats = st;
  return KNNJ_OK;
}

}  // extern "C"
