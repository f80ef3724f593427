// server.hip -- continuous-batching request loop over a resident index.
//
// The reference serves one query set per process run (src/knnjoin.cu:983-1172,
// with the small-batch kernels KNNQuery_theta / KNNQuery / final at
// knnjoin.cu:859-979 for |Q| <= 13,312).  Here the reference side stays
// resident (knnj_index) and requests of 1..max_batch_rows queries arrive at
// any time from any thread:
//   - a request is queued with its arrival time;
//   - each of `slots` workers (own thread, own knnj_ctx: own HIP stream and
//     workspace) takes the oldest waiting request and then every further
//     waiting request while the fused batch stays within max_batch_rows
//     (the memory-footprint admission rule: a slot's workspace is sized
//     for max_batch_rows queries, so a batch that fits rows fits memory);
//   - the fused batch runs as one knnj_index_query on the worker's stream,
//     so batches of different slots overlap on the device;
//   - results are scattered back to each request's buffers and the request
//     is completed.
// Each query's answer is the exact k-NN, independent of what it was batched
// with: the same result alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/knnj.h"

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
  uint64_t ticket = 0;
  const float *q = nullptr;
  int64_t n = 0, ld = 0;
  int32_t *out_idx = nullptr;
  float *out_dist = nullptr;
  Clock::time_point t_submit, t_start, t_done;
  knnj_status status = KNNJ_OK;
  std::string error;
  int32_t batch_rows = 0, batch_requests = 0, slot = -1;
  bool done = false;
  bool claimed = false;  // a waiter owns the request (it alone frees it)
};

}  // namespace

struct knnj_server {
  const knnj_index *index = nullptr;
  knnj_server_config cfg{};
  int d = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request *> queue;
  std::unordered_map<uint64_t, Request *> all;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::vector<std::thread> workers;
  std::vector<knnj_ctx *> ctxs;
  std::atomic<int64_t> batches{0}, rows{0};
  Clock::time_point t_created = Clock::now();  // origin of the request timestamps
};

namespace {

// Pinned staging of one slot: the fused batch's rows and results.
struct Staging {
  float *q = nullptr;
  int32_t *idx = nullptr;
  float *dist = nullptr;
  ~Staging() {
    if (q) hipHostFree(q);
    if (idx) hipHostFree(idx);
    if (dist) hipHostFree(dist);
  }
};

void worker(knnj_server *S, int slot) {
  knnj_ctx *ctx = S->ctxs[slot];
  const int d = S->d, k = S->cfg.k;
  const int64_t maxr = S->cfg.max_batch_rows;
  Staging st;
  const bool pinned = hipHostMalloc((void **)&st.q, sizeof(float) * maxr * d) == hipSuccess &&
                      hipHostMalloc((void **)&st.idx, sizeof(int32_t) * maxr * k) == hipSuccess &&
                      hipHostMalloc((void **)&st.dist, sizeof(float) * maxr * k) == hipSuccess;
  std::vector<Request *> batch;
  for (;;) {
    batch.clear();
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_work.wait(lk, [&] { return S->stop || !S->queue.empty(); });
      if (S->queue.empty()) return;  // stopping and drained
      int64_t total = 0;
      while (!S->queue.empty()) {
        Request *r = S->queue.front();
        if (!batch.empty() && total + r->n > maxr) break;
        S->queue.pop_front();
        batch.push_back(r);
        total += r->n;
      }
    }
    const Clock::time_point t0 = Clock::now();
    int64_t total = 0;
    for (Request *r : batch) total += r->n;
    knnj_status status = pinned ? KNNJ_OK : KNNJ_ENOMEM;
    std::string err = pinned ? "" : "pinned staging allocation failed";
    if (status == KNNJ_OK) {
      int64_t row = 0;
      for (Request *r : batch) {
        for (int64_t i = 0; i < r->n; ++i) std::memcpy(st.q + (row + i) * d, r->q + i * r->ld, sizeof(float) * d);
        row += r->n;
      }
      knnj_matrix Q{st.q, total, d, d, 0};
      status = knnj_index_query(ctx, S->index, Q, k, &S->cfg.opts, st.idx, st.dist, nullptr);
      if (status != KNNJ_OK) err = knnj_last_error();
    }
    if (status == KNNJ_OK) {
      int64_t row = 0;
      for (Request *r : batch) {
        std::memcpy(r->out_idx, st.idx + row * k, sizeof(int32_t) * r->n * k);
        std::memcpy(r->out_dist, st.dist + row * k, sizeof(float) * r->n * k);
        row += r->n;
      }
    }
    const Clock::time_point t1 = Clock::now();
    S->batches += 1;
    S->rows += total;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      for (Request *r : batch) {
        r->t_start = t0;
        r->t_done = t1;
        r->status = status;
        r->error = err;
        r->batch_rows = (int32_t)total;
        r->batch_requests = (int32_t)batch.size();
        r->slot = slot;
        r->done = true;
      }
    }
    S->cv_done.notify_all();
  }
}

}  // namespace

extern "C" {

knnj_status knnj_server_create(const knnj_index *index, const knnj_server_config *cfg, knnj_server **out) {
  if (!index || !cfg || !out) return KNNJ_EINVAL;
  *out = nullptr;
  knnj_index_desc info;
  if (knnj_index_info(index, &info) != KNNJ_OK) return KNNJ_EINVAL;
  if (cfg->k < 1 || cfg->k > info.n || cfg->slots < 1 || cfg->slots > 16 || cfg->max_batch_rows < 1)
    return KNNJ_EINVAL;
  knnj_server *S = new knnj_server();
  S->index = index;
  S->cfg = *cfg;
  S->d = info.d;
  for (int i = 0; i < cfg->slots; ++i) {
    knnj_ctx *c = nullptr;
    const knnj_status e = knnj_ctx_create(info.device, &c);
    if (e != KNNJ_OK) {
      for (knnj_ctx *x : S->ctxs) knnj_ctx_destroy(x);
      delete S;
      return e;
    }
    if (cfg->slot_budget > 0) knnj_ctx_set_budget(c, cfg->slot_budget);
    S->ctxs.push_back(c);
  }
  for (int i = 0; i < cfg->slots; ++i) S->workers.emplace_back(worker, S, i);
  *out = S;
  return KNNJ_OK;
}

knnj_status knnj_server_submit(knnj_server *S, const float *Q, int64_t n, int64_t ld, int32_t *out_idx,
                               float *out_dist, uint64_t *ticket) {
  if (!S || !ticket || n < 1 || !Q || !out_idx || !out_dist) return KNNJ_EINVAL;
  if (ld == 0) ld = S->d;
  if (ld < S->d) return KNNJ_EINVAL;
  if (n > S->cfg.max_batch_rows) return KNNJ_EUNSUPPORTED;
  Request *r = new Request();
  r->q = Q;
  r->n = n;
  r->ld = ld;
  r->out_idx = out_idx;
  r->out_dist = out_dist;
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) {
      delete r;
      return KNNJ_EINVAL;
    }
    r->ticket = S->next_ticket++;
    S->all[r->ticket] = r;
    S->queue.push_back(r);
    *ticket = r->ticket;
  }
  S->cv_work.notify_one();
  return KNNJ_OK;
}

knnj_status knnj_server_wait(knnj_server *S, uint64_t ticket, knnj_request_info *info) {
  if (!S) return KNNJ_EINVAL;
  Request *r = nullptr;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(ticket);
    // unknown, already waited for, or being waited for by another thread
    if (it == S->all.end() || it->second->claimed) return KNNJ_EINVAL;
    r = it->second;
    r->claimed = true;  // later waiters on this ticket are turned away above
    S->cv_done.wait(lk, [&] { return r->done; });
    S->all.erase(ticket);  // by key: inserts while waiting may have rehashed the map
  }
  const knnj_status st = r->status;
  if (info) {
    knnj_request_info v{};
    v.status = st;
    v.batch_rows = r->batch_rows;
    v.batch_requests = r->batch_requests;
    v.slot = r->slot;
    v.queue_ms = ms_between(r->t_submit, r->t_start);
    v.latency_ms = ms_between(r->t_submit, r->t_done);
    v.t_submit_ms = ms_between(S->t_created, r->t_submit);
    v.t_start_ms = ms_between(S->t_created, r->t_start);
    v.t_end_ms = ms_between(S->t_created, r->t_done);
    std::snprintf(v.message, sizeof(v.message), "%s", r->error.c_str());
    *info = v;
  }
  delete r;
  return st;
}

knnj_status knnj_server_counters(knnj_server *S, int64_t *batches, int64_t *rows) {
  if (!S) return KNNJ_EINVAL;
  if (batches) *batches = S->batches.load();
  if (rows) *rows = S->rows.load();
  return KNNJ_OK;
}

void knnj_server_destroy(knnj_server *S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  for (auto &t : S->workers) t.join();  // workers drain the queue first
  for (auto &kv : S->all) delete kv.second;
  for (knnj_ctx *c : S->ctxs) knnj_ctx_destroy(c);
  delete S;
}

}  // extern "C"
