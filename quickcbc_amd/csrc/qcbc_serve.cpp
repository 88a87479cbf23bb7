// Serving loop: many analyses resident on one GPU, likelihood requests from
// all of them batched continuously into persistent queue launches.
//
// Requests that arrive while a batch is on the GPU wait in the queue and go
// into the next batch together, whatever analysis they belong to.  Each
// analysis keeps a fixed bin-split length and the kernels' tile grid is fixed,
// so a point's result is bit-for-bit the same alone or in any mix.
// Admission charges each analysis its resident footprint against the budget.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/qcbc.h"
#include "qcbc_api_internal.h"
#include "qcbc_internal.h"

using namespace qcbc;
using Clock = std::chrono::steady_clock;

namespace {

constexpr int kCpb = 8;
constexpr int kTile = 512;
constexpr int kSplitBins = 4096;

struct Request {
  int64_t id = 0;
  int aid = 0;
  int B = 0;
  std::vector<double> params, logL, rho;
  std::vector<int> status;
  int state = 0;  // 0 queued, 1 on the GPU, 2 done, 3 canceled
  int result = QCBC_OK;
  Clock::time_point t_submit, t_done;
};

struct Served {
  qcbc_ctx* ctx = nullptr;
  size_t bytes = 0;
  int nd = 1;
  int split_len_splits = 1;
  bool evicting = false;
  int in_flight = 0;
};

}  // namespace

struct qcbc_server {
  size_t budget = 0, in_use = 0;
  int precision = 0;
  int device = 0;
  int max_points = 1 << 16;  // per batch
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::map<int, Served> analyses;
  std::deque<std::shared_ptr<Request>> queue;
  std::unordered_map<int64_t, std::shared_ptr<Request>> reqs;
  int next_aid = 1;
  int64_t next_rid = 1;
  bool stop = false;
  std::thread worker;
  // device staging for one batch
  hipStream_t stream = nullptr;
  double *d_params = nullptr, *d_logL = nullptr, *d_rho = nullptr, *d_partial = nullptr;
  int* d_status = nullptr;
  Point* d_pts = nullptr;
  LLInst* d_insts = nullptr;
  LLItem* d_items = nullptr;
  int* d_counter = nullptr;
  int cap_items = 0, cap_insts = 0;
  double *h_params = nullptr, *h_logL = nullptr, *h_rho = nullptr;  // pinned
  int* h_status = nullptr;
  size_t scratch_bytes = 0;
  long long evals = 0, batches = 0;
  int blocks = 1024;
};

#define HIPCHK(x)                \
  do {                           \
    hipError_t e_ = (x);         \
    if (e_ != hipSuccess) {      \
      qcbc::t_last_hip = e_;     \
      return QCBC_EHIP;          \
    }                            \
  } while (0)

static int grow_items(qcbc_server* s, int n_items, int n_insts) {
  if (n_items > s->cap_items) {
    if (s->d_items) hipFree(s->d_items);
    s->d_items = nullptr;
    int cap = std::max(n_items, 2 * s->cap_items);
    HIPCHK(hipMalloc(&s->d_items, sizeof(LLItem) * cap));
    s->cap_items = cap;
  }
  if (n_insts > s->cap_insts) {
    if (s->d_insts) hipFree(s->d_insts);
    s->d_insts = nullptr;
    int cap = std::max(n_insts, 2 * s->cap_insts);
    HIPCHK(hipMalloc(&s->d_insts, sizeof(LLInst) * cap));
    s->cap_insts = cap;
  }
  return QCBC_OK;
}

// Evaluate one batch: requests grouped by analysis, one set-up launch per
// analysis, one persistent queue launch for everything, finalize per analysis.
static int run_batch(qcbc_server* s, std::vector<std::shared_ptr<Request>>& batch) {
  std::map<int, std::vector<Request*>> by_a;
  for (auto& r : batch) by_a[r->aid].push_back(r.get());
  std::vector<LLInst> insts;
  std::vector<std::pair<double, LLItem>> items;
  std::vector<int> offs;  // point offset of each analysis in the staging
  std::vector<qcbc_ctx*> ctxs;
  int P = 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto& kv : by_a) {
      Served& a = s->analyses.at(kv.first);
      ctxs.push_back(a.ctx);
    }
  }
  int ai = 0;
  for (auto& kv : by_a) {
    qcbc_ctx* c = ctxs[ai];
    int Ba = 0;
    for (Request* r : kv.second) {
      std::memcpy(s->h_params + (size_t)11 * (P + Ba), r->params.data(), sizeof(double) * 11 * r->B);
      Ba += r->B;
    }
    int splits = std::min(QCBC_MAX_SPLITS, std::max(1, (c->nb + kSplitBins - 1) / kSplitBins));
    LLInst in = make_inst(c, kTile, Ba, splits, s->d_pts + P, s->d_partial + (size_t)6 * QCBC_MAX_SPLITS * P);
    const int groups = (Ba + kCpb - 1) / kCpb;
    for (int g = 0; g < groups; g++)
      for (int sp = 0; sp < in.splits; sp++) {
        int bins = std::min(in.split_len, c->nb - sp * in.split_len);
        int npts = std::min(kCpb, Ba - g * kCpb);
        items.push_back({(double)bins * npts * (1.0 + 0.5 * in.nd), LLItem{ai, g, sp, 0}});
      }
    insts.push_back(in);
    offs.push_back(P);
    P += Ba;
    ai++;
  }
  std::stable_sort(items.begin(), items.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
  std::vector<LLItem> flat(items.size());
  for (size_t k = 0; k < items.size(); k++) flat[k] = items[k].second;
  int st = grow_items(s, (int)flat.size(), (int)insts.size());
  if (st) return st;
  hipStream_t q = s->stream;
  int max_nd = 1;
  for (auto& in : insts) max_nd = std::max(max_nd, in.nd);
  HIPCHK(hipMemcpyAsync(s->d_params, s->h_params, sizeof(double) * 11 * P, hipMemcpyHostToDevice, q));
  HIPCHK(hipMemcpyAsync(s->d_insts, insts.data(), sizeof(LLInst) * insts.size(), hipMemcpyHostToDevice, q));
  HIPCHK(hipMemcpyAsync(s->d_items, flat.data(), sizeof(LLItem) * flat.size(), hipMemcpyHostToDevice, q));
  for (size_t i = 0; i < insts.size(); i++)
    HIPCHK(launch_setup(s->d_params + (size_t)11 * offs[i], insts[i].B, ctxs[i]->an, s->d_pts + offs[i],
                        s->d_status + offs[i], q));
  HIPCHK(launch_fill_i32(s->d_counter, 1, 0, q));
  HIPCHK(launch_ll_queue(s->precision, s->d_insts, s->d_items, (int)flat.size(), s->d_counter, max_nd, kCpb, kTile,
                         std::min(s->blocks, std::max(1, (int)flat.size())), q));
  for (size_t i = 0; i < insts.size(); i++)
    HIPCHK(launch_finalize(insts[i].partial, insts[i].splits, insts[i].nd, s->d_pts + offs[i], insts[i].B,
                           s->precision == QCBC_MIXED ? FIN_MIXED : FIN_FP64, ctxs[i]->d_wref, s->d_logL + offs[i],
                           s->d_rho + (size_t)3 * offs[i], q));
  // results: logL (P), rho packed 3 per point slot, status (P)
  HIPCHK(hipMemcpyAsync(s->h_logL, s->d_logL, sizeof(double) * P, hipMemcpyDeviceToHost, q));
  HIPCHK(hipMemcpyAsync(s->h_rho, s->d_rho, sizeof(double) * 3 * P, hipMemcpyDeviceToHost, q));
  HIPCHK(hipMemcpyAsync(s->h_status, s->d_status, sizeof(int) * P, hipMemcpyDeviceToHost, q));
  HIPCHK(hipStreamSynchronize(q));
  // scatter back (finalize wrote rho with stride nd from each analysis offset*3)
  ai = 0;
  for (auto& kv : by_a) {
    int off = offs[ai], nd = insts[ai].nd, k = 0;
    for (Request* r : kv.second) {
      r->logL.assign(s->h_logL + off + k, s->h_logL + off + k + r->B);
      r->rho.assign(s->h_rho + (size_t)3 * off + (size_t)nd * k, s->h_rho + (size_t)3 * off + (size_t)nd * (k + r->B));
      r->status.assign(s->h_status + off + k, s->h_status + off + k + r->B);
      k += r->B;
    }
    ai++;
  }
  s->evals += P;
  s->batches++;
  return QCBC_OK;
}

static void worker_main(qcbc_server* s) {
  hipSetDevice(s->device);
  for (;;) {
    std::vector<std::shared_ptr<Request>> batch;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty(); });
      if (s->stop && s->queue.empty()) break;
      int pts = 0;
      while (!s->queue.empty()) {
        auto r = s->queue.front();
        if (pts > 0 && pts + r->B > s->max_points) break;
        s->queue.pop_front();
        if (r->state == 3) continue;
        r->state = 1;
        pts += r->B;
        s->analyses[r->aid].in_flight++;
        batch.push_back(r);
      }
    }
    if (batch.empty()) continue;
    int st = run_batch(s, batch);
    std::vector<int> done_aids;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      auto now = Clock::now();
      for (auto& r : batch) {
        r->result = st;
        r->state = 2;
        r->t_done = now;
        s->analyses[r->aid].in_flight--;
      }
      // analyses evicted while their requests were on the GPU
      for (auto it = s->analyses.begin(); it != s->analyses.end();) {
        if (it->second.evicting && it->second.in_flight == 0) {
          qcbc_ctx_destroy(it->second.ctx);
          s->in_use -= it->second.bytes;
          it = s->analyses.erase(it);
        } else {
          ++it;
        }
      }
    }
    s->cv_done.notify_all();
  }
}

extern "C" {

int qcbc_server_create(size_t budget_bytes, int precision, qcbc_server** out) {
  if (!out || (precision != QCBC_FP64 && precision != QCBC_MIXED)) return QCBC_EINVAL;
  *out = nullptr;
  int st = ensure_device();
  if (st) return st;
  qcbc_server* s = new qcbc_server();
  s->budget = budget_bytes;
  s->precision = precision;
  hipGetDevice(&s->device);
  const size_t P = s->max_points + 4096;  // a batch may exceed max_points by one request
  s->scratch_bytes = P * (sizeof(double) * (11 + 1 + 3) + sizeof(int) + sizeof(Point) +
                          sizeof(double) * 6 * QCBC_MAX_SPLITS);
  if (s->scratch_bytes > budget_bytes ||
      hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&s->d_params, sizeof(double) * 11 * P) != hipSuccess ||
      hipMalloc(&s->d_logL, sizeof(double) * P) != hipSuccess ||
      hipMalloc(&s->d_rho, sizeof(double) * 3 * P) != hipSuccess ||
      hipMalloc(&s->d_status, sizeof(int) * P) != hipSuccess ||
      hipMalloc(&s->d_pts, sizeof(Point) * P) != hipSuccess ||
      hipMalloc(&s->d_partial, sizeof(double) * 6 * QCBC_MAX_SPLITS * P) != hipSuccess ||
      hipMalloc(&s->d_counter, sizeof(int)) != hipSuccess ||
      hipHostMalloc(&s->h_params, sizeof(double) * 11 * P, 0) != hipSuccess ||
      hipHostMalloc(&s->h_logL, sizeof(double) * P, 0) != hipSuccess ||
      hipHostMalloc(&s->h_rho, sizeof(double) * 3 * P, 0) != hipSuccess ||
      hipHostMalloc(&s->h_status, sizeof(int) * P, 0) != hipSuccess) {
    bool over = s->scratch_bytes > budget_bytes;
    qcbc_server_destroy(s);
    return over ? QCBC_EBUSY : QCBC_ENOMEM;
  }
  s->in_use = s->scratch_bytes;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, s->device) == hipSuccess) s->blocks = prop.multiProcessorCount * (16 / kCpb);
  s->worker = std::thread(worker_main, s);
  *out = s;
  return QCBC_OK;
}

int qcbc_server_admit(qcbc_server* s, const qcbc_desc* desc, const double* D, const double* SN, int* analysis_id) {
  if (!s || !desc || !D || !SN || !analysis_id) return QCBC_EINVAL;
  size_t need = qcbc_ctx_footprint(desc);
  if (need == 0) return QCBC_EINVAL;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->in_use + need > s->budget) return QCBC_EBUSY;
    s->in_use += need;  // reserve before the (slow) copy
  }
  qcbc_ctx* c = nullptr;
  int st = qcbc_ctx_create(desc, &c);
  if (!st) st = qcbc_bind_data(c, D, SN);
  if (!st && (c->nb + kSplitBins - 1) / kSplitBins > QCBC_MAX_SPLITS) st = QCBC_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  if (st) {
    s->in_use -= need;
    if (c) qcbc_ctx_destroy(c);
    return st;
  }
  Served a;
  a.ctx = c;
  a.bytes = need;
  a.nd = desc->n_det;
  int id = s->next_aid++;
  s->analyses[id] = a;
  *analysis_id = id;
  return QCBC_OK;
}

int qcbc_server_evict(qcbc_server* s, int aid) {
  if (!s) return QCBC_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->analyses.find(aid);
  if (it == s->analyses.end() || it->second.evicting) return QCBC_EINVAL;
  for (auto& r : s->queue)
    if (r->aid == aid && r->state == 0) {
      r->state = 3;
      r->result = QCBC_ECANCELED;
      r->t_done = Clock::now();
    }
  if (it->second.in_flight == 0) {
    qcbc_ctx_destroy(it->second.ctx);
    s->in_use -= it->second.bytes;
    s->analyses.erase(it);
  } else {
    it->second.evicting = true;
  }
  s->cv_done.notify_all();
  return QCBC_OK;
}

int qcbc_server_submit(qcbc_server* s, int aid, const double* params, int B, int64_t* request_id) {
  if (!s || !params || B < 1 || B > s->max_points || !request_id) return QCBC_EINVAL;
  auto r = std::make_shared<Request>();
  r->aid = aid;
  r->B = B;
  r->params.assign(params, params + (size_t)11 * B);
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->analyses.find(aid);
    if (it == s->analyses.end() || it->second.evicting) return QCBC_EINVAL;
    r->id = s->next_rid++;
    s->reqs[r->id] = r;
    s->queue.push_back(r);
    *request_id = r->id;
  }
  s->cv_work.notify_one();
  return QCBC_OK;
}

int qcbc_server_wait(qcbc_server* s, int64_t rid, double* logL, double* rho, int* status, double* latency_ms,
                     int timeout_ms) {
  if (!s) return QCBC_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(rid);
  if (it == s->reqs.end()) return QCBC_EINVAL;
  std::shared_ptr<Request> r = it->second;
  auto ready = [&] { return r->state >= 2; };
  if (timeout_ms < 0) {
    s->cv_done.wait(lk, ready);
  } else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
    return QCBC_ETIMEDOUT;
  }
  s->reqs.erase(rid);
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  if (r->state == 3) return QCBC_ECANCELED;
  if (r->result != QCBC_OK) return r->result;
  const int nd = s->analyses.count(r->aid) ? s->analyses[r->aid].nd : (int)(r->rho.size() / std::max(1, r->B));
  if (logL) std::memcpy(logL, r->logL.data(), sizeof(double) * r->B);
  if (rho) std::memcpy(rho, r->rho.data(), sizeof(double) * (size_t)nd * r->B);
  if (status) std::memcpy(status, r->status.data(), sizeof(int) * r->B);
  return QCBC_OK;
}

int qcbc_server_stats(qcbc_server* s, size_t* bytes_in_use, int* n_analyses, long long* evals, long long* batches) {
  if (!s) return QCBC_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  if (bytes_in_use) *bytes_in_use = s->in_use;
  if (n_analyses) *n_analyses = (int)s->analyses.size();
  if (evals) *evals = s->evals;
  if (batches) *batches = s->batches;
  return QCBC_OK;
}

int qcbc_server_destroy(qcbc_server* s) {
  if (!s) return QCBC_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    for (auto& r : s->queue)
      if (r->state == 0) {
        r->state = 3;
        r->result = QCBC_ECANCELED;
      }
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  s->cv_done.notify_all();
  hipSetDevice(s->device);
  for (auto& kv : s->analyses) qcbc_ctx_destroy(kv.second.ctx);
  void* dbufs[] = {s->d_params, s->d_logL, s->d_rho, s->d_partial, s->d_status, s->d_pts, s->d_insts, s->d_items,
                   s->d_counter};
  for (void* b : dbufs)
    if (b) hipFree(b);
  void* hbufs[] = {s->h_params, s->h_logL, s->h_rho, s->h_status};
  for (void* b : hbufs)
    if (b) hipHostFree(b);
  if (s->stream) hipStreamDestroy(s->stream);
  delete s;
  return QCBC_OK;
}

}  // extern "C"
