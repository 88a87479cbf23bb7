// Serving loop: many analyses resident on one GPU, likelihood requests from
// all of them batched continuously into persistent queue launches.
//
// An analysis may also own a resident parallel-tempered ensemble (its
// states, logL, pending proposals, temperatures and random-number step
// counter stay in HBM).  A step request advances it by n PT steps: each
// batch gives every analysis with a pending step one step, its proposals'
// set-up before the batch's queue launch and its post kernel (acceptance,
// swaps, next proposals) after it.  An analysis is in at most one batch on
// the GPU at a time, so its steps are ordered and its chain is the same
// alone and in any mix.
//
// Requests that arrive while a batch is on the GPU wait in the queue and go
// into the next batch together, whatever analysis they belong to.  The worker
// keeps two batch slots (device and pinned staging, a stream each): it builds
// and enqueues the next batch while the previous one runs, then completes the
// previous one, so the GPU does not idle while the host gathers, packs and
// unpacks.  Each analysis keeps a fixed bin-split length and the kernels' tile
// grid is fixed, so a point's result is bit-for-bit the same alone or in any
// mix.  Admission charges each analysis its resident footprint against the
// budget.
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
  int steps_left = 0;  // > 0: a PT-step request (B = 0), steps still to run
  std::vector<double> params, logL, rho;
  std::vector<int> status;
  int state = 0;  // 0 queued, 1 on the GPU, 2 done, 3 canceled
  int result = QCBC_OK;
  Clock::time_point t_submit, t_done;
};

// A resident PT ensemble (jiggle proposals, Metropolis acceptance,
// adjacent-temperature swaps: the kernels of the resident PT step), in point
// order t * n_chains + c
struct Ensemble {
  int n_temps = 0, n_chains = 0, P = 0, splits = 1;
  uint64_t seed = 0;
  double *d_x = nullptr, *d_y = nullptr, *d_L = nullptr, *d_Ly = nullptr, *d_rho = nullptr, *d_sigma = nullptr,
         *d_tc = nullptr, *d_partial = nullptr;
  int* d_st = nullptr;
  unsigned long long *d_acc = nullptr, *d_swaps = nullptr, *d_step = nullptr;
  unsigned int* d_done = nullptr;
  Point* d_pts = nullptr;
  long long steps = 0;
  ~Ensemble() {
    void* b[] = {d_x, d_y, d_L, d_Ly, d_rho, d_sigma, d_tc, d_partial, d_st, d_acc, d_swaps, d_step, d_done, d_pts};
    for (void* p : b)
      if (p) (void)hipFree(p);
  }
};

size_t ensemble_bytes(int n_temps, int n_chains, int splits) {
  const size_t P = (size_t)n_temps * n_chains;
  return P * (sizeof(double) * (11 + 11 + 1 + 1 + 3 + 6 * (size_t)splits) + sizeof(int) + sizeof(Point) +
              sizeof(unsigned long long)) +
         sizeof(double) * (11 + 3 * (size_t)n_temps) + sizeof(unsigned long long) * (n_chains + 1) +
         sizeof(unsigned int);
}

struct Served {
  qcbc_ctx* ctx = nullptr;
  size_t bytes = 0;
  int nd = 1;
  int split_len_splits = 1;
  bool evicting = false;
  int in_flight = 0;
  std::shared_ptr<Ensemble> ens;
  bool stepping = false;  // a step of the ensemble is in a batch on the GPU
};

// Device and pinned staging of one batch, and the batch's layout.
struct Slot {
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
  // the batch on this slot: its requests by analysis, point offsets, detectors
  std::vector<std::shared_ptr<Request>> batch;
  std::vector<std::shared_ptr<Request>> steps;  // one step each of these requests' ensembles
  std::vector<std::vector<Request*>> groups;
  std::vector<int> offs, nds;
  int P = 0;
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
  Slot slots[2];
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

static int grow_items(Slot* sl, int n_items, int n_insts) {
  if (n_items > sl->cap_items) {
    if (sl->d_items) (void)hipFree(sl->d_items);
    sl->d_items = nullptr;
    int cap = std::max(n_items, 2 * sl->cap_items);
    HIPCHK(hipMalloc(&sl->d_items, sizeof(LLItem) * cap));
    sl->cap_items = cap;
  }
  if (n_insts > sl->cap_insts) {
    if (sl->d_insts) (void)hipFree(sl->d_insts);
    sl->d_insts = nullptr;
    int cap = std::max(n_insts, 2 * sl->cap_insts);
    HIPCHK(hipMalloc(&sl->d_insts, sizeof(LLInst) * cap));
    sl->cap_insts = cap;
  }
  return QCBC_OK;
}

// Pack one batch into slot sl and enqueue it on the slot's stream: requests
// grouped by analysis, one set-up launch per analysis (for a stepping
// ensemble, of its pending proposals), one persistent queue launch for
// everything (items longest first), finalize per analysis (for an ensemble,
// its post kernel: acceptance, swaps and the next proposals), and the copies
// of the likelihood results back to pinned memory.  Returns without waiting.
static int enqueue_batch(qcbc_server* s, Slot& sl) {
  std::map<int, std::vector<Request*>> by_a;
  for (auto& r : sl.batch) by_a[r->aid].push_back(r.get());
  std::vector<LLInst> insts;
  std::vector<qcbc_ctx*> ctxs, sctxs;
  std::vector<std::shared_ptr<Ensemble>> ens;
  sl.groups.clear();
  sl.offs.clear();
  sl.nds.clear();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto& kv : by_a) ctxs.push_back(s->analyses.at(kv.first).ctx);
    for (auto& r : sl.steps) {
      const Served& a = s->analyses.at(r->aid);
      sctxs.push_back(a.ctx);
      ens.push_back(a.ens);
    }
  }
  // per analysis: the item cost (bins x points x detectors) and its item count
  std::vector<std::pair<double, int>> order;
  int P = 0, ai = 0;
  for (auto& kv : by_a) {
    qcbc_ctx* c = ctxs[ai];
    int Ba = 0;
    for (Request* r : kv.second) {
      std::memcpy(sl.h_params + (size_t)11 * (P + Ba), r->params.data(), sizeof(double) * 11 * r->B);
      Ba += r->B;
    }
    int splits = std::min(QCBC_MAX_SPLITS, std::max(1, (c->nb + kSplitBins - 1) / kSplitBins));
    LLInst in = make_inst(c, kTile, Ba, splits, sl.d_pts + P, sl.d_partial + (size_t)6 * QCBC_MAX_SPLITS * P);
    order.push_back({(double)in.split_len * kCpb * (1.0 + 0.5 * in.nd), ai});
    insts.push_back(in);
    sl.groups.push_back(kv.second);
    sl.offs.push_back(P);
    sl.nds.push_back(in.nd);
    P += Ba;
    ai++;
  }
  sl.P = P;
  const int n_ll = (int)insts.size();
  for (size_t e = 0; e < ens.size(); e++) {
    const Ensemble& E = *ens[e];
    LLInst in = make_inst(sctxs[e], kTile, E.P, E.splits, E.d_pts, E.d_partial);
    order.push_back({(double)in.split_len * kCpb * (1.0 + 0.5 * in.nd), (int)insts.size()});
    insts.push_back(in);
  }
  // longest items first: analyses by per-item cost, then their (group, split) items
  std::stable_sort(order.begin(), order.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
  std::vector<LLItem> flat;
  for (auto& o : order) {
    const LLInst& in = insts[o.second];
    const int groups = (in.B + kCpb - 1) / kCpb;
    for (int g = 0; g < groups; g++)
      for (int sp = 0; sp < in.splits; sp++) flat.push_back(LLItem{o.second, g, sp, 0});
  }
  int st = grow_items(&sl, (int)flat.size(), (int)insts.size());
  if (st) return st;
  hipStream_t q = sl.stream;
  int max_nd = 1;
  for (auto& in : insts) max_nd = std::max(max_nd, in.nd);
  if (P > 0) HIPCHK(hipMemcpyAsync(sl.d_params, sl.h_params, sizeof(double) * 11 * P, hipMemcpyHostToDevice, q));
  HIPCHK(hipMemcpyAsync(sl.d_insts, insts.data(), sizeof(LLInst) * insts.size(), hipMemcpyHostToDevice, q));
  HIPCHK(hipMemcpyAsync(sl.d_items, flat.data(), sizeof(LLItem) * flat.size(), hipMemcpyHostToDevice, q));
  for (int i = 0; i < n_ll; i++)
    HIPCHK(launch_setup(sl.d_params + (size_t)11 * sl.offs[i], insts[i].B, ctxs[i]->an, sl.d_pts + sl.offs[i],
                        sl.d_status + sl.offs[i], q));
  for (size_t e = 0; e < ens.size(); e++)
    HIPCHK(launch_setup(ens[e]->d_y, ens[e]->P, sctxs[e]->an, ens[e]->d_pts, ens[e]->d_st, q));
  HIPCHK(launch_fill_i32(sl.d_counter, 1, 0, q));
  HIPCHK(launch_ll_queue(s->precision, sl.d_insts, sl.d_items, (int)flat.size(), sl.d_counter, max_nd, kCpb, kTile,
                         std::min(s->blocks, std::max(1, (int)flat.size())), q));
  const int fin = s->precision == QCBC_MIXED ? FIN_MIXED : FIN_FP64;
  for (int i = 0; i < n_ll; i++)
    HIPCHK(launch_finalize(insts[i].partial, insts[i].splits, insts[i].nd, sl.d_pts + sl.offs[i], insts[i].B, fin,
                           ctxs[i]->d_wref, sl.d_logL + sl.offs[i], sl.d_rho + (size_t)3 * sl.offs[i], q));
  for (size_t e = 0; e < ens.size(); e++) {
    Ensemble& E = *ens[e];
    HIPCHK(launch_pt_post(insts[n_ll + e], fin, sctxs[e]->d_wref, E.d_x, E.d_L, E.d_y, E.d_Ly, E.d_rho, E.d_st,
                          E.d_sigma, E.d_tc, E.n_temps, E.n_chains, 0, E.n_chains, E.seed, E.d_step, E.d_acc,
                          E.d_swaps, E.d_done, q));
  }
  // results: logL (P), rho packed 3 per point slot, status (P)
  if (P > 0) {
    HIPCHK(hipMemcpyAsync(sl.h_logL, sl.d_logL, sizeof(double) * P, hipMemcpyDeviceToHost, q));
    HIPCHK(hipMemcpyAsync(sl.h_rho, sl.d_rho, sizeof(double) * 3 * P, hipMemcpyDeviceToHost, q));
    HIPCHK(hipMemcpyAsync(sl.h_status, sl.d_status, sizeof(int) * P, hipMemcpyDeviceToHost, q));
  }
  return QCBC_OK;
}

// Wait for slot sl's batch and hand each request its results (st: the
// enqueue's status).
static void complete_batch(qcbc_server* s, Slot& sl, int st) {
  if (!st && hipStreamSynchronize(sl.stream) != hipSuccess) st = QCBC_EHIP;
  if (!st) {
    // finalize wrote rho with stride nd from each analysis offset * 3
    for (size_t ai = 0; ai < sl.groups.size(); ai++) {
      int off = sl.offs[ai], nd = sl.nds[ai], k = 0;
      for (Request* r : sl.groups[ai]) {
        r->logL.assign(sl.h_logL + off + k, sl.h_logL + off + k + r->B);
        r->rho.assign(sl.h_rho + (size_t)3 * off + (size_t)nd * k, sl.h_rho + (size_t)3 * off + (size_t)nd * (k + r->B));
        r->status.assign(sl.h_status + off + k, sl.h_status + off + k + r->B);
        k += r->B;
      }
    }
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto now = Clock::now();
    for (auto& r : sl.batch) {
      r->result = st;
      r->state = 2;
      r->t_done = now;
      s->analyses[r->aid].in_flight--;
    }
    long long step_evals = 0;
    for (auto& r : sl.steps) {
      Served& a = s->analyses[r->aid];
      a.in_flight--;
      a.stepping = false;
      if (st) {
        r->result = st;
        r->steps_left = 0;
      } else {
        a.ens->steps++;
        step_evals += a.ens->P;
        r->steps_left--;
      }
      if (r->steps_left > 0 && (a.evicting || s->stop)) {  // the rest is canceled
        r->state = 3;
        r->result = QCBC_ECANCELED;
        r->steps_left = 0;
      }
      if (r->steps_left == 0) {  // done: out of the queue
        if (r->state != 3) r->state = 2;
        r->t_done = now;
        for (auto it = s->queue.begin(); it != s->queue.end(); ++it)
          if (it->get() == r.get()) {
            s->queue.erase(it);
            break;
          }
      } else if (r->state == 1) {
        r->state = 0;  // the next step goes into a later batch
      }
    }
    if (!st) {
      s->evals += sl.P + step_evals;
      s->batches++;
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
  sl.batch.clear();
  sl.steps.clear();
  s->cv_done.notify_all();
  s->cv_work.notify_one();
}

// Two-deep pipeline: gather a batch (blocking only when nothing is on the
// GPU), enqueue it on the free slot, then complete the batch that was already
// running.
static void worker_main(qcbc_server* s) {
  hipSetDevice(s->device);
  int inflight = -1, inflight_st = QCBC_OK;  // slot of the running batch
  for (;;) {
    std::vector<std::shared_ptr<Request>> batch, steps;
    bool stopping;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      // something that can go on the GPU now: a likelihood request, or a step
      // request whose ensemble is not already stepping
      auto ready = [&] {
        for (auto& r : s->queue)
          if (r->state == 0 && (r->steps_left == 0 || !s->analyses[r->aid].stepping)) return true;
        return false;
      };
      if (inflight < 0) s->cv_work.wait(lk, [&] { return s->stop || ready(); });
      int pts = 0;
      for (auto it = s->queue.begin(); it != s->queue.end();) {
        auto r = *it;
        if (r->state == 3) {
          it = s->queue.erase(it);
          continue;
        }
        if (r->state != 0) {
          ++it;
          continue;
        }
        Served& a = s->analyses[r->aid];
        const int need = r->steps_left > 0 ? a.ens->P : r->B;
        if (r->steps_left > 0 && a.stepping) {
          ++it;
          continue;
        }
        if (pts > 0 && pts + need > s->max_points) break;
        pts += need;
        a.in_flight++;
        r->state = 1;
        if (r->steps_left > 0) {
          a.stepping = true;  // one step of this ensemble per batch
          steps.push_back(r);
          ++it;  // stays queued until its last step is done
        } else {
          batch.push_back(r);
          it = s->queue.erase(it);
        }
      }
      stopping = s->stop && s->queue.empty();
    }
    int next = -1, next_st = QCBC_OK;
    if (!batch.empty() || !steps.empty()) {
      next = inflight == 0 ? 1 : 0;
      s->slots[next].batch = std::move(batch);
      s->slots[next].steps = std::move(steps);
      next_st = enqueue_batch(s, s->slots[next]);
    }
    if (inflight >= 0) complete_batch(s, s->slots[inflight], inflight_st);
    inflight = next;
    inflight_st = next_st;
    if (stopping && inflight < 0) break;
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
  s->scratch_bytes = 2 * P * (sizeof(double) * (11 + 1 + 3) + sizeof(int) + sizeof(Point) +
                              sizeof(double) * 6 * QCBC_MAX_SPLITS);
  bool ok = s->scratch_bytes <= budget_bytes;
  for (Slot& sl : s->slots)
    ok = ok && hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) == hipSuccess &&
         hipMalloc(&sl.d_params, sizeof(double) * 11 * P) == hipSuccess &&
         hipMalloc(&sl.d_logL, sizeof(double) * P) == hipSuccess &&
         hipMalloc(&sl.d_rho, sizeof(double) * 3 * P) == hipSuccess &&
         hipMalloc(&sl.d_status, sizeof(int) * P) == hipSuccess &&
         hipMalloc(&sl.d_pts, sizeof(Point) * P) == hipSuccess &&
         hipMalloc(&sl.d_partial, sizeof(double) * 6 * QCBC_MAX_SPLITS * P) == hipSuccess &&
         hipMalloc(&sl.d_counter, sizeof(int)) == hipSuccess &&
         hipHostMalloc(&sl.h_params, sizeof(double) * 11 * P, 0) == hipSuccess &&
         hipHostMalloc(&sl.h_logL, sizeof(double) * P, 0) == hipSuccess &&
         hipHostMalloc(&sl.h_rho, sizeof(double) * 3 * P, 0) == hipSuccess &&
         hipHostMalloc(&sl.h_status, sizeof(int) * P, 0) == hipSuccess;
  if (!ok) {
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

}  // extern "C"

// The ensemble of an admitted analysis: its buffers, its states' logL, the
// temperatures and counters, and the proposals of its first step.
static int ensemble_create(qcbc_ctx* c, int precision, int n_temps, int n_chains, double ladder, const double* init,
                           const double* sigma, uint64_t seed, int splits, std::shared_ptr<Ensemble>* out) {
  auto E = std::make_shared<Ensemble>();
  E->n_temps = n_temps;
  E->n_chains = n_chains;
  E->P = n_temps * n_chains;
  E->seed = seed;
  E->splits = splits;
  const size_t P = E->P;
  bool ok = hipMalloc(&E->d_x, sizeof(double) * 11 * P) == hipSuccess &&
            hipMalloc(&E->d_y, sizeof(double) * 11 * P) == hipSuccess &&
            hipMalloc(&E->d_L, sizeof(double) * P) == hipSuccess &&
            hipMalloc(&E->d_Ly, sizeof(double) * P) == hipSuccess &&
            hipMalloc(&E->d_rho, sizeof(double) * 3 * P) == hipSuccess &&
            hipMalloc(&E->d_partial, sizeof(double) * 6 * (size_t)splits * P) == hipSuccess &&
            hipMalloc(&E->d_sigma, sizeof(double) * 11) == hipSuccess &&
            hipMalloc(&E->d_tc, sizeof(double) * 3 * n_temps) == hipSuccess &&
            hipMalloc(&E->d_st, sizeof(int) * P) == hipSuccess &&
            hipMalloc(&E->d_acc, sizeof(unsigned long long) * P) == hipSuccess &&
            hipMalloc(&E->d_swaps, sizeof(unsigned long long) * n_chains) == hipSuccess &&
            hipMalloc(&E->d_step, sizeof(unsigned long long)) == hipSuccess &&
            hipMalloc(&E->d_done, sizeof(unsigned int)) == hipSuccess &&
            hipMalloc(&E->d_pts, sizeof(Point) * P) == hipSuccess;
  if (!ok) return QCBC_ENOMEM;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t q = c->stream;
  HIPCHK(hipMemcpyAsync(E->d_x, init, sizeof(double) * 11 * P, hipMemcpyHostToDevice, q));
  HIPCHK(hipMemcpyAsync(E->d_sigma, sigma, sizeof(double) * 11, hipMemcpyHostToDevice, q));
  HIPCHK(launch_pt_temps(E->d_tc, n_temps, ladder, q));
  HIPCHK(launch_fill_u64(E->d_acc, (int)P, 0ULL, q));
  HIPCHK(launch_fill_u64(E->d_swaps, n_chains, 0ULL, q));
  HIPCHK(launch_fill_u64(E->d_step, 1, 0ULL, q));
  HIPCHK(launch_fill_i32(reinterpret_cast<int*>(E->d_done), 1, 0, q));
  int st = loglike_core(c, E->d_x, E->P, precision, E->d_L, E->d_rho, E->d_st);
  if (st) return st;
  HIPCHK(launch_pt_propose(E->d_x, E->d_y, E->d_sigma, E->d_tc, n_temps, n_chains, 0, n_chains, seed, E->d_step, q));
  HIPCHK(hipStreamSynchronize(q));
  *out = E;
  return QCBC_OK;
}

static int admit_impl(qcbc_server* s, const qcbc_desc* desc, const double* D, const double* SN, int n_temps,
                      int n_chains, double ladder, const double* init, const double* sigma, uint64_t seed,
                      int* analysis_id) {
  if (!s || !desc || !D || !SN || !analysis_id) return QCBC_EINVAL;
  const bool pt = n_temps > 0;
  if (pt && (!init || !sigma || n_chains < 1 || n_chains > 1024 || !(ladder >= 1.0) ||
             (size_t)n_temps * n_chains > (size_t)s->max_points))
    return QCBC_EINVAL;
  size_t need = qcbc_ctx_footprint(desc);
  if (need == 0) return QCBC_EINVAL;
  AnalysisConst an{};
  int nh = 0, nb = 0, nbp = 0;
  if (make_analysis(desc, &an, &nh, &nb, &nbp)) return QCBC_EINVAL;
  const int splits = std::max(1, (nb + kSplitBins - 1) / kSplitBins);
  if (splits > QCBC_MAX_SPLITS) return QCBC_EINVAL;
  if (pt) need += ensemble_bytes(n_temps, n_chains, splits);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->in_use + need > s->budget) return QCBC_EBUSY;
    s->in_use += need;  // reserve before the (slow) copy
  }
  qcbc_ctx* c = nullptr;
  std::shared_ptr<Ensemble> E;
  int st = qcbc_ctx_create(desc, &c);
  if (!st) st = qcbc_bind_data(c, D, SN);
  if (!st && pt) st = ensemble_create(c, s->precision, n_temps, n_chains, ladder, init, sigma, seed, splits, &E);
  std::lock_guard<std::mutex> lk(s->mu);
  if (st) {
    s->in_use -= need;
    E.reset();
    if (c) qcbc_ctx_destroy(c);
    return st;
  }
  Served a;
  a.ctx = c;
  a.bytes = need;
  a.nd = desc->n_det;
  a.ens = E;
  int id = s->next_aid++;
  s->analyses[id] = a;
  *analysis_id = id;
  return QCBC_OK;
}

extern "C" {

int qcbc_server_admit(qcbc_server* s, const qcbc_desc* desc, const double* D, const double* SN, int* analysis_id) {
  return admit_impl(s, desc, D, SN, 0, 0, 1.0, nullptr, nullptr, 0, analysis_id);
}

int qcbc_server_admit_pt(qcbc_server* s, const qcbc_desc* desc, const double* D, const double* SN, int n_temps,
                         int n_chains, double ladder, const double* init, const double* sigma, uint64_t seed,
                         int* analysis_id) {
  if (n_temps < 1) return QCBC_EINVAL;
  return admit_impl(s, desc, D, SN, n_temps, n_chains, ladder, init, sigma, seed, analysis_id);
}

int qcbc_server_submit_steps(qcbc_server* s, int aid, int nsteps, int64_t* request_id) {
  if (!s || nsteps < 1 || !request_id) return QCBC_EINVAL;
  auto r = std::make_shared<Request>();
  r->aid = aid;
  r->steps_left = nsteps;
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->analyses.find(aid);
    if (it == s->analyses.end() || it->second.evicting || !it->second.ens) return QCBC_EINVAL;
    r->id = s->next_rid++;
    s->reqs[r->id] = r;
    s->queue.push_back(r);
    *request_id = r->id;
  }
  s->cv_work.notify_one();
  return QCBC_OK;
}

int qcbc_server_get_pt(qcbc_server* s, int aid, double* params, double* logL, long long* accepted, long long* swaps,
                       long long* steps) {
  if (!s) return QCBC_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->analyses.find(aid);
  if (it == s->analyses.end() || it->second.evicting || !it->second.ens) return QCBC_EINVAL;
  // between steps: no step of this ensemble is on the GPU while the lock is held
  s->cv_done.wait(lk, [&] {
    auto jt = s->analyses.find(aid);
    return jt == s->analyses.end() || !jt->second.stepping;
  });
  it = s->analyses.find(aid);
  if (it == s->analyses.end() || it->second.evicting) return QCBC_EINVAL;
  const Ensemble& E = *it->second.ens;
  const size_t P = E.P;
  std::vector<unsigned long long> acc(P), sw(E.n_chains);
  hipSetDevice(s->device);
  if (params) HIPCHK(hipMemcpy(params, E.d_x, sizeof(double) * 11 * P, hipMemcpyDeviceToHost));
  if (logL) HIPCHK(hipMemcpy(logL, E.d_L, sizeof(double) * P, hipMemcpyDeviceToHost));
  if (accepted) {
    HIPCHK(hipMemcpy(acc.data(), E.d_acc, sizeof(unsigned long long) * P, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < P; i++) accepted[i] = (long long)acc[i];
  }
  if (swaps) {
    HIPCHK(hipMemcpy(sw.data(), E.d_swaps, sizeof(unsigned long long) * E.n_chains, hipMemcpyDeviceToHost));
    for (int i = 0; i < E.n_chains; i++) swaps[i] = (long long)sw[i];
  }
  if (steps) *steps = E.steps;
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
  if (r->B > 0) {  // a likelihood request (a step request has no per-point results)
    if (logL) std::memcpy(logL, r->logL.data(), sizeof(double) * r->B);
    if (rho) std::memcpy(rho, r->rho.data(), sizeof(double) * (size_t)nd * r->B);
    if (status) std::memcpy(status, r->status.data(), sizeof(int) * r->B);
  }
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
  for (Slot& sl : s->slots) {
    void* dbufs[] = {sl.d_params, sl.d_logL, sl.d_rho, sl.d_partial, sl.d_status, sl.d_pts, sl.d_insts, sl.d_items,
                     sl.d_counter};
    for (void* b : dbufs)
      if (b) (void)hipFree(b);
    void* hbufs[] = {sl.h_params, sl.h_logL, sl.h_rho, sl.h_status};
    for (void* b : hbufs)
      if (b) (void)hipHostFree(b);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
  }
  delete s;
  return QCBC_OK;
}

}  // extern "C"
