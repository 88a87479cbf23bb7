// Many analyses of different shapes (observation time, sample rate,
// detector set, number of points) evaluated by one persistent launch that
// drains an (analysis, point group, bin split) work queue.  A plan fixes the
// queue; running it is set-up per analysis, one queue launch, finalize per
// analysis, all on the plan's stream, so a run can be captured as a hipGraph.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "../../include/qcbc.h"
#include "qcbc_api_internal.h"
#include "qcbc_internal.h"

using namespace qcbc;

#define HIPCHK(x)                \
  do {                           \
    hipError_t e_ = (x);         \
    if (e_ != hipSuccess) {      \
      qcbc::t_last_hip = e_;     \
      return QCBC_EHIP;          \
    }                            \
  } while (0)

struct qcbc_multi {
  std::vector<qcbc_ctx*> ctxs;
  std::vector<int> B;
  int precision = 0;
  int cpb = 8, tile = 512, split_bins = 4096, blocks = 0;
  int max_nd = 1;
  hipStream_t stream = nullptr;
  int device = 0;
  // per-analysis device scratch (points, partials), owned by the plan
  std::vector<Point*> pts;
  std::vector<double*> partial;
  std::vector<LLInst> insts;
  LLInst* d_insts = nullptr;
  LLItem* d_items = nullptr;
  int n_items = 0;
  int* d_counter = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<const void*> graph_ptrs;  // pointers baked into the graph
  double last_ms = 0.0;
};

static void multi_free_queue(qcbc_multi* m) {
  if (m->exec) hipGraphExecDestroy(m->exec);
  if (m->graph) hipGraphDestroy(m->graph);
  m->exec = nullptr;
  m->graph = nullptr;
  m->graph_ptrs.clear();
  if (m->d_insts) hipFree(m->d_insts);
  if (m->d_items) hipFree(m->d_items);
  m->d_insts = nullptr;
  m->d_items = nullptr;
}

// Work items sorted longest first (bins x points), so the tail of the queue
// holds the short items and the blocks finish together.
static int multi_build(qcbc_multi* m) {
  multi_free_queue(m);
  const int n = (int)m->ctxs.size();
  m->insts.assign(n, LLInst{});
  std::vector<std::pair<double, LLItem>> items;
  m->max_nd = 1;
  for (int i = 0; i < n; i++) {
    qcbc_ctx* c = m->ctxs[i];
    int splits = std::max(1, (c->nb + m->split_bins - 1) / m->split_bins);
    splits = std::min(splits, QCBC_MAX_SPLITS);
    m->insts[i] = make_inst(c, m->tile, m->B[i], splits, m->pts[i], m->partial[i]);
    m->max_nd = std::max(m->max_nd, c->desc.n_det);
    const int groups = (m->B[i] + m->cpb - 1) / m->cpb;
    for (int g = 0; g < groups; g++)
      for (int s = 0; s < m->insts[i].splits; s++) {
        int bins = std::min(m->insts[i].split_len, c->nb - s * m->insts[i].split_len);
        int npts = std::min(m->cpb, m->B[i] - g * m->cpb);
        double cost = (double)bins * npts * (1.0 + 0.5 * c->desc.n_det);
        items.push_back({cost, LLItem{i, g, s, 0}});
      }
  }
  std::stable_sort(items.begin(), items.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
  std::vector<LLItem> flat(items.size());
  for (size_t k = 0; k < items.size(); k++) flat[k] = items[k].second;
  m->n_items = (int)flat.size();
  HIPCHK(hipMalloc(&m->d_insts, sizeof(LLInst) * std::max(n, 1)));
  HIPCHK(hipMalloc(&m->d_items, sizeof(LLItem) * std::max(m->n_items, 1)));
  HIPCHK(hipMemcpy(m->d_insts, m->insts.data(), sizeof(LLInst) * n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(m->d_items, flat.data(), sizeof(LLItem) * m->n_items, hipMemcpyHostToDevice));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, m->device));
  // persistent grid: a few resident blocks per CU, never more than the items
  const int per_cu = std::max(1, 16 / m->cpb);
  m->blocks = std::max(1, std::min(m->n_items, prop.multiProcessorCount * per_cu));
  return QCBC_OK;
}

static int multi_enqueue(qcbc_multi* m, const double* const* d_params, double* const* d_logL, double* const* d_rho,
                         int* const* d_status) {
  const int n = (int)m->ctxs.size();
  hipStream_t s = m->stream;
  for (int i = 0; i < n; i++)
    HIPCHK(launch_setup(d_params[i], m->B[i], m->ctxs[i]->an, m->pts[i], d_status ? d_status[i] : nullptr, s));
  HIPCHK(launch_fill_i32(m->d_counter, 1, 0, s));
  HIPCHK(launch_ll_queue(m->precision, m->d_insts, m->d_items, m->n_items, m->d_counter, m->max_nd, m->cpb, m->tile,
                         m->blocks, s));
  for (int i = 0; i < n; i++)
    HIPCHK(launch_finalize(m->partial[i], m->insts[i].splits, m->ctxs[i]->desc.n_det, m->pts[i], m->B[i],
                           m->precision == QCBC_MIXED ? FIN_MIXED : FIN_FP64, m->ctxs[i]->d_wref, d_logL[i], d_rho[i], s));
  return QCBC_OK;
}

extern "C" {

int qcbc_multi_create(qcbc_ctx* const* ctxs, const int* B, int n, int precision, qcbc_multi** out) {
  if (!ctxs || !B || n < 1 || !out || (precision != QCBC_FP64 && precision != QCBC_MIXED)) return QCBC_EINVAL;
  *out = nullptr;
  qcbc_multi* m = new qcbc_multi();
  m->precision = precision;
  hipGetDevice(&m->device);
  for (int i = 0; i < n; i++) {
    if (!ctxs[i] || !ctxs[i]->bound || B[i] < 0) {
      delete m;
      return ctxs[i] && !ctxs[i]->bound ? QCBC_ENOTBOUND : QCBC_EINVAL;
    }
    m->ctxs.push_back(ctxs[i]);
    m->B.push_back(B[i]);
  }
  // the largest split count is nb / split_bins; keep it within the partial layout
  int maxnb = 0;
  for (auto* c : m->ctxs) maxnb = std::max(maxnb, c->nb);
  m->split_bins = std::max(4096, (maxnb + QCBC_MAX_SPLITS - 1) / QCBC_MAX_SPLITS);
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&m->d_counter, sizeof(int)) != hipSuccess) {
    qcbc_multi_destroy(m);
    return QCBC_ENOMEM;
  }
  for (int i = 0; i < n; i++) {
    Point* p = nullptr;
    double* q = nullptr;
    size_t nb = std::max(B[i], 1);
    if (hipMalloc(&p, sizeof(Point) * nb) != hipSuccess ||
        hipMalloc(&q, sizeof(double) * 6 * QCBC_MAX_SPLITS * nb) != hipSuccess) {
      m->pts.push_back(p);
      m->partial.push_back(q);
      qcbc_multi_destroy(m);
      return QCBC_ENOMEM;
    }
    m->pts.push_back(p);
    m->partial.push_back(q);
  }
  int st = multi_build(m);
  if (st) {
    qcbc_multi_destroy(m);
    return st;
  }
  *out = m;
  return QCBC_OK;
}

int qcbc_multi_set_shape(qcbc_multi* m, int cpb, int tile, int split_bins) {
  if (!m || cpb < 1 || cpb > 8 || !ll_tile_ok(tile) || split_bins < tile) return QCBC_EINVAL;
  int maxnb = 0, maxnd = 1;
  for (auto* c : m->ctxs) {
    maxnb = std::max(maxnb, c->nb);
    maxnd = std::max(maxnd, c->desc.n_det);
  }
  if ((maxnb + split_bins - 1) / split_bins > QCBC_MAX_SPLITS) return QCBC_EINVAL;
  if (ll_lds_bytes(m->precision, maxnd, tile) > 64 * 1024) return QCBC_EINVAL;
  hipSetDevice(m->device);
  m->cpb = cpb;
  m->tile = tile;
  m->split_bins = split_bins;
  return multi_build(m);
}

int qcbc_multi_get_shape(qcbc_multi* m, int* cpb, int* tile, int* split_bins, int* n_items) {
  if (!m) return QCBC_EINVAL;
  if (cpb) *cpb = m->cpb;
  if (tile) *tile = m->tile;
  if (split_bins) *split_bins = m->split_bins;
  if (n_items) *n_items = m->n_items;
  return QCBC_OK;
}

int qcbc_multi_run(qcbc_multi* m, const double* const* d_params, double* const* d_logL, double* const* d_rho,
                   int* const* d_status, int use_graph) {
  if (!m || !d_params || !d_logL || !d_rho) return QCBC_EINVAL;
  hipSetDevice(m->device);
  const int n = (int)m->ctxs.size();
  auto t0 = std::chrono::steady_clock::now();
  if (use_graph) {
    std::vector<const void*> ptrs;
    for (int i = 0; i < n; i++) {
      ptrs.push_back(d_params[i]);
      ptrs.push_back(d_logL[i]);
      ptrs.push_back(d_rho[i]);
      ptrs.push_back(d_status ? d_status[i] : nullptr);
    }
    if (m->exec && ptrs != m->graph_ptrs) {  // different buffers: capture again
      hipGraphExecDestroy(m->exec);
      hipGraphDestroy(m->graph);
      m->exec = nullptr;
      m->graph = nullptr;
    }
    if (!m->exec) {
      HIPCHK(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
      int st = multi_enqueue(m, d_params, d_logL, d_rho, d_status);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(m->stream, &g);
      if (st) return st;
      HIPCHK(e);
      m->graph = g;
      HIPCHK(hipGraphInstantiate(&m->exec, g, nullptr, nullptr, 0));
      m->graph_ptrs = ptrs;
    }
    HIPCHK(hipGraphLaunch(m->exec, m->stream));
  } else {
    int st = multi_enqueue(m, d_params, d_logL, d_rho, d_status);
    if (st) return st;
  }
  HIPCHK(hipStreamSynchronize(m->stream));
  m->last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return QCBC_OK;
}

// Autotuner for the persistent launch: times candidate (points per block,
// tile, split length) shapes on the plan's own buffers and keeps the fastest.
int qcbc_multi_autotune(qcbc_multi* m, const double* const* d_params, double* const* d_logL, double* const* d_rho,
                        int reps, int* best_cpb, int* best_tile, int* best_split, double* best_ms) {
  if (!m || reps < 1) return QCBC_EINVAL;
  int cpbs[] = {2, 4, 8};
  int tiles[] = {128, 256, 512};
  int splits[] = {2048, 4096, 8192, 16384};
  double best = 1e30;
  int bc = m->cpb, bt = m->tile, bs = m->split_bins;
  for (int cpb : cpbs)
    for (int tile : tiles)
      for (int sb : splits) {
        if (qcbc_multi_set_shape(m, cpb, tile, sb) != QCBC_OK) continue;
        int st = qcbc_multi_run(m, d_params, d_logL, d_rho, nullptr, 0);  // warm
        if (st) return st;
        double t = 0.0;
        for (int r = 0; r < reps; r++) {
          st = qcbc_multi_run(m, d_params, d_logL, d_rho, nullptr, 0);
          if (st) return st;
          t += m->last_ms;
        }
        t /= reps;
        if (t < best) {
          best = t;
          bc = cpb; bt = tile; bs = sb;
        }
      }
  int st = qcbc_multi_set_shape(m, bc, bt, bs);
  if (best_cpb) *best_cpb = bc;
  if (best_tile) *best_tile = bt;
  if (best_split) *best_split = bs;
  if (best_ms) *best_ms = best;
  return st;
}

int qcbc_multi_last_ms(qcbc_multi* m, double* ms) {
  if (!m || !ms) return QCBC_EINVAL;
  *ms = m->last_ms;
  return QCBC_OK;
}

int qcbc_multi_destroy(qcbc_multi* m) {
  if (!m) return QCBC_OK;
  hipSetDevice(m->device);
  if (m->stream) hipStreamSynchronize(m->stream);
  multi_free_queue(m);
  for (auto* p : m->pts)
    if (p) hipFree(p);
  for (auto* q : m->partial)
    if (q) hipFree(q);
  if (m->d_counter) hipFree(m->d_counter);
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
  return QCBC_OK;
}

int qcbc_multi_loglike(qcbc_ctx* const* ctxs, const double* const* d_params, const int* B, int n, int precision,
                       double* const* d_logL, double* const* d_rho, int* const* d_status) {
  qcbc_multi* m = nullptr;
  int st = qcbc_multi_create(ctxs, B, n, precision, &m);
  if (st) return st;
  st = qcbc_multi_run(m, d_params, d_logL, d_rho, d_status, 0);
  qcbc_multi_destroy(m);
  return st;
}

}  // extern "C"
