// Parallel-tempered ensemble resident on the GPU: the PT step around the
// likelihood (QuickCBC.c:2029-2092, :2238-2405 in shape: Gaussian jiggle
// proposals, Metropolis acceptance per temperature, adjacent-temperature
// swaps).  State, proposals and random counters never leave the device, so a
// step is a fixed sequence of launches that can be captured as a hipGraph.
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/qcbc.h"
#include "qcbc_api_internal.h"
#include "qcbc_internal.h"

using namespace qcbc;

struct qcbc_pt {
  qcbc_ctx* ctx = nullptr;
  int n_temps = 0, n_chains = 0, P = 0, precision = 0;
  double ladder = 1.0;
  uint64_t seed = 0;
  double *d_x = nullptr, *d_L = nullptr, *d_y = nullptr, *d_Ly = nullptr, *d_rho = nullptr, *d_sigma = nullptr;
  int* d_st = nullptr;
  unsigned long long *d_step = nullptr, *d_acc = nullptr, *d_swaps = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool step_timing = false;
  std::vector<double> lat_ms;
};

#define HIPCHK(x)                \
  do {                           \
    hipError_t e_ = (x);         \
    if (e_ != hipSuccess) {      \
      qcbc::t_last_hip = e_;     \
      return QCBC_EHIP;          \
    }                            \
  } while (0)

static int pt_enqueue_step(qcbc_pt* pt) {
  qcbc_ctx* c = pt->ctx;
  hipStream_t s = c->stream;
  HIPCHK(launch_pt_propose(pt->d_x, pt->d_y, pt->d_sigma, pt->n_chains, pt->P, pt->ladder, pt->seed, pt->d_step, s));
  int st = loglike_core(c, pt->d_y, pt->P, pt->precision, pt->d_Ly, pt->d_rho, pt->d_st);
  if (st) return st;
  HIPCHK(launch_pt_accept(pt->d_x, pt->d_L, pt->d_y, pt->d_Ly, pt->d_st, pt->n_chains, pt->P, pt->ladder, pt->seed,
                          pt->d_step, pt->d_acc, s));
  HIPCHK(launch_pt_swap(pt->d_x, pt->d_L, pt->n_temps, pt->n_chains, pt->ladder, pt->seed, pt->d_step, pt->d_swaps,
                        s));
  return QCBC_OK;
}

extern "C" {

int qcbc_pt_create(qcbc_ctx* ctx, int n_temps, int n_chains, double ladder, const double* init, const double* sigma,
                   uint64_t seed, int precision, qcbc_pt** out) {
  if (!ctx || !out || !init || !sigma || n_temps < 1 || n_chains < 1 || n_chains > 1024 || !(ladder >= 1.0) ||
      (precision != QCBC_FP64 && precision != QCBC_MIXED))
    return QCBC_EINVAL;
  if (!ctx->bound) return QCBC_ENOTBOUND;
  *out = nullptr;
  std::lock_guard<std::mutex> lk(ctx->mu);
  hipSetDevice(ctx->device);
  qcbc_pt* pt = new qcbc_pt();
  pt->ctx = ctx;
  pt->n_temps = n_temps;
  pt->n_chains = n_chains;
  pt->P = n_temps * n_chains;
  pt->ladder = ladder;
  pt->seed = seed;
  pt->precision = precision;
  const size_t P = pt->P;
  if (hipMalloc(&pt->d_x, sizeof(double) * 11 * P) || hipMalloc(&pt->d_y, sizeof(double) * 11 * P) ||
      hipMalloc(&pt->d_L, sizeof(double) * P) || hipMalloc(&pt->d_Ly, sizeof(double) * P) ||
      hipMalloc(&pt->d_rho, sizeof(double) * 3 * P) || hipMalloc(&pt->d_sigma, sizeof(double) * 11) ||
      hipMalloc(&pt->d_st, sizeof(int) * P) || hipMalloc(&pt->d_step, sizeof(unsigned long long)) ||
      hipMalloc(&pt->d_acc, sizeof(unsigned long long) * P) ||
      hipMalloc(&pt->d_swaps, sizeof(unsigned long long) * n_chains)) {
    qcbc_pt_destroy(pt);
    return QCBC_ENOMEM;
  }
  hipStream_t s = ctx->stream;
  HIPCHK(hipMemcpyAsync(pt->d_x, init, sizeof(double) * 11 * P, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(pt->d_sigma, sigma, sizeof(double) * 11, hipMemcpyHostToDevice, s));
  HIPCHK(launch_fill_u64(pt->d_step, 1, 0ULL, s));
  HIPCHK(launch_fill_u64(pt->d_acc, (int)P, 0ULL, s));
  HIPCHK(launch_fill_u64(pt->d_swaps, n_chains, 0ULL, s));
  int st = loglike_core(ctx, pt->d_x, pt->P, precision, pt->d_L, pt->d_rho, pt->d_st);
  if (st) return st;
  HIPCHK(hipStreamSynchronize(s));
  *out = pt;
  return QCBC_OK;
}

int qcbc_pt_set_step_timing(qcbc_pt* pt, int on) {
  if (!pt) return QCBC_EINVAL;
  pt->step_timing = on != 0;
  return QCBC_OK;
}

int qcbc_pt_run(qcbc_pt* pt, int nsteps, int use_graph) {
  if (!pt || nsteps < 0) return QCBC_EINVAL;
  qcbc_ctx* c = pt->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  hipStream_t s = c->stream;
  pt->lat_ms.clear();
  if (use_graph && !pt->exec) {
    // capture one steady-state step; no memset nodes (fills are kernels)
    bool timing = c->timing;
    c->timing = false;  // event records stay out of the graph
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int st = pt_enqueue_step(pt);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    c->timing = timing;
    if (st) return st;
    HIPCHK(e);
    pt->graph = g;
    HIPCHK(hipGraphInstantiate(&pt->exec, g, nullptr, nullptr, 0));
  }
  for (int i = 0; i < nsteps; i++) {
    auto t0 = std::chrono::steady_clock::now();
    if (use_graph) {
      HIPCHK(hipGraphLaunch(pt->exec, s));
    } else {
      int st = pt_enqueue_step(pt);
      if (st) return st;
    }
    if (pt->step_timing) {
      HIPCHK(hipStreamSynchronize(s));
      auto t1 = std::chrono::steady_clock::now();
      pt->lat_ms.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
    }
  }
  HIPCHK(hipStreamSynchronize(s));
  return QCBC_OK;
}

int qcbc_pt_get(qcbc_pt* pt, double* params, double* logL, long long* accepted, long long* swaps) {
  if (!pt) return QCBC_EINVAL;
  qcbc_ctx* c = pt->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  hipStream_t s = c->stream;
  const size_t P = pt->P;
  if (params) HIPCHK(hipMemcpyAsync(params, pt->d_x, sizeof(double) * 11 * P, hipMemcpyDeviceToHost, s));
  if (logL) HIPCHK(hipMemcpyAsync(logL, pt->d_L, sizeof(double) * P, hipMemcpyDeviceToHost, s));
  std::vector<unsigned long long> acc(P), sw(pt->n_chains);
  HIPCHK(hipMemcpyAsync(acc.data(), pt->d_acc, sizeof(unsigned long long) * P, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(sw.data(), pt->d_swaps, sizeof(unsigned long long) * pt->n_chains, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (accepted)
    for (size_t i = 0; i < P; i++) accepted[i] = (long long)acc[i];
  if (swaps) {
    long long t = 0;
    for (auto v : sw) t += (long long)v;
    *swaps = t;
  }
  return QCBC_OK;
}

int qcbc_pt_latencies(qcbc_pt* pt, double* ms, int max, int* n) {
  if (!pt || !n) return QCBC_EINVAL;
  int k = (int)pt->lat_ms.size();
  if (ms)
    for (int i = 0; i < k && i < max; i++) ms[i] = pt->lat_ms[i];
  *n = k;
  return QCBC_OK;
}

int qcbc_pt_destroy(qcbc_pt* pt) {
  if (!pt) return QCBC_OK;
  hipSetDevice(pt->ctx->device);
  hipStreamSynchronize(pt->ctx->stream);
  if (pt->exec) hipGraphExecDestroy(pt->exec);
  if (pt->graph) hipGraphDestroy(pt->graph);
  void* bufs[] = {pt->d_x, pt->d_L, pt->d_y, pt->d_Ly, pt->d_rho, pt->d_sigma, pt->d_st, pt->d_step, pt->d_acc,
                  pt->d_swaps};
  for (void* b : bufs)
    if (b) hipFree(b);
  delete pt;
  return QCBC_OK;
}

}  // extern "C"
