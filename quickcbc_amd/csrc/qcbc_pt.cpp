// Parallel-tempered ensemble resident on the GPU: the PT step around the
// likelihood (QuickCBC.c:2029-2092, :2238-2405 in shape: Gaussian jiggle
// proposals, Metropolis acceptance per temperature, adjacent-temperature
// swaps).  State, proposals and random counters never leave the device, so a
// step is a fixed sequence of launches that can be captured as a hipGraph.
//
// Chain columns only meet through swaps between temperatures of the same
// column, so the ensemble splits into groups of columns that step
// independently, each on a stream of its own with nothing ordering one group
// against another.  The GPU then overlaps them freely: one group's
// latency-bound kernels (set-up, and the post-likelihood kernel: acceptance,
// swaps, next proposals) and the tail of its likelihood kernel run beside
// another group's likelihood kernel.  Random counters use the whole
// ensemble's point numbers, so every number of groups gives the same chains.
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/qcbc.h"
#include "qcbc_api_internal.h"
#include "qcbc_internal.h"
#include "qcbc_update_internal.h"

using namespace qcbc;

namespace {
struct Group {
  int c0 = 0, ncols = 0, P = 0, off = 0;  // columns [c0, c0 + ncols), points [off, off + P) of the storage
  hipStream_t s = nullptr;
  bool own_stream = false;
  Point* d_pts = nullptr;
  double* d_partial = nullptr;
  unsigned long long* d_step = nullptr;
  unsigned int* d_done = nullptr;  // blocks of k_pt_post finished this step
};
}  // namespace

struct qcbc_pt {
  qcbc_ctx* ctx = nullptr;
  qcbc_het* het = nullptr;  // the heterodyned likelihood instead of the full one
  int n_temps = 0, n_chains = 0, P = 0, precision = 0, G = 1;
  double ladder = 1.0;
  uint64_t seed = 0;
  Group g[QCBC_PT_MAX_GROUPS];
  // storage order: group by group, t * ncols + (c - c0) within a group
  double *d_x = nullptr, *d_L = nullptr, *d_y = nullptr, *d_Ly = nullptr, *d_rho = nullptr, *d_sigma = nullptr;
  double* d_tc = nullptr;  // per-temperature constants
  int* d_st = nullptr;
  unsigned long long *d_acc = nullptr, *d_swaps = nullptr;
  std::vector<int> perm;  // storage index -> point number t * n_chains + c
  hipEvent_t ev_fork = nullptr;  // a run's groups start after the work queued on the context stream
  // captured steps, one graph per column group, replayed on the group's own
  // stream (so the groups overlap as the free-running streams do): [0] one
  // step, [1] kGraphSteps steps, [2] the last run's remainder of rem_steps
  // (< kGraphSteps) steps
  hipGraph_t graph[3][QCBC_PT_MAX_GROUPS] = {};
  hipGraphExec_t exec[3][QCBC_PT_MAX_GROUPS] = {};
  int rem_steps = 0;
  bool step_timing = false;
  std::vector<double> lat_ms;
  // QuickCBC's own step (update() and MCMC_all's swaps) instead of the
  // jiggle step: Fisher jump sizes and directions per point (storage order,
  // moved with the state on swaps), the history of each (column,
  // temperature) (column-major: column c's rows at c n_temps n_hist), its
  // fill count per column, the proposal type of each point and per-column
  // statistics
  bool upd = false;
  UpdateConst uc{};
  double efix_zs = 0.0;  // > 0: efix of the Fisher jumps after each refresh
  double *d_ej = nullptr, *d_ev = nullptr, *d_hist = nullptr;
  unsigned long long *d_hcount = nullptr, *d_ustats = nullptr;
  int* d_typ = nullptr;
  double* d_jac = nullptr;  // log density of each pending sky-ring proposal (0 for the others)
};

#define HIPCHK(x)                \
  do {                           \
    hipError_t e_ = (x);         \
    if (e_ != hipSuccess) {      \
      qcbc::t_last_hip = e_;     \
      return QCBC_EHIP;          \
    }                            \
  } while (0)

// n of QuickCBC's steps of group k in one launch of the fused step kernel:
// per step the set-up and heterodyned logL of the pending proposals y, the
// acceptance or swap, the history and the next proposals.
static int group_update_steps(qcbc_pt* pt, int k, int n) {
  Group& gr = pt->g[k];
  const size_t o = gr.off;
  HetGridView hv{};
  int st = het_grid_view(pt->het, &hv);
  if (st) return st;
  const HetArgs h{hv.freq, hv.pl, hv.amp, hv.phase, hv.wf, hv.M, hv.nd, hv.logLb};
  HIPCHK(launch_pt_steps(pt->d_x + 11 * o, pt->d_L + o, pt->d_ej + 11 * o, pt->d_ev + 121 * o,
                         pt->d_hist + (size_t)gr.c0 * pt->n_temps * pt->uc.n_hist * 11, pt->d_hcount,
                         pt->d_y + 11 * o, pt->d_typ + o, pt->d_jac + o, pt->d_tc, pt->uc, hv.an, h, pt->n_temps,
                         pt->n_chains, gr.c0, gr.ncols, pt->seed, n, gr.d_step, pt->d_ustats, gr.d_done, gr.s));
  return QCBC_OK;
}

// One PT step of group k: the proposals y are already on the device (made by
// the previous step, or at creation).  Set-up and likelihood partial sums of
// y, then k_pt_post: logL of y, acceptance, swaps, the next proposals and the
// group's step counter.
static int group_step(qcbc_pt* pt, int k) {
  qcbc_ctx* c = pt->ctx;
  Group& gr = pt->g[k];
  const int prec = pt->precision;
  const size_t o = gr.off;
  if (pt->upd) return group_update_steps(pt, k, 1);  // QuickCBC's step, fused
  if (pt->het) {  // set-up and heterodyned logL of y, then the same post kernel
    int st = het_enqueue(pt->het, pt->d_y + 11 * o, gr.P, gr.d_pts, pt->d_st + o, pt->d_Ly + o, gr.s);
    if (st) return st;
    LLInst in{};
    in.pts = gr.d_pts;
    in.nd = c->desc.n_det;
    HIPCHK(launch_pt_post(in, FIN_DONE, c->d_wref, pt->d_x + 11 * o, pt->d_L + o, pt->d_y + 11 * o, pt->d_Ly + o,
                          pt->d_rho + 3 * o, pt->d_st + o, pt->d_sigma, pt->d_tc, pt->n_temps, pt->n_chains, gr.c0,
                          gr.ncols, pt->seed, gr.d_step, pt->d_acc + o, pt->d_swaps, gr.d_done, gr.s));
    return QCBC_OK;
  }
  HIPCHK(launch_setup(pt->d_y + 11 * o, gr.P, c->an, gr.d_pts, pt->d_st + o, gr.s));
  const LLInst in = make_inst(c, c->tile[prec], gr.P, effective_splits(c, prec, pt->P), gr.d_pts, gr.d_partial);
  int st = ll_launch_timed(c, prec, in, gr.s);
  if (st) return st;
  HIPCHK(launch_pt_post(in, prec == QCBC_MIXED ? FIN_MIXED : FIN_FP64, c->d_wref, pt->d_x + 11 * o, pt->d_L + o,
                        pt->d_y + 11 * o, pt->d_Ly + o, pt->d_rho + 3 * o, pt->d_st + o, pt->d_sigma, pt->d_tc,
                        pt->n_temps, pt->n_chains, gr.c0, gr.ncols, pt->seed, gr.d_step, pt->d_acc + o, pt->d_swaps,
                        gr.d_done, gr.s));
  return QCBC_OK;
}

// One step of the whole ensemble.
static int pt_enqueue_step(qcbc_pt* pt) {
  for (int k = 0; k < pt->G; k++) {
    int st = group_step(pt, k);
    if (st) return st;
  }
  return QCBC_OK;
}

static void pt_free(qcbc_pt* pt) {
  for (Group& gr : pt->g) {
    if (gr.own_stream && gr.s) (void)hipStreamDestroy(gr.s);
    void* gb[] = {gr.d_pts, gr.d_partial, gr.d_step, gr.d_done};
    for (void* b : gb)
      if (b) (void)hipFree(b);
  }
  if (pt->ev_fork) (void)hipEventDestroy(pt->ev_fork);
  for (int i = 0; i < 3; i++)
    for (int k = 0; k < QCBC_PT_MAX_GROUPS; k++) {
      if (pt->exec[i][k]) (void)hipGraphExecDestroy(pt->exec[i][k]);
      if (pt->graph[i][k]) (void)hipGraphDestroy(pt->graph[i][k]);
    }
  void* bufs[] = {pt->d_x,     pt->d_L,  pt->d_y,  pt->d_Ly,   pt->d_rho,    pt->d_sigma,  pt->d_tc,
                  pt->d_st,    pt->d_acc, pt->d_swaps, pt->d_ej, pt->d_ev, pt->d_hist, pt->d_hcount,
                  pt->d_ustats, pt->d_typ, pt->d_jac};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  delete pt;
}

extern "C" {

static int pt_sync(qcbc_pt* pt);

// Proposals of the current step of every group, QuickCBC's kind (on stream s)
static int update_propose_all(qcbc_pt* pt, hipStream_t s) {
  for (int k = 0; k < pt->G; k++) {
    const Group& gr = pt->g[k];
    const size_t o = gr.off;
    HIPCHK(launch_update_propose(pt->d_x + 11 * o, pt->d_ej + 11 * o, pt->d_ev + 121 * o,
                                 pt->d_hist + (size_t)gr.c0 * pt->n_temps * pt->uc.n_hist * 11, pt->d_tc, pt->uc,
                                 pt->n_temps, pt->n_chains, gr.c0, gr.ncols, pt->seed, gr.d_step, pt->d_y + 11 * o,
                                 pt->d_typ + o, pt->d_jac + o, s));
  }
  return QCBC_OK;
}

static int pt_create(qcbc_ctx* ctx, qcbc_het* het, int n_temps, int n_chains, double ladder, const double* init,
                     const double* sigma, uint64_t seed, int precision, int groups, qcbc_pt** out,
                     const qcbc_update_desc* ud = nullptr) {
  if (ud && (!het || !(ud->a >= ud->b && ud->b >= ud->c && ud->c >= 0.0 && ud->a <= 1.0) || ud->n_hist < 2 ||
             ud->hist_every < 1 || !(ud->etamin >= 0.0) || !(ud->efix_zs >= 0.0 && ud->efix_zs < 1e300) ||
             (ud->efix_zs > 0.0 && !(ud->etamin > 0.0)) || (ud->c > 0.0 && ctx && ctx->desc.n_det < 2)))
    return QCBC_EINVAL;  // (a sky ring needs two detectors)
  if (ud)
    for (int i = 0; i < 11; i++)
      if (!(ud->min[i] < ud->max[i])) return QCBC_EINVAL;
  if (!ctx || !out || !init || !sigma || n_temps < 1 || n_chains < 1 || n_chains > 1024 || !(ladder >= 1.0) ||
      (precision != QCBC_FP64 && precision != QCBC_MIXED) || groups < 0 || groups > QCBC_PT_MAX_GROUPS ||
      groups > n_chains)
    return QCBC_EINVAL;
  if (het && het_device(het) != ctx->device) return QCBC_EINVAL;
  if (!het && !ctx->bound) return QCBC_ENOTBOUND;
  *out = nullptr;
  std::lock_guard<std::mutex> lk(ctx->mu);
  hipSetDevice(ctx->device);
  qcbc_pt* pt = new qcbc_pt();
  pt->ctx = ctx;
  pt->het = het;
  pt->n_temps = n_temps;
  pt->n_chains = n_chains;
  pt->P = n_temps * n_chains;
  pt->ladder = ladder;
  pt->seed = seed;
  pt->precision = precision;
  // groups by default: three for the full likelihood (one group's set-up and
  // post kernels run beside another group's likelihood kernel), one on the
  // heterodyned likelihood, whose kernels are all short and latency-bound
  // (more groups only add launches)
  pt->G = groups ? groups : (!het && n_chains >= 3 && pt->P >= 1536 ? 3 : 1);
  if (ud) {
    pt->upd = true;
    for (int i = 0; i < 11; i++) {
      pt->uc.min[i] = ud->min[i];
      pt->uc.max[i] = ud->max[i];
    }
    pt->uc.a = ud->a;
    pt->uc.b = ud->b;
    pt->uc.c = ud->c;
    pt->uc.etamin = ud->etamin;
    pt->uc.n_hist = ud->n_hist;
    pt->uc.hist_every = ud->hist_every;
    pt->efix_zs = ud->efix_zs;
    pt->uc.gmst = ctx->desc.gmst;
    pt->uc.n_det = ctx->desc.n_det;
    for (int d = 0; d < 3; d++) pt->uc.labels[d] = d < ctx->desc.n_det ? ctx->desc.labels[d] : 0;
  }
  const size_t P = pt->P;
  int c0 = 0, off = 0;
  for (int k = 0; k < pt->G; k++) {
    Group& gr = pt->g[k];
    gr.c0 = c0;
    gr.ncols = n_chains / pt->G + (k < n_chains % pt->G ? 1 : 0);
    gr.P = gr.ncols * n_temps;
    gr.off = off;
    c0 += gr.ncols;
    off += gr.P;
  }
  pt->perm.resize(P);
  for (int k = 0; k < pt->G; k++) {
    const Group& gr = pt->g[k];
    for (int t = 0; t < n_temps; t++)
      for (int j = 0; j < gr.ncols; j++) pt->perm[gr.off + t * gr.ncols + j] = t * n_chains + gr.c0 + j;
  }
  bool ok = hipMalloc(&pt->d_x, sizeof(double) * 11 * P) == hipSuccess &&
            hipMalloc(&pt->d_y, sizeof(double) * 11 * P) == hipSuccess &&
            hipMalloc(&pt->d_L, sizeof(double) * P) == hipSuccess &&
            hipMalloc(&pt->d_Ly, sizeof(double) * P) == hipSuccess &&
            hipMalloc(&pt->d_rho, sizeof(double) * 3 * P) == hipSuccess &&
            hipMalloc(&pt->d_sigma, sizeof(double) * 11) == hipSuccess &&
            hipMalloc(&pt->d_tc, sizeof(double) * 3 * n_temps) == hipSuccess &&
            hipMalloc(&pt->d_st, sizeof(int) * P) == hipSuccess &&
            hipMalloc(&pt->d_acc, sizeof(unsigned long long) * P) == hipSuccess &&
            hipMalloc(&pt->d_swaps, sizeof(unsigned long long) * n_chains) == hipSuccess;
  if (ok && pt->upd)
    ok = hipMalloc(&pt->d_ej, sizeof(double) * 11 * P) == hipSuccess &&
         hipMalloc(&pt->d_ev, sizeof(double) * 121 * P) == hipSuccess &&
         hipMalloc(&pt->d_hist, sizeof(double) * 11 * P * (size_t)pt->uc.n_hist) == hipSuccess &&
         hipMalloc(&pt->d_hcount, sizeof(unsigned long long) * n_chains) == hipSuccess &&
         hipMalloc(&pt->d_ustats, sizeof(unsigned long long) * UPD_STATS * n_chains) == hipSuccess &&
         hipMalloc(&pt->d_typ, sizeof(int) * P) == hipSuccess &&
         hipMalloc(&pt->d_jac, sizeof(double) * P) == hipSuccess;
  for (int k = 0; k < pt->G && ok; k++) {
    Group& gr = pt->g[k];
    ok = hipMalloc(&gr.d_pts, sizeof(Point) * gr.P) == hipSuccess &&
         hipMalloc(&gr.d_partial, sizeof(double) * 6 * (size_t)QCBC_MAX_SPLITS * gr.P) == hipSuccess &&
         hipMalloc(&gr.d_step, sizeof(unsigned long long)) == hipSuccess &&
         hipMalloc(&gr.d_done, sizeof(unsigned int)) == hipSuccess;
    if (ok && k == 0) {
      gr.s = ctx->stream;
    } else if (ok) {
      ok = hipStreamCreateWithFlags(&gr.s, hipStreamNonBlocking) == hipSuccess;
      gr.own_stream = ok;
    }
  }
  ok = ok && hipEventCreateWithFlags(&pt->ev_fork, hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    pt_free(pt);
    return QCBC_ENOMEM;
  }
  std::vector<double> x0(11 * P);
  for (size_t i = 0; i < P; i++) std::memcpy(&x0[11 * i], init + 11 * (size_t)pt->perm[i], sizeof(double) * 11);
  hipStream_t s = ctx->stream;
  auto fail = [&](int code) {
    hipStreamSynchronize(s);
    pt_free(pt);
    return code;
  };
  if (hipMemcpyAsync(pt->d_x, x0.data(), sizeof(double) * 11 * P, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(pt->d_sigma, sigma, sizeof(double) * 11, hipMemcpyHostToDevice, s) != hipSuccess ||
      launch_pt_temps(pt->d_tc, n_temps, ladder, s) != hipSuccess ||
      launch_fill_u64(pt->d_acc, (int)P, 0ULL, s) != hipSuccess ||
      launch_fill_u64(pt->d_swaps, n_chains, 0ULL, s) != hipSuccess)
    return fail(QCBC_EHIP);
  for (int k = 0; k < pt->G; k++)
    if (launch_fill_u64(pt->g[k].d_step, 1, 0ULL, s) != hipSuccess ||
        launch_fill_i32(reinterpret_cast<int*>(pt->g[k].d_done), 1, 0, s) != hipSuccess)
      return fail(QCBC_EHIP);
  int st = QCBC_OK;
  if (het) {  // each group's points on the group's own set-up records
    for (int k = 0; k < pt->G && !st; k++) {
      const Group& gr = pt->g[k];
      st = het_enqueue(het, pt->d_x + 11 * (size_t)gr.off, gr.P, gr.d_pts, pt->d_st + gr.off, pt->d_L + gr.off, s);
    }
  } else {
    st = loglike_core(ctx, pt->d_x, pt->P, precision, pt->d_L, pt->d_rho, pt->d_st);
  }
  if (st) return fail(st);
  if (pt->upd) {
    // MCMC_all's start (QuickCBC.c:1918-1954): Fisher jumps at every state,
    // each column's history from its own states, then the proposals of step 0
    if (launch_fill_u64(pt->d_hcount, n_chains, 0ULL, s) != hipSuccess ||
        launch_fill_u64(pt->d_ustats, UPD_STATS * n_chains, 0ULL, s) != hipSuccess)
      return fail(QCBC_EHIP);
    for (int k = 0; k < pt->G; k++) {
      const Group& gr = pt->g[k];
      if (launch_update_hist_init(pt->d_x + 11 * (size_t)gr.off, n_temps, n_chains, gr.c0, gr.ncols,
                                  pt->uc.n_hist, seed,
                                  pt->d_hist + (size_t)gr.c0 * n_temps * pt->uc.n_hist * 11, s) != hipSuccess)
        return fail(QCBC_EHIP);
    }
    if (hipStreamSynchronize(s) != hipSuccess) return fail(QCBC_EHIP);
    st = het_fisher_enqueue(het, pt->d_x, pt->P, pt->d_ej, pt->d_ev, s);
    if (st) return fail(st);
    if (pt->efix_zs > 0.0) {
      st = het_efix_device(het, pt->d_x, pt->P, pt->d_ej, pt->d_ev, pt->uc.min, pt->uc.max, pt->uc.etamin,
                           pt->efix_zs, s);
      if (st) return fail(st);
    }
    st = update_propose_all(pt, s);
    if (st) return fail(st);
    if (hipStreamSynchronize(s) != hipSuccess) return fail(QCBC_EHIP);
    *out = pt;
    return QCBC_OK;
  }
  // proposals of step 0
  for (int k = 0; k < pt->G; k++) {
    const Group& gr = pt->g[k];
    if (launch_pt_propose(pt->d_x + 11 * (size_t)gr.off, pt->d_y + 11 * (size_t)gr.off, pt->d_sigma, pt->d_tc,
                          n_temps, n_chains, gr.c0, gr.ncols, seed, gr.d_step, s) != hipSuccess)
      return fail(QCBC_EHIP);
  }
  if (hipStreamSynchronize(s) != hipSuccess) return fail(QCBC_EHIP);
  *out = pt;
  return QCBC_OK;
}

int qcbc_pt_create_groups(qcbc_ctx* ctx, int n_temps, int n_chains, double ladder, const double* init,
                          const double* sigma, uint64_t seed, int precision, int groups, qcbc_pt** out) {
  return pt_create(ctx, nullptr, n_temps, n_chains, ladder, init, sigma, seed, precision, groups, out);
}

int qcbc_pt_create_het(qcbc_ctx* ctx, qcbc_het* het, int n_temps, int n_chains, double ladder, const double* init,
                       const double* sigma, uint64_t seed, int groups, qcbc_pt** out) {
  if (!het) return QCBC_EINVAL;
  return pt_create(ctx, het, n_temps, n_chains, ladder, init, sigma, seed, QCBC_FP64, groups, out);
}

int qcbc_pt_create_update(qcbc_ctx* ctx, qcbc_het* het, int n_temps, int n_chains, double ladder,
                          const double* init, const qcbc_update_desc* desc, uint64_t seed, int groups,
                          qcbc_pt** out) {
  if (!het || !desc) return QCBC_EINVAL;
  static const double unit[11] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};  // unused by this step
  return pt_create(ctx, het, n_temps, n_chains, ladder, init, unit, seed, QCBC_FP64, groups, out, desc);
}

int qcbc_pt_refresh_fisher(qcbc_pt* pt) {
  if (!pt || !pt->upd) return QCBC_EINVAL;
  qcbc_ctx* c = pt->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  int st = pt_sync(pt);
  if (st) return st;
  st = het_fisher_enqueue(pt->het, pt->d_x, pt->P, pt->d_ej, pt->d_ev, c->stream);
  if (st) return st;
  if (pt->efix_zs > 0.0) {
    st = het_efix_device(pt->het, pt->d_x, pt->P, pt->d_ej, pt->d_ev, pt->uc.min, pt->uc.max, pt->uc.etamin,
                         pt->efix_zs, c->stream);
    if (st) return st;
  }
  // the pending proposals were made with the old directions (the reference
  // refreshes before its update, QuickCBC.c:2019-2027)
  st = update_propose_all(pt, c->stream);
  if (st) return st;
  HIPCHK(hipStreamSynchronize(c->stream));
  return QCBC_OK;
}

int qcbc_pt_get_update(qcbc_pt* pt, double* ej, double* ev, double* proposals, int* types, double* jac,
                       long long* stats) {
  if (!pt || !pt->upd) return QCBC_EINVAL;
  qcbc_ctx* c = pt->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  int st = pt_sync(pt);
  if (st) return st;
  hipStream_t s = c->stream;
  const size_t P = pt->P;
  std::vector<double> e(11 * P), v(121 * P), y(11 * P), jc(P);
  std::vector<int> ty(P);
  std::vector<unsigned long long> us((size_t)UPD_STATS * pt->n_chains);
  HIPCHK(hipMemcpyAsync(e.data(), pt->d_ej, sizeof(double) * 11 * P, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(v.data(), pt->d_ev, sizeof(double) * 121 * P, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(y.data(), pt->d_y, sizeof(double) * 11 * P, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(ty.data(), pt->d_typ, sizeof(int) * P, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(jc.data(), pt->d_jac, sizeof(double) * P, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(us.data(), pt->d_ustats, sizeof(unsigned long long) * us.size(), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (size_t i = 0; i < P; i++) {
    const size_t p = pt->perm[i];
    if (ej) std::memcpy(ej + 11 * p, &e[11 * i], sizeof(double) * 11);
    if (ev) std::memcpy(ev + 121 * p, &v[121 * i], sizeof(double) * 121);
    if (proposals) std::memcpy(proposals + 11 * p, &y[11 * i], sizeof(double) * 11);
    if (types) types[p] = ty[i];
    if (jac) jac[p] = jc[i];
  }
  if (stats)
    for (size_t i = 0; i < us.size(); i++) stats[i] = (long long)us[i];
  return QCBC_OK;
}

int qcbc_pt_get_history(qcbc_pt* pt, double* hist, long long* counts) {
  if (!pt || !pt->upd) return QCBC_EINVAL;
  qcbc_ctx* c = pt->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  int st = pt_sync(pt);
  if (st) return st;
  hipStream_t s = c->stream;
  std::vector<unsigned long long> hc(pt->n_chains);
  if (hist)
    HIPCHK(hipMemcpyAsync(hist, pt->d_hist, sizeof(double) * 11 * (size_t)pt->P * pt->uc.n_hist,
                          hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(hc.data(), pt->d_hcount, sizeof(unsigned long long) * pt->n_chains, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (counts)
    for (int i = 0; i < pt->n_chains; i++) counts[i] = (long long)hc[i];
  return QCBC_OK;
}

int qcbc_pt_create(qcbc_ctx* ctx, int n_temps, int n_chains, double ladder, const double* init, const double* sigma,
                   uint64_t seed, int precision, qcbc_pt** out) {
  return qcbc_pt_create_groups(ctx, n_temps, n_chains, ladder, init, sigma, seed, precision, 0, out);
}

int qcbc_pt_groups(qcbc_pt* pt, int* groups) {
  if (!pt || !groups) return QCBC_EINVAL;
  *groups = pt->G;
  return QCBC_OK;
}

int qcbc_pt_set_step_timing(qcbc_pt* pt, int on) {
  if (!pt) return QCBC_EINVAL;
  pt->step_timing = on != 0;
  return QCBC_OK;
}

static int pt_sync(qcbc_pt* pt) {
  for (int k = 0; k < pt->G; k++) HIPCHK(hipStreamSynchronize(pt->g[k].s));
  return QCBC_OK;
}

// Steps per launch of the multi-step graphs.
constexpr int kGraphSteps = 32;

// Steps per launch of QuickCBC's fused step kernel in a run (a launch of
// 128 steps of a 32 x 64 ensemble takes a few ms)
constexpr int kUpdateStepsPerLaunch = 128;

// Captures n consecutive steps of each column group into
// pt->graph/exec[slot][k], on the group's own stream.
static int pt_capture(qcbc_pt* pt, int slot, int n) {
  qcbc_ctx* c = pt->ctx;
  bool timing = c->timing;
  c->timing = false;  // event records stay out of the graphs
  int st = QCBC_OK;
  for (int k = 0; k < pt->G && !st; k++) {
    hipStream_t s = pt->g[k].s;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      st = QCBC_EHIP;
      break;
    }
    for (int i = 0; i < n && !st; i++) st = group_step(pt, k);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (!st && e != hipSuccess) st = QCBC_EHIP;
    if (st) {
      if (g) hipGraphDestroy(g);
      break;
    }
    pt->graph[slot][k] = g;
    if (hipGraphInstantiate(&pt->exec[slot][k], g, nullptr, nullptr, 0) != hipSuccess) st = QCBC_EHIP;
  }
  c->timing = timing;
  return st;
}

static void pt_drop_graphs(qcbc_pt* pt, int slot) {
  for (int k = 0; k < QCBC_PT_MAX_GROUPS; k++) {
    if (pt->exec[slot][k]) (void)hipGraphExecDestroy(pt->exec[slot][k]);
    if (pt->graph[slot][k]) (void)hipGraphDestroy(pt->graph[slot][k]);
    pt->exec[slot][k] = nullptr;
    pt->graph[slot][k] = nullptr;
  }
}

static int pt_launch_graphs(qcbc_pt* pt, int slot) {
  for (int k = 0; k < pt->G; k++) HIPCHK(hipGraphLaunch(pt->exec[slot][k], pt->g[k].s));
  return QCBC_OK;
}

int qcbc_pt_run(qcbc_pt* pt, int nsteps, int use_graph) {
  if (!pt || nsteps < 0) return QCBC_EINVAL;
  qcbc_ctx* c = pt->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  hipStream_t s = c->stream;
  pt->lat_ms.clear();
  // no memset nodes in any graph: every fill is a kernel.  Timed steps
  // replay the one-step graph; otherwise the run is kGraphSteps-step graphs
  // and one graph of the remaining steps, captured for that count.
  // QuickCBC's step runs whole runs of steps per launch of its fused kernel,
  // each group on its own stream: graphs would capture the same launches
  if (pt->upd && !pt->step_timing) {
    for (int done = 0; done < nsteps;) {
      const int n = nsteps - done < kUpdateStepsPerLaunch ? nsteps - done : kUpdateStepsPerLaunch;
      for (int k = 0; k < pt->G; k++) {
        int st = group_update_steps(pt, k, n);
        if (st) return st;
      }
      done += n;
    }
    return pt_sync(pt);
  }
  const bool multi = use_graph && !pt->step_timing;
  const int full = multi ? nsteps / kGraphSteps : 0, rem = multi ? nsteps % kGraphSteps : 0;
  if (use_graph && !multi && !pt->exec[0][0]) {
    int st = pt_capture(pt, 0, 1);
    if (st) {
      pt_drop_graphs(pt, 0);
      return st;
    }
  }
  if (full > 0 && !pt->exec[1][0]) {
    int st = pt_capture(pt, 1, kGraphSteps);
    if (st) {
      pt_drop_graphs(pt, 1);
      return st;
    }
  }
  if (rem > 0 && (!pt->exec[2][0] || pt->rem_steps != rem)) {
    pt_drop_graphs(pt, 2);
    pt->rem_steps = 0;
    int st = pt_capture(pt, 2, rem);
    if (st) {
      pt_drop_graphs(pt, 2);
      return st;
    }
    pt->rem_steps = rem;
  }
  // the groups' streams start after the work queued on the context stream
  if (pt->G > 1) {
    HIPCHK(hipEventRecord(pt->ev_fork, s));
    for (int k = 1; k < pt->G; k++) HIPCHK(hipStreamWaitEvent(pt->g[k].s, pt->ev_fork, 0));
  }
  if (multi) {
    for (int i = 0; i < full; i++) {
      int st = pt_launch_graphs(pt, 1);
      if (st) return st;
    }
    if (rem > 0) {
      int st = pt_launch_graphs(pt, 2);
      if (st) return st;
    }
    return pt_sync(pt);
  }
  for (int i = 0; i < nsteps; i++) {
    auto t0 = std::chrono::steady_clock::now();
    if (use_graph) {
      int st = pt_launch_graphs(pt, 0);
      if (st) return st;
    } else {
      int st = pt_enqueue_step(pt);
      if (st) return st;
    }
    if (pt->step_timing) {
      int st = pt_sync(pt);
      if (st) return st;
      auto t1 = std::chrono::steady_clock::now();
      pt->lat_ms.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
    }
  }
  return pt_sync(pt);
}

int qcbc_pt_get(qcbc_pt* pt, double* params, double* logL, long long* accepted, long long* swaps) {
  if (!pt) return QCBC_EINVAL;
  qcbc_ctx* c = pt->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  int st = pt_sync(pt);
  if (st) return st;
  hipStream_t s = c->stream;
  const size_t P = pt->P;
  std::vector<double> x(11 * P), L(P);
  std::vector<unsigned long long> acc(P), sw(pt->n_chains);
  HIPCHK(hipMemcpyAsync(x.data(), pt->d_x, sizeof(double) * 11 * P, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(L.data(), pt->d_L, sizeof(double) * P, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(acc.data(), pt->d_acc, sizeof(unsigned long long) * P, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(sw.data(), pt->d_swaps, sizeof(unsigned long long) * pt->n_chains, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (size_t i = 0; i < P; i++) {
    const size_t p = pt->perm[i];
    if (params) std::memcpy(params + 11 * p, &x[11 * i], sizeof(double) * 11);
    if (logL) logL[p] = L[i];
    if (accepted) accepted[p] = (long long)acc[i];
  }
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
  pt_sync(pt);
  pt_free(pt);
  return QCBC_OK;
}

}  // extern "C"
