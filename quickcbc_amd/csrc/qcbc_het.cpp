// Heterodyned likelihood (QuickCBC.c:5112-5444) on the GPU: the set-up of a
// heterodyne (heterodyne(), on a grid the caller chose, as het_space does)
// from a bound analysis, or imported from set-up arrays computed elsewhere,
// then batched evaluation of log_likelihood_het.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <utility>
#include <vector>

#include "../../include/qcbc.h"
#include "qcbc_api_internal.h"
#include "qcbc_fisher_internal.h"
#include "qcbc_het_internal.h"
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

struct qcbc_het {
  int device = 0;
  int M = 0, J = 0, NR = 0, nd = 0;
  double logLb = 0.0;
  AnalysisConst an{};  // the analysis's, with freq[0] of the grid as the model's f_min
  hipStream_t stream = nullptr;
  double* d_freq = nullptr;   // M
  double* d_amp = nullptr;    // nd x M
  double* d_phase = nullptr;  // nd x M
  double* d_w = nullptr;      // 3 x nd x M
  double* d_wf = nullptr;     // 3 x nd x M, folded with the reference amplitude
  double* d_pl = nullptr;     // 4 x M sample powers
  double* d_raw = nullptr;    // IP | SL | lc | ls | rc | rs | aa
  size_t off[7] = {}, len[7] = {};
  // batch scratch
  int cap = 0;
  double* d_params = nullptr;
  Point* d_pts = nullptr;
  int* d_status = nullptr;
  double* d_logL = nullptr;
  // Fisher-matrix scratch: FISHER_NV parameter sets, records and statuses per
  // point, then the matrices, jump sizes, eigenvectors and point statuses
  int fcap = 0;
  double* d_fq = nullptr;
  Point* d_fpts = nullptr;
  int* d_fst = nullptr;
  double* d_fp = nullptr;
  double* d_fish = nullptr;
  double* d_ej = nullptr;
  double* d_ev = nullptr;
  int* d_fstatus = nullptr;
  // kernel timing
  bool timing = false;
  double kernel_ms = 0.0;
  long long launches = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  std::mutex mu;
};

namespace {

enum { R_IP = 0, R_SL, R_LC, R_LS, R_RC, R_RS, R_AA };

void het_layout(qcbc_het* h) {
  const size_t J1 = h->J + 1, NR = h->NR, nd = h->nd;
  const size_t L[7] = {NR * J1 * J1, NR * nd * J1, NR * nd * J1, NR * nd * J1, nd * NR, nd * NR, nd * NR};
  size_t o = 0;
  for (int i = 0; i < 7; i++) {
    h->off[i] = o;
    h->len[i] = L[i];
    o += L[i];
  }
}

size_t raw_total(const qcbc_het* h) { return h->off[6] + h->len[6]; }

void het_free(qcbc_het* h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (auto& e : h->ev) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  void* bufs[] = {h->d_freq, h->d_amp,  h->d_phase, h->d_w,    h->d_wf,   h->d_pl,   h->d_raw,
                  h->d_params, h->d_pts, h->d_status, h->d_logL, h->d_fq, h->d_fpts, h->d_fst,
                  h->d_fp,   h->d_fish, h->d_ej,    h->d_ev,   h->d_fstatus};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

bool shape_ok(int M, int J) { return J >= 1 && J <= HET_MAX_J && M >= J + 1 && (M - 1) % J == 0; }

int het_alloc(qcbc_ctx* ctx, int M, int J, qcbc_het** out) {
  qcbc_het* h = new qcbc_het();
  h->device = ctx->device;
  h->M = M;
  h->J = J;
  h->NR = (M - 1) / J;
  h->nd = ctx->desc.n_det;
  h->an = ctx->an;
  het_layout(h);
  const size_t nd = h->nd;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&h->d_freq, sizeof(double) * M) != hipSuccess ||
      hipMalloc(&h->d_amp, sizeof(double) * nd * M) != hipSuccess ||
      hipMalloc(&h->d_phase, sizeof(double) * nd * M) != hipSuccess ||
      hipMalloc(&h->d_w, sizeof(double) * 3 * nd * M) != hipSuccess ||
      hipMalloc(&h->d_wf, sizeof(double) * 3 * nd * M) != hipSuccess ||
      hipMalloc(&h->d_pl, sizeof(double) * 4 * M) != hipSuccess ||
      hipMalloc(&h->d_raw, sizeof(double) * raw_total(h)) != hipSuccess) {
    het_free(h);
    return QCBC_ENOMEM;
  }
  *out = h;
  return QCBC_OK;
}

int het_reserve(qcbc_het* h, int B) {
  if (B <= h->cap) return QCBC_OK;
  const int nb = std::max(B, 2 * h->cap);
  (void)hipFree(h->d_params);
  (void)hipFree(h->d_pts);
  (void)hipFree(h->d_status);
  (void)hipFree(h->d_logL);
  h->d_params = nullptr;
  h->d_pts = nullptr;
  h->d_status = nullptr;
  h->d_logL = nullptr;
  h->cap = 0;
  HIPCHK(hipMalloc(&h->d_params, sizeof(double) * 11 * nb));
  HIPCHK(hipMalloc(&h->d_pts, sizeof(Point) * nb));
  HIPCHK(hipMalloc(&h->d_status, sizeof(int) * nb));
  HIPCHK(hipMalloc(&h->d_logL, sizeof(double) * nb));
  h->cap = nb;
  return QCBC_OK;
}

int fisher_reserve(qcbc_het* h, int B) {
  if (B <= h->fcap) return QCBC_OK;
  const int nb = std::max(B, 2 * h->fcap);
  void** bufs[] = {(void**)&h->d_fq,   (void**)&h->d_fpts, (void**)&h->d_fst, (void**)&h->d_fp,
                   (void**)&h->d_fish, (void**)&h->d_ej,   (void**)&h->d_ev,  (void**)&h->d_fstatus};
  for (void** b : bufs) {
    if (*b) (void)hipFree(*b);
    *b = nullptr;
  }
  h->fcap = 0;
  const size_t n = nb, nv = (size_t)nb * FISHER_NV, np = FISHER_NP;
  HIPCHK(hipMalloc(&h->d_fq, sizeof(double) * np * nv));
  HIPCHK(hipMalloc(&h->d_fpts, sizeof(Point) * nv));
  HIPCHK(hipMalloc(&h->d_fst, sizeof(int) * nv));
  HIPCHK(hipMalloc(&h->d_fp, sizeof(double) * np * n));
  HIPCHK(hipMalloc(&h->d_fish, sizeof(double) * np * np * n));
  HIPCHK(hipMalloc(&h->d_ej, sizeof(double) * np * n));
  HIPCHK(hipMalloc(&h->d_ev, sizeof(double) * np * np * n));
  HIPCHK(hipMalloc(&h->d_fstatus, sizeof(int) * n));
  h->fcap = nb;
  return QCBC_OK;
}

// Fisher_Het of B points (device params) into d_fish, and FisherEvec of each
// into d_ej / d_ev when they are given, on stream s (caller holds mu)
int fisher_eval(qcbc_het* h, const double* d_params, int B, double* d_fish, double* d_ej, double* d_ev,
                int* d_status, hipStream_t s) {
  int st = fisher_reserve(h, B);
  if (st) return st;
  HIPCHK(launch_fisher_params(d_params, B, h->d_fq, s));
  HIPCHK(launch_setup(h->d_fq, B * FISHER_NV, h->an, h->d_fpts, h->d_fst, s));
  HIPCHK(launch_het_fisher(h->d_fpts, d_params, B, h->nd, h->d_freq, h->d_pl, h->d_amp, h->d_w, h->M, d_fish,
                           d_status, s));
  if (d_ej && d_ev) HIPCHK(launch_fisher_evec(d_fish, B, FISHER_NP, d_ej, d_ev, nullptr, s));
  return QCBC_OK;
}

void harvest(qcbc_het* h) {
  for (auto& e : h->ev) {
    float ms = 0.f;
    if (hipEventSynchronize(e.second) == hipSuccess && hipEventElapsedTime(&ms, e.first, e.second) == hipSuccess)
      h->kernel_ms += ms;
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  h->ev.clear();
}

// set-up of B points then the evaluation kernel, on the het's stream (caller holds mu)
int het_eval(qcbc_het* h, const double* d_params, int B, double* d_logL, int* d_status) {
  int st = het_reserve(h, B);
  if (st) return st;
  HIPCHK(launch_setup(d_params, B, h->an, h->d_pts, d_status, h->stream));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (h->timing) {
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, h->stream));
  }
  HIPCHK(launch_het_ll(h->d_pts, B, h->nd, h->d_freq, h->d_pl, h->d_amp, h->d_phase, h->d_wf, h->M, h->logLb, d_logL,
                       h->stream));
  if (h->timing) {
    HIPCHK(hipEventRecord(e1, h->stream));
    h->ev.emplace_back(e0, e1);
    h->launches++;
    if (h->ev.size() >= 512) harvest(h);
  }
  return QCBC_OK;
}

}  // namespace

int qcbc::het_enqueue(qcbc_het* h, const double* d_params, int B, Point* pts, int* d_status, double* d_logL,
                      hipStream_t s) {
  HIPCHK(launch_setup(d_params, B, h->an, pts, d_status, s));
  HIPCHK(launch_het_ll(pts, B, h->nd, h->d_freq, h->d_pl, h->d_amp, h->d_phase, h->d_wf, h->M, h->logLb, d_logL, s));
  return QCBC_OK;
}

int qcbc::het_device(const qcbc_het* h) { return h->device; }

int qcbc::het_fisher_enqueue(qcbc_het* h, const double* d_params, int B, double* d_ej, double* d_ev, hipStream_t s) {
  std::lock_guard<std::mutex> lk(h->mu);
  int st = fisher_reserve(h, B);
  if (st) return st;
  // the het's own Fisher scratch: wait for any earlier use of it on the het's stream
  HIPCHK(hipStreamSynchronize(h->stream));
  st = fisher_eval(h, d_params, B, h->d_fish, d_ej, d_ev, h->d_fstatus, s);
  if (st) return st;
  HIPCHK(hipStreamSynchronize(s));  // the scratch is free again when this returns
  return QCBC_OK;
}

int qcbc::het_grid_view(const qcbc_het* h, HetGridView* out) {
  if (!h || !out) return QCBC_EINVAL;
  out->device = h->device;
  out->M = h->M;
  out->J = h->J;
  out->NR = h->NR;
  out->nd = h->nd;
  out->an = h->an;
  out->freq = h->d_freq;
  out->pl = h->d_pl;
  out->wf = h->d_wf;
  out->amp = h->d_amp;
  out->phase = h->d_phase;
  out->logLb = h->logLb;
  return QCBC_OK;
}

extern "C" {

int qcbc_het_create(qcbc_ctx* ctx, const double* params, const int* fgrid, int M, int J, int MN, int MM,
                    qcbc_het** out) {
  if (!out) return QCBC_EINVAL;
  *out = nullptr;
  if (!ctx || !params || !fgrid || !shape_ok(M, J)) return QCBC_EINVAL;
  const int n2 = ctx->nhalf, nd = ctx->desc.n_det, N = ctx->desc.N;
  // the bins heterodyne() reads: the grid inside (0, N/2), strictly increasing
  // (every region spans at least J bins), and the logLb band inside the data
  if (fgrid[0] < 1 || fgrid[M - 1] >= n2 || MN < 1 || MM > n2 || MN >= MM) return QCBC_EINVAL;
  for (int m = 1; m < M; m++)
    if (fgrid[m] <= fgrid[m - 1]) return QCBC_EINVAL;
  std::lock_guard<std::mutex> lk(ctx->mu);
  hipSetDevice(ctx->device);
  if (!ctx->bound) return QCBC_ENOTBOUND;
  qcbc_het* h = nullptr;
  int st = het_alloc(ctx, M, J, &h);
  if (st) return st;
  // scratch: params, point, status, logLb, h22 amp/phase, per-bin planes, partials, grid
  const size_t nbins = 3 * (size_t)nd * n2, npart = (size_t)HET_SETUP_BLOCKS * 2 * nd;
  const size_t ndbl = 11 + 1 + 2 * (size_t)n2 + nbins + npart;
  double* d_tmp = nullptr;
  Point* d_pt = nullptr;
  int* d_i = nullptr;
  auto cleanup = [&](int code) {
    if (d_tmp) hipFree(d_tmp);
    if (d_pt) hipFree(d_pt);
    if (d_i) hipFree(d_i);
    if (code != QCBC_OK) het_free(h);
    return code;
  };
  if (hipMalloc(&d_tmp, sizeof(double) * ndbl) != hipSuccess || hipMalloc(&d_pt, sizeof(Point)) != hipSuccess ||
      hipMalloc(&d_i, sizeof(int) * (1 + (size_t)M)) != hipSuccess)
    return cleanup(QCBC_ENOMEM);
  double* d_p = d_tmp;
  double* d_lb = d_p + 11;
  double* d_wa = d_lb + 1;
  double* d_wp = d_wa + n2;
  double* d_bins = d_wp + n2;
  double* d_part = d_bins + nbins;
  int* d_st = d_i;
  int* d_fg = d_i + 1;
  hipStream_t s = ctx->stream;
  double* r = h->d_raw;
  int pst = 0;
  double lb = 0.0;
  if (hipMemcpyAsync(d_p, params, sizeof(double) * 11, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(d_fg, fgrid, sizeof(int) * M, hipMemcpyHostToDevice, s) != hipSuccess ||
      launch_setup(d_p, 1, ctx->an, d_pt, d_st, s) != hipSuccess ||
      launch_waveform(d_pt, 1, n2, ctx->desc.Tobs, d_wa, d_wp, s) != hipSuccess ||
      launch_het_setup(d_pt, d_wa, d_wp, ctx->d_D, ctx->d_SN, nd, N, ctx->desc.Tobs, ctx->an.sqT, d_fg, M, J, MN, MM,
                       d_bins, d_part, d_lb, h->d_freq, h->d_amp, h->d_phase, r + h->off[R_IP], r + h->off[R_SL],
                       r + h->off[R_LC], r + h->off[R_LS], r + h->off[R_RC], r + h->off[R_RS], r + h->off[R_AA],
                       h->d_w, h->d_wf, h->d_pl, s) != hipSuccess ||
      hipMemcpyAsync(&pst, d_st, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(&lb, d_lb, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return cleanup(QCBC_EHIP);
  if (pst != QCBC_OK) return cleanup(pst);
  h->logLb = lb;
  h->an.fmin0 = ((double)fgrid[0]) / ctx->desc.Tobs;  // the grid's first frequency, as het->freq holds it
  cleanup(QCBC_OK);
  *out = h;
  return QCBC_OK;
}

int qcbc_het_import(qcbc_ctx* ctx, int M, int J, const double* freq, const double* amp, const double* phase,
                    const double* IP, const double* SL, const double* lc, const double* ls, const double* rc,
                    const double* rs, const double* aa, double logLb, qcbc_het** out) {
  if (!out) return QCBC_EINVAL;
  *out = nullptr;
  if (!ctx || !shape_ok(M, J) || !freq || !amp || !phase || !IP || !SL || !lc || !ls || !rc || !rs || !aa)
    return QCBC_EINVAL;
  if (!(freq[0] > 0)) return QCBC_EINVAL;
  hipSetDevice(ctx->device);
  qcbc_het* h = nullptr;
  int st = het_alloc(ctx, M, J, &h);
  if (st) return st;
  h->logLb = logLb;
  h->an.fmin0 = freq[0];
  const size_t nd = h->nd;
  const double* src[7] = {IP, SL, lc, ls, rc, rs, aa};
  hipStream_t s = h->stream;
  bool ok = hipMemcpyAsync(h->d_freq, freq, sizeof(double) * M, hipMemcpyHostToDevice, s) == hipSuccess &&
            hipMemcpyAsync(h->d_amp, amp, sizeof(double) * nd * M, hipMemcpyHostToDevice, s) == hipSuccess &&
            hipMemcpyAsync(h->d_phase, phase, sizeof(double) * nd * M, hipMemcpyHostToDevice, s) == hipSuccess;
  for (int i = 0; ok && i < 7; i++)
    ok = hipMemcpyAsync(h->d_raw + h->off[i], src[i], sizeof(double) * h->len[i], hipMemcpyHostToDevice, s) ==
         hipSuccess;
  double* r = h->d_raw;
  ok = ok &&
       launch_het_weights(h->d_freq, h->d_amp, r + h->off[R_IP], r + h->off[R_SL], r + h->off[R_LC],
                          r + h->off[R_LS], r + h->off[R_RC], r + h->off[R_RS], r + h->off[R_AA], M, J, h->nd, h->d_w,
                          h->d_wf, h->d_pl, s) == hipSuccess &&
       hipStreamSynchronize(s) == hipSuccess;
  if (!ok) {
    het_free(h);
    return QCBC_EHIP;
  }
  *out = h;
  return QCBC_OK;
}

int qcbc_het_info(qcbc_het* h, int* M, int* J, int* NR, int* n_det, double* logLb) {
  if (!h) return QCBC_EINVAL;
  if (M) *M = h->M;
  if (J) *J = h->J;
  if (NR) *NR = h->NR;
  if (n_det) *n_det = h->nd;
  if (logLb) *logLb = h->logLb;
  return QCBC_OK;
}

int qcbc_het_export(qcbc_het* h, double* freq, double* amp, double* phase, double* IP, double* SL, double* lc,
                    double* ls, double* rc, double* rs, double* aa) {
  if (!h) return QCBC_EINVAL;
  std::lock_guard<std::mutex> lk(h->mu);
  hipSetDevice(h->device);
  const size_t nd = h->nd, M = h->M;
  hipStream_t s = h->stream;
  if (freq) HIPCHK(hipMemcpyAsync(freq, h->d_freq, sizeof(double) * M, hipMemcpyDeviceToHost, s));
  if (amp) HIPCHK(hipMemcpyAsync(amp, h->d_amp, sizeof(double) * nd * M, hipMemcpyDeviceToHost, s));
  if (phase) HIPCHK(hipMemcpyAsync(phase, h->d_phase, sizeof(double) * nd * M, hipMemcpyDeviceToHost, s));
  double* dst[7] = {IP, SL, lc, ls, rc, rs, aa};
  for (int i = 0; i < 7; i++)
    if (dst[i])
      HIPCHK(hipMemcpyAsync(dst[i], h->d_raw + h->off[i], sizeof(double) * h->len[i], hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return QCBC_OK;
}

int qcbc_het_weights(qcbc_het* h, double* w) {
  if (!h || !w) return QCBC_EINVAL;
  std::lock_guard<std::mutex> lk(h->mu);
  hipSetDevice(h->device);
  HIPCHK(hipMemcpyAsync(w, h->d_w, sizeof(double) * 3 * h->nd * h->M, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return QCBC_OK;
}

int qcbc_het_loglike(qcbc_het* h, const double* params, int B, double* logL, int* status) {
  if (!h || B < 0 || (B > 0 && (!params || !logL))) return QCBC_EINVAL;
  if (B == 0) return QCBC_OK;
  std::lock_guard<std::mutex> lk(h->mu);
  hipSetDevice(h->device);
  int st = het_reserve(h, B);
  if (st) return st;
  HIPCHK(hipMemcpyAsync(h->d_params, params, sizeof(double) * 11 * B, hipMemcpyHostToDevice, h->stream));
  st = het_eval(h, h->d_params, B, h->d_logL, h->d_status);
  if (st) return st;
  HIPCHK(hipMemcpyAsync(logL, h->d_logL, sizeof(double) * B, hipMemcpyDeviceToHost, h->stream));
  if (status) HIPCHK(hipMemcpyAsync(status, h->d_status, sizeof(int) * B, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return QCBC_OK;
}

int qcbc_het_loglike_device(qcbc_het* h, const double* d_params, int B, double* d_logL, int* d_status) {
  if (!h || B < 0 || (B > 0 && (!d_params || !d_logL))) return QCBC_EINVAL;
  if (B == 0) return QCBC_OK;
  std::lock_guard<std::mutex> lk(h->mu);
  hipSetDevice(h->device);
  int st = het_reserve(h, B);
  if (st) return st;
  return het_eval(h, d_params, B, d_logL, d_status ? d_status : h->d_status);
}

int qcbc_het_fisher(qcbc_het* h, const double* params, int B, double* fish, double* ej, double* ev, int* status) {
  if (!h || B < 0 || (B > 0 && (!params || !fish)) || (!ej) != (!ev)) return QCBC_EINVAL;
  if (B == 0) return QCBC_OK;
  std::lock_guard<std::mutex> lk(h->mu);
  hipSetDevice(h->device);
  int st = fisher_reserve(h, B);
  if (st) return st;
  const size_t np = FISHER_NP;
  HIPCHK(hipMemcpyAsync(h->d_fp, params, sizeof(double) * np * B, hipMemcpyHostToDevice, h->stream));
  st = fisher_eval(h, h->d_fp, B, h->d_fish, ej ? h->d_ej : nullptr, ev ? h->d_ev : nullptr, h->d_fstatus,
                   h->stream);
  if (st) return st;
  HIPCHK(hipMemcpyAsync(fish, h->d_fish, sizeof(double) * np * np * B, hipMemcpyDeviceToHost, h->stream));
  if (ej) {
    HIPCHK(hipMemcpyAsync(ej, h->d_ej, sizeof(double) * np * B, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(ev, h->d_ev, sizeof(double) * np * np * B, hipMemcpyDeviceToHost, h->stream));
  }
  if (status) HIPCHK(hipMemcpyAsync(status, h->d_fstatus, sizeof(int) * B, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return QCBC_OK;
}

int qcbc_het_fisher_device(qcbc_het* h, const double* d_params, int B, double* d_fish, double* d_ej, double* d_ev,
                           int* d_status) {
  if (!h || B < 0 || (B > 0 && (!d_params || !d_fish)) || (!d_ej) != (!d_ev)) return QCBC_EINVAL;
  if (B == 0) return QCBC_OK;
  std::lock_guard<std::mutex> lk(h->mu);
  hipSetDevice(h->device);
  int st = fisher_reserve(h, B);
  if (st) return st;
  return fisher_eval(h, d_params, B, d_fish, d_ej, d_ev, d_status ? d_status : h->d_fstatus, h->stream);
}

int qcbc_het_fisher_time(qcbc_het* h, const double* d_params, int B, int reps, double* ms) {
  if (!h || B < 1 || !d_params || reps < 1 || !ms) return QCBC_EINVAL;
  std::lock_guard<std::mutex> lk(h->mu);
  hipSetDevice(h->device);
  int st = fisher_reserve(h, B);
  if (st) return st;
  hipEvent_t e[4];
  for (auto& x : e) HIPCHK(hipEventCreate(&x));
  HIPCHK(launch_fisher_params(d_params, B, h->d_fq, h->stream));
  HIPCHK(hipEventRecord(e[0], h->stream));
  for (int r = 0; r < reps; r++) HIPCHK(launch_setup(h->d_fq, B * FISHER_NV, h->an, h->d_fpts, h->d_fst, h->stream));
  HIPCHK(hipEventRecord(e[1], h->stream));
  for (int r = 0; r < reps; r++)
    HIPCHK(launch_het_fisher(h->d_fpts, d_params, B, h->nd, h->d_freq, h->d_pl, h->d_amp, h->d_w, h->M, h->d_fish,
                             h->d_fstatus, h->stream));
  HIPCHK(hipEventRecord(e[2], h->stream));
  for (int r = 0; r < reps; r++)
    HIPCHK(launch_fisher_evec(h->d_fish, B, FISHER_NP, h->d_ej, h->d_ev, nullptr, h->stream));
  HIPCHK(hipEventRecord(e[3], h->stream));
  HIPCHK(hipEventSynchronize(e[3]));
  for (int i = 0; i < 3; i++) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e[i], e[i + 1]));
    ms[i] = t / reps;
  }
  for (auto& x : e) hipEventDestroy(x);
  return QCBC_OK;
}

int qcbc_fisher_evec(qcbc_ctx* ctx, const double* fish, int B, int d, double* ej, double* ev) {
  if (!ctx || B < 0 || d < 1 || d > FISHER_EVEC_MAX_D || (B > 0 && (!fish || !ej || !ev))) return QCBC_EINVAL;
  if (B == 0) return QCBC_OK;
  std::lock_guard<std::mutex> lk(ctx->mu);
  hipSetDevice(ctx->device);
  const size_t dd = (size_t)d * d, n = (size_t)B * (2 * dd + d);
  double* buf = nullptr;
  HIPCHK(hipMalloc(&buf, sizeof(double) * n));
  double *d_f = buf, *d_ev = buf + dd * B, *d_ej = buf + 2 * dd * B;
  hipError_t e = hipMemcpyAsync(d_f, fish, sizeof(double) * dd * B, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = launch_fisher_evec(d_f, B, d, d_ej, d_ev, nullptr, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(ej, d_ej, sizeof(double) * d * B, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(ev, d_ev, sizeof(double) * dd * B, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(buf);
  HIPCHK(e);
  return QCBC_OK;
}

int qcbc_het_sync(qcbc_het* h) {
  if (!h) return QCBC_EINVAL;
  hipSetDevice(h->device);
  HIPCHK(hipStreamSynchronize(h->stream));
  return QCBC_OK;
}

int qcbc_het_timing(qcbc_het* h, int enable, double* kernel_ms, long long* launches) {
  if (!h) return QCBC_EINVAL;
  std::lock_guard<std::mutex> lk(h->mu);
  hipSetDevice(h->device);
  harvest(h);
  if (kernel_ms) *kernel_ms = h->kernel_ms;
  if (launches) *launches = h->launches;
  if (enable >= 0) {
    h->timing = enable != 0;
    h->kernel_ms = 0.0;
    h->launches = 0;
  }
  return QCBC_OK;
}

int qcbc_het_time_kernel(qcbc_het* h, const double* d_params, int B, int reps, double* ms_per_launch) {
  if (!h || B < 1 || !d_params || reps < 1 || !ms_per_launch) return QCBC_EINVAL;
  std::lock_guard<std::mutex> lk(h->mu);
  hipSetDevice(h->device);
  int st = het_reserve(h, B);
  if (st) return st;
  HIPCHK(launch_setup(d_params, B, h->an, h->d_pts, h->d_status, h->stream));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, h->stream));
  for (int r = 0; r < reps; r++)
    HIPCHK(launch_het_ll(h->d_pts, B, h->nd, h->d_freq, h->d_pl, h->d_amp, h->d_phase, h->d_wf, h->M, h->logLb,
                         h->d_logL, h->stream));
  HIPCHK(hipEventRecord(e1, h->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  *ms_per_launch = ms / reps;
  return QCBC_OK;
}

int qcbc_het_destroy(qcbc_het* h) {
  het_free(h);
  return QCBC_OK;
}

}  // extern "C"
