// Host side of libqcbc: resident analyses, the batch C ABI and the QuickCBC
// drop-in entry points.  Every numerical step runs in the HIP kernels; this
// file only moves data, sizes launches and maps errors.
#include <hip/hip_runtime_api.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/qcbc.h"
#include "qcbc_api_internal.h"
#include "qcbc_update_internal.h"
#include "qcbc_condition_internal.h"
#include "qcbc_internal.h"

using namespace qcbc;

namespace qcbc {

thread_local hipError_t t_last_hip = hipSuccess;

// ------------------------------------------------------------ device init
static std::mutex g_init_mu;
static int g_ready_dev[64];

int ensure_device() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return QCBC_ENODEV;
  if (dev < 0 || dev >= 64) return QCBC_ENODEV;
  if (g_ready_dev[dev] == 1) return QCBC_OK;
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_ready_dev[dev] == 1) return QCBC_OK;
  if (g_ready_dev[dev] == -1) return QCBC_ENODEV;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess || std::strstr(prop.gcnArchName, "gfx950") == nullptr) {
    g_ready_dev[dev] = -1;
    return QCBC_ENODEV;
  }
  if (launch_qnm_init(nullptr) != hipSuccess || launch_update_qnm_init(nullptr) != hipSuccess ||
      hipDeviceSynchronize() != hipSuccess) {
    g_ready_dev[dev] = -1;
    return QCBC_ENODEV;
  }
  // the no-allocation entry point's array; a failure here leaves that one
  // entry point reporting QCBC_ENODEV and the rest of the library working
  (void)noalloc_init_device(dev);
  g_ready_dev[dev] = 1;
  return QCBC_OK;
}

}  // namespace qcbc

#define HIPCHK(x)                     \
  do {                                \
    hipError_t e_ = (x);              \
    if (e_ != hipSuccess) {           \
      qcbc::t_last_hip = e_;          \
      return QCBC_EHIP;               \
    }                                 \
  } while (0)

// ------------------------------------------------------------------ ctx
static int ctx_reserve(qcbc_ctx* c, int B) {
  if (B <= c->cap_B) return QCBC_OK;
  int nb = std::max(B, 2 * c->cap_B);
  hipFree(c->d_params);
  hipFree(c->d_pts);
  hipFree(c->d_status);
  hipFree(c->d_logL);
  hipFree(c->d_rho);
  hipFree(c->d_partial);
  c->d_params = nullptr; c->d_pts = nullptr; c->d_status = nullptr;
  c->d_logL = nullptr; c->d_rho = nullptr; c->d_partial = nullptr;
  c->cap_B = 0;
  HIPCHK(hipMalloc(&c->d_params, sizeof(double) * 13 * nb));  // 11, or templates' 7 + 3 (n_det - 1)
  HIPCHK(hipMalloc(&c->d_pts, sizeof(Point) * nb));
  HIPCHK(hipMalloc(&c->d_status, sizeof(int) * nb));
  HIPCHK(hipMalloc(&c->d_logL, sizeof(double) * nb));
  HIPCHK(hipMalloc(&c->d_rho, sizeof(double) * 3 * nb));
  HIPCHK(hipMalloc(&c->d_partial, sizeof(double) * 6 * (size_t)QCBC_MAX_SPLITS * nb));
  c->cap_B = nb;
  return QCBC_OK;
}

int qcbc::make_analysis(const qcbc_desc* d, AnalysisConst* an, int* nhalf, int* nb, int* nbp) {
  if (!d || d->n_det < 1 || d->n_det > 3 || d->N < 4 || (d->N & 1) || !(d->Tobs > 0)) return QCBC_EINVAL;
  for (int i = 0; i < d->n_det; i++)
    if (d->labels[i] < 0 || d->labels[i] > 2) return QCBC_EINVAL;
  std::memset(an, 0, sizeof(*an));
  an->Tobs = d->Tobs;
  an->inv_T = 1.0 / d->Tobs;
  an->sqT = std::sqrt(d->Tobs);
  an->gmst = d->gmst;
  an->fref = d->fref;
  an->fmin0 = 1.0 / d->Tobs;  // freq[0] = 1/Tobs (QuickCBC.c:418)
  an->n_det = d->n_det;
  for (int i = 0; i < 3; i++) an->labels[i] = i < d->n_det ? d->labels[i] : 0;
  an->N = d->N;
  *nhalf = d->N / 2;
  int imin = (int)(d->Tobs * d->fmin);  // QuickCBC.c:4250-4252
  int imax = (int)(d->Tobs * d->fmax);
  if (imax > d->N / 2) imax = d->N / 2;
  if (imin < 1) imin = 1;
  if (imax < imin) imax = imin;
  an->imin = imin;
  an->imax = imax;
  *nb = imax - imin;
  *nbp = ((*nb + 1023) / 1024) * 1024;
  if (*nbp == 0) *nbp = 1024;
  return QCBC_OK;
}

size_t qcbc_ctx_footprint(const qcbc_desc* d) {
  AnalysisConst an;
  int nhalf, nb, nbp;
  if (make_analysis(d, &an, &nhalf, &nb, &nbp) != QCBC_OK) return 0;
  size_t nd = d->n_det;
  size_t bytes = sizeof(double) * (4 + 3 * nd) * nbp + sizeof(float) * 3 * nd * nbp;  // band planes
  bytes += sizeof(double) * nd * ((size_t)d->N + nhalf);                               // raw D, SN
  bytes += sizeof(double) * 2 * nhalf + sizeof(double) * nd * d->N;                    // plain-path scratch
  return bytes;
}

int qcbc_ctx_create(const qcbc_desc* desc, qcbc_ctx** out) {
  if (!out) return QCBC_EINVAL;
  *out = nullptr;
  int st = ensure_device();
  if (st) return st;
  qcbc_ctx* c = new qcbc_ctx();
  st = make_analysis(desc, &c->an, &c->nhalf, &c->nb, &c->nbp);
  if (st) {
    delete c;
    return st;
  }
  c->desc = *desc;
  hipGetDevice(&c->device);
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess && prop.multiProcessorCount > 0)
      c->n_cu = prop.multiProcessorCount;
  }
  size_t nd = desc->n_det;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&c->d_p64, sizeof(double) * (4 + 3 * nd) * c->nbp) != hipSuccess ||
      hipMalloc(&c->d_p32, sizeof(float) * 3 * nd * c->nbp) != hipSuccess ||
      hipMalloc(&c->d_D, sizeof(double) * nd * desc->N) != hipSuccess ||
      hipMalloc(&c->d_SN, sizeof(double) * nd * c->nhalf) != hipSuccess ||
      hipMalloc(&c->d_wref, sizeof(double) * 3) != hipSuccess) {
    qcbc_ctx_destroy(c);
    return QCBC_ENOMEM;
  }
  for (int p = 0; p < 2; p++) {
    c->cpb[p] = 8;
    c->tile[p] = 512;
    c->splits[p] = 0;
  }
  *out = c;
  return QCBC_OK;
}

int qcbc_ctx_destroy(qcbc_ctx* c) {
  if (!c) return QCBC_OK;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  for (auto& e : c->ev_pairs) {
    hipEventDestroy(e.first);
    hipEventDestroy(e.second);
  }
  if (c->cond_plan) hipfftDestroy(static_cast<hipfftHandle>(c->cond_plan));
  if (c->cond_iplan) hipfftDestroy(static_cast<hipfftHandle>(c->cond_iplan));
  void* bufs[] = {c->d_p64, c->d_p32, c->d_D, c->d_SN, c->d_wref, c->d_params, c->d_pts, c->d_status,
                  c->d_logL, c->d_rho, c->d_partial, c->d_amp, c->d_phase, c->d_hw, c->d_plain_partial,
                  c->d_cond_x, c->d_cond_X, c->d_spec};
  for (void* b : bufs)
    if (b) hipFree(b);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return QCBC_OK;
}

// The band planes from the data already in c->d_D and the PSD rows SN (host
// memory): SN is copied to the device, the FP32 scales chosen, the planes
// built.
static int bind_finish(qcbc_ctx* c, const double* const* SN) {
  const int nd = c->desc.n_det, N = c->desc.N, nh = c->nhalf;
  for (int d = 0; d < nd; d++)
    HIPCHK(hipMemcpyAsync(c->d_SN + (size_t)d * nh, SN[d], sizeof(double) * nh, hipMemcpyHostToDevice, c->stream));
  // per-detector FP32 scale: 1/S at the middle of the band
  double wref[3] = {1.0, 1.0, 1.0};
  int kmid = (c->an.imin + c->an.imax) / 2;
  for (int d = 0; d < nd; d++) {
    double s = SN[d][kmid];
    if (s > 0 && std::isfinite(s)) wref[d] = 1.0 / s;
    c->wref[d] = wref[d];
  }
  HIPCHK(hipMemcpyAsync(c->d_wref, wref, sizeof(wref), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(launch_bind(c->d_D, c->d_SN, nd, N, c->desc.Tobs, c->an.imin, c->nb, c->nbp, c->d_wref, c->d_p64, c->d_p32,
                     c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->bound = true;
  return QCBC_OK;
}

static int bind_impl(qcbc_ctx* c, const double* const* D, const double* const* SN) {
  if (!c || !D || !SN) return QCBC_EINVAL;
  hipSetDevice(c->device);
  const int nd = c->desc.n_det, N = c->desc.N;
  for (int d = 0; d < nd; d++)
    if (!D[d] || !SN[d]) return QCBC_EINVAL;
  c->bound = false;
  for (int d = 0; d < nd; d++)
    HIPCHK(hipMemcpyAsync(c->d_D + (size_t)d * N, D[d], sizeof(double) * N, hipMemcpyHostToDevice, c->stream));
  return bind_finish(c, SN);
}

int qcbc_bind_rows(qcbc_ctx* c, double* const* D, double* const* SN) {
  if (!c) return QCBC_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  return bind_impl(c, (const double* const*)D, (const double* const*)SN);
}

namespace {
struct CondTemp {
  std::vector<void*> bufs;
  ~CondTemp() {
    for (void* p : bufs) (void)hipFree(p);
  }
  int alloc(size_t bytes, void** out) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 1)) != hipSuccess) return QCBC_ENOMEM;
    bufs.push_back(p);
    *out = p;
    return QCBC_OK;
  }
};
}  // namespace

int qcbc_condition(qcbc_ctx* c, const double* x, const double* SN, double t_rise, double fmin_floor, double* D,
                   double* DW, double* pspec, double* wcheck) {
  if (!c || !x || !SN || !(t_rise >= 0.0) || !(fmin_floor >= 0.0)) return QCBC_EINVAL;
  const int nd = c->desc.n_det, N = c->desc.N, nh = c->nhalf;
  const double Tobs = c->desc.Tobs;
  if (N & (N - 1)) return QCBC_EINVAL;  // radix 2, as gsl_fft_real_radix2_transform
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  // scratch and plans are made on the first call and kept with the analysis
  if (!c->d_cond_x) HIPCHK(hipMalloc(&c->d_cond_x, sizeof(double) * nd * N));
  if (!c->d_cond_X) HIPCHK(hipMalloc(&c->d_cond_X, sizeof(double) * 2 * nd * (N / 2 + 1)));
  if (!c->cond_plan) {
    hipfftHandle plan = nullptr;
    if (hipfftPlan1d(&plan, N, HIPFFT_D2Z, nd) != HIPFFT_SUCCESS) return QCBC_EHIP;
    c->cond_plan = plan;
  }
  const bool spectra = DW || pspec || wcheck;
  if (DW && !c->cond_iplan) {
    hipfftHandle plan = nullptr;
    if (hipfftPlan1d(&plan, N, HIPFFT_Z2D, nd) != HIPFFT_SUCCESS) return QCBC_EHIP;
    c->cond_iplan = plan;
  }
  hipfftHandle plan = static_cast<hipfftHandle>(c->cond_plan);
  if (hipfftSetStream(plan, c->stream) != HIPFFT_SUCCESS) return QCBC_EHIP;
  c->bound = false;
  HIPCHK(hipMemcpyAsync(c->d_cond_x, x, sizeof(double) * nd * N, hipMemcpyHostToDevice, c->stream));
  // QuickCBC.c:257-265: alpha = 2 t_rise / Tobs, tukey, the forward transform
  HIPCHK(launch_tukey(c->d_cond_x, nd, N, 2.0 * t_rise / Tobs, c->stream));
  if (hipfftExecD2Z(plan, c->d_cond_x, static_cast<hipfftDoubleComplex*>(c->d_cond_X)) != HIPFFT_SUCCESS)
    return QCBC_EHIP;
  // D = the transform in the half-complex layout times sqrt(Tobs)/N (:309-313)
  HIPCHK(launch_pack_halfcomplex(c->d_cond_X, nd, N, sqrt(Tobs) / (double)N, c->d_D, c->stream));
  if (D) HIPCHK(hipMemcpyAsync(D, c->d_D, sizeof(double) * nd * N, hipMemcpyDeviceToHost, c->stream));
  CondTemp tmp;
  if (spectra) {
    // main's whitened series and spectra (:267-306) of the unscaled transform
    void *d_S, *d_W, *d_p = nullptr, *d_w = nullptr;
    int st;
    if ((st = tmp.alloc(sizeof(double) * nd * nh, &d_S)) ||
        (st = tmp.alloc(sizeof(double) * 2 * nd * (N / 2 + 1), &d_W)) ||
        (pspec && (st = tmp.alloc(sizeof(double) * nd * nh, &d_p))) ||
        (wcheck && (st = tmp.alloc(sizeof(double) * nd * nh, &d_w))))
      return st;
    HIPCHK(hipMemcpyAsync(d_S, SN, sizeof(double) * nd * nh, hipMemcpyHostToDevice, c->stream));
    HIPCHK(launch_spectra(c->d_cond_X, static_cast<double*>(d_S), nd, N, Tobs, d_W, static_cast<double*>(d_p),
                          static_cast<double*>(d_w), c->stream));
    if (pspec) HIPCHK(hipMemcpyAsync(pspec, d_p, sizeof(double) * nd * nh, hipMemcpyDeviceToHost, c->stream));
    if (wcheck) HIPCHK(hipMemcpyAsync(wcheck, d_w, sizeof(double) * nd * nh, hipMemcpyDeviceToHost, c->stream));
    if (DW) {
      hipfftHandle ip = static_cast<hipfftHandle>(c->cond_iplan);
      if (hipfftSetStream(ip, c->stream) != HIPFFT_SUCCESS) return QCBC_EHIP;
      // the time series is no longer needed: the backward transform lands there
      if (hipfftExecZ2D(ip, static_cast<hipfftDoubleComplex*>(d_W), c->d_cond_x) != HIPFFT_SUCCESS) return QCBC_EHIP;
      HIPCHK(launch_whitened_series(c->d_cond_x, nd, N, Tobs, c->d_cond_x, c->stream));
      HIPCHK(hipMemcpyAsync(DW, c->d_cond_x, sizeof(double) * nd * N, hipMemcpyDeviceToHost, c->stream));
    }
  }
  // then bind as qcbc_bind_data does, with D already on the device; CBC_start's
  // re-conditioning sets the PSD to 1 below fmin (:817-823)
  std::vector<double> floored;
  const double* rs[3];
  for (int d = 0; d < nd; d++) rs[d] = SN + (size_t)d * nh;
  if (fmin_floor > 0.0) {
    floored.assign(SN, SN + (size_t)nd * nh);
    const int j = std::min((int)(fmin_floor * Tobs), nh);
    for (int d = 0; d < nd; d++) {
      for (int i = 0; i < j; i++) floored[(size_t)d * nh + i] = 1.0;
      rs[d] = floored.data() + (size_t)d * nh;
    }
  }
  return bind_finish(c, rs);
}

int qcbc_bind_time_series(qcbc_ctx* c, const double* x, const double* SN, double t_rise, double* D) {
  return qcbc_condition(c, x, SN, t_rise, 0.0, D, nullptr, nullptr, nullptr);
}

int qcbc_makespec(qcbc_ctx* c, int Nspline, int Nlines, const double* ffit, const double* Sspline,
                  const double* linef, const double* lineh, const double* lineQ, const double* linew,
                  const double* deltafmax, double* SN) {
  if (!c || Nspline < 2 || Nlines < 0 || !ffit || !Sspline || !SN ||
      (Nlines > 0 && (!linef || !lineh || !lineQ || !linew || !deltafmax)))
    return QCBC_EINVAL;
  const int N = c->desc.N, h = N / 2;
  const double Tobs = c->desc.Tobs;
  for (int i = 1; i < Nspline; i++)
    if (!(ffit[i] > ffit[i - 1])) return QCBC_EINVAL;
  // gsl_spline_eval is defined on [ffit[0], ffit[n-1]] only (the reference
  // aborts outside it); bins 1 .. N/2 - 2 are evaluated
  if (h > 3 && (ffit[0] > 1.0 / Tobs || ffit[Nspline - 1] < (double)(h - 2) / Tobs)) return QCBC_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  const size_t nl = (size_t)Nlines, ns = (size_t)Nspline;
  const size_t total = std::max<size_t>(2 * ns + 5 * nl + ns + 3 * ns + (size_t)h, 1);
  if (total > c->spec_cap) {  // grown on demand and kept with the analysis
    if (c->d_spec) hipFree(c->d_spec);
    c->d_spec = nullptr;
    c->spec_cap = 0;
    HIPCHK(hipMalloc(&c->d_spec, sizeof(double) * total));
    c->spec_cap = total;
  }
  double* buf = c->d_spec;
  double *xa = buf, *ya = xa + ns, *lines = ya + ns, *cc = lines + 5 * nl, *w = cc + ns, *out = w + 3 * ns;
  hipError_t e = hipMemcpyAsync(xa, ffit, sizeof(double) * ns, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(ya, Sspline, sizeof(double) * ns, hipMemcpyHostToDevice, c->stream);
  const double* lrows[5] = {linef, lineh, lineQ, linew, deltafmax};
  for (int r = 0; r < 5 && e == hipSuccess && nl > 0; r++)
    e = hipMemcpyAsync(lines + r * nl, lrows[r], sizeof(double) * nl, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = launch_makespec(xa, ya, Nspline, lines, Nlines, Tobs, N, cc, w, out, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(SN, out, sizeof(double) * h, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  HIPCHK(e);
  return QCBC_OK;
}

int qcbc_bind_data(qcbc_ctx* c, const double* D, const double* SN) {
  if (!c || !D || !SN) return QCBC_EINVAL;
  const double* rd[3];
  const double* rs[3];
  for (int d = 0; d < c->desc.n_det; d++) {
    rd[d] = D + (size_t)d * c->desc.N;
    rs[d] = SN + (size_t)d * c->nhalf;
  }
  std::lock_guard<std::mutex> lk(c->mu);
  return bind_impl(c, rd, rs);
}

// Launch shape: default splits give >= ~12 blocks per CU for B points (B
// counts every point in flight at once: a PT ensemble's column groups run
// concurrently, so their launches share that target).  Measured on the
// 16 s / 3-detector PT step: 12 per CU beats 8 (tail) and 16 or more
// (per-block set-up and reductions).
int qcbc::effective_splits(const qcbc_ctx* c, int prec, int B) {
  int s = c->splits[prec];
  if (s <= 0) {
    const int target = 12 * c->n_cu;
    int blocks = (B + c->cpb[prec] - 1) / c->cpb[prec];
    s = (target + blocks - 1) / std::max(blocks, 1);
  }
  int maxs = std::max(1, c->nb / 256);
  s = std::min(s, std::min(maxs, QCBC_MAX_SPLITS));
  return std::max(s, 1);
}

// Kernel view of a context for B points: split length a multiple of 64 bins
// (the kernels' chunk), at least one tile.
LLInst qcbc::make_inst(const qcbc_ctx* c, int tile, int B, int splits, const Point* pts, double* partial) {
  LLInst in{};
  in.p64 = c->d_p64;
  in.p32 = c->d_p32;
  in.pts = pts;
  in.partial = partial;
  in.inv_T = c->an.inv_T;
  in.nbp = c->nbp;
  in.imin = c->an.imin;
  in.nb = c->nb;
  in.nd = c->desc.n_det;
  in.B = B;
  int len = (c->nb + splits - 1) / std::max(splits, 1);
  len = ((len + 63) / 64) * 64;
  in.split_len = std::max(len, tile);
  in.splits = (c->nb + in.split_len - 1) / in.split_len;
  if (in.splits < 1) in.splits = 1;
  return in;
}

static void harvest_events(qcbc_ctx* c) {
  for (auto& e : c->ev_pairs) {
    float ms = 0.f;
    if (hipEventSynchronize(e.second) == hipSuccess && hipEventElapsedTime(&ms, e.first, e.second) == hipSuccess)
      c->kernel_ms += ms;
    hipEventDestroy(e.first);
    hipEventDestroy(e.second);
  }
  c->ev_pairs.clear();
}

int qcbc::loglike_partials(qcbc_ctx* c, const double* d_params, int B, int prec, int* d_status, LLInst* out) {
  if (!c->bound) return QCBC_ENOTBOUND;
  if (prec != QCBC_FP64 && prec != QCBC_MIXED) return QCBC_EINVAL;
  int st = ctx_reserve(c, B);
  if (st) return st;
  HIPCHK(launch_setup(d_params, B, c->an, c->d_pts, d_status, c->stream));
  const LLInst in = make_inst(c, c->tile[prec], B, effective_splits(c, prec, B), c->d_pts, c->d_partial);
  *out = in;
  return ll_launch_timed(c, prec, in, c->stream);
}

int qcbc::ll_launch_timed(qcbc_ctx* c, int prec, const LLInst& in, hipStream_t s) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c->timing) {
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, s));
  }
  HIPCHK(launch_ll(prec, in, c->cpb[prec], c->tile[prec], s));
  if (c->timing) {
    HIPCHK(hipEventRecord(e1, s));
    c->ev_pairs.emplace_back(e0, e1);
    c->launches++;
    if (c->ev_pairs.size() >= 512) harvest_events(c);
  }
  return QCBC_OK;
}

int qcbc::loglike_core(qcbc_ctx* c, const double* d_params, int B, int prec, double* d_logL, double* d_rho,
                       int* d_status) {
  LLInst in;
  int st = loglike_partials(c, d_params, B, prec, d_status, &in);
  if (st) return st;
  HIPCHK(launch_finalize(c->d_partial, in.splits, c->desc.n_det, c->d_pts, B, prec == QCBC_MIXED ? FIN_MIXED : FIN_FP64,
                         c->d_wref, d_logL, d_rho, c->stream));
  return QCBC_OK;
}

int qcbc_loglike_device(qcbc_ctx* c, const double* d_params, int B, int prec, double* d_logL, double* d_rho,
                        int* d_status) {
  if (!c || B < 0) return QCBC_EINVAL;
  if (B == 0) return QCBC_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  return loglike_core(c, d_params, B, prec, d_logL, d_rho, d_status);
}

// The fused likelihood kernel alone: set-up of B points (device memory) once,
// then reps launches back to back on the context's stream between two HIP
// events; their average duration, launch gaps included.
int qcbc_time_loglike_kernel(qcbc_ctx* c, const double* d_params, int B, int prec, int reps, double* ms_per_launch) {
  if (!c || B < 1 || !d_params || reps < 1 || !ms_per_launch) return QCBC_EINVAL;
  if (prec != QCBC_FP64 && prec != QCBC_MIXED) return QCBC_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  if (!c->bound) return QCBC_ENOTBOUND;
  int st = ctx_reserve(c, B);
  if (st) return st;
  HIPCHK(launch_setup(d_params, B, c->an, c->d_pts, c->d_status, c->stream));
  const LLInst in = make_inst(c, c->tile[prec], B, effective_splits(c, prec, B), c->d_pts, c->d_partial);
  HIPCHK(launch_ll(prec, in, c->cpb[prec], c->tile[prec], c->stream));  // warm
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return QCBC_EHIP;
  }
  hipError_t e = hipEventRecord(e0, c->stream);
  for (int r = 0; r < reps && e == hipSuccess; r++) e = launch_ll(prec, in, c->cpb[prec], c->tile[prec], c->stream);
  if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (e != hipSuccess) {
    qcbc::t_last_hip = e;
    return QCBC_EHIP;
  }
  *ms_per_launch = ms / reps;
  return QCBC_OK;
}

int qcbc_loglike(qcbc_ctx* c, const double* params, int B, int prec, double* logL, double* rho, int* status) {
  if (!c || B < 0 || (B > 0 && (!params || !logL))) return QCBC_EINVAL;
  if (B == 0) return QCBC_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  int st = ctx_reserve(c, B);
  if (st) return st;
  const int nd = c->desc.n_det;
  HIPCHK(hipMemcpyAsync(c->d_params, params, sizeof(double) * 11 * B, hipMemcpyHostToDevice, c->stream));
  st = loglike_core(c, c->d_params, B, prec, c->d_logL, c->d_rho, c->d_status);
  if (st) return st;
  HIPCHK(hipMemcpyAsync(logL, c->d_logL, sizeof(double) * B, hipMemcpyDeviceToHost, c->stream));
  if (rho) HIPCHK(hipMemcpyAsync(rho, c->d_rho, sizeof(double) * nd * B, hipMemcpyDeviceToHost, c->stream));
  if (status) HIPCHK(hipMemcpyAsync(status, c->d_status, sizeof(int) * B, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return QCBC_OK;
}

// ---------------------------------------------------------- plain path
static int plain_reserve(qcbc_ctx* c) {
  if (c->d_amp) return QCBC_OK;
  HIPCHK(hipMalloc(&c->d_amp, sizeof(double) * c->nhalf));
  HIPCHK(hipMalloc(&c->d_phase, sizeof(double) * c->nhalf));
  HIPCHK(hipMalloc(&c->d_hw, sizeof(double) * c->desc.n_det * c->desc.N));
  HIPCHK(hipMalloc(&c->d_plain_partial, sizeof(double) * 6 * QCBC_PLAIN_BLOCKS));
  return QCBC_OK;
}

int qcbc_loglike_plain(qcbc_ctx* c, const double* params, int B, double* logL, double* rho, int* status) {
  if (!c || B < 0 || (B > 0 && (!params || !logL))) return QCBC_EINVAL;
  if (B == 0) return QCBC_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  if (!c->bound) return QCBC_ENOTBOUND;
  int st = ctx_reserve(c, B);
  if (st) return st;
  st = plain_reserve(c);
  if (st) return st;
  const int nd = c->desc.n_det, N = c->desc.N;
  HIPCHK(hipMemcpyAsync(c->d_params, params, sizeof(double) * 11 * B, hipMemcpyHostToDevice, c->stream));
  for (int b = 0; b < B; b++) {
    // one point at a time, as each OpenMP thread of the reference runs one
    // log_likelihood_full: set-up, waveform loop, projection loop, inner products
    HIPCHK(launch_setup(c->d_params + 11 * b, 1, c->an, c->d_pts + b, c->d_status + b, c->stream));
    HIPCHK(launch_waveform(c->d_pts + b, 1, c->nhalf, c->desc.Tobs, c->d_amp, c->d_phase, c->stream));
    HIPCHK(launch_project(c->d_pts + b, 1, c->d_amp, c->d_phase, nd, N, c->desc.Tobs, c->an.sqT, c->d_hw, c->stream));
    HIPCHK(launch_nwip(c->d_hw, c->d_D, c->d_SN, 1, nd, N, c->an.imin, c->an.imax, QCBC_PLAIN_BLOCKS,
                       c->d_plain_partial, c->stream));
    HIPCHK(launch_finalize(c->d_plain_partial, QCBC_PLAIN_BLOCKS, nd, c->d_pts + b, 1, FIN_RAW, c->d_wref, c->d_logL + b,
                           c->d_rho + (size_t)nd * b, c->stream));
  }
  HIPCHK(hipMemcpyAsync(logL, c->d_logL, sizeof(double) * B, hipMemcpyDeviceToHost, c->stream));
  if (rho) HIPCHK(hipMemcpyAsync(rho, c->d_rho, sizeof(double) * nd * B, hipMemcpyDeviceToHost, c->stream));
  if (status) HIPCHK(hipMemcpyAsync(status, c->d_status, sizeof(int) * B, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return QCBC_OK;
}

// ------------------------------------------------- waveform / templates
namespace {
// Device temporaries of one call, freed on every return path.
struct DevTemp {
  void* p = nullptr;
  DevTemp() = default;
  DevTemp(const DevTemp&) = delete;
  DevTemp& operator=(const DevTemp&) = delete;
  ~DevTemp() {
    if (p) hipFree(p);
  }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};
}  // namespace

// ips = 0: the 11 parameters of fulltemplates; ips > 0: templates' ips
// reference-frame parameters per point
static int wave_batch(qcbc_ctx* c, const double* params, int B, double* amp, double* phase, double* hwave,
                      int* status, int ips = 0) {
  if (!c || B < 0 || (B > 0 && !params)) return QCBC_EINVAL;
  if (B == 0) return QCBC_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  const int nd = c->desc.n_det, N = c->desc.N, nh = c->nhalf;
  const int chunk = std::max(1, std::min(B, (int)((size_t)1 << 28) / std::max(1, (nd + 2) * N)));
  int st = ctx_reserve(c, B);
  if (st) return st;
  DevTemp t_amp, t_ph, t_hw;
  HIPCHK(hipMalloc(&t_amp.p, sizeof(double) * nh * chunk));
  HIPCHK(hipMalloc(&t_ph.p, sizeof(double) * nh * chunk));
  if (hwave) HIPCHK(hipMalloc(&t_hw.p, sizeof(double) * nd * N * chunk));
  double *d_amp = t_amp.as<double>(), *d_ph = t_ph.as<double>(), *d_hw = t_hw.as<double>();
  HIPCHK(hipMemcpyAsync(c->d_params, params, sizeof(double) * (ips ? ips : 11) * B, hipMemcpyHostToDevice,
                        c->stream));
  if (ips)
    HIPCHK(launch_setup_intrinsic(c->d_params, ips, B, c->an, c->d_pts, c->d_status, c->stream));
  else
    HIPCHK(launch_setup(c->d_params, B, c->an, c->d_pts, c->d_status, c->stream));
  for (int b0 = 0; b0 < B; b0 += chunk) {
    int n = std::min(chunk, B - b0);
    HIPCHK(launch_waveform(c->d_pts + b0, n, nh, c->desc.Tobs, d_amp, d_ph, c->stream));
    if (amp) HIPCHK(hipMemcpyAsync(amp + (size_t)b0 * nh, d_amp, sizeof(double) * nh * n, hipMemcpyDeviceToHost, c->stream));
    if (phase) HIPCHK(hipMemcpyAsync(phase + (size_t)b0 * nh, d_ph, sizeof(double) * nh * n, hipMemcpyDeviceToHost, c->stream));
    if (hwave) {
      HIPCHK(launch_project(c->d_pts + b0, n, d_amp, d_ph, nd, N, c->desc.Tobs, c->an.sqT, d_hw, c->stream));
      HIPCHK(hipMemcpyAsync(hwave + (size_t)b0 * nd * N, d_hw, sizeof(double) * nd * N * n, hipMemcpyDeviceToHost,
                            c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  if (status) HIPCHK(hipMemcpyAsync(status, c->d_status, sizeof(int) * B, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return QCBC_OK;
}

int qcbc_waveform(qcbc_ctx* c, const double* params, int B, double* amp, double* phase, int* status) {
  return wave_batch(c, params, B, amp, phase, nullptr, status);
}

int qcbc_templates(qcbc_ctx* c, const double* params, int B, double* hwave, int* status) {
  if (!hwave) return QCBC_EINVAL;
  return wave_batch(c, params, B, nullptr, nullptr, hwave, status);
}

int qcbc_templates_intrinsic(qcbc_ctx* c, const double* params, int B, double* hwave, int* status) {
  if (!c || !hwave) return QCBC_EINVAL;
  return wave_batch(c, params, B, nullptr, nullptr, hwave, status, 7 + 3 * (c->desc.n_det - 1));
}

int qcbc_loglike_intrinsic(qcbc_ctx* c, const double* params, int B, double* logL, double* snrsq, int* status) {
  if (!c || B < 0 || (B > 0 && (!params || !logL))) return QCBC_EINVAL;
  if (B == 0) return QCBC_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  if (!c->bound) return QCBC_ENOTBOUND;
  hipSetDevice(c->device);
  int st = ctx_reserve(c, B);
  if (st) return st;
  const int nd = c->desc.n_det, ips = 7 + 3 * (nd - 1);
  HIPCHK(hipMemcpyAsync(c->d_params, params, sizeof(double) * ips * B, hipMemcpyHostToDevice, c->stream));
  HIPCHK(launch_setup_intrinsic(c->d_params, ips, B, c->an, c->d_pts, c->d_status, c->stream));
  const LLInst in = make_inst(c, c->tile[QCBC_FP64], B, effective_splits(c, QCBC_FP64, B), c->d_pts, c->d_partial);
  st = ll_launch_timed(c, QCBC_FP64, in, c->stream);
  if (st) return st;
  HIPCHK(launch_finalize_intrinsic(c->d_partial, in.splits, nd, c->d_pts, B, c->d_logL, c->d_rho, c->stream));
  HIPCHK(hipMemcpyAsync(logL, c->d_logL, sizeof(double) * B, hipMemcpyDeviceToHost, c->stream));
  if (snrsq) HIPCHK(hipMemcpyAsync(snrsq, c->d_rho, sizeof(double) * nd * B, hipMemcpyDeviceToHost, c->stream));
  if (status) HIPCHK(hipMemcpyAsync(status, c->d_status, sizeof(int) * B, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return QCBC_OK;
}

// ------------------------------------------------------ shapes / timing
int qcbc_ctx_set_shape(qcbc_ctx* c, int prec, int cpb, int tile, int splits) {
  if (!c || (prec != 0 && prec != 1)) return QCBC_EINVAL;
  if (cpb < 1 || cpb > 16 || !ll_tile_ok(tile) || splits < 0 || splits > QCBC_MAX_SPLITS) return QCBC_EINVAL;
  if (ll_lds_bytes(prec, c->desc.n_det, tile) > 64 * 1024) return QCBC_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  c->cpb[prec] = cpb;
  c->tile[prec] = tile;
  c->splits[prec] = splits;
  return QCBC_OK;
}

int qcbc_ctx_get_shape(qcbc_ctx* c, int prec, int* cpb, int* tile, int* splits) {
  if (!c || (prec != 0 && prec != 1)) return QCBC_EINVAL;
  if (cpb) *cpb = c->cpb[prec];
  if (tile) *tile = c->tile[prec];
  if (splits) *splits = c->splits[prec];
  return QCBC_OK;
}

int qcbc_ctx_timing(qcbc_ctx* c, int enable, double* kernel_ms, long long* launches) {
  if (!c) return QCBC_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  harvest_events(c);
  if (kernel_ms) *kernel_ms = c->kernel_ms;
  if (launches) *launches = c->launches;
  if (enable >= 0) {
    c->timing = enable != 0;
    c->kernel_ms = 0.0;
    c->launches = 0;
  }
  return QCBC_OK;
}

int qcbc_ctx_sync(qcbc_ctx* c) {
  if (!c) return QCBC_EINVAL;
  hipSetDevice(c->device);
  HIPCHK(hipStreamSynchronize(c->stream));
  return QCBC_OK;
}

// Autotuner: times the fused kernel for candidate (chains per block, tile,
// splits) on a synthetic batch and keeps the fastest.  The batch repeats the
// centre of a typical prior so every candidate does the same work.
int qcbc_autotune(qcbc_ctx* c, int prec, int B) {
  if (!c || (prec != 0 && prec != 1) || B <= 0) return QCBC_EINVAL;
  if (!c->bound) return QCBC_ENOTBOUND;
  std::vector<double> prm((size_t)11 * B);
  // a spread of masses from 1.2+1.2 to 60+50 Msun at tc = Tobs - 2 s
  for (int b = 0; b < B; b++) {
    double u = (b + 0.5) / B;
    double m1 = 1.3 * std::pow(60.0 / 1.3, u), m2 = 0.85 * m1;
    double mc = std::pow(m1 * m2, 0.6) / std::pow(m1 + m2, 0.2);
    const double MS = 1.988546954961461467461011951140572744e30, MPC = 3.085677581491367278913937957796471611e22;
    double* p = &prm[(size_t)11 * b];
    p[0] = std::log(mc * MS); p[1] = std::log((m1 + m2) * MS); p[2] = 0.1; p[3] = -0.1; p[4] = 1.0;
    p[5] = c->desc.Tobs - 2.0; p[6] = std::log(400.0 * MPC); p[7] = 1.0; p[8] = 0.2; p[9] = 0.5; p[10] = 0.3;
  }
  std::vector<double> logL(B);
  int cpbs[] = {1, 2, 4, 8, 16};
  int tiles[] = {128, 256, 512};
  int best_cpb = c->cpb[prec], best_tile = c->tile[prec], best_split = c->splits[prec];
  double best = 1e30;
  int old_timing;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    old_timing = c->timing;
  }
  for (int cpb : cpbs)
    for (int tile : tiles) {
      if (ll_lds_bytes(prec, c->desc.n_det, tile) > 64 * 1024) continue;
      int blocks = (B + cpb - 1) / cpb;
      int base = std::max(1, (12 * c->n_cu + blocks - 1) / blocks);
      int cands[] = {std::max(1, base / 2), base, base * 2};
      for (int sp : cands) {
        if (sp > QCBC_MAX_SPLITS) continue;
        if (qcbc_ctx_set_shape(c, prec, cpb, tile, sp) != QCBC_OK) continue;
        qcbc_ctx_timing(c, 1, nullptr, nullptr);
        int st = QCBC_OK;
        for (int r = 0; r < 4 && st == QCBC_OK; r++) st = qcbc_loglike(c, prm.data(), B, prec, logL.data(), nullptr, nullptr);
        double ms = 0;
        long long n = 0;
        qcbc_ctx_timing(c, 0, &ms, &n);
        if (st != QCBC_OK) return st;
        if (n > 1) {
          // first launch warms caches; rate of the rest
          double per = ms / n;
          if (per < best) {
            best = per;
            best_cpb = cpb; best_tile = tile; best_split = sp;
          }
        }
      }
    }
  qcbc_ctx_set_shape(c, prec, best_cpb, best_tile, best_split);
  std::lock_guard<std::mutex> lk(c->mu);
  c->timing = old_timing;
  return QCBC_OK;
}

// ------------------------------------------------------------ misc
const char* qcbc_strerror(int code) {
  switch (code) {
    case QCBC_OK: return "success";
    case QCBC_EINVAL: return "invalid argument";
    case QCBC_ENOMEM: return "out of device memory";
    case QCBC_EFAULT: return "invalid pointer";
    case QCBC_EDOM: return "parameter point outside the IMRPhenomD domain";
    case QCBC_EHIP: return hipGetErrorString(qcbc::t_last_hip);
    case QCBC_ENODEV: return "no usable gfx950 device";
    case QCBC_ENOTBOUND: return "no data bound to this analysis";
    case QCBC_EBUSY: return "admission refused: memory budget or slots exhausted";
    case QCBC_ECANCELED: return "request canceled";
    case QCBC_ETIMEDOUT: return "timed out waiting for a request";
    case QCBC_ERANGE: return "batch or data larger than the capacity of this entry point";
    case QCBC_EIO: return "file cannot be read or written";
    default: return "unknown error";
  }
}

int qcbc_device_info(char* name, size_t len, int* cu_count, double* mem_gb) {
  int st = ensure_device();
  if (st) return st;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, dev));
  if (name && len) {
    std::snprintf(name, len, "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (cu_count) *cu_count = prop.multiProcessorCount;
  if (mem_gb) *mem_gb = prop.totalGlobalMem / 1e9;
  return QCBC_OK;
}

// ------------------------------------------------------ device memory
int qcbc_dev_alloc(size_t bytes, void** ptr) {
  if (!ptr) return QCBC_EINVAL;
  int st = ensure_device();
  if (st) return st;
  if (hipMalloc(ptr, bytes) != hipSuccess) return QCBC_ENOMEM;
  return QCBC_OK;
}

int qcbc_dev_free(void* ptr) {
  if (ptr) HIPCHK(hipFree(ptr));
  return QCBC_OK;
}

int qcbc_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return QCBC_OK;
}

int qcbc_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return QCBC_OK;
}

// ------------------------------------------------ set-up records
size_t qcbc_point_record_size(void) { return sizeof(Point); }

int qcbc_point_records(qcbc_ctx* c, const double* params, int B, void* records) {
  if (!c || B < 0 || (B > 0 && (!params || !records))) return QCBC_EINVAL;
  if (B == 0) return QCBC_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  int st = ctx_reserve(c, B);
  if (st) return st;
  HIPCHK(hipMemcpyAsync(c->d_params, params, sizeof(double) * 11 * B, hipMemcpyHostToDevice, c->stream));
  HIPCHK(launch_setup(c->d_params, B, c->an, c->d_pts, c->d_status, c->stream));
  HIPCHK(hipMemcpyAsync(records, c->d_pts, sizeof(Point) * B, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return QCBC_OK;
}

int qcbc_dev_copy(void* dst, const void* src, size_t bytes) {
  if (!dst || !src || (bytes & 15)) return QCBC_EINVAL;
  HIPCHK(launch_copy(src, dst, bytes, nullptr));
  HIPCHK(hipDeviceSynchronize());
  return QCBC_OK;
}
