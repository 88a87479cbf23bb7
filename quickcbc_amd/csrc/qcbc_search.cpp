// Host side of QuickCBC's search likelihoods (QuickCBC.c:5672-6408): the
// templates, the matched-filter transforms, the peak search and every
// penalized likelihood run on the GPU (qcbc_search.hip, the plain template
// kernels, hipFFT); what runs here is the per-point bookkeeping around them:
// APmax's parameter moves and its second pass, and log_likelihood_max's
// network maximisation over the detectors' peaks.
#include <hip/hip_runtime_api.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/qcbc.h"
#include "qcbc_api_internal.h"
#include "qcbc_internal.h"
#include "qcbc_search_internal.h"

using namespace qcbc;

#define HIPCHK(x)                \
  do {                           \
    hipError_t e_ = (x);         \
    if (e_ != hipSuccess) {      \
      qcbc::t_last_hip = e_;     \
      return QCBC_EHIP;          \
    }                            \
  } while (0)

namespace {

constexpr int NX = 7;              // quasi-intrinsic parameters (ConstCBC.h:67)
constexpr double FBW = 8.0;        // penalized band width, Hz (ConstCBC.h:26)
constexpr double LPMIN = 9.0;      // least logL of a peak (ConstCBC.h:25)
constexpr double DTP = 0.05;       // least time between peaks, s (ConstCBC.h:24)
constexpr double PI_ = 3.1415926535897931159979634685442;
constexpr double TWO_PI = 6.2831853071795862319959269370884;

struct Mem {
  std::vector<void*> bufs;
  ~Mem() {
    for (void* p : bufs) (void)hipFree(p);
  }
  template <class T>
  int alloc(size_t n, T** out) {
    void* p = nullptr;
    if (hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) return QCBC_ENOMEM;
    bufs.push_back(p);
    *out = static_cast<T*>(p);
    return QCBC_OK;
  }
};

struct Plan {
  hipfftHandle h = nullptr;
  ~Plan() {
    if (h) hipfftDestroy(h);
  }
};

int band_start(double fs, double fmin, double Tobs) { return (fmin > fs) ? (int)(Tobs * fmin) : (int)(Tobs * fs); }

// log_likelihood_penalized for nq jobs on the analysis's data: out nq x 3
int penalized_run(qcbc_ctx* c, const std::vector<PenJob>& jobs, const double* d_hw, const double* d_h,
                  const int* d_imin, std::vector<double>& out, hipStream_t s) {
  const int nq = (int)jobs.size();
  out.assign((size_t)nq * 3, 0.0);
  if (nq == 0) return QCBC_OK;
  const double Tobs = c->desc.Tobs;
  int imax = (int)(Tobs * c->desc.fmax);
  if (imax > c->nhalf) imax = c->nhalf;
  Mem mem;
  PenJob* d_jobs;
  double* d_out;
  int st;
  if ((st = mem.alloc(jobs.size(), &d_jobs)) || (st = mem.alloc(out.size(), &d_out))) return st;
  HIPCHK(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(PenJob) * nq, hipMemcpyHostToDevice, s));
  HIPCHK(launch_penalized(d_jobs, nq, d_hw, d_h, c->d_D, c->d_SN, d_imin, c->desc.n_det, c->desc.N, Tobs, imax,
                          (int)(FBW * Tobs), d_out, s));
  HIPCHK(hipMemcpyAsync(out.data(), d_out, sizeof(double) * out.size(), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return QCBC_OK;
}

// the penalized likelihood's first bin per point: max(fmin, fbegin) Tobs
int penalized_starts(qcbc_ctx* c, const double* d_par, int ips, int B, double tc_at, int* d_imin, hipStream_t s) {
  Mem mem;
  double* d_fs;
  int st = mem.alloc(B, &d_fs);
  if (st) return st;
  HIPCHK(launch_fbegin_at(d_par, ips, B, tc_at, d_fs, s));
  std::vector<double> fs(B);
  HIPCHK(hipMemcpyAsync(fs.data(), d_fs, sizeof(double) * B, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  std::vector<int> lo(B);
  for (int b = 0; b < B; b++) lo[b] = band_start(fs[b], c->desc.fmin, c->desc.Tobs);
  HIPCHK(hipMemcpyAsync(d_imin, lo.data(), sizeof(int) * B, hipMemcpyHostToDevice, s));
  return QCBC_OK;
}

// log_likelihood_APmax (QuickCBC.c:5672-5762) for B points, params in place
int apmax_impl(qcbc_ctx* c, double* params, int B, double* logL, int* status) {
  const int nd = c->desc.n_det, N = c->desc.N, nh = c->nhalf, ips = NX + 3 * (nd - 1);
  hipStream_t s = c->stream;
  std::vector<int> active(B), st_h(B);
  for (int b = 0; b < B; b++) active[b] = b;
  for (int b = 0; b < B; b++) logL[b] = 0.0;
  for (int round = 0; round < 2 && !active.empty(); round++) {
    const int K = (int)active.size();
    Mem mem;
    double *d_par, *d_amp, *d_ph, *d_hw;
    Point* d_pts;
    int *d_st, *d_imin;
    int st;
    if ((st = mem.alloc((size_t)K * ips, &d_par)) || (st = mem.alloc((size_t)K, &d_pts)) ||
        (st = mem.alloc((size_t)K, &d_st)) || (st = mem.alloc((size_t)K, &d_imin)) ||
        (st = mem.alloc((size_t)K * nh, &d_amp)) || (st = mem.alloc((size_t)K * nh, &d_ph)) ||
        (st = mem.alloc((size_t)K * nd * N, &d_hw)))
      return st;
    std::vector<double> p((size_t)K * ips);
    for (int k = 0; k < K; k++) std::memcpy(&p[(size_t)k * ips], params + (size_t)active[k] * ips, sizeof(double) * ips);
    HIPCHK(hipMemcpyAsync(d_par, p.data(), sizeof(double) * p.size(), hipMemcpyHostToDevice, s));
    // templates() of each point (QuickCBC.c:3527-3611) on the plain kernels
    HIPCHK(launch_setup_intrinsic(d_par, ips, K, c->an, d_pts, d_st, s));
    HIPCHK(launch_waveform(d_pts, K, nh, c->desc.Tobs, d_amp, d_ph, s));
    HIPCHK(launch_project(d_pts, K, d_amp, d_ph, nd, N, c->desc.Tobs, c->an.sqT, d_hw, s));
    if ((st = penalized_starts(c, d_par, ips, K, NAN, d_imin, s))) return st;
    std::vector<int> pst(K);
    HIPCHK(hipMemcpyAsync(pst.data(), d_st, sizeof(int) * K, hipMemcpyDeviceToHost, s));
    std::vector<PenJob> jobs;
    for (int k = 0; k < K; k++)
      for (int d = 0; d < nd; d++) jobs.push_back(PenJob{k, d, 0, 0, 0.0, 0.0, 0.0});
    std::vector<double> out;
    if ((st = penalized_run(c, jobs, d_hw, nullptr, d_imin, out, s))) return st;
    std::vector<int> next;
    for (int k = 0; k < K; k++) {
      const int b = active[k];
      st_h[b] = pst[k];
      if (pst[k] != ST_OK) {
        logL[b] = 0.0;
        continue;
      }
      double* q = params + (size_t)b * ips;
      double A[3], P[3], L = 0.0;
      for (int d = 0; d < nd; d++) {
        L += out[((size_t)k * nd + d) * 3 + 0];
        A[d] = out[((size_t)k * nd + d) * 3 + 1];
        P[d] = out[((size_t)k * nd + d) * 3 + 2];
      }
      logL[b] = L;
      q[6] -= std::log(A[0]);
      q[4] -= P[0] / 2.0;  // the template phase is the GW phase, twice the orbital one
      if (q[4] < 0.0) q[4] += PI_;
      if (q[4] > PI_) q[4] -= PI_;
      for (int d = 1; d < nd; d++) {
        double* off = q + NX + (d - 1) * 3;
        off[0] -= P[d] - P[0];
        off[2] *= A[d] / A[0];
        if (off[0] < 0.0) off[0] += TWO_PI;
        if (off[0] > TWO_PI) off[0] -= TWO_PI;
      }
      bool moved = false;
      for (int d = 0; d < nd; d++)
        if (A[d] > 1.1 || A[d] < 0.9 || std::fabs(P[d]) > 0.1) moved = true;
      if (moved) next.push_back(b);  // one more pass from the moved point
    }
    active.swap(next);
  }
  if (status)
    for (int b = 0; b < B; b++) status[b] = st_h[b];
  return QCBC_OK;
}

struct Peak {
  double L, t, P, A;
};

// log_likelihood_max (QuickCBC.c:5765-6231) for B points, params in place
int llmax_impl(qcbc_ctx* c, double* params, int B, double tmin, double tmax, double net_tmax, const double* delays,
               double* logL, int* status) {
  const int nd = c->desc.n_det, N = c->desc.N, ips = NX + 3 * (nd - 1);
  const double Tobs = c->desc.Tobs, dt = Tobs / (double)N;
  hipStream_t s = c->stream;
  for (int b = 0; b < B; b++) {
    logL[b] = 0.0;
    if (status) status[b] = ST_OK;
  }
  // points whose fbegin (their own tc) is below fmax take part (QuickCBC.c:5809)
  std::vector<int> act;
  {
    Mem mem;
    double *d_par, *d_fs;
    int st;
    if ((st = mem.alloc((size_t)B * ips, &d_par)) || (st = mem.alloc((size_t)B, &d_fs))) return st;
    HIPCHK(hipMemcpyAsync(d_par, params, sizeof(double) * ips * B, hipMemcpyHostToDevice, s));
    HIPCHK(launch_fbegin_at(d_par, ips, B, NAN, d_fs, s));
    std::vector<double> fs(B);
    HIPCHK(hipMemcpyAsync(fs.data(), d_fs, sizeof(double) * B, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int b = 0; b < B; b++)
      if (fs[b] < c->desc.fmax) act.push_back(b);
  }
  const int K = (int)act.size();
  if (K == 0) return QCBC_OK;
  int imin = (int)(Tobs * c->desc.fmin), imax = (int)(Tobs * c->desc.fmax);
  if (imin < 1) imin = 1;
  if (imax > N / 2) imax = N / 2;
  int mn = (int)(tmin / Tobs * (double)N) - 1, mx = (int)(tmax / Tobs * (double)N) + 1;
  if (mn < 0) mn = 0;
  if (mx > N) mx = N;
  const int M = mx - mn;
  if (M < 1) return QCBC_EINVAL;
  const int maxp = std::min(M, (int)((tmax - tmin) / DTP) + 3);
  const int rows = K * nd;
  Mem mem;
  double *d_par, *d_fs, *d_ts, *d_h, *d_hhp, *d_hh, *d_pk;
  void* d_z;
  Point* d_pts;
  int *d_st, *d_npk, *d_imin;
  unsigned char* d_mask;
  int st;
  if ((st = mem.alloc((size_t)K * ips, &d_par)) || (st = mem.alloc((size_t)K, &d_fs)) ||
      (st = mem.alloc((size_t)K, &d_ts)) || (st = mem.alloc((size_t)K, &d_pts)) || (st = mem.alloc((size_t)K, &d_st)) ||
      (st = mem.alloc((size_t)K * N, &d_h)) || (st = mem.alloc((size_t)rows * 2 * N, (double**)&d_z)) ||
      (st = mem.alloc((size_t)rows * xspec_blocks(N), &d_hhp)) || (st = mem.alloc((size_t)rows, &d_hh)) ||
      (st = mem.alloc((size_t)rows * M, &d_mask)) || (st = mem.alloc((size_t)rows, &d_npk)) ||
      (st = mem.alloc((size_t)rows * maxp * 4, &d_pk)) || (st = mem.alloc((size_t)K, &d_imin)))
    return st;
  // tc moved to the centre; the template PDwave with fbegin at tc = net_tmax
  std::vector<double> p((size_t)K * ips), ts(K, Tobs - Tobs / 2.0);
  for (int k = 0; k < K; k++) {
    std::memcpy(&p[(size_t)k * ips], params + (size_t)act[k] * ips, sizeof(double) * ips);
    p[(size_t)k * ips + 5] = Tobs / 2.0;
  }
  HIPCHK(hipMemcpyAsync(d_par, p.data(), sizeof(double) * p.size(), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_ts, ts.data(), sizeof(double) * K, hipMemcpyHostToDevice, s));
  HIPCHK(launch_setup_intrinsic(d_par, ips, K, c->an, d_pts, d_st, s));
  HIPCHK(launch_fbegin_at(d_par, ips, K, net_tmax, d_fs, s));
  HIPCHK(launch_pdwave(d_pts, d_fs, K, N, Tobs, d_ts, d_h, s));
  // the matched filter over every lag: one backward transform per (point, detector)
  HIPCHK(launch_xspec(d_h, c->d_D, c->d_SN, K, nd, N, imin, imax, d_z, d_hhp, d_hh, s));
  Plan plan;
  if (hipfftPlan1d(&plan.h, N, HIPFFT_Z2Z, rows) != HIPFFT_SUCCESS || hipfftSetStream(plan.h, s) != HIPFFT_SUCCESS)
    return QCBC_EHIP;
  hipfftDoubleComplex* zc = static_cast<hipfftDoubleComplex*>(d_z);
  if (hipfftExecZ2Z(plan.h, zc, zc, HIPFFT_BACKWARD) != HIPFFT_SUCCESS) return QCBC_EHIP;
  HIPCHK(launch_peaks(d_z, d_hh, rows, N, dt, mn, mx, LPMIN, DTP, maxp, d_mask, d_npk, d_pk, s));
  std::vector<int> npk(rows), pst(K);
  std::vector<double> pk((size_t)rows * maxp * 4);
  HIPCHK(hipMemcpyAsync(npk.data(), d_npk, sizeof(int) * rows, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(pk.data(), d_pk, sizeof(double) * pk.size(), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(pst.data(), d_st, sizeof(int) * K, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  // every peak's template, shifted and scaled, refined by the penalized likelihood
  if ((st = penalized_starts(c, d_par, ips, K, NAN, d_imin, s))) return st;
  std::vector<PenJob> jobs;
  for (int k = 0; k < K; k++)
    for (int d = 0; d < nd; d++)
      for (int j = 0; j < npk[(size_t)k * nd + d]; j++) {
        const double* q = &pk[(((size_t)k * nd + d) * maxp + j) * 4];
        jobs.push_back(PenJob{k, d, 1, 0, q[3], q[2], q[1]});
      }
  std::vector<double> out;
  if ((st = penalized_run(c, jobs, nullptr, d_h, d_imin, out, s))) return st;
  size_t jq = 0;
  for (int k = 0; k < K; k++) {
    const int b = act[k];
    if (status) status[b] = pst[k];
    std::vector<Peak> pkv[3];
    for (int d = 0; d < nd; d++)
      for (int j = 0; j < npk[(size_t)k * nd + d]; j++, jq++) {
        const double* q = &pk[(((size_t)k * nd + d) * maxp + j) * 4];
        pkv[d].push_back(Peak{out[jq * 3 + 0], q[1], q[2] - out[jq * 3 + 2], q[3] * out[jq * 3 + 1]});
      }
    if (pst[k] != ST_OK) continue;
    double* prm = params + (size_t)b * ips;
    prm[5] = Tobs / 2.0;
    double L = 0.0, delt[3] = {0, 0, 0}, psh[3] = {0, 0, 0}, nrm[3] = {1, 1, 1};
    if (nd > 1) {
      // the best pair of peaks within the pair's light travel time, a third
      // detector folded in with its best consistent peak
      double maxL[3][3] = {}, tx[3][3][3] = {}, px[3][3][3] = {}, ax[3][3][3] = {};
      for (int a = 0; a < nd; a++)
        for (int e = a + 1; e < nd; e++)
          for (const Peak& u : pkv[a]) {
            if (!(u.L > LPMIN)) continue;
            for (const Peak& v : pkv[e]) {
              if (!(v.L > LPMIN) || !(std::fabs(u.t - v.t) < delays[a * nd + e])) continue;
              const double x = u.L + v.L;
              if (x > maxL[a][e]) {
                maxL[a][e] = x;
                tx[a][e][a] = u.t;
                px[a][e][a] = u.P;
                ax[a][e][a] = u.A;
                tx[a][e][e] = v.t;
                px[a][e][e] = v.P;
                ax[a][e][e] = v.A;
              }
            }
          }
      if (nd > 2)
        for (int a = 0; a < nd; a++)
          for (int e = a + 1; e < nd; e++) {
            if (!(maxL[a][e] > 0.0)) continue;
            const double t0 = tx[a][e][a], t1 = tx[a][e][e];
            for (int g = 0; g < nd; g++) {
              if (g == a || g == e) continue;
              double x = 0.0;
              tx[a][e][g] = (t0 + t1) / 2.0;
              px[a][e][g] = 0.0;
              ax[a][e][g] = 1.0;
              // the reference's loop stops before the detector's last peak
              for (int i = 0; i + 1 < (int)pkv[g].size(); i++) {
                const Peak& w = pkv[g][i];
                if (std::fabs(t0 - w.t) < delays[a * nd + g] && std::fabs(t1 - w.t) < delays[e * nd + g] && w.L > x) {
                  x = w.L;
                  tx[a][e][g] = w.t;
                  px[a][e][g] = w.P;
                  ax[a][e][g] = w.A;
                }
              }
              maxL[a][e] += x;
            }
          }
      double best = 0.0;
      int bi = 0, bj = 1;
      for (int a = 0; a < nd; a++)
        for (int e = a + 1; e < nd; e++)
          if (maxL[a][e] > best) {
            best = maxL[a][e];
            bi = a;
            bj = e;
          }
      if (best > 0.0) {
        L = maxL[bi][bj];
        for (int d = 0; d < nd; d++) {
          delt[d] = tx[bi][bj][d];
          psh[d] = px[bi][bj][d];
          nrm[d] = ax[bi][bj][d];
        }
      }
    } else {
      double best = 0.0;
      int bk = -1;
      for (int i = 0; i < (int)pkv[0].size(); i++)
        if (pkv[0][i].L > best) {
          best = pkv[0][i].L;
          bk = i;
        }
      if (bk >= 0) {
        delt[0] = pkv[0][bk].t;
        nrm[0] = pkv[0][bk].A;
        psh[0] = pkv[0][bk].P;
        L = pkv[0][bk].L;
      }
    }
    logL[b] = L;
    prm[6] -= std::log(nrm[0]);
    prm[5] += delt[0];
    prm[4] += psh[0] / 2.0;
    if (prm[4] < 0.0) prm[4] += PI_;
    if (prm[4] > PI_) prm[4] -= PI_;
    for (int d = 1; d < nd; d++) {
      double* off = prm + NX + (d - 1) * 3;
      off[0] = psh[d] - psh[0];
      off[1] = delt[d] - delt[0];
      off[2] = nrm[d] / nrm[0];
      if (off[0] < 0.0) off[0] += TWO_PI;
      if (off[0] > TWO_PI) off[0] -= TWO_PI;
    }
  }
  return QCBC_OK;
}

}  // namespace

namespace qcbc {
int search_apmax(qcbc_ctx* c, double* params, int B, double* logL, int* status) {
  return apmax_impl(c, params, B, logL, status);
}
int search_llmax(qcbc_ctx* c, double* params, int B, double tmin, double tmax, double net_tmax, const double* delays,
                 double* logL, int* status) {
  return llmax_impl(c, params, B, tmin, tmax, net_tmax, delays, logL, status);
}
}  // namespace qcbc

extern "C" {

int qcbc_loglike_apmax(qcbc_ctx* c, double* params, int B, double* logL, int* status) {
  if (!c || B < 0 || (B > 0 && (!params || !logL))) return QCBC_EINVAL;
  if (B == 0) return QCBC_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  if (!c->bound) return QCBC_ENOTBOUND;
  hipSetDevice(c->device);
  return apmax_impl(c, params, B, logL, status);
}

int qcbc_loglike_max(qcbc_ctx* c, double* params, int B, double tmin, double tmax, double net_tmax,
                     const double* delays, double* logL, int* status) {
  if (!c || B < 0 || (B > 0 && (!params || !logL || !delays)) || !(tmin < tmax)) return QCBC_EINVAL;
  if (B == 0) return QCBC_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  if (!c->bound) return QCBC_ENOTBOUND;
  hipSetDevice(c->device);
  return llmax_impl(c, params, B, tmin, tmax, net_tmax, delays, logL, status);
}

}  // extern "C"

namespace qcbc {

// log_likelihood_penalized of one detector's arrays (host memory): D, twave
// half-complex N, SN N/2; params: at least 7 reference-frame parameters
int search_penalized_one(const double* params, const double* D, const double* twave, const double* SN, int N,
                         double Tobs, double fmin, double fmax, double* logL, double* A, double* P, hipStream_t s) {
  Mem mem;
  double *d_par, *d_D, *d_tw, *d_S, *d_fs, *d_out;
  int *d_imin;
  PenJob* d_job;
  int st;
  if ((st = mem.alloc(NX, &d_par)) || (st = mem.alloc((size_t)N, &d_D)) || (st = mem.alloc((size_t)N, &d_tw)) ||
      (st = mem.alloc((size_t)N / 2, &d_S)) || (st = mem.alloc(1, &d_fs)) || (st = mem.alloc(3, &d_out)) ||
      (st = mem.alloc(1, &d_imin)) || (st = mem.alloc(1, &d_job)))
    return st;
  HIPCHK(hipMemcpyAsync(d_par, params, sizeof(double) * NX, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_D, D, sizeof(double) * N, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_tw, twave, sizeof(double) * N, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_S, SN, sizeof(double) * (N / 2), hipMemcpyHostToDevice, s));
  HIPCHK(launch_fbegin_at(d_par, NX, 1, NAN, d_fs, s));
  double fs = 0.0;
  HIPCHK(hipMemcpyAsync(&fs, d_fs, sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  const int lo = band_start(fs, fmin, Tobs);
  int imax = (int)(Tobs * fmax);
  if (imax > N / 2) imax = N / 2;
  const PenJob job{0, 0, 0, 0, 0.0, 0.0, 0.0};
  HIPCHK(hipMemcpyAsync(d_imin, &lo, sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_job, &job, sizeof(PenJob), hipMemcpyHostToDevice, s));
  HIPCHK(launch_penalized(d_job, 1, d_tw, nullptr, d_D, d_S, d_imin, 1, N, Tobs, imax, (int)(FBW * Tobs), d_out, s));
  double out[3];
  HIPCHK(hipMemcpyAsync(out, d_out, sizeof(out), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  *logL = out[0];
  *A = out[1];
  *P = out[2];
  return QCBC_OK;
}

// PDwave of one point (host memory): wavef half-complex N
int search_pdwave(const double* params, double tcmin, int N, double Tobs, double fref, double fmin0, double* wavef,
                  hipStream_t s) {
  Mem mem;
  double *d_par, *d_fs, *d_ts, *d_h;
  Point* d_pt;
  int* d_st;
  int st;
  if ((st = mem.alloc(NX, &d_par)) || (st = mem.alloc(1, &d_fs)) || (st = mem.alloc(1, &d_ts)) ||
      (st = mem.alloc((size_t)N, &d_h)) || (st = mem.alloc(1, &d_pt)) || (st = mem.alloc(1, &d_st)))
    return st;
  AnalysisConst an{};
  an.Tobs = Tobs;
  an.inv_T = 1.0 / Tobs;
  an.sqT = std::sqrt(Tobs);
  an.fref = fref;
  an.fmin0 = fmin0;
  an.n_det = 1;
  an.N = N;
  an.imin = 1;
  an.imax = N / 2;
  const double ts = Tobs - params[5];
  HIPCHK(hipMemcpyAsync(d_par, params, sizeof(double) * NX, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_ts, &ts, sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(launch_setup_intrinsic(d_par, NX, 1, an, d_pt, d_st, s));
  HIPCHK(launch_fbegin_at(d_par, NX, 1, tcmin, d_fs, s));
  HIPCHK(launch_pdwave(d_pt, d_fs, 1, N, Tobs, d_ts, d_h, s));
  int pst = 0;
  HIPCHK(hipMemcpyAsync(&pst, d_st, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(wavef, d_h, sizeof(double) * N, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return pst == ST_OK ? QCBC_OK : QCBC_EDOM;
}

// pbt_shift (host memory): corr and corrf, N a power of 2
int search_pbt_shift(double* corr, double* corrf, const double* data1, const double* data2, const double* Sn, int imin,
                     int imax, int N, hipStream_t s) {
  Mem mem;
  double *d_a, *d_b, *d_S, *d_hhp, *d_hh;
  double* d_z;
  int st;
  if ((st = mem.alloc((size_t)N, &d_a)) || (st = mem.alloc((size_t)N, &d_b)) || (st = mem.alloc((size_t)N / 2, &d_S)) ||
      (st = mem.alloc((size_t)2 * N, &d_z)) || (st = mem.alloc((size_t)xspec_blocks(N), &d_hhp)) ||
      (st = mem.alloc(1, &d_hh)))
    return st;
  HIPCHK(hipMemcpyAsync(d_a, data1, sizeof(double) * N, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_b, data2, sizeof(double) * N, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_S, Sn, sizeof(double) * (N / 2), hipMemcpyHostToDevice, s));
  HIPCHK(launch_xspec(d_b, d_a, d_S, 1, 1, N, imin, imax, d_z, d_hhp, d_hh, s));
  Plan plan;
  if (hipfftPlan1d(&plan.h, N, HIPFFT_Z2Z, 1) != HIPFFT_SUCCESS || hipfftSetStream(plan.h, s) != HIPFFT_SUCCESS)
    return QCBC_EHIP;
  hipfftDoubleComplex* zc = reinterpret_cast<hipfftDoubleComplex*>(d_z);
  if (hipfftExecZ2Z(plan.h, zc, zc, HIPFFT_BACKWARD) != HIPFFT_SUCCESS) return QCBC_EHIP;
  std::vector<double> z((size_t)2 * N);
  HIPCHK(hipMemcpyAsync(z.data(), d_z, sizeof(double) * z.size(), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  // the half-complex inverse transforms are (2/N) Re z and (2/N) Im z
  for (int i = 0; i < N; i++) {
    corr[i] = z[2 * (size_t)i] * (2.0 / (double)N);
    corrf[i] = z[2 * (size_t)i + 1] * (2.0 / (double)N);
  }
  return QCBC_OK;
}

}  // namespace qcbc
