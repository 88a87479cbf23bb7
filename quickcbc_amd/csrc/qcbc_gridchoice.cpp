// Host side of the heterodyne's grid choice (QuickCBC.c:4275-5053) and of
// Fisher_All (:2695-2817).  The waveforms, the Fisher sums and every
// distance efix evaluates run on the GPU (qcbc_gridchoice.hip,
// qcbc_fisher.hip); what runs here is the serial logic around them: limits'
// running SNR sum, efix's search over the jump scale (all directions of all
// points advance in lockstep, one batched GPU evaluation per step), and
// het_space's scan for linear spans and its flat grid.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/qcbc.h"
#include "qcbc_api_internal.h"
#include "qcbc_fisher_internal.h"
#include "qcbc_gridchoice_internal.h"
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

namespace qcbc {
// Device allocations of one call, freed on every return path.
struct Scratch {
  std::vector<void*> bufs;
  ~Scratch() {
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
}  // namespace qcbc

namespace {

constexpr int NP = 11;
constexpr double TSUN = 4.92569043916e-6;  // ConstCBC.h:1
constexpr double MSUN_SI = 1.988546954961461467461011951140572744e30;
constexpr double TWO_PI = 6.2831853071795862319959269370884;  // ConstCBC.h:4
constexpr double PI_ = 3.1415926535897931159979634685442;

// fQNM (Utilities.c:313-354): the ringdown frequency of the final-mass and
// final-spin fits, Berti's fit of the fundamental mode (gr-qc/0512160);
// masses in seconds
double fqnm(double m1, double m2, double chi1, double chi2) {
  const double p0 = 0.04827, p1 = 0.01707, scri = 0.41616;
  const double k00 = -3.82, k01 = -1.2019, k02 = -1.20764, k10 = 3.79245, k11 = 1.18385, k12 = 4.90494;
  const double q = (m1 < m2) ? m1 / m2 : m2 / m1;
  const double eta = q / ((1.0 + q) * (1.0 + q));
  const double atot = (chi1 + q * q * chi2) / ((1.0 + q) * (1.0 + q));
  const double aeff = atot + scri * eta * (chi1 + chi2);
  const double Z1 =
      1.0 + std::pow((1.0 - aeff * aeff), (1. / 3.)) * (std::pow((1. + aeff), (1. / 3.)) + std::pow((1. - aeff), (1. / 3.)));
  const double Z2 = std::sqrt(3.0 * aeff * aeff + Z1 * Z1);
  const double risco = 3.0 + Z2 - (aeff / (std::fabs(aeff) + 1.0e-20)) * std::sqrt((3.0 - Z1) * (3.0 + Z1 + 2.0 * Z2));
  const double Eisco = std::sqrt(1.0 - 2. / (3.0 * risco));
  const double Lisco = (2. / (3.0 * std::sqrt(3.))) * (1. + 2.0 * std::sqrt(3.0 * risco - 2.0));
  const double Mf =
      (1.0 - eta * (1.0 - Eisco) - 4.0 * eta * eta * (4.0 * p0 + 16.0 * p1 * atot * (atot + 1.0) + Eisco - 1.0)) *
      (m1 + m2);
  const double e2 = eta * eta, e3 = e2 * eta;
  const double af = atot + eta * (Lisco - 2 * atot * (Eisco - 1.0)) + k00 * e2 + k01 * e2 * aeff + k02 * e2 * aeff * aeff +
                    k10 * e3 + k11 * e3 * aeff + k12 * e3 * aeff * aeff;
  return (1.5251 - 1.1568 * std::pow((1.0 - af), 0.1292)) / (TWO_PI * Mf);
}

// fringdown (Utilities.c:567-584)
double fringdown(const double* p) {
  const double Mchirp = std::exp(p[0]) / MSUN_SI * TSUN;
  const double M = std::exp(p[1]) / MSUN_SI * TSUN;
  double eta = std::pow((Mchirp / M), (5.0 / 3.0));
  if (eta > 0.25) eta = 0.25;
  const double dm = std::sqrt(1.0 - 4.0 * eta);
  return fqnm(M * (1.0 + dm) / 2.0, M * (1.0 - dm) / 2.0, p[2], p[3]);
}

// bins limits() and diff() cover: twice the ringdown frequency, at least
// 64 Hz, times Tobs (QuickCBC.c:4288-4290, :4933-4935)
int signal_bins(const double* p, double Tobs) {
  double fstop = 2.0 * fringdown(p);
  if (fstop < 64.0) fstop = 64.0;
  return (int)(fstop * Tobs);
}

// first bin of limits() and diff(): max(fmin, fbegin) Tobs (QuickCBC.c:4296-4303)
int first_bin(double fs, double fmin, double Tobs) { return (fmin > fs) ? (int)(Tobs * fmin) : (int)(Tobs * fs); }

// The displaced point of efix and het_space (QuickCBC.c:4418-4450,
// :4628-4662): params + step dir (- if that puts eta above 1/4); phi, alpha
// (period 2 pi) and psi (pi) reduced as the reference does (one radian added
// after the reduction of a negative angle, then whole periods until in
// [0, P]); clipped to the prior box; ln Mc reset to keep eta in [etamin, 0.2499].
void displace(const double* params, const double* dir, double step, const double* mn, const double* mx, double etamin,
              double* px) {
  static const int idx[3] = {4, 7, 9};
  const double per[3] = {TWO_PI, TWO_PI, PI_};
  for (int j = 0; j < NP; j++) px[j] = params[j] + step * dir[j];
  double eta = std::exp((5.0 / 3.0) * (px[0] - px[1]));
  if (eta > 0.25)
    for (int j = 0; j < NP; j++) px[j] = params[j] - step * dir[j];
  for (int a = 0; a < 3; a++) {
    const double x = px[idx[a]] / per[a];
    px[idx[a]] = (x - std::floor(x)) * per[a];
    if (x < 0.0) px[idx[a]] += 1.0;
  }
  for (int a = 0; a < 3; a++) {
    // a non-finite angle would never leave these loops; it fails the set-up instead
    if (!std::isfinite(px[idx[a]])) continue;
    while (px[idx[a]] > per[a]) px[idx[a]] -= per[a];
    while (px[idx[a]] < 0.0) px[idx[a]] += per[a];
  }
  for (int j = 0; j < NP; j++)
    if (px[j] > mx[j]) px[j] = mx[j];
  for (int j = 0; j < NP; j++)
    if (px[j] < mn[j]) px[j] = mn[j];
  eta = std::exp((5.0 / 3.0) * (px[0] - px[1]));
  if (eta > 0.25) px[0] = px[1] + 3.0 / 5.0 * std::log(0.2499);
  if (eta < etamin) px[0] = px[1] + 3.0 / 5.0 * std::log(etamin);
}

// Batched distances for efix: B start points (set up once) and, per call,
// K displaced points each paired with its start point.  Full-data mode: diff
// on the analysis's data (SN on the device); heterodyne mode: het_diff on
// the heterodyne's grid.
struct Distances {
  hipStream_t s = nullptr;
  AnalysisConst an{};
  int nd = 1, n2 = 0;
  double Tobs = 0;
  // full-data mode
  const double* d_SN = nullptr;
  double fmin = 0;
  // heterodyne mode
  bool het = false;
  HetGridView hv{};
  // start points
  int B = 0;
  std::vector<int> lo, hi;  // diff's range per start point
  Scratch mem;
  double* d_par = nullptr;  // (B + cap) x 11
  Point* d_pts = nullptr;
  int* d_st = nullptr;
  DiffPair* d_pairs = nullptr;
  double* d_partial = nullptr;
  double* d_z = nullptr;
  int cap = 0, splits = 1, max_bins = 1;

  int init(const double* params, int nb_points, int kcap) {
    B = nb_points;
    cap = kcap;
    int st;
    if ((st = mem.alloc((size_t)(B + cap) * NP, &d_par)) || (st = mem.alloc((size_t)(B + cap), &d_pts)) ||
        (st = mem.alloc((size_t)(B + cap), &d_st)) || (st = mem.alloc((size_t)cap, &d_pairs)) ||
        (st = mem.alloc((size_t)cap, &d_z)))
      return st;
    HIPCHK(hipMemcpyAsync(d_par, params, sizeof(double) * NP * B, hipMemcpyHostToDevice, s));
    HIPCHK(launch_setup(d_par, B, an, d_pts, d_st, s));
    if (!het) {
      // diff's bin range from each start point: its fbegin (the set-up
      // record's fs) and its ringdown frequency
      std::vector<Point> rec(B);
      HIPCHK(hipMemcpyAsync(rec.data(), d_pts, sizeof(Point) * B, hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      lo.resize(B);
      hi.resize(B);
      for (int b = 0; b < B; b++) {
        int Mw = 2 * signal_bins(params + (size_t)NP * b, Tobs);
        if (Mw > 2 * n2) Mw = 2 * n2;
        lo[b] = std::max(1, first_bin(rec[b].fs, fmin, Tobs));
        hi[b] = Mw / 2;
        max_bins = std::max(max_bins, hi[b] - lo[b]);
      }
      splits = diff_splits(max_bins);
      if ((st = mem.alloc((size_t)cap * splits, &d_partial))) return st;
    }
    return QCBC_OK;
  }

  // z[k] = distance between start point of[k] and px[k]
  int eval(const std::vector<double>& px, const std::vector<int>& of, std::vector<double>& z) {
    const int K = (int)of.size();
    z.assign(K, 0.0);
    if (K == 0) return QCBC_OK;
    if (K > cap) return QCBC_EINVAL;
    std::vector<DiffPair> pr(K);
    for (int k = 0; k < K; k++) {
      pr[k].a = of[k];
      pr[k].b = B + k;
      pr[k].lo = het ? 0 : lo[of[k]];
      pr[k].hi = het ? 0 : std::max(lo[of[k]], hi[of[k]]);
    }
    HIPCHK(hipMemcpyAsync(d_par + (size_t)NP * B, px.data(), sizeof(double) * NP * K, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d_pairs, pr.data(), sizeof(DiffPair) * K, hipMemcpyHostToDevice, s));
    HIPCHK(launch_setup(d_par + (size_t)NP * B, K, an, d_pts + B, d_st + B, s));
    if (het)
      HIPCHK(launch_diff_het(d_pts, d_pairs, K, nd, hv.freq, hv.pl, hv.wf, hv.M, d_z, s));
    else
      HIPCHK(launch_diff_full(d_pts, d_pairs, K, nd, d_SN, n2, Tobs, max_bins, splits, d_partial, d_z, s));
    HIPCHK(hipMemcpyAsync(z.data(), d_z, sizeof(double) * K, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return QCBC_OK;
  }
};

// efix (QuickCBC.c:4387-4567) for B points, every (point, direction) a
// small state machine advanced in lockstep: each round displaces every
// active direction by its current scale and evaluates all their distances
// in one batch.  ej (B x 11) scaled in place; trials (B x 11, may be null)
// the distances evaluated per direction.
int efix_run(Distances& E, const double* params, double* ej, const double* ev, const double* mn, const double* mx,
             double etamin, double zs, int* trials) {
  struct Dir {
    int phase = 0, k = 0, n = 0;
    double alpha0 = 1.0, alpha = 0.0, z0 = 0.0, dzmin = 1.0e20, alpham = 1.0;
  };
  const int B = E.B;
  const double zmx = zs * 2.0, zmn = zs / 2.0;
  std::vector<Dir> st((size_t)B * NP);
  std::vector<double> px, z;
  std::vector<int> of, which;
  for (;;) {
    px.clear();
    of.clear();
    which.clear();
    for (int q = 0; q < B * NP; q++) {
      Dir& d = st[q];
      if (d.phase == 2) continue;
      const int b = q / NP, i = q % NP;
      const double scale = (d.phase == 0 ? d.alpha0 : d.alpha) * ej[q];
      double y[NP];
      displace(params + (size_t)NP * b, ev + (size_t)NP * NP * b + (size_t)NP * i, scale, mn, mx, etamin, y);
      px.insert(px.end(), y, y + NP);
      of.push_back(b);
      which.push_back(q);
    }
    if (which.empty()) break;
    int rc = E.eval(px, of, z);
    if (rc) return rc;
    for (size_t n = 0; n < which.size(); n++) {
      Dir& d = st[which[n]];
      const double zz = z[n];
      d.n++;
      if (d.phase == 0) {
        d.z0 = zz;
        const double dz = std::fabs(zz - zs);
        if (dz < d.dzmin) {
          d.dzmin = dz;
          d.alpham = d.alpha0;
        }
        if (zz < zmn) d.alpha0 *= 2.0;
        if (zz > zmx) d.alpha0 /= 1.9;
        d.k++;
        if (!((zz > zmx || zz < zmn) && d.k < 15 && d.alpha0 < 1.0e4)) {
          d.phase = 1;
          d.alpha = d.alpha0 * 1.1;
          d.k = 0;
        }
      } else {
        double alphanew;
        if (d.alpha > d.alpha0)
          alphanew = d.alpha0 + (zs - d.z0) / (zz - d.z0) * (d.alpha - d.alpha0);
        else
          alphanew = d.alpha + (zs - zz) / (d.z0 - zz) * (d.alpha0 - d.alpha);
        const double dz = std::fabs(zz - zs);
        if (dz < d.dzmin) {
          d.dzmin = dz;
          d.alpham = d.alpha;
        }
        d.z0 = zz;
        d.alpha0 = d.alpha;
        d.alpha = alphanew;
        d.k++;
        if (!(std::fabs(zz - zs) > 0.2 && d.k < 10 && d.alpha < 1.0e4)) d.phase = 2;
      }
    }
  }
  for (int q = 0; q < B * NP; q++) {
    Dir& d = st[q];
    if (d.dzmin / zs > 1.0) d.alpham = 0.0;  // a direction that never came near zs
    ej[q] *= d.alpham;
    if (trials) trials[q] = d.n;
  }
  return QCBC_OK;
}

int check_box(const double* mn, const double* mx, double etamin) {
  if (!mn || !mx || !(etamin > 0.0 && etamin <= 0.25)) return QCBC_EINVAL;
  for (int j = 0; j < NP; j++)
    if (!(mn[j] <= mx[j])) return QCBC_EINVAL;
  return QCBC_OK;
}

}  // namespace

namespace qcbc {

// Fisher_All of B points (device params d_par) on the band bins of c, SN on
// the device, into d_fish (B x 11 x 11) and d_status (B), on stream s
static int fisher_all_enqueue(qcbc_ctx* c, const double* d_SN, const double* d_par, int B, double* d_fish, int* d_status,
                       hipStream_t s, Scratch& mem) {
  const int nd = c->desc.n_det, nb = c->nb, imin = c->an.imin;
  if (nb < 1) return QCBC_EINVAL;
  const int splits = std::max(1, std::min((nb + 255) / 256, (1024 + B - 1) / B));
  double *freq, *pl, *w, *q, *partial = nullptr;
  Point* pts;
  int* pst;
  int st;
  if ((st = mem.alloc((size_t)nb, &freq)) || (st = mem.alloc((size_t)4 * nb, &pl)) ||
      (st = mem.alloc((size_t)nd * nb, &w)) || (st = mem.alloc((size_t)B * FISHER_NV * NP, &q)) ||
      (st = mem.alloc((size_t)B * FISHER_NV, &pts)) || (st = mem.alloc((size_t)B * FISHER_NV, &pst)))
    return st;
  if (splits > 1 && (st = mem.alloc((size_t)B * splits * FISHER_NPAIR, &partial))) return st;
  HIPCHK(launch_band_grid(d_SN, nd, c->nhalf, imin, nb, c->desc.Tobs, freq, w, s));
  HIPCHK(launch_het_planes(freq, nb, pl, s));
  HIPCHK(launch_fisher_params(d_par, B, q, s));
  HIPCHK(launch_setup(q, B * FISHER_NV, c->an, pts, pst, s));
  HIPCHK(launch_het_fisher(pts, d_par, B, nd, freq, pl, nullptr, w, nb, d_fish, d_status, s, splits, partial));
  return QCBC_OK;
}

// limits (QuickCBC.c:4275-4385) at params on c's grid with the PSD rows SN
// (host) and on the device: stst, and ratio (n2 entries, 1 outside 1 .. M-1)
int limits_impl(qcbc_ctx* c, const double* const* SN, const double* params, int* stst, double* ratio,
                hipStream_t s) {
  const int nd = c->desc.n_det, n2 = c->nhalf;
  const double Tobs = c->desc.Tobs;
  int M = std::min(signal_bins(params, Tobs), n2);
  if (M < 1) return QCBC_EINVAL;
  Scratch mem;
  double *d_par, *d_amp, *d_phs;
  Point* d_pt;
  int* d_st;
  int st;
  if ((st = mem.alloc(NP, &d_par)) || (st = mem.alloc(1, &d_pt)) || (st = mem.alloc(1, &d_st)) ||
      (st = mem.alloc((size_t)nd * M, &d_amp)) || (st = mem.alloc((size_t)nd * M, &d_phs)))
    return st;
  Point rec;
  std::vector<double> amp((size_t)nd * M);
  HIPCHK(hipMemcpyAsync(d_par, params, sizeof(double) * NP, hipMemcpyHostToDevice, s));
  HIPCHK(launch_setup(d_par, 1, c->an, d_pt, d_st, s));
  HIPCHK(launch_phaseamp(d_pt, nullptr, M, nd, Tobs, std::sqrt(Tobs), d_amp, d_phs, s));
  HIPCHK(hipMemcpyAsync(&rec, d_pt, sizeof(Point), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(amp.data(), d_amp, sizeof(double) * nd * M, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (rec.status != ST_OK) return QCBC_EDOM;
  const int jmin = first_bin(rec.fs, c->desc.fmin, Tobs);
  // running network SNR^2 from jmin, and its density per bin
  std::vector<double> dens(M, 0.0), cum(M, 0.0);
  double run = 0.0;
  for (int j = std::max(jmin, 0); j < M; j++) {
    for (int d = 0; d < nd; d++) {
      const double a = amp[(size_t)d * M + j];
      const double x = 4.0 * (a * a) / SN[d][j];
      dens[j] += x;
      run += x;
    }
    cum[j] = run;
  }
  const double total = cum[M - 1];
  stst[0] = jmin;
  stst[1] = M - 1;
  bool f0 = false, f1 = false;
  for (int j = std::max(jmin, 0); j < M; j++) {
    if (!f0 && cum[j] > 0.001) {
      f0 = true;
      stst[0] = j;
    }
    if (!f1 && total - cum[j] < 0.001) {
      f1 = true;
      stst[1] = j;
    }
  }
  if (stst[0] < jmin) stst[0] = jmin;
  if (ratio) {
    const double avg = total / (double)(stst[1] - stst[0]);  // the band's average SNR^2 per bin
    for (int j = 0; j < n2; j++) ratio[j] = 1.0;
    for (int j = 1; j < M; j++) {
      const double r = dens[j] / avg;
      if (r > 0.0) ratio[j] = 1.0 / r;
    }
  }
  return QCBC_OK;
}

// diff (QuickCBC.c:4913-4964) or, with h, het_diff (:4966-5053) of K
// displaced points against their start points, host memory
int distances_impl(qcbc_ctx* c, const double* d_SN, qcbc_het* h, const double* params, int B, const double* px,
                   const int* of, int K, double* z, hipStream_t s) {
  if (!c && !h) return QCBC_EINVAL;
  Distances E;
  E.s = s;
  E.nd = c ? c->desc.n_det : 1;
  E.n2 = c ? c->nhalf : 0;
  E.Tobs = c ? c->desc.Tobs : 0.0;
  E.fmin = c ? c->desc.fmin : 0.0;
  E.d_SN = d_SN;
  if (h) {
    E.het = true;
    int st = het_grid_view(h, &E.hv);
    if (st) return st;
    E.an = E.hv.an;
    E.nd = E.hv.nd;
  } else {
    E.an = c->an;
  }
  int st = E.init(params, B, K);
  if (st) return st;
  std::vector<double> p(px, px + (size_t)NP * K), zz;
  std::vector<int> o(of, of + K);
  st = E.eval(p, o, zz);
  if (st) return st;
  std::memcpy(z, zz.data(), sizeof(double) * K);
  return QCBC_OK;
}

// efix on B points: with h the heterodyne distance, else the full-data one
// (SN on the device)
int efix_impl(qcbc_ctx* c, const double* d_SN, qcbc_het* h, const double* params, int B, double* ej, const double* ev,
              const double* mn, const double* mx, double etamin, double zs, int* trials, hipStream_t s) {
  if (!c && !h) return QCBC_EINVAL;
  Distances E;
  E.s = s;
  E.nd = c ? c->desc.n_det : 1;
  E.n2 = c ? c->nhalf : 0;
  E.Tobs = c ? c->desc.Tobs : 0.0;
  E.fmin = c ? c->desc.fmin : 0.0;
  E.d_SN = d_SN;
  if (h) {
    E.het = true;
    int st = het_grid_view(h, &E.hv);
    if (st) return st;
    E.an = E.hv.an;
    E.nd = E.hv.nd;
  } else {
    E.an = c->an;
  }
  int st = E.init(params, B, B * NP);
  if (st) return st;
  return efix_run(E, params, ej, ev, mn, mx, etamin, zs, trials);
}

// efix (zs, heterodyne distance) of B points whose states, jump sizes and
// directions are on the device; d_ej scaled in place.  Synchronous on s.
int het_efix_device(qcbc_het* h, const double* d_x, int B, double* d_ej, const double* d_ev, const double* mn,
                    const double* mx, double etamin, double zs, hipStream_t s) {
  std::vector<double> x((size_t)NP * B), ej((size_t)NP * B), ev((size_t)NP * NP * B);
  HIPCHK(hipMemcpyAsync(x.data(), d_x, sizeof(double) * x.size(), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(ej.data(), d_ej, sizeof(double) * ej.size(), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(ev.data(), d_ev, sizeof(double) * ev.size(), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  int st = efix_impl(nullptr, nullptr, h, x.data(), B, ej.data(), ev.data(), mn, mx, etamin, zs, nullptr, s);
  if (st) return st;
  HIPCHK(hipMemcpyAsync(d_ej, ej.data(), sizeof(double) * ej.size(), hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  return QCBC_OK;
}

// het_space (QuickCBC.c:4570-4910) at params: SN rows on the host and on
// the device.  fgrid has room for cap entries.
int het_space_impl(qcbc_ctx* c, const double* const* SN, const double* d_SN, const double* params, const double* mn,
                   const double* mx, double etamin, int* MN_out, int* MM_out, int* M_out, int* NR_out, int* fgrid,
                   int cap, double* eval_out, double* evec_out, double* dpar_out, hipStream_t s) {
  constexpr int J = 2;
  const double tol = 0.003;
  const int nd = c->desc.n_det, n2 = c->nhalf;
  const double Tobs = c->desc.Tobs;
  // 1. the band and the SNR weighting
  int stst[2];
  std::vector<double> ratio(n2, 1.0);
  int st = limits_impl(c, SN, params, stst, ratio.data(), s);
  if (st) return st;
  const int MN = stst[0], MM = stst[1];
  if (MM < 1 || MN >= MM) return QCBC_EDOM;
  // 2. Fisher_All, FisherEvec, efix to zs = 30 on the full data
  double ej[NP], ev[NP * NP], dpar[NP * NP];
  {
    Scratch mem;
    double *d_par, *d_fish, *d_ej, *d_ev;
    int* d_fst;
    if ((st = mem.alloc(NP, &d_par)) || (st = mem.alloc(NP * NP, &d_fish)) || (st = mem.alloc(NP, &d_ej)) ||
        (st = mem.alloc(NP * NP, &d_ev)) || (st = mem.alloc(1, &d_fst)))
      return st;
    HIPCHK(hipMemcpyAsync(d_par, params, sizeof(double) * NP, hipMemcpyHostToDevice, s));
    st = fisher_all_enqueue(c, d_SN, d_par, 1, d_fish, d_fst, s, mem);
    if (st) return st;
    HIPCHK(launch_fisher_evec(d_fish, 1, NP, d_ej, d_ev, nullptr, s));
    HIPCHK(hipMemcpyAsync(ej, d_ej, sizeof(ej), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(ev, d_ev, sizeof(ev), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  st = efix_impl(c, d_SN, nullptr, params, 1, ej, ev, mn, mx, etamin, 30.0, nullptr, s);
  if (st) return st;
  if (eval_out) std::memcpy(eval_out, ej, sizeof(ej));
  if (evec_out) std::memcpy(evec_out, ev, sizeof(ev));
  for (int i = 0; i < NP; i++) displace(params, ev + NP * i, ej[i], mn, mx, etamin, dpar + NP * i);
  if (dpar_out) std::memcpy(dpar_out, dpar, sizeof(dpar));
  // 3. the reference point's and the displaced points' waveforms on [0, MM)
  std::vector<double> A((size_t)(NP + 1) * nd * MM), P((size_t)(NP + 1) * nd * MM);
  {
    Scratch mem;
    double *d_par, *d_amp, *d_phs;
    Point* d_pts;
    int* d_st;
    if ((st = mem.alloc((NP + 1) * NP, &d_par)) || (st = mem.alloc(NP + 1, &d_pts)) ||
        (st = mem.alloc(NP + 1, &d_st)) || (st = mem.alloc(A.size(), &d_amp)) || (st = mem.alloc(P.size(), &d_phs)))
      return st;
    std::vector<double> pp((size_t)(NP + 1) * NP);
    std::memcpy(pp.data(), params, sizeof(double) * NP);
    std::memcpy(pp.data() + NP, dpar, sizeof(dpar));
    std::vector<int> pst(NP + 1);
    HIPCHK(hipMemcpyAsync(d_par, pp.data(), sizeof(double) * pp.size(), hipMemcpyHostToDevice, s));
    HIPCHK(launch_setup(d_par, NP + 1, c->an, d_pts, d_st, s));
    for (int v = 0; v <= NP; v++)
      HIPCHK(launch_phaseamp(d_pts + v, nullptr, MM, nd, Tobs, std::sqrt(Tobs), d_amp + (size_t)v * nd * MM,
                             d_phs + (size_t)v * nd * MM, s));
    HIPCHK(hipMemcpyAsync(pst.data(), d_st, sizeof(int) * (NP + 1), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(A.data(), d_amp, sizeof(double) * A.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(P.data(), d_phs, sizeof(double) * P.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int v = 0; v <= NP; v++)
      if (pst[v] != ST_OK) return QCBC_EDOM;
  }
  // ratio terms per displaced point and detector: Cs = 1 - (A_i/A) cos dphi,
  // Sn = (A_i/A) sin dphi, DH = Cs^2 + Sn^2 ([i][d][j], j in [MN, MM))
  const size_t plane = (size_t)nd * MM;
  std::vector<double> Cs((size_t)NP * plane), Sn((size_t)NP * plane), DH((size_t)NP * plane);
  for (int i = 0; i < NP; i++)
    for (int d = 0; d < nd; d++)
      for (int j = MN; j < MM; j++) {
        const size_t o = (size_t)i * plane + (size_t)d * MM + j;
        const size_t r = (size_t)d * MM + j, q = (size_t)(i + 1) * plane + r;
        const double rr = A[q] / A[r], dphi = P[r] - P[q];
        Cs[o] = 1.0 - rr * std::cos(dphi);
        Sn[o] = rr * std::sin(dphi);
        DH[o] = Cs[o] * Cs[o] + Sn[o] * Sn[o];
      }
  // 4. spans on which every ratio term is linear to the tolerance
  std::vector<int> marks(1, MN);
  {
    int jmin = MN;
    const int djmax = (int)(100.0 * Tobs);
    for (int j = MN; j < MM; j++) {
      bool flag = false;
      double tl = tol * ratio[j];
      if (tl < tol) tl = tol;
      const int jmid = (j + jmin) / 2;
      // at j = jmin the line through the ends is 0/0, NaN, and no test fires
      const double wa = (double)(j - jmid), wb = (double)(jmid - jmin), wd = (double)(j - jmin);
      for (int d = 0; d < nd && !flag; d++)
        for (int i = 0; i < NP && !flag; i++) {
          const size_t o = (size_t)i * plane + (size_t)d * MM;
          const double cm = (Cs[o + jmin] * wa + Cs[o + j] * wb) / wd;
          const double sm = (Sn[o + jmin] * wa + Sn[o + j] * wb) / wd;
          const double dm = (DH[o + jmin] * wa + DH[o + j] * wb) / wd;
          if (std::fabs(Cs[o + jmid] - cm) > tl || std::fabs(Sn[o + jmid] - sm) > tl ||
              std::fabs(DH[o + jmid] - dm) > tl)
            flag = true;
        }
      if ((j - jmin) > djmax) flag = true;
      if (flag) {
        jmin = j;
        marks.push_back(jmin);
      }
    }
  }
  // 5. every bin's span length, then the flat grid: J equal steps per
  // region, each the shortest span the region crosses (a crossing past N/2
  // ends the search, as in the reference)
  std::vector<int> span(n2, 100);
  const int count = (int)marks.size() - 1;
  for (int ii = 0; ii < count; ii++)
    for (int i = marks[ii]; i < marks[ii + 1]; i++) span[i] = marks[ii + 1] - marks[ii];
  for (int i = marks[count]; i < MM; i++) span[i] = MM - marks[count];
  std::vector<int> g(1, MN);
  {
    int j = MN, k = MN;
    do {
      int step = span[j];
      bool again;
      do {
        again = false;
        k = j + step * J;
        for (int i = j; i <= k; i++) {
          if (i < n2) {
            if (span[i] < step) {
              step = span[i];
              again = true;
            }
          } else {
            again = false;
          }
        }
      } while (again);
      j = k;
      for (int i = 0; i < J; i++) g.push_back(g.back() + step);
      if ((int)g.size() > n2) return QCBC_EINVAL;
    } while (j < MM);
    // the last region refitted to end at MM
    const int last = (int)g.size() - 1;
    const int step = (MM - g[last - J]) / J;
    if (step >= 1)
      for (int i = 0; i < J; i++) g[last - J + 1 + i] = g[last - J + i] + step;
  }
  const int M = (int)g.size();
  if (M > cap) return QCBC_ERANGE;
  std::memcpy(fgrid, g.data(), sizeof(int) * M);
  *MN_out = MN;
  *MM_out = MM;
  *M_out = M;
  *NR_out = (M - 1) / J;
  return QCBC_OK;
}

// Fisher_All of B points (host memory): fish B x 11 x 11, status B (may be null)
int fisher_all_impl(qcbc_ctx* c, const double* d_SN, const double* params, int B, double* fish, int* status,
                    hipStream_t s) {
  Scratch mem;
  double *d_par, *d_fish;
  int* d_st;
  int st;
  if ((st = mem.alloc((size_t)B * NP, &d_par)) || (st = mem.alloc((size_t)B * NP * NP, &d_fish)) ||
      (st = mem.alloc((size_t)B, &d_st)))
    return st;
  HIPCHK(hipMemcpyAsync(d_par, params, sizeof(double) * NP * B, hipMemcpyHostToDevice, s));
  st = fisher_all_enqueue(c, d_SN, d_par, B, d_fish, d_st, s, mem);
  if (st) return st;
  HIPCHK(hipMemcpyAsync(fish, d_fish, sizeof(double) * NP * NP * B, hipMemcpyDeviceToHost, s));
  if (status) HIPCHK(hipMemcpyAsync(status, d_st, sizeof(int) * B, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return QCBC_OK;
}

// Fisher_Full (QuickCBC.c:2551-2688) of B reference-frame points (host
// memory, 7 + 3 (n_det - 1) each) on the band bins of c, SN on the device:
// fish B x d x d (the top-left d x d block of the seven-parameter matrix),
// status B (may be null)
int fisher_full_impl(qcbc_ctx* c, const double* d_SN, const double* params, int B, int d, double* fish, int* status,
                     hipStream_t s) {
  const int nd = c->desc.n_det, ips = 7 + 3 * (nd - 1), nb = c->nb, imin = c->an.imin;
  if (nb < 1 || d < 1 || d > FISHER_FULL_NP) return QCBC_EINVAL;
  const int splits = std::max(1, std::min((nb + 1023) / 1024, (2048 + B - 1) / B));
  Scratch mem;
  double *d_par, *q, *partial, *d_fish;
  Point* pts;
  int *pst, *d_st;
  int st;
  if ((st = mem.alloc((size_t)B * ips, &d_par)) || (st = mem.alloc((size_t)B * FISHER_FULL_NV * FISHER_FULL_NP, &q)) ||
      (st = mem.alloc((size_t)B * FISHER_FULL_NV, &pts)) || (st = mem.alloc((size_t)B * FISHER_FULL_NV, &pst)) ||
      (st = mem.alloc((size_t)B * splits * FISHER_FULL_NPAIR, &partial)) ||
      (st = mem.alloc((size_t)B * FISHER_FULL_NP * FISHER_FULL_NP, &d_fish)) || (st = mem.alloc((size_t)B, &d_st)))
    return st;
  HIPCHK(hipMemcpyAsync(d_par, params, sizeof(double) * ips * B, hipMemcpyHostToDevice, s));
  HIPCHK(launch_fisher_full_params(d_par, ips, B, q, s));
  // AmpPhase's model: phi0 = params[4] itself, as templates() hands it over
  HIPCHK(launch_setup_intrinsic(q, FISHER_FULL_NP, B * FISHER_FULL_NV, c->an, pts, pst, s));
  HIPCHK(launch_fisher_full(pts, q, d_par, ips, B, nd, d_SN, c->nhalf, imin, nb, c->desc.Tobs, splits, partial, d_fish,
                            d_st, s));
  std::vector<double> F((size_t)B * FISHER_FULL_NP * FISHER_FULL_NP);
  HIPCHK(hipMemcpyAsync(F.data(), d_fish, sizeof(double) * F.size(), hipMemcpyDeviceToHost, s));
  if (status) HIPCHK(hipMemcpyAsync(status, d_st, sizeof(int) * B, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int b = 0; b < B; b++)
    for (int i = 0; i < d; i++)
      for (int j = 0; j < d; j++)
        fish[((size_t)b * d + i) * d + j] = F[((size_t)b * FISHER_FULL_NP + i) * FISHER_FULL_NP + j];
  return QCBC_OK;
}

}  // namespace qcbc

extern "C" {

int qcbc_fisher_full(qcbc_ctx* c, const double* params, int B, int d, double* fish, int* status) {
  if (!c || B < 0 || (B > 0 && (!params || !fish)) || d < 1 || d > FISHER_FULL_NP) return QCBC_EINVAL;
  if (B == 0) return QCBC_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  if (!c->bound) return QCBC_ENOTBOUND;
  hipSetDevice(c->device);
  return fisher_full_impl(c, c->d_SN, params, B, d, fish, status, c->stream);
}

int qcbc_fisher_all(qcbc_ctx* c, const double* params, int B, double* fish, int* status) {
  if (!c || B < 0 || (B > 0 && (!params || !fish))) return QCBC_EINVAL;
  if (B == 0) return QCBC_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  if (!c->bound) return QCBC_ENOTBOUND;
  hipSetDevice(c->device);
  return fisher_all_impl(c, c->d_SN, params, B, fish, status, c->stream);
}

static int host_psd(qcbc_ctx* c, std::vector<double>& SNh, std::vector<const double*>& rows) {
  const int nd = c->desc.n_det, n2 = c->nhalf;
  SNh.resize((size_t)nd * n2);
  HIPCHK(hipMemcpyAsync(SNh.data(), c->d_SN, sizeof(double) * SNh.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  rows.resize(nd);
  for (int d = 0; d < nd; d++) rows[d] = SNh.data() + (size_t)d * n2;
  return QCBC_OK;
}

int qcbc_limits(qcbc_ctx* c, const double* params, int* band, double* ratio) {
  if (!c || !params || !band) return QCBC_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  if (!c->bound) return QCBC_ENOTBOUND;
  hipSetDevice(c->device);
  std::vector<double> SNh;
  std::vector<const double*> rows;
  int st = host_psd(c, SNh, rows);
  if (st) return st;
  return limits_impl(c, rows.data(), params, band, ratio, c->stream);
}

int qcbc_diff(qcbc_ctx* c, const double* params, int B, const double* px, const int* of, int K, double* z) {
  if (!c || B < 1 || K < 0 || !params || (K > 0 && (!px || !of || !z))) return QCBC_EINVAL;
  for (int k = 0; k < K; k++)
    if (of[k] < 0 || of[k] >= B) return QCBC_EINVAL;
  if (K == 0) return QCBC_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  if (!c->bound) return QCBC_ENOTBOUND;
  hipSetDevice(c->device);
  return distances_impl(c, c->d_SN, nullptr, params, B, px, of, K, z, c->stream);
}

int qcbc_efix(qcbc_ctx* c, const double* params, int B, double* ej, const double* ev, const double* pmin,
              const double* pmax, double etamin, double zs, int* trials) {
  if (!c || B < 0 || (B > 0 && (!params || !ej || !ev)) || !(zs > 0.0)) return QCBC_EINVAL;
  int st = check_box(pmin, pmax, etamin);
  if (st) return st;
  if (B == 0) return QCBC_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  if (!c->bound) return QCBC_ENOTBOUND;
  hipSetDevice(c->device);
  return efix_impl(c, c->d_SN, nullptr, params, B, ej, ev, pmin, pmax, etamin, zs, trials, c->stream);
}

int qcbc_het_space(qcbc_ctx* c, const double* params, const double* pmin, const double* pmax, double etamin,
                   int* MN, int* MM, int* M, int* NR, int* fgrid, int cap, double* eval, double* evec,
                   double* dparams) {
  if (!c || !params || !MN || !MM || !M || !NR || !fgrid || cap < 3) return QCBC_EINVAL;
  int st = check_box(pmin, pmax, etamin);
  if (st) return st;
  std::lock_guard<std::mutex> lk(c->mu);
  if (!c->bound) return QCBC_ENOTBOUND;
  hipSetDevice(c->device);
  std::vector<double> SNh;
  std::vector<const double*> rows;
  st = host_psd(c, SNh, rows);
  if (st) return st;
  return het_space_impl(c, rows.data(), c->d_SN, params, pmin, pmax, etamin, MN, MM, M, NR, fgrid, cap, eval, evec,
                        dparams, c->stream);
}

int qcbc_het_diff(qcbc_het* h, const double* params, int B, const double* px, const int* of, int K, double* z) {
  if (!h || B < 1 || K < 0 || !params || (K > 0 && (!px || !of || !z))) return QCBC_EINVAL;
  for (int k = 0; k < K; k++)
    if (of[k] < 0 || of[k] >= B) return QCBC_EINVAL;
  if (K == 0) return QCBC_OK;
  HetGridView hv;
  int st = het_grid_view(h, &hv);
  if (st) return st;
  hipSetDevice(hv.device);
  hipStream_t s = nullptr;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  Distances E;
  E.s = s;
  E.het = true;
  E.hv = hv;
  E.an = hv.an;
  E.nd = hv.nd;
  st = E.init(params, B, K);
  std::vector<double> p(px, px + (size_t)NP * K), zz;
  std::vector<int> o(of, of + K);
  if (!st) st = E.eval(p, o, zz);
  if (!st) std::memcpy(z, zz.data(), sizeof(double) * K);
  (void)hipStreamSynchronize(s);
  (void)hipStreamDestroy(s);
  return st;
}

int qcbc_het_efix(qcbc_het* h, const double* params, int B, double* ej, const double* ev, const double* pmin,
                  const double* pmax, double etamin, double zs, int* trials) {
  if (!h || B < 0 || (B > 0 && (!params || !ej || !ev)) || !(zs > 0.0)) return QCBC_EINVAL;
  int st = check_box(pmin, pmax, etamin);
  if (st) return st;
  if (B == 0) return QCBC_OK;
  HetGridView hv;
  st = het_grid_view(h, &hv);
  if (st) return st;
  hipSetDevice(hv.device);
  hipStream_t s = nullptr;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  st = efix_impl(nullptr, nullptr, h, params, B, ej, ev, pmin, pmax, etamin, zs, trials, s);
  (void)hipStreamSynchronize(s);
  (void)hipStreamDestroy(s);
  return st;
}

}  // extern "C"
