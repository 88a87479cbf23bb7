// HIP kernels of QuickCBC's search likelihoods (QuickCBC.c:5672-6408,
// :3448-3525; Utilities.c:687-713), compiled with -ffp-contract=off.
//
//   k_fbegin_at     fbegin of each point with tc replaced (PDwave's mask,
//                   the penalized likelihood's band start)
//   k_pdwave        PDwave's template per point and bin, its arithmetic
//   k_xspec         pbt_shift's cross spectra D conj(h) / S per detector as
//                   complex rows for one backward FFT each, and (h|h)
//   k_peaks         per (point, detector): logL over the allowed lags from
//                   the transformed rows and its peaks, in descending order,
//                   more than dtp apart, above Lpmin (the first always)
//   k_penalized     log_likelihood_penalized per (point, detector, peak):
//                   band sums (one band per lane, its bins in order), the
//                   band selection on one lane, the amplitude and phase
//                   maximised over the kept bands
#include "qcbc_device.h"
#include "qcbc_het_device.h"
#include "qcbc_search_internal.h"

namespace qcbc {

namespace {

constexpr int PEN_THREADS = 256;
constexpr int PEN_MAX_BANDS = 1024;
constexpr double RHOCUT = 4.0;  // ConstCBC.h:27

// fbegin (Utilities.c:587-629) of each point's reference-frame parameters
// (its first seven), tc replaced by tc_at unless that is NaN
__global__ void k_fbegin_at(const double* __restrict__ params, int ips, int B, double tc_at, double* __restrict__ fs) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double prm[11];
  for (int i = 0; i < 11; i++) prm[i] = i < 7 ? params[(size_t)ips * b + i] : 0.0;
  if (tc_at == tc_at) prm[5] = tc_at;
  fs[b] = fbegin(prm);
}

// PDwave (QuickCBC.c:3448-3525) for B points set up in the reference frame
// (phi0 = params[4]): h[b] half-complex N, zero at f <= fs[b]
__global__ void k_pdwave(const Point* __restrict__ pts, const double* __restrict__ fs, int N, double Tobs,
                         double sqT, const double* __restrict__ ts, double* __restrict__ h) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  const int nh = N / 2;
  if (i >= nh) return;
  double* w = h + (size_t)b * N;
  if (i == 0) {
    w[0] = 0.0;
    w[nh] = 0.0;
    return;
  }
  const Point& c = pts[b];
  const double f = (double)i / Tobs;
  double re = 0.0, im = 0.0;
  if (c.status == ST_OK && f > fs[b]) {
    double a, ph;
    h22_at(c, f, &a, &ph);
    const double p = TPI * f * ts[b], cp = cos(p), sp = sin(p);
    const double x = cos(ph), y = sin(ph), Amp = H22FAC * a / sqT;
    re = Amp * (x * cp + y * sp);
    im = Amp * (x * sp - y * cp);
  }
  w[i] = re;
  w[N - i] = im;
}

// The cross spectra C_k = D_k conj(h_k) / S_k, k in [imin, imax), as
// complex rows z[b][d] of length N (zero elsewhere), and the partial sums of
// (h|h) = 4 sum |h|^2 / S (fourier_nwip) per (point, detector, block)
__global__ void k_xspec(const double* __restrict__ h, const double* __restrict__ D, const double* __restrict__ SN,
                        int nd, int N, int imin, int imax, double2* __restrict__ z, double* __restrict__ hh_part) {
  __shared__ double red[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y, d = blockIdx.z;
  const int nh = N / 2;
  const double* w = h + (size_t)b * N;
  double2* row = z + ((size_t)b * nd + d) * N;
  double hh = 0.0;
  if (i < N) {
    double2 v = make_double2(0.0, 0.0);
    if (i >= imin && i < imax) {
      const int k = N - i;
      const double* Dd = D + (size_t)d * N;
      const double S = SN[(size_t)d * nh + i];
      v.x = (Dd[i] * w[i] + Dd[k] * w[k]) / S;
      v.y = (Dd[k] * w[i] - Dd[i] * w[k]) / S;
      hh = (w[i] * w[i] + w[k] * w[k]) / S;
    }
    row[i] = v;
  }
  hh = het_wave_sum(hh);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = hh;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) s += red[q];
    hh_part[((size_t)b * nd + d) * gridDim.x + blockIdx.x] = s;
  }
}

__global__ void k_hh_reduce(const double* __restrict__ hh_part, int nparts, int n, double* __restrict__ hh) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double s = 0.0;
  for (int q = 0; q < nparts; q++) s += hh_part[(size_t)r * nparts + q];
  hh[r] = 4.0 * s;
}

// One block per (point, detector): the lag indices [mn, mx) of the backward
// transform z (lag (i - N/2) dt at z[i - N/2] for i >= N/2, else
// z[i + N/2]), logL = (4|z|)^2 / (h|h) / 2 (the reference's (2 N |corr +
// i corrf|)^2 / HH / 2), and its peaks: repeatedly the largest lag not
// within dtp of a kept one (lowest index on a tie), kept if above Lpmin or
// if it is the first, until none is left; at most maxp.  Kept lags' logL,
// time, phase atan2(Im z, Re z) and scale 4|z| / (h|h).
__global__ __launch_bounds__(256) void k_peaks(const double2* __restrict__ z, const double* __restrict__ hh, int nd,
                                               int N, double dt, int mn, int mx, double lpmin, double dtp, int maxp,
                                               unsigned char* __restrict__ mask, int* __restrict__ npk,
                                               double* __restrict__ pk) {
  __shared__ double bv[4];
  __shared__ int bi[4];
  __shared__ int s_best;
  __shared__ double s_bestv;
  const int r = blockIdx.x;  // (point, detector)
  const int M = mx - mn;
  const double2* row = z + (size_t)r * N;
  unsigned char* m = mask + (size_t)r * M;
  const double H = hh[r];
  for (int j = threadIdx.x; j < M; j += blockDim.x) m[j] = 0;
  __syncthreads();
  int n = 0;
  for (;;) {
    double best = -1.0;
    int bidx = 0x7fffffff;
    for (int j = threadIdx.x; j < M; j += blockDim.x) {
      if (m[j]) continue;
      const int i = mn + j;
      const int src = (i >= N / 2) ? i - N / 2 : i + N / 2;
      const double2 v = row[src];
      const double HD = 4.0 * sqrt(v.x * v.x + v.y * v.y);
      const double L = (HD * HD / H) / 2.0;
      if (L > best || (L == best && j < bidx)) {
        best = L;
        bidx = j;
      }
    }
    // block argmax (largest logL, lowest index on a tie)
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(best, o, 64);
      const int oi = __shfl_xor(bidx, o, 64);
      if (ov > best || (ov == best && oi < bidx)) {
        best = ov;
        bidx = oi;
      }
    }
    if ((threadIdx.x & 63) == 0) {
      bv[threadIdx.x >> 6] = best;
      bi[threadIdx.x >> 6] = bidx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double v = bv[0];
      int k = bi[0];
      for (int q = 1; q < (int)(blockDim.x >> 6); q++)
        if (bv[q] > v || (bv[q] == v && bi[q] < k)) {
          v = bv[q];
          k = bi[q];
        }
      s_best = (k != 0x7fffffff && (n == 0 || v > lpmin) && n < maxp) ? k : -1;
      s_bestv = v;
    }
    __syncthreads();
    const int k = s_best;
    if (k < 0) break;  // block-uniform
    const double tk = dt * (double)(mn + k - N / 2);
    if (threadIdx.x == 0) {
      const int i = mn + k;
      const int src = (i >= N / 2) ? i - N / 2 : i + N / 2;
      const double2 v = row[src];
      const double HD = 4.0 * sqrt(v.x * v.x + v.y * v.y);
      double* p = pk + ((size_t)r * maxp + n) * 4;
      p[0] = s_bestv;
      p[1] = tk;
      p[2] = atan2(v.y, v.x);
      p[3] = HD / H;
    }
    // suppress every lag within dtp of the new peak (and the peak itself)
    const int w = (int)(dtp / dt) + 2;
    for (int j = k - w + (int)threadIdx.x; j <= k + w; j += blockDim.x) {
      if (j < 0 || j >= M) continue;
      const double tj = dt * (double)(mn + j - N / 2);
      if (j == k || fabs(tj - tk) < dtp) m[j] = 1;
    }
    n++;
    __syncthreads();
  }
  if (threadIdx.x == 0) npk[r] = n;
}

// log_likelihood_penalized (QuickCBC.c:6233-6408) of one template per block.
// Job q: point b = job[q].b, detector d = job[q].d; its template is the
// half-complex row tw (hw mode) or nm h e^{i (psh - 2 pi f tsh)} from the
// point's PDwave row (peak mode, QuickCBC.c:5968-5981).  Band sums one band
// per lane, bins in order; lane 0 chooses the bands.
__global__ __launch_bounds__(PEN_THREADS) void k_penalized(const PenJob* __restrict__ jobs, int nq,
                                                           const double* __restrict__ hw, const double* __restrict__ h,
                                                           const double* __restrict__ D,
                                                           const double* __restrict__ SN, const int* __restrict__ imin,
                                                           int nd, int N, double Tobs, int imax, int bw,
                                                           double* __restrict__ out) {
  __shared__ double sx[PEN_MAX_BANDS], ss[PEN_MAX_BANDS], sy[PEN_MAX_BANDS];
  const int q = blockIdx.x;
  if (q >= nq) return;
  const PenJob J = jobs[q];
  const int lo = imin[J.b];
  int nb = (imax - lo) / bw;
  if (nb < 1) nb = 1;
  const bool peak = J.mode != 0;
  const double* w = peak ? h + (size_t)J.b * N : hw + ((size_t)J.b * nd + J.d) * N;
  const double* Dd = D + (size_t)J.d * N;
  const double* S = SN + (size_t)J.d * (N / 2);
  // a band's sums count only if the scan leaves it (bands 0 .. nb-1 whose
  // end lies before imax); the last band never closes
  for (int k = threadIdx.x; k < nb && k < PEN_MAX_BANDS; k += blockDim.x) {
    const int b0 = lo + k * bw, b1 = b0 + bw;
    double x = 0.0, xs = 0.0, y = 0.0;
    if (b1 < imax) {
      for (int i = b0; i < b1; i++) {
        double tr, ti;
        if (peak) {
          const double f = (double)(i) / Tobs, ph = J.psh - TPI * f * J.tsh, cph = cos(ph), sph = sin(ph);
          tr = J.nm * (w[i] * cph - w[N - i] * sph);
          ti = J.nm * (w[N - i] * cph + w[i] * sph);
        } else {
          tr = w[i];
          ti = w[N - i];
        }
        x += 4.0 * (Dd[i] * tr + Dd[N - i] * ti) / S[i];
        xs += 4.0 * (Dd[i] * ti - Dd[N - i] * tr) / S[i];
        y += 4.0 * (tr * tr + ti * ti) / S[i];
      }
    }
    sx[k] = x;
    ss[k] = xs;
    sy[k] = y;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int n = nb < PEN_MAX_BANDS ? nb : PEN_MAX_BANDS;
    int kstart = 0, kend = n;
    for (int k = 0; k < n; k++)
      if (sy[k] > 0.1) {
        kstart = k;
        break;
      }
    for (int k = kstart; k < n; k++)
      if (sy[k] < 0.1) {
        kend = k;
        break;
      }
    double HH = 0.0001, HC = 0.0, HS = 0.0;
    for (int k = kstart; k < kend; k++) {
      const double rho = sy[k] > 0.0 ? (sqrt(sx[k] * sx[k] + ss[k] * ss[k]) - sy[k]) / sqrt(sy[k]) : 0.0;
      if (rho > RHOCUT) continue;
      HH += sy[k];
      HC += sx[k];
      HS += ss[k];
    }
    const double HD = sqrt(HC * HC + HS * HS);
    out[(size_t)q * 3 + 0] = (HD * HD / HH) / 2.0;
    out[(size_t)q * 3 + 1] = HD / HH;
    out[(size_t)q * 3 + 2] = atan2(HS, HC);
  }
}

}  // namespace

hipError_t launch_fbegin_at(const double* params, int ips, int B, double tc_at, double* fs, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fbegin_at, dim3((B + 63) / 64), dim3(64), 0, s, params, ips, B, tc_at, fs);
  return hipGetLastError();
}

hipError_t launch_pdwave(const Point* pts, const double* fs, int B, int N, double Tobs, const double* ts, double* h,
                         hipStream_t s) {
  if (B <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pdwave, dim3((N / 2 + 255) / 256, B), dim3(256), 0, s, pts, fs, N, Tobs, sqrt(Tobs), ts, h);
  return hipGetLastError();
}

int xspec_blocks(int N) { return (N + 255) / 256; }

hipError_t launch_xspec(const double* h, const double* D, const double* SN, int B, int nd, int N, int imin, int imax,
                        void* z, double* hh_part, double* hh, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  const int nbk = xspec_blocks(N);
  hipLaunchKernelGGL(k_xspec, dim3(nbk, B, nd), dim3(256), 0, s, h, D, SN, nd, N, imin, imax,
                     static_cast<double2*>(z), hh_part);
  hipLaunchKernelGGL(k_hh_reduce, dim3((B * nd + 255) / 256), dim3(256), 0, s, hh_part, nbk, B * nd, hh);
  return hipGetLastError();
}

hipError_t launch_peaks(const void* z, const double* hh, int rows, int N, double dt, int mn, int mx, double lpmin,
                        double dtp, int maxp, unsigned char* mask, int* npk, double* pk, hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_peaks, dim3(rows), dim3(256), 0, s, static_cast<const double2*>(z), hh, 0, N, dt, mn, mx,
                     lpmin, dtp, maxp, mask, npk, pk);
  return hipGetLastError();
}

int penalized_max_bands() { return PEN_MAX_BANDS; }

hipError_t launch_penalized(const PenJob* jobs, int nq, const double* hw, const double* h, const double* D,
                            const double* SN, const int* imin, int nd, int N, double Tobs, int imax, int bw,
                            double* out, hipStream_t s) {
  if (nq <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_penalized, dim3(nq), dim3(PEN_THREADS), 0, s, jobs, nq, hw, h, D, SN, imin, nd, N, Tobs, imax,
                     bw, out);
  return hipGetLastError();
}

}  // namespace qcbc
