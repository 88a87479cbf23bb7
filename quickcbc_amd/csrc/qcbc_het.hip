// HIP kernels of the heterodyned likelihood (QuickCBC.c:3780-3869,
// 5055-5444), compiled with -ffp-contract=off like the parity path: the
// set-up (heterodyne) and the evaluation (log_likelihood_het) perform the
// reference's double arithmetic operation for operation, except where a sum
// is split over lanes (set-up moments and logLb) or regrouped (the
// evaluation's per-sample weights, below).
//
// Layouts, row-major, nd detectors, grid of M = NR J + 1 samples:
//   amp, phase (reference on the grid)      nd x M
//   IP                                      NR x (J+1) x (J+1)
//   SL, lc, ls                              NR x nd x (J+1)
//   rc, rs, aa                              nd x NR
//   w (evaluation weights)                  3 x nd x M: wH, wC, wS
//   wf (the same over a_ref^2, a_ref, a_ref) 3 x nd x M
//   pl (sample powers)                       4 x M: f^(1/6), f^(-1/6), 1/f, log f
#include "qcbc_device.h"
#include "qcbc_het_device.h"
#include "qcbc_het_internal.h"

namespace qcbc {

namespace {

// legendre_maker's recurrence (QuickCBC.c:5066-5088) at one point k of 0..U
__device__ inline void legendre_at(int J, int U, int k, double* P) {
  double c0 = 0.0, c1 = (double)(U - 2 * k), c2 = (double)(U);
  double d0 = c2, d1 = c1 + c1, d2 = d0;
  P[0] = 1.0;
  P[1] = c1 / (double)(U);
  for (int i = 1; i < J; ++i) {
    d0 += 2.0;
    d2 -= 2.0;
    c0 += d0;
    c1 += d1;
    c2 += d2;
    P[i + 1] = (c1 * P[i] - c0 * P[i - 1]) / c2;
  }
}

// Inverse (Utilities.c:1862-1903): IM = transpose(inverse(A)), d x d, by
// elimination with row pivoting on the largest |a_ij| and substitution of the
// permuted unit vectors (the method of GSL's LU decomposition and inversion)
__device__ inline void inverse_t(const double* A, double* IM, int d) {
  double a[HET_MAX_J1 * HET_MAX_J1], x[HET_MAX_J1];
  int perm[HET_MAX_J1];
  for (int i = 0; i < d * d; i++) a[i] = A[i];
  for (int i = 0; i < d; i++) perm[i] = i;
  for (int j = 0; j < d - 1; j++) {
    int ip = j;
    double amax = fabs(a[j * d + j]);
    for (int i = j + 1; i < d; i++)
      if (fabs(a[i * d + j]) > amax) {
        amax = fabs(a[i * d + j]);
        ip = i;
      }
    if (ip != j) {
      for (int k = 0; k < d; k++) {
        double t = a[j * d + k];
        a[j * d + k] = a[ip * d + k];
        a[ip * d + k] = t;
      }
      int t = perm[j];
      perm[j] = perm[ip];
      perm[ip] = t;
    }
    if (a[j * d + j] != 0.0)
      for (int i = j + 1; i < d; i++) {
        double l = a[i * d + j] / a[j * d + j];
        a[i * d + j] = l;
        for (int k = j + 1; k < d; k++) a[i * d + k] -= l * a[j * d + k];
      }
  }
  for (int c = 0; c < d; c++) {
    for (int i = 0; i < d; i++) x[i] = (perm[i] == c) ? 1.0 : 0.0;
    for (int i = 1; i < d; i++)
      for (int k = 0; k < i; k++) x[i] -= a[i * d + k] * x[k];
    for (int i = d - 1; i >= 0; i--) {
      for (int k = i + 1; k < d; k++) x[i] -= a[i * d + k] * x[k];
      x[i] /= a[i * d + i];
    }
    for (int i = 0; i < d; i++) IM[c * d + i] = x[i];  // IM[c][i] = inv[i][c]
  }
}

}  // namespace

// fullphaseamp (QuickCBC.c:3780-3869) for point 0 on a frequency list (freq
// null: the analysis grid f_i = i/Tobs, f_0 = 1/Tobs): amp and phase rows of n
__global__ void k_phaseamp(const Point* __restrict__ pts, const double* __restrict__ freq, int n, int nd,
                           double Tobs, double sqT, double* __restrict__ amp, double* __restrict__ phs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Point& c = pts[0];
  const double f = freq ? freq[i] : ((i == 0) ? 1.0 / Tobs : (double)i / Tobs);
  double a, p;
  h22_at(c, f, &a, &p);
  const double y = taper(f, c.fs);
  for (int d = 0; d < nd; d++) {
    const double A = c.Fs[d] * H22FAC * a / sqT;
    const double x = TPI * f * c.tshift[d] + c.lambda[d] - p;
    amp[(size_t)d * n + i] = A * y;
    phs[(size_t)d * n + i] = x;
  }
}

// heterodyne's per-bin work (QuickCBC.c:5150-5200) at the reference point,
// from the h22 amplitude and phase of every bin (k_waveform): the template
// hb (fulltemplates), residual rb = D - hb and, with fullphaseamp's amplitude
// and phase of the bin, the products RC, RS, AA; and per block the partial
// sums of (h|h) and (r|h) over [MN, MM) (fourier_nwip).  Planes nd x N/2:
// out = [RC | RS | AA].
__global__ void k_het_bins(const Point* __restrict__ pts, const double* __restrict__ wa, const double* __restrict__ wp,
                           const double* __restrict__ D, const double* __restrict__ SN, int nd, int N, double Tobs,
                           double sqT, int MN, int MM, double* __restrict__ out, double* __restrict__ partial) {
  __shared__ double red[4][6];
  const Point& c = pts[0];
  const int n2 = N / 2;
  const size_t P = (size_t)nd * n2;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n2; j += gridDim.x * blockDim.x) {
    const double f = (j == 0) ? 1.0 / Tobs : (double)j / Tobs;
    const double y = taper(f, c.fs);
    for (int d = 0; d < nd; d++) {
      const size_t o = (size_t)d * n2 + j;
      const double A = c.Fs[d] * H22FAC * wa[j] / sqT;
      const double x = TPI * f * c.tshift[d] + c.lambda[d] - wp[j];
      const double cp = cos(x), sp = sin(x);
      if (j == 0) {  // RC[0] = RS[0] = 0 (:5188-5189); AA[0] is never read
        out[o] = out[P + o] = out[2 * P + o] = 0.0;
        continue;
      }
      double hr = 0.0, hi = 0.0;
      if (f > c.fs) {
        hr = A * cp;
        hi = A * sp;
      }
      const double rr = D[(size_t)d * N + j] - hr, ri = D[(size_t)d * N + (N - j)] - hi;
      const double S = SN[o];
      const double ab = A * y;
      const double xs = ab / S;
      out[o] = 2.0 * (rr * cp + ri * sp) * xs;
      out[P + o] = 2.0 * (-rr * sp + ri * cp) * xs;
      out[2 * P + o] = ab * ab / S;
      if (j >= MN && j < MM) {
        acc[2 * d] += (hr * hr + hi * hi) / S;
        acc[2 * d + 1] += (rr * hr + ri * hi) / S;
      }
    }
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int q = 0; q < 2 * nd; q++) {
    const double v = het_wave_sum(acc[q]);
    if (lane == 0) red[wave][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < 2 * nd) {
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += red[w][threadIdx.x];
    partial[(size_t)blockIdx.x * (2 * nd) + threadIdx.x] = s;
  }
}

// logLb = sum_d [(r|h) + (h|h)/2] (QuickCBC.c:5170-5176); the block partials
// summed in block order, then scaled by fourier_nwip's 4
__global__ void k_het_logLb(const double* __restrict__ partial, int nblk, int nd, double* __restrict__ logLb) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double logL = 0.0;
  for (int d = 0; d < nd; d++) {
    double hh = 0.0, hr = 0.0;
    for (int b = 0; b < nblk; b++) {
      hh += partial[(size_t)b * (2 * nd) + 2 * d];
      hr += partial[(size_t)b * (2 * nd) + 2 * d + 1];
    }
    logL += 4.0 * hr + 0.5 * (4.0 * hh);
  }
  logLb[0] = logL;
}

// The grid frequencies fgrid / Tobs (het_space, QuickCBC.c:4901)
__global__ void k_het_freq(const int* __restrict__ fgrid, int M, double Tobs, double* __restrict__ freq) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  freq[m] = ((double)fgrid[m]) / Tobs;
}

// The reference amplitude and phase on the grid (:5154-5163).  The
// reference takes them from fullphaseamp of every bin; here they come from
// the evaluation kernel's own per-sample code, so that, as in the reference,
// the heterodyne ratios of the reference point are exactly 0.
__global__ void k_het_ref(const Point* __restrict__ pts, const double* __restrict__ freq,
                          const double* __restrict__ pl, int M, int nd, double* __restrict__ amp,
                          double* __restrict__ phs) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const Point& c = pts[0];
  const double f = freq[m];
  double a, p;
  h22_planes(c, 1.0 / c.Msec, f, pl, M, m, &a, &p);
  const double y = taper(f, c.fs);
  for (int d = 0; d < nd; d++) project(c, d, f, a, p, y, &amp[(size_t)d * M + m], &phs[(size_t)d * M + m]);
}

// The sample powers of the evaluation (4 x M)
__global__ void k_het_planes(const double* __restrict__ freq, int M, double* __restrict__ pl) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const double f = freq[m];
  pl[m] = pow(f, 1.0 / 6.0);
  pl[M + m] = pow(f, -1.0 / 6.0);
  pl[2 * M + m] = 1.0 / f;
  pl[3 * M + m] = log(f);
}

// One 64-lane block per region i (:5202-5290): the Legendre moments SL, lc,
// ls of AA, RC, RS over the region's bins, the transposed inverse of the
// downsampled Legendre matrix, and RC, RS, AA at the region start.
__global__ __launch_bounds__(64) void k_het_regions(const double* __restrict__ bins, const int* __restrict__ fgrid,
                                                    int J, int NR, int nd, int n2, double* __restrict__ IP,
                                                    double* __restrict__ SL, double* __restrict__ lc,
                                                    double* __restrict__ ls, double* __restrict__ rc,
                                                    double* __restrict__ rs, double* __restrict__ aa) {
  const int i = blockIdx.x, lane = threadIdx.x;
  if (i >= NR) return;
  const size_t P = (size_t)nd * n2;
  const int base = fgrid[i * J], U = fgrid[(i + 1) * J] - base;
  double Pn[HET_MAX_J1];
  for (int d = 0; d < nd; d++) {
    const double* RCd = bins + (size_t)d * n2;
    const double* RSd = bins + P + (size_t)d * n2;
    const double* AAd = bins + 2 * P + (size_t)d * n2;
    double s[HET_MAX_J1], l[HET_MAX_J1], t[HET_MAX_J1];
    for (int j = 0; j <= J; j++) s[j] = l[j] = t[j] = 0.0;
    for (int n = lane; n <= U; n += 64) {
      legendre_at(J, U, n, Pn);
      const int k = base + n;
      const double a = AAd[k], r = RCd[k], q = RSd[k];
      for (int j = 0; j <= J; j++) {
        s[j] += a * Pn[j];
        l[j] += r * Pn[j];
        t[j] += q * Pn[j];
      }
    }
    const size_t o = ((size_t)i * nd + d) * (J + 1);
    for (int j = 0; j <= J; j++) {
      const double vs = het_wave_sum(s[j]), vl = het_wave_sum(l[j]), vt = het_wave_sum(t[j]);
      if (lane == 0) {
        SL[o + j] = vs;
        lc[o + j] = vl;
        ls[o + j] = vt;
      }
    }
    if (lane == 0) {
      rc[(size_t)d * NR + i] = RCd[base];
      rs[(size_t)d * NR + i] = RSd[base];
      aa[(size_t)d * NR + i] = AAd[base];
    }
  }
  if (lane == 0) {
    const int J1 = J + 1;
    double PP[HET_MAX_J1 * HET_MAX_J1];
    for (int k = 0; k <= J; k++) {
      legendre_at(J, U, fgrid[i * J + k] - base, Pn);
      for (int j = 0; j <= J; j++) PP[j * J1 + k] = Pn[j];
    }
    inverse_t(PP, IP + (size_t)i * J1 * J1, J1);
  }
}

// The evaluation's weights.  log_likelihood_het (:5352-5420) is linear in the
// per-sample quantities cc, hc, hs: sample m = iJ + k of region i enters
// HH with sum_j IP[i][j][k] SL[i][d][j], and the boundary sample m = iJ of
// regions i = 1 .. NR-2 is taken off once with aa[d][i] (the overcount
// correction); likewise HR with lc, rc (hc) and ls, rs (hs).  A boundary
// sample belongs to two regions: the earlier region's term comes first.
// wf folds the ratios' 1/a_ref in: wH/a_ref^2, wC/a_ref, wS/a_ref.
__global__ void k_het_weights(const double* __restrict__ IP, const double* __restrict__ SL,
                              const double* __restrict__ lc, const double* __restrict__ ls,
                              const double* __restrict__ rc, const double* __restrict__ rs,
                              const double* __restrict__ aa, const double* __restrict__ amp, int M, int J, int NR,
                              int nd, double* __restrict__ w, double* __restrict__ wf) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const int J1 = J + 1;
  for (int d = 0; d < nd; d++) {
    double wh = 0.0, wc = 0.0, ws = 0.0;
    const int i1 = m / J, k1 = m - i1 * J;
    for (int pass = 0; pass < 2; pass++) {
      int i = -1, k = 0;
      if (pass == 0 && k1 == 0 && i1 >= 1 && i1 - 1 < NR) {
        i = i1 - 1;
        k = J;
      } else if (pass == 1 && i1 < NR) {
        i = i1;
        k = k1;
      }
      if (i < 0) continue;
      const double* IPi = IP + (size_t)i * J1 * J1;
      const size_t o = ((size_t)i * nd + d) * J1;
      double h = 0.0, c = 0.0, s = 0.0;
      for (int j = 0; j <= J; j++) {
        h += IPi[j * J1 + k] * SL[o + j];
        c += IPi[j * J1 + k] * lc[o + j];
        s += IPi[j * J1 + k] * ls[o + j];
      }
      wh += h;
      wc += c;
      ws += s;
    }
    if (k1 == 0 && i1 >= 1 && i1 <= NR - 2) {
      wh -= aa[(size_t)d * NR + i1];
      wc -= rc[(size_t)d * NR + i1];
      ws -= rs[(size_t)d * NR + i1];
    }
    w[(size_t)(0 * nd + d) * M + m] = wh;
    w[(size_t)(1 * nd + d) * M + m] = wc;
    w[(size_t)(2 * nd + d) * M + m] = ws;
    const double ar = amp[(size_t)d * M + m];
    wf[(size_t)(0 * nd + d) * M + m] = wh / (ar * ar);
    wf[(size_t)(1 * nd + d) * M + m] = wc / ar;
    wf[(size_t)(2 * nd + d) * M + m] = ws / ar;
  }
}

// log_likelihood_het (:5303-5444) for a batch: WPP waves per point, the
// grid samples on their lanes.  Per sample and detector fullphaseamp's
// amplitude a and phase phi (:3849-3862) and the heterodyne differences
// (:5340-5345) cc = 2 (a_ref - a cos(phi_ref - phi)), ss = 2 a sin(phi_ref - phi);
// the ratios hc = cc / a_ref, hs = ss / a_ref enter through the folded
// weights: HR_d = sum_m (wC hc + wS hs) = sum_m (wfC cc + wfS ss), HH_d =
// sum_m wH (hc^2 + hs^2) = sum_m wfH (cc^2 + ss^2), and
// logL = logLb - sum_d (HR_d + HH_d / 2).  256-thread blocks, 4 / WPP points each.
template <int ND, int WPP>
__global__ __launch_bounds__(256) void k_het_ll(const Point* __restrict__ pts, int B, const double* __restrict__ freq,
                                                const double* __restrict__ pl, const double* __restrict__ ha,
                                                const double* __restrict__ hp, const double* __restrict__ wf, int M,
                                                double logLb, double* __restrict__ logL) {
  __shared__ double red[4][2 * ND];
  // the wave index made visibly uniform: the point's constants are then scalar loads
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int b = blockIdx.x * (4 / WPP) + wave / WPP, sub = wave % WPP;
  double hr[ND], hh[ND];
#pragma unroll
  for (int d = 0; d < ND; d++) hr[d] = hh[d] = 0.0;
  const bool live = b < B && pts[b < B ? b : 0].status == ST_OK;
  if (live) {
    const Point& c = pts[b];
    const double i_Msec = 1.0 / c.Msec;
    for (int m = sub * 64 + lane; m < M; m += 64 * WPP) {
      const double f = freq[m];
      double a, p;
      h22_planes(c, i_Msec, f, pl, M, m, &a, &p);
      const double y = taper(f, c.fs);
#pragma unroll
      for (int d = 0; d < ND; d++) {
        double ad, x, sn, cs;
        project(c, d, f, a, p, y, &ad, &x);
        sincos(hp[(size_t)d * M + m] - x, &sn, &cs);
        const double cc = 2.0 * (ha[(size_t)d * M + m] - ad * cs);
        const double ss = 2.0 * ad * sn;
        hr[d] += wf[(size_t)(1 * ND + d) * M + m] * cc + wf[(size_t)(2 * ND + d) * M + m] * ss;
        hh[d] += wf[(size_t)(0 * ND + d) * M + m] * (cc * cc + ss * ss);
      }
    }
  }
#pragma unroll
  for (int d = 0; d < ND; d++) {
    const double r = het_wave_sum(hr[d]), q = het_wave_sum(hh[d]);
    if (lane == 0) {
      red[wave][2 * d] = r;
      red[wave][2 * d + 1] = q;
    }
  }
  __syncthreads();
  if (sub == 0 && lane == 0 && b < B) {
    double L = 0.0;
    if (live) {
      L = logLb;
#pragma unroll
      for (int d = 0; d < ND; d++) {
        double HR = 0.0, HH = 0.0;
        for (int k = 0; k < WPP; k++) {
          HR += red[wave + k][2 * d];
          HH += red[wave + k][2 * d + 1];
        }
        L -= (HR + 0.5 * HH);
      }
    }
    logL[b] = L;
  }
}

// ------------------------------------------------------------ launchers
hipError_t launch_phaseamp(const Point* pts, const double* freq, int n, int nd, double Tobs, double sqT, double* amp,
                           double* phs, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_phaseamp, dim3((n + 255) / 256), dim3(256), 0, s, pts, freq, n, nd, Tobs, sqT, amp, phs);
  return hipGetLastError();
}

hipError_t launch_het_setup(const Point* pts, const double* wa, const double* wp, const double* D, const double* SN,
                            int nd, int N, double Tobs, double sqT, const int* fgrid, int M, int J, int MN, int MM,
                            double* bins, double* partial, double* logLb, double* freq, double* amp, double* phs,
                            double* IP, double* SL, double* lc, double* ls, double* rc, double* rs, double* aa,
                            double* w, double* wf, double* pl, hipStream_t s) {
  const int NR = (M - 1) / J;
  hipLaunchKernelGGL(k_het_bins, dim3(HET_SETUP_BLOCKS), dim3(256), 0, s, pts, wa, wp, D, SN, nd, N, Tobs, sqT, MN,
                     MM, bins, partial);
  hipLaunchKernelGGL(k_het_logLb, dim3(1), dim3(64), 0, s, partial, HET_SETUP_BLOCKS, nd, logLb);
  hipLaunchKernelGGL(k_het_freq, dim3((M + 255) / 256), dim3(256), 0, s, fgrid, M, Tobs, freq);
  hipLaunchKernelGGL(k_het_planes, dim3((M + 255) / 256), dim3(256), 0, s, freq, M, pl);
  hipLaunchKernelGGL(k_het_ref, dim3((M + 255) / 256), dim3(256), 0, s, pts, freq, pl, M, nd, amp, phs);
  hipLaunchKernelGGL(k_het_regions, dim3(NR), dim3(64), 0, s, bins, fgrid, J, NR, nd, N / 2, IP, SL, lc, ls, rc, rs,
                     aa);
  hipLaunchKernelGGL(k_het_weights, dim3((M + 255) / 256), dim3(256), 0, s, IP, SL, lc, ls, rc, rs, aa, amp, M, J,
                     NR, nd, w, wf);
  return hipGetLastError();
}

hipError_t launch_het_planes(const double* freq, int M, double* pl, hipStream_t s) {
  if (M <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_het_planes, dim3((M + 255) / 256), dim3(256), 0, s, freq, M, pl);
  return hipGetLastError();
}

hipError_t launch_het_weights(const double* freq, const double* amp, const double* IP, const double* SL,
                              const double* lc, const double* ls, const double* rc, const double* rs,
                              const double* aa, int M, int J, int nd, double* w, double* wf, double* pl,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_het_planes, dim3((M + 255) / 256), dim3(256), 0, s, freq, M, pl);
  hipLaunchKernelGGL(k_het_weights, dim3((M + 255) / 256), dim3(256), 0, s, IP, SL, lc, ls, rc, rs, aa, amp, M, J,
                     (M - 1) / J, nd, w, wf);
  return hipGetLastError();
}

// One wave per point once the batch gives every SIMD two waves (the kernel's
// ~150 VGPRs allow no more than three); smaller batches split each point's
// samples over up to four waves.
int het_waves_per_point(int M, int B) {
  int wpp = 1;
  while (wpp < 4 && B * wpp < 2048 && M > 64 * wpp) wpp *= 2;
  return wpp;
}

template <int ND>
static void het_ll_nd(int wpp, dim3 g, hipStream_t s, const Point* pts, int B, const double* freq, const double* pl,
                      const double* ha, const double* hp, const double* wf, int M, double logLb,
                      double* logL) {
  if (wpp == 1)
    hipLaunchKernelGGL((k_het_ll<ND, 1>), g, dim3(256), 0, s, pts, B, freq, pl, ha, hp, wf, M, logLb, logL);
  else if (wpp == 2)
    hipLaunchKernelGGL((k_het_ll<ND, 2>), g, dim3(256), 0, s, pts, B, freq, pl, ha, hp, wf, M, logLb, logL);
  else
    hipLaunchKernelGGL((k_het_ll<ND, 4>), g, dim3(256), 0, s, pts, B, freq, pl, ha, hp, wf, M, logLb, logL);
}

hipError_t launch_het_ll(const Point* pts, int B, int nd, const double* freq, const double* pl, const double* ha,
                         const double* hp, const double* wf, int M, double logLb, double* logL,
                         hipStream_t s) {
  if (B <= 0) return hipSuccess;
  const int wpp = het_waves_per_point(M, B), ppb = 4 / wpp;
  const dim3 g((B + ppb - 1) / ppb);
  if (nd == 1)
    het_ll_nd<1>(wpp, g, s, pts, B, freq, pl, ha, hp, wf, M, logLb, logL);
  else if (nd == 2)
    het_ll_nd<2>(wpp, g, s, pts, B, freq, pl, ha, hp, wf, M, logLb, logL);
  else
    het_ll_nd<3>(wpp, g, s, pts, B, freq, pl, ha, hp, wf, M, logLb, logL);
  return hipGetLastError();
}

}  // namespace qcbc
