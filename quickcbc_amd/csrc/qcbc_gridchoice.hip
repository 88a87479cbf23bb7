// HIP kernels of the heterodyne's grid choice (QuickCBC.c:4275-5053),
// compiled with -ffp-contract=off: each per-bin product is the reference's
// double arithmetic; the sums over bins are regrouped (a tree per block, then
// the blocks in order).
//
//   k_diff_full   diff (QuickCBC.c:4913-4964) for K pairs of set-up points:
//                 sum_d sum_j 4 |h_d(a) - h_d(b)|^2 / S_d(j) over a bin range
//                 of the analysis grid, fullphaseamp's amplitude and phase.
//   k_diff_het    het_diff (QuickCBC.c:4966-5053) for K pairs on a
//                 heterodyne's grid, the Legendre moments folded into one
//                 weight per sample (the heterodyned likelihood's wH / a_ref^2).
//   k_band_grid   the grid of every band bin for Fisher_All (QuickCBC.c:
//                 2695-2817): frequencies, and 1/S per detector and bin as the
//                 per-sample weight of the heterodyned Fisher kernel.
#include "qcbc_device.h"
#include "qcbc_gridchoice_internal.h"
#include "qcbc_het_device.h"

namespace qcbc {

namespace {

constexpr int DIFF_THREADS = 256;

__device__ __forceinline__ double block_sum(double v, double* red) {
  v = het_wave_sum(v);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < DIFF_THREADS / 64; w++) s += red[w];
  return s;
}

// fullphaseamp's amplitude and phase of one point at bin frequency f for
// detector d, as k_phaseamp forms them (QuickCBC.c:3849-3862)
__device__ __forceinline__ void phaseamp_at(const Point& c, double f, double a, double p, double y, double sqT, int d,
                                            double* A, double* x) {
  const double Ad = c.Fs[d] * H22FAC * a / sqT;
  *x = TPI * f * c.tshift[d] + c.lambda[d] - p;
  *A = Ad * y;
}

// Block (k, s) of pair k sums bins [lo_k + s len, min(hi_k, lo_k + (s + 1) len))
// into partial[k splits + s].
__global__ __launch_bounds__(DIFF_THREADS) void k_diff_full(const Point* __restrict__ pts,
                                                            const DiffPair* __restrict__ pairs, int K, int nd,
                                                            const double* __restrict__ SN, int n2, double Tobs,
                                                            double sqT, int len, double* __restrict__ partial) {
  __shared__ double red[DIFF_THREADS / 64];
  const int k = blockIdx.x, s = blockIdx.y;
  if (k >= K) return;
  const DiffPair pr = pairs[k];
  const Point& ca = pts[pr.a];
  const Point& cb = pts[pr.b];
  double acc = 0.0;
  if (ca.status == ST_OK && cb.status == ST_OK) {
    const int j0 = pr.lo + s * len, j1 = min(pr.hi, j0 + len);
    for (int j = j0 + (int)threadIdx.x; j < j1; j += DIFF_THREADS) {
      const double f = (double)j / Tobs;  // bins j >= 1 of the grid f_j = j / Tobs
      double aa, pa, ab, pb;
      h22_at(ca, f, &aa, &pa);
      h22_at(cb, f, &ab, &pb);
      const double ya = taper(f, ca.fs), yb = taper(f, cb.fs);
      for (int d = 0; d < nd; d++) {
        double Ar, Pr, Ap, Pp;
        phaseamp_at(ca, f, aa, pa, ya, sqT, d, &Ar, &Pr);
        phaseamp_at(cb, f, ab, pb, yb, sqT, d, &Ap, &Pp);
        const double Re = (Ar - Ap * cos(Pr - Pp));
        const double Im = Ap * sin(Pr - Pp);
        acc += 4.0 * (Re * Re + Im * Im) / SN[(size_t)d * n2 + j];
      }
    }
  }
  const double tot = block_sum(acc, red);
  if (threadIdx.x == 0) partial[(size_t)k * gridDim.y + s] = tot;
}

// z[k] = the sum of pair k's partials in order, NaN when either point's
// set-up failed (the reference's arithmetic on an unset waveform)
__global__ void k_diff_reduce(const Point* __restrict__ pts, const DiffPair* __restrict__ pairs, int K, int splits,
                              const double* __restrict__ partial, double* __restrict__ z) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  const DiffPair pr = pairs[k];
  if (pts[pr.a].status != ST_OK || pts[pr.b].status != ST_OK) {
    z[k] = __builtin_nan("");
    return;
  }
  double x = 0.0;
  for (int s = 0; s < splits; s++) x += partial[(size_t)k * splits + s];
  z[k] = x;
}

// One wave per pair over the heterodyne's M samples: sum_d sum_m
// wf[d][m] 4 |h_d(a) - h_d(b)|^2, wf the folded slow-term weight wH / a_ref^2
// (QuickCBC.c:5000-5040: cc = 4 |.|^2 / a_ref^2, Legendre coefficients
// IP cc against SL, less the region-start overcount, is this linear form)
__global__ __launch_bounds__(256) void k_diff_het(const Point* __restrict__ pts, const DiffPair* __restrict__ pairs,
                                                  int K, int nd, const double* __restrict__ freq,
                                                  const double* __restrict__ pl, const double* __restrict__ wf, int M,
                                                  double* __restrict__ z) {
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (k >= K) return;  // wave-uniform
  const DiffPair pr = pairs[k];
  const Point& ca = pts[pr.a];
  const Point& cb = pts[pr.b];
  if (ca.status != ST_OK || cb.status != ST_OK) {
    if (lane == 0) z[k] = __builtin_nan("");
    return;
  }
  const double ia = 1.0 / ca.Msec, ib = 1.0 / cb.Msec;
  double acc = 0.0;
  for (int m = lane; m < M; m += 64) {
    const double f = freq[m];
    double aa, pa, ab, pb;
    h22_planes(ca, ia, f, pl, M, m, &aa, &pa);
    h22_planes(cb, ib, f, pl, M, m, &ab, &pb);
    const double ya = taper(f, ca.fs), yb = taper(f, cb.fs);
    for (int d = 0; d < nd; d++) {
      double Ar, Pr, Ap, Pp;
      project(ca, d, f, aa, pa, ya, &Ar, &Pr);
      project(cb, d, f, ab, pb, yb, &Ap, &Pp);
      const double Re = (Ar - Ap * cos(Pr - Pp));
      const double Im = Ap * sin(Pr - Pp);
      acc += wf[(size_t)d * M + m] * (4.0 * (Re * Re + Im * Im));
    }
  }
  acc = het_wave_sum(acc);
  if (lane == 0) z[k] = acc;
}

// The band bins [imin, imin + nb) as a grid: freq[m] = (imin + m) / Tobs and
// w[d][m] = 1 / S_d(imin + m)
__global__ void k_band_grid(const double* __restrict__ SN, int nd, int n2, int imin, int nb, double Tobs,
                            double* __restrict__ freq, double* __restrict__ w) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nb) return;
  const int j = imin + m;
  freq[m] = (double)j / Tobs;
  for (int d = 0; d < nd; d++) w[(size_t)d * nb + m] = 1.0 / SN[(size_t)d * n2 + j];
}

}  // namespace

int diff_splits(int max_bins) {
  const int len = 2048;
  int s = (max_bins + len - 1) / len;
  return s < 1 ? 1 : s;
}

hipError_t launch_diff_full(const Point* pts, const DiffPair* pairs, int K, int nd, const double* SN, int n2,
                            double Tobs, int max_bins, int splits, double* partial, double* z, hipStream_t s) {
  if (K <= 0) return hipSuccess;
  if (splits < 1) return hipErrorInvalidValue;
  const int len = (max_bins + splits - 1) / splits;
  hipLaunchKernelGGL(k_diff_full, dim3(K, splits), dim3(DIFF_THREADS), 0, s, pts, pairs, K, nd, SN, n2, Tobs,
                     sqrt(Tobs), len < 1 ? 1 : len, partial);
  hipLaunchKernelGGL(k_diff_reduce, dim3((K + 255) / 256), dim3(256), 0, s, pts, pairs, K, splits, partial, z);
  return hipGetLastError();
}

hipError_t launch_diff_het(const Point* pts, const DiffPair* pairs, int K, int nd, const double* freq,
                           const double* pl, const double* wf, int M, double* z, hipStream_t s) {
  if (K <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_diff_het, dim3((K + 3) / 4), dim3(256), 0, s, pts, pairs, K, nd, freq, pl, wf, M, z);
  return hipGetLastError();
}

hipError_t launch_band_grid(const double* SN, int nd, int n2, int imin, int nb, double Tobs, double* freq, double* w,
                            hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_band_grid, dim3((nb + 255) / 256), dim3(256), 0, s, SN, nd, n2, imin, nb, Tobs, freq, w);
  return hipGetLastError();
}

}  // namespace qcbc
