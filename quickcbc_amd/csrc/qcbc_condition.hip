// Data conditioning on the GPU (QuickCBC.c:268-318): each detector's time
// series windowed by tukey() (Utilities.c:635-654), Fourier transformed,
// written in GSL's half-complex layout (gsl_fft_real_radix2_transform: the
// real part of frequency i at i, the imaginary part at N - i, i = 1 .. N/2 -
// 1, the DC and Nyquist reals at 0 and N/2) and scaled by sqrt(Tobs)/N, the
// D the likelihood reads.  The transform is hipFFT's batched real-to-complex
// FFT (forward sign, e^{-2 pi i jk/N}, as GSL's).  Compiled with
// -ffp-contract=off, so the window is the reference's arithmetic.
#include <hip/hip_runtime.h>

#include "qcbc_condition_internal.h"

namespace qcbc {

namespace {

// tukey(): flat for (1 - alpha) of the samples, cosine tapers at the ends
__global__ void k_tukey(double* __restrict__ x, int nd, int N, double alpha) {
  const size_t n = (size_t)nd * N;
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int i = (int)(r % N);
  const int imin = (int)(alpha * (double)(N - 1) / 2.0);
  const int imax = (int)((double)(N - 1) * (1.0 - alpha / 2.0));
  const int Nwin = N - imax;
  double filter = 1.0;
  if (i < imin) filter = 0.5 * (1.0 + cos(M_PI * ((double)(i) / (double)(imin)-1.0)));
  if (i > imax) filter = 0.5 * (1.0 + cos(M_PI * ((double)(i - imax) / (double)(Nwin))));
  x[r] *= filter;
}

// the half-complex layout of each detector's N/2 + 1 complex coefficients,
// times fac
__global__ void k_pack_halfcomplex(const double2* __restrict__ X, int nd, int N, double fac, double* __restrict__ D) {
  const int h = N / 2;
  const size_t n = (size_t)nd * (h + 1);
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int d = (int)(r / (h + 1)), i = (int)(r % (h + 1));
  const double2 c = X[r];
  double* row = D + (size_t)d * N;
  if (i == 0) {
    row[0] = c.x * fac;
  } else if (i == h) {
    row[h] = c.x * fac;
  } else {
    row[i] = c.x * fac;
    row[N - i] = c.y * fac;
  }
}

}  // namespace

hipError_t launch_tukey(double* x, int nd, int N, double alpha, hipStream_t s) {
  const size_t n = (size_t)nd * N;
  hipLaunchKernelGGL(k_tukey, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, nd, N, alpha);
  return hipGetLastError();
}

hipError_t launch_pack_halfcomplex(const void* X, int nd, int N, double fac, double* D, hipStream_t s) {
  const size_t n = (size_t)nd * (N / 2 + 1);
  hipLaunchKernelGGL(k_pack_halfcomplex, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     static_cast<const double2*>(X), nd, N, fac, D);
  return hipGetLastError();
}

}  // namespace qcbc

// ------------------------------------------------------------- makespec
// makespec (Utilities.c:1908-1959): a PSD draw from a natural cubic spline
// of ln S through (ffit, Sspline) times (1 + the Lorentzian lines).
namespace qcbc {
namespace {

// The spline's c = M/2 (second derivative over 2) at the n knots: the
// natural end conditions, the interior equations
//   h_{i-1} M_{i-1} + 2 (h_{i-1} + h_i) M_i + h_i M_{i+1} = 6 (s_i - s_{i-1})
// by the Thomas algorithm, one thread (n is tens of knots).  w: 3 n scratch.
__global__ void k_spline_solve(const double* __restrict__ x, const double* __restrict__ y, int n,
                               double* __restrict__ c, double* __restrict__ w) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double* diag = w;
  double* rhs = w + n;
  double* off = w + 2 * n;
  for (int i = 1; i < n - 1; i++) {
    const double h0 = x[i] - x[i - 1], h1 = x[i + 1] - x[i];
    diag[i] = 2.0 * (h0 + h1);
    off[i] = h1;
    rhs[i] = 6.0 * ((y[i + 1] - y[i]) / h1 - (y[i] - y[i - 1]) / h0);
  }
  for (int i = 2; i < n - 1; i++) {
    const double h0 = x[i] - x[i - 1];
    const double m = h0 / diag[i - 1];
    diag[i] -= m * off[i - 1];
    rhs[i] -= m * rhs[i - 1];
  }
  c[0] = 0.0;
  c[n - 1] = 0.0;
  double Mnext = 0.0;
  for (int i = n - 2; i >= 1; i--) {
    const double M = (rhs[i] - off[i] * Mnext) / diag[i];
    c[i] = 0.5 * M;
    Mnext = M;
  }
}

// One thread per bin: GSL's cspline evaluation (the interval by bisection,
// y_lo + dx (b + dx (c_lo + dx d))), exp, and the lines
__global__ void k_makespec(const double* __restrict__ xa, const double* __restrict__ ya,
                           const double* __restrict__ c, int n, const double* __restrict__ linef,
                           const double* __restrict__ lineh, const double* __restrict__ lineQ,
                           const double* __restrict__ linew, const double* __restrict__ dfm, int nlines,
                           double Tobs, int N, double* __restrict__ SN) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int h = N / 2;
  if (i >= h) return;
  if (i == 0) {
    SN[0] = exp(ya[0]);
    return;
  }
  if (i == h - 1) {
    SN[h - 1] = exp(ya[n - 1]);
    return;
  }
  const double f = (double)(i) / Tobs;
  int lo = 0, hi = n - 1;
  while (hi > lo + 1) {
    const int mid = (lo + hi) / 2;
    if (xa[mid] > f)
      hi = mid;
    else
      lo = mid;
  }
  const double hh = xa[lo + 1] - xa[lo];
  const double dy = ya[lo + 1] - ya[lo];
  const double dx = f - xa[lo];
  const double b = dy / hh - hh * (c[lo + 1] + 2.0 * c[lo]) / 3.0;
  const double d = (c[lo + 1] - c[lo]) / (3.0 * hh);
  const double sm = exp(ya[lo] + dx * (b + dx * (c[lo] + dx * d)));
  double y = 0.0;
  for (int j = 0; j < nlines; j++) {
    const double x = fabs(f - linef[j]);
    if (x < linew[j]) {
      double z = 1.0;
      if (x > dfm[j]) z = exp(-(x - dfm[j]) / dfm[j]);
      const double f2 = linef[j] * linef[j];
      const double ff2 = f * f;
      const double df2 = lineQ[j] * (1.0 - ff2 / f2);
      const double df4 = df2 * df2;
      y += z * lineh[j] / (ff2 / f2 + df4);
    }
  }
  SN[i] = sm * (1.0 + y);
}

}  // namespace

hipError_t launch_makespec(const double* xa, const double* ya, int n, const double* lines, int nlines, double Tobs,
                           int N, double* c, double* w, double* SN, hipStream_t s) {
  hipLaunchKernelGGL(k_spline_solve, dim3(1), dim3(64), 0, s, xa, ya, n, c, w);
  const int h = N / 2;
  hipLaunchKernelGGL(k_makespec, dim3((h + 255) / 256), dim3(256), 0, s, xa, ya, c, n, lines, lines + nlines,
                     lines + 2 * nlines, lines + 3 * nlines, lines + 4 * nlines, nlines, Tobs, N, SN);
  return hipGetLastError();
}

}  // namespace qcbc
