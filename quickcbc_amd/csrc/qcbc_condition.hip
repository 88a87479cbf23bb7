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

// main's spectra of the windowed data (QuickCBC.c:286-306) from the
// transform X (nd rows of N/2 + 1, unscaled): per bin i < N/2 the
// periodogram pspec = (2 Tobs / N^2) (Re^2 + Im^2) and wcheck = pspec / S
// (the Nyquist slot N - 0 of bin 0 is the row's spare element, 0 here); and
// W, whiten() (Utilities.c:294-309) of X: bins 1 .. N/2 - 1 times 1/sqrt(S),
// the DC and Nyquist bins 0
__global__ void k_spectra(const double2* __restrict__ X, const double* __restrict__ SN, int nd, int N, double pfac,
                          double2* __restrict__ W, double* __restrict__ pspec, double* __restrict__ wcheck) {
  const int h = N / 2;
  const size_t n = (size_t)nd * (h + 1);
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int d = (int)(r / (h + 1)), i = (int)(r % (h + 1));
  const double2 c = X[r];
  if (i < h) {
    const double re = c.x, im = (i == 0) ? 0.0 : c.y;
    const double p = pfac * (re * re + im * im);
    const size_t k = (size_t)d * h + i;
    if (pspec) pspec[k] = p;
    if (wcheck) wcheck[k] = p / SN[k];
  }
  double2 w = make_double2(0.0, 0.0);
  if (i > 0 && i < h) {
    const double x = 1.0 / sqrt(SN[(size_t)d * h + i]);
    w = make_double2(c.x * x, c.y * x);
  }
  W[r] = w;
}

// the whitened series: GSL's half-complex inverse (the backward sum times
// 1/N) of W, divided by sqrt(N/2)/sqrt(Tobs) (QuickCBC.c:270-276)
__global__ void k_whitened_series(const double* __restrict__ w, size_t n, double invN, double fac,
                                  double* __restrict__ DW) {
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  DW[r] = (w[r] * invN) / fac;
}

// whiten() (Utilities.c:294-309) in place on one half-complex row of N
__global__ void k_whiten(double* __restrict__ data, const double* __restrict__ Sn, int N) {
  const int h = N / 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > h) return;
  if (i == 0 || i == h) {
    data[i] = 0.0;
    return;
  }
  const double x = 1.0 / sqrt(Sn[i]);
  data[i] *= x;
  data[N - i] *= x;
}

// tukey_scale() (Utilities.c:656-683): the window's mean and root mean
// square over N samples; one block, lane-strided sums reduced in LDS
__global__ __launch_bounds__(256) void k_tukey_scale(int N, double alpha, double* __restrict__ s12) {
  __shared__ double a[256], b[256];
  const int imin = (int)(alpha * (double)(N - 1) / 2.0);
  const int imax = (int)((double)(N - 1) * (1.0 - alpha / 2.0));
  const int Nwin = N - imax;
  double x1 = 0.0, x2 = 0.0;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    double filter = 1.0;
    if (i < imin) filter = 0.5 * (1.0 + cos(M_PI * ((double)(i) / (double)(imin)-1.0)));
    if (i > imax) filter = 0.5 * (1.0 + cos(M_PI * ((double)(i - imax) / (double)(Nwin))));
    x1 += filter;
    x2 += filter * filter;
  }
  a[threadIdx.x] = x1;
  b[threadIdx.x] = x2;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      a[threadIdx.x] += a[threadIdx.x + o];
      b[threadIdx.x] += b[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    s12[0] = a[0] / (double)N;
    s12[1] = sqrt(b[0] / (double)N);
  }
}

}  // namespace

hipError_t launch_spectra(const void* X, const double* SN, int nd, int N, double Tobs, void* W, double* pspec,
                          double* wcheck, hipStream_t s) {
  const size_t n = (size_t)nd * (N / 2 + 1);
  // fac 2 of the reference's fac * 2.0 * (...), fac = Tobs / (N N)
  const double pfac = (Tobs / ((double)N * (double)N)) * 2.0;
  hipLaunchKernelGGL(k_spectra, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, static_cast<const double2*>(X),
                     SN, nd, N, pfac, static_cast<double2*>(W), pspec, wcheck);
  return hipGetLastError();
}

hipError_t launch_whitened_series(const double* w, int nd, int N, double Tobs, double* DW, hipStream_t s) {
  const size_t n = (size_t)nd * N;
  const double fac = sqrt((double)(N / 2)) / sqrt(Tobs);
  hipLaunchKernelGGL(k_whitened_series, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, w, n, 1.0 / (double)N,
                     fac, DW);
  return hipGetLastError();
}

hipError_t launch_whiten(double* data, const double* Sn, int N, hipStream_t s) {
  const int n = N / 2 + 1;
  hipLaunchKernelGGL(k_whiten, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, data, Sn, N);
  return hipGetLastError();
}

hipError_t launch_tukey_scale(int N, double alpha, double* s12, hipStream_t s) {
  hipLaunchKernelGGL(k_tukey_scale, dim3(1), dim3(256), 0, s, N, alpha, s12);
  return hipGetLastError();
}

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
