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
