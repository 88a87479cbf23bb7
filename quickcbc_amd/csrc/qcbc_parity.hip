// HIP kernels of the parity path, compiled with -ffp-contract=off: the
// per-point set-up, the reference-structured waveform / projection / inner
// product loops, the data binding and the detector utilities.  Without fused
// multiply-adds these reproduce the reference's x86-64 double arithmetic, so
// amplitude and phase agree with it to the last few ulps (the returned phase
// near f_ref is a difference of two ~1e4 rad values, where contraction alone
// moves it by ~1e-12 rad).
#include "qcbc_device.h"
#include "qcbc_internal.h"

namespace qcbc {

// ------------------------------------------------------------ utilities
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// this unit's copy of the QNM spline slopes (the set-up kernels below)
__global__ void k_qnm_init() { qnm_init_tables(threadIdx.x); }

// ------------------------------------------------------------- set-up
// 64 points per block, four waves: waves 0 and 3 the model (both the masses
// and IMRPhenomD's prefix, then wave 0 the amplitude and wave 3 the phase
// coefficients), wave 1 fbegin, wave 2 the projection, each writing its own
// fields of the record in place in global memory (a ~1 KB local copy would not
// fit in registers; the prefix fields both model waves write hold the same
// values); after the barrier wave 0 adds the time shift and reference phase
// and finishes the record.  The parts are independent, so a point's set-up
// latency is the longest part's instead of their sum.
__global__ __launch_bounds__(256) void k_setup(const double* __restrict__ params, int B, AnalysisConst an,
                                               Point* __restrict__ pts, int* __restrict__ status) {
  const int lane = threadIdx.x & 63, role = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int b = blockIdx.x * 64 + lane;
  const bool on = b < B;
  int st = ST_OK;
  double prm[11];
  if (on) {
#pragma unroll
    for (int i = 0; i < 11; i++) prm[i] = params[11 * b + i];
    Point& c = pts[b];
    if (role == 0)
      st = point_setup_model_half(prm, an, c, 0);
    else if (role == 3)
      point_setup_model_half(prm, an, c, 1);
    else if (role == 1)
      c.fs = fbegin(prm);
    else
      point_setup_projection(prm, an, c);
  }
  __syncthreads();
  if (on && role == 0) {
    Point& c = pts[b];
    if (st == ST_OK) {
      point_setup_refphase(prm, an, c);
      point_setup_finish(an, c);
    } else {
      c.kstart = c.kend = an.imin;
    }
    c.status = st;
    if (status) status[b] = st;
  }
}

// k_setup for templates' parameters (QuickCBC.c:3527-3611): ips doubles per
// point, the quasi-intrinsic seven (ln Mc, ln Mt, chi1, chi2, phi0, tc,
// ln DL, with phi0 = params[4] itself: the model sees 2 params[4], which
// k_setup's model halves exactly) then (phase, time, amplitude) offsets of
// detectors 1, 2.  Same four-wave split as k_setup.
__global__ __launch_bounds__(256) void k_setup_intrinsic(const double* __restrict__ params, int ips, int B,
                                                         AnalysisConst an, Point* __restrict__ pts,
                                                         int* __restrict__ status) {
  const int lane = threadIdx.x & 63, role = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int b = blockIdx.x * 64 + lane;
  const bool on = b < B;
  int st = ST_OK;
  double prm[11], q[13];
  if (on) {
#pragma unroll
    for (int i = 0; i < 13; i++) q[i] = i < ips ? params[(size_t)ips * b + i] : 0.0;
#pragma unroll
    for (int i = 0; i < 11; i++) prm[i] = i < 7 ? q[i] : 0.0;
    prm[4] = 2.0 * q[4];
    Point& c = pts[b];
    if (role == 0)
      st = point_setup_model_half(prm, an, c, 0);
    else if (role == 3)
      point_setup_model_half(prm, an, c, 1);
    else if (role == 1)
      c.fs = fbegin(prm);
    else
      point_setup_projection_offsets(q, an, c);
  }
  __syncthreads();
  if (on && role == 0) {
    Point& c = pts[b];
    if (st == ST_OK) {
      point_setup_refphase(prm, an, c);
      point_setup_finish(an, c);
    } else {
      c.kstart = c.kend = an.imin;
    }
    c.status = st;
    if (status) status[b] = st;
  }
}

// One point from explicit IMRPhenomD arguments (the h22 drop-in).
__global__ __launch_bounds__(64) void k_setup_h22(double m1_SI, double m2_SI, double chi1, double chi2,
                                                  double distance, double phi0, double fref, double fmin0,
                                                  Point* pts, int* status) {
  if (threadIdx.x != 0) return;
  Point& c = pts[0];
  int st = phenomd_setup(m1_SI, m2_SI, chi1, chi2, distance, phi0, fref, fmin0, c);
  c.status = st;
  status[0] = st;
}

// Build the band planes from QuickCBC-layout D (nd x N) and SN (nd x N/2).
__global__ void k_bind(const double* __restrict__ D, const double* __restrict__ SN, int nd, int N, double Tobs,
                       int imin, int nb, int nbp, const double* __restrict__ wref, double* __restrict__ p64,
                       float* __restrict__ p32) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nbp) return;
  if (j >= nb) {  // padding: zero weight so padded bins contribute nothing
    for (int p = 0; p < 4 + 3 * nd; p++) p64[(size_t)p * nbp + j] = (p == 0) ? 1.0 : 0.0;
    for (int p = 0; p < 3 * nd; p++) p32[(size_t)p * nbp + j] = 0.0f;
    return;
  }
  int k = imin + j;
  double f = (k == 0) ? 1.0 / Tobs : (double)k / Tobs;
  p64[0 * (size_t)nbp + j] = f;
  p64[1 * (size_t)nbp + j] = cbrt(f);
  p64[2 * (size_t)nbp + j] = pow(f, -1.0 / 6.0);
  p64[3 * (size_t)nbp + j] = log(f);
  for (int d = 0; d < nd; d++) {
    double S = SN[(size_t)d * (N / 2) + k];
    double w = 1.0 / S;
    double dr = D[(size_t)d * N + k], di = D[(size_t)d * N + (N - k)];
    p64[(size_t)(4 + 3 * d) * nbp + j] = dr / S;
    p64[(size_t)(5 + 3 * d) * nbp + j] = di / S;
    p64[(size_t)(6 + 3 * d) * nbp + j] = w;
    double sw = sqrt(wref[d]);
    p32[(size_t)(3 * d) * nbp + j] = (float)((dr / S) / sw);
    p32[(size_t)(3 * d + 1) * nbp + j] = (float)((di / S) / sw);
    p32[(size_t)(3 * d + 2) * nbp + j] = (float)(w / wref[d]);
  }
}

// ------------------------------------------------------ plain FP64 path
// The reference's three loops as three kernels.  Loop 1 (IMRPhenomD.c:527-556):
// amplitude and phase of every bin of the N/2 grid into HBM.
__global__ void k_waveform(const Point* __restrict__ pts, int nhalf, double Tobs, double* __restrict__ amp,
                           double* __restrict__ phs) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (i >= nhalf) return;
  const Point& c = pts[b];
  double f = (i == 0) ? 1.0 / Tobs : (double)i / Tobs;
  double a = 0.0, ph = 0.0;
  if (c.status == ST_OK && !(f > c.fCut)) {
    Pow p = pow_direct(c.Msec * f);
    a = c.amp0 * amplitude(c, p);
    ph = phase(c, p);
  }
  amp[(size_t)b * nhalf + i] = a;
  phs[(size_t)b * nhalf + i] = ph;
}

// Same on an arbitrary frequency list (IMRPhenomDGenerateh22FDAmpPhase).
__global__ void k_waveform_freqs(const Point* __restrict__ pts, const double* __restrict__ freq, int n,
                                 double* __restrict__ amp, double* __restrict__ phs) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Point& c = pts[0];
  double f = freq[i];
  double a = 0.0, ph = 0.0;
  if (!(f > c.fCut)) {
    Pow p = pow_direct(c.Msec * f);
    a = c.amp0 * amplitude(c, p);
    ph = phase(c, p);
  }
  amp[i] = a;
  phs[i] = ph;
}

// Loop 2 (QuickCBC.c:3752-3771): projection onto each detector, half-complex.
__global__ void k_project(const Point* __restrict__ pts, const double* __restrict__ amp,
                          const double* __restrict__ phs, int nd, int N, double Tobs, double sqT,
                          double* __restrict__ hw) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  int nhalf = N / 2;
  if (i >= nhalf) return;
  const Point& c = pts[b];
  double f = (double)i / Tobs;
  double A = amp[(size_t)b * nhalf + i], p = phs[(size_t)b * nhalf + i];
  for (int d = 0; d < nd; d++) {
    double* h = hw + ((size_t)b * nd + d) * N;
    if (i == 0) {
      h[0] = 0.0;
      h[nhalf] = 0.0;
      continue;
    }
    double re = 0.0, im = 0.0;
    if (c.status == ST_OK && f > c.fs) {
      double Ad = c.Fs[d] * H22FAC * A / sqT;
      double x = TPI * f * c.tshift[d] + c.lambda[d] - p;
      re = Ad * cos(x);
      im = Ad * sin(x);
    }
    h[i] = re;
    h[N - i] = im;
  }
}

// Loop 3 (QuickCBC.c:4259-4266 with fourier_nwip, Utilities.c:716-736):
// (h|h) and (d|h) per detector over [imin, imax), block tree reduction to one
// partial per block; k_finalize sums the blocks in order.
__global__ void k_nwip(const double* __restrict__ hw, const double* __restrict__ D, const double* __restrict__ SN,
                       int nd, int N, int imin, int imax, double* __restrict__ partial) {
  __shared__ double red[8][6];
  int b = blockIdx.y, nblk = gridDim.x;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int i = imin + blockIdx.x * blockDim.x + threadIdx.x; i < imax; i += nblk * blockDim.x) {
    for (int d = 0; d < nd; d++) {
      const double* h = hw + ((size_t)b * nd + d) * N;
      const double* dd = D + (size_t)d * N;
      double S = SN[(size_t)d * (N / 2) + i];
      double hr = h[i], hi = h[N - i];
      acc[2 * d] += (hr * hr + hi * hi) / S;
      acc[2 * d + 1] += (dd[i] * hr + dd[N - i] * hi) / S;
    }
  }
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  for (int q = 0; q < 2 * nd; q++) {
    double v = wave_sum(acc[q]);
    if (lane == 0) red[wave][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < 2 * nd) {
    double s = 0.0;
    for (int w = 0; w < nw; w++) s += red[w][threadIdx.x];
    partial[((size_t)b * nblk + blockIdx.x) * (2 * nd) + threadIdx.x] = s;
  }
}

// fourier_nwip on arbitrary arrays (the drop-in): partial sums per block.
__global__ void k_nwip_pair(const double* __restrict__ a, const double* __restrict__ bb,
                            const double* __restrict__ Sn, int imin, int imax, int N, double* __restrict__ partial) {
  __shared__ double red[8];
  double acc = 0.0;
  for (int i = imin + blockIdx.x * blockDim.x + threadIdx.x; i < imax; i += gridDim.x * blockDim.x)
    acc += (a[i] * bb[i] + a[N - i] * bb[N - i]) / Sn[i];
  double v = wave_sum(acc);
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < nw; w++) s += red[w];
    partial[blockIdx.x] = s;
  }
}

__global__ void k_sum_partials(const double* __restrict__ partial, int n, double scale, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double s = 0.0;
  for (int i = 0; i < n; i++) s += partial[i];
  out[0] = scale * s;
}

// Detector utilities for the scalar drop-ins: rows of (psi, alpha, sindelta,
// gmst, id) -> (F+, Fx), (alpha, sindelta, gmst) -> 3 delays, params -> fs.
__global__ void k_geometry(int what, const double* __restrict__ in, int n, double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (what == 0) {
    const double* r = in + 5 * i;
    double fp = 0.0, fc = 0.0;
    int id = (int)r[4];
    if (id >= 0 && id <= 2) antenna(r[0], r[1], r[2], r[3], id, &fp, &fc);
    out[2 * i] = fp;
    out[2 * i + 1] = fc;
  } else if (what == 1) {
    const double* r = in + 3 * i;
    for (int d = 0; d < 3; d++) out[3 * i + d] = delay(r[0], r[1], r[2], d);
  } else {
    out[i] = fbegin(in + 11 * i);
  }
}

// ------------------------------------------------------------ launchers
hipError_t launch_qnm_init(hipStream_t s) {
  hipLaunchKernelGGL(k_qnm_init, dim3(1), dim3(64), 0, s);
  return hipGetLastError();
}

hipError_t launch_setup_intrinsic(const double* d_params, int ips, int B, const AnalysisConst& an, Point* d_pts,
                                  int* d_status, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_setup_intrinsic, dim3((B + 63) / 64), dim3(256), 0, s, d_params, ips, B, an, d_pts, d_status);
  return hipGetLastError();
}

hipError_t launch_setup(const double* d_params, int B, const AnalysisConst& an, Point* d_pts, int* d_status,
                        hipStream_t s) {
  if (B <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_setup, dim3((B + 63) / 64), dim3(256), 0, s, d_params, B, an, d_pts, d_status);
  return hipGetLastError();
}

hipError_t launch_setup_h22(double m1, double m2, double c1, double c2, double dist, double phi0, double fref,
                            double fmin0, Point* d_pts, int* d_status, hipStream_t s) {
  hipLaunchKernelGGL(k_setup_h22, dim3(1), dim3(64), 0, s, m1, m2, c1, c2, dist, phi0, fref, fmin0, d_pts, d_status);
  return hipGetLastError();
}

hipError_t launch_bind(const double* D, const double* SN, int nd, int N, double Tobs, int imin, int nb, int nbp,
                       const double* wref, double* p64, float* p32, hipStream_t s) {
  hipLaunchKernelGGL(k_bind, dim3((nbp + 255) / 256), dim3(256), 0, s, D, SN, nd, N, Tobs, imin, nb, nbp, wref, p64, p32);
  return hipGetLastError();
}

hipError_t launch_waveform(const Point* pts, int B, int nhalf, double Tobs, double* amp, double* phs, hipStream_t s) {
  hipLaunchKernelGGL(k_waveform, dim3((nhalf + 255) / 256, B), dim3(256), 0, s, pts, nhalf, Tobs, amp, phs);
  return hipGetLastError();
}

hipError_t launch_waveform_freqs(const Point* pts, const double* freq, int n, double* amp, double* phs, hipStream_t s) {
  hipLaunchKernelGGL(k_waveform_freqs, dim3((n + 255) / 256), dim3(256), 0, s, pts, freq, n, amp, phs);
  return hipGetLastError();
}

hipError_t launch_project(const Point* pts, int B, const double* amp, const double* phs, int nd, int N, double Tobs,
                          double sqT, double* hw, hipStream_t s) {
  hipLaunchKernelGGL(k_project, dim3((N / 2 + 255) / 256, B), dim3(256), 0, s, pts, amp, phs, nd, N, Tobs, sqT, hw);
  return hipGetLastError();
}

hipError_t launch_nwip(const double* hw, const double* D, const double* SN, int B, int nd, int N, int imin, int imax,
                       int nblk, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_nwip, dim3(nblk, B), dim3(256), 0, s, hw, D, SN, nd, N, imin, imax, partial);
  return hipGetLastError();
}

hipError_t launch_nwip_pair(const double* a, const double* b, const double* Sn, int imin, int imax, int N, int nblk,
                            double* partial, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_nwip_pair, dim3(nblk), dim3(256), 0, s, a, b, Sn, imin, imax, N, partial);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, s, partial, nblk, 4.0, out);
  return hipGetLastError();
}

hipError_t launch_geometry(int what, const double* in, int n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_geometry, dim3((n + 63) / 64), dim3(64), 0, s, what, in, n, out);
  return hipGetLastError();
}

}  // namespace qcbc
