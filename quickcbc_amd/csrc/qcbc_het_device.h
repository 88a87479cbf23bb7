// Per-sample device code shared by the heterodyned-likelihood and the
// Fisher-matrix kernels: fullphaseamp's amplitude and phase at one grid
// sample (QuickCBC.c:3780-3869) from a point's set-up record and the grid's
// sample powers.  Both translation units are compiled with
// -ffp-contract=off, so the two evaluate the same double operations.
#pragma once

#include "qcbc_device.h"

namespace qcbc {
namespace {

__device__ __forceinline__ double het_wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// fullphaseamp's taper below fbegin (QuickCBC.c:3858-3859)
__device__ __forceinline__ double taper(double f, double fs) {
  double y = 1.0;
  if (f < fs) y = exp(10.0 * (f - fs) / fs);
  return y;
}

// IMRPhenomD amplitude (amp0 included) and phase at one frequency, zero above
// fCut (IMRPhenomD.c:529-555)
__device__ __forceinline__ void h22_at(const Point& c, double f, double* a, double* ph) {
  *a = 0.0;
  *ph = 0.0;
  if (!(f > c.fCut)) {
    Pow p = pow_direct(c.Msec * f);
    *a = c.amp0 * amplitude(c, p);
    *ph = phase(c, p);
  }
}

// The frequency powers of one sample from its planes and the point's
// Msec^(1/6), Msec^(-1/6), 1/Msec and log Msec: the products and sums
// pow_direct forms with pow, log and divisions, to a few ulps of its values.
__device__ __forceinline__ Pow pow_planes(const Point& c, double i_Msec, double f, const double* __restrict__ pl,
                                          int M, int m) {
  Pow p;
  p.Mf = c.Msec * f;
  p.s6 = c.s6_scale * pl[m];
  p.third = p.s6 * p.s6;
  p.two_thirds = p.third * p.third;
  p.four_thirds = p.Mf * p.third;
  p.five_thirds = p.four_thirds * p.third;
  p.two = p.Mf * p.Mf;
  p.seven_thirds = p.third * p.two;
  p.eight_thirds = p.two_thirds * p.two;
  p.i_s6 = c.is6_scale * pl[M + m];
  p.i_third = p.i_s6 * p.i_s6;
  p.i_two_thirds = p.i_third * p.i_third;
  p.i_Mf = i_Msec * pl[2 * M + m];
  p.i_five_thirds = p.i_Mf * p.i_two_thirds;
  p.logMf = c.logMsec + pl[3 * M + m];
  p.logv = (LOG_PI + p.logMf) / 3.0;  // log v, v = (pi Mf)^(1/3)
  return p;
}

// The h22 amplitude without amp0 (a) and phase (p) of point c at grid
// sample m, zero above fCut (IMRPhenomD.c:529-555), powers from the planes
__device__ __forceinline__ void h22_planes(const Point& c, double i_Msec, double f, const double* __restrict__ pl,
                                           int M, int m, double* a, double* p) {
  *a = 0.0;
  *p = 0.0;
  if (!(f > c.fCut)) {
    const Pow pw = pow_planes(c, i_Msec, f, pl, M, m);
    *a = amplitude(c, pw);
    *p = phase(c, pw);
  }
}

// fullphaseamp's projected amplitude ad and phase x (QuickCBC.c:3849-3862)
// for detector d: ad = Fs h22fac amp0 a / sqrt(Tobs) y with the point's
// detector scale Fs h22fac amp0 / sqrt(Tobs) (Ascale) and the taper y
__device__ __forceinline__ void project(const Point& c, int d, double f, double a, double p, double y, double* ad,
                                        double* x) {
  *x = TPI * f * c.tshift[d] + c.lambda[d] - p;
  *ad = c.Ascale[d] * a * y;
}

}  // namespace
}  // namespace qcbc
