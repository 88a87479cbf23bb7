// Device-side physics of the waveform + likelihood path (gfx950).
//
// One thread builds the frequency-independent state of one parameter point
// (qcbc_point_setup); the per-bin evaluators then run with bins on lanes.
// The per-point record qcbc_point is read wave-uniformly by the likelihood
// kernels (every lane of a wave works on the same point), so it is an
// array-of-structs: the compiler turns its loads into scalar loads.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qcbc_types.h"

#define QCBC_TABLE_STORAGE static __constant__
#define QCBC_CONST_STORAGE static constexpr
#include "qcbc_tables.inc"

namespace qcbc {

// IMRPhenomD_internals.h:102-161
constexpr double PI = 3.141592653589793238462643383279502884;
constexpr double PI_4 = 0.785398163397448309615660845819875721;
constexpr double EULER_GAMMA = 0.577215664901532860606512090082402431;
constexpr double MSUN_SI = 1.988546954961461467461011951140572744e30;
constexpr double MTSUN_SI = 4.925491025543575903411922162094833998e-6;
constexpr double MRSUN_SI = 1.476625061404649406193430731479084713e3;
constexpr double F_CUT = 0.2;
constexpr double AMP_JOIN = 0.014;
constexpr double PHI_JOIN = 0.018;
// ConstCBC.h:1-12
constexpr double TSUN = 4.92569043916e-6;
constexpr double TPI = 6.2831853071795862319959269370884;
constexpr double CLIGHT = 299792458.0;
constexpr double H22FAC = 0.31539156525252;
constexpr double LOG_PI = 1.1447298858494001741434273513530587;
constexpr double INV_TPI = 0.15915494309189533576888376337251436;
// powers of pi as init_useful_powers(&powers_of_pi, PI) forms them
// (IMRPhenomD_internals.c:515-533): pow(pi, 1/6) rounded to double, then
// products and quotients in double
constexpr double PI_S6 = 0x1.35cfe132fe3abp+0;
constexpr double PI_THIRD = 0x1.76ef7e73104b8p+0;
constexpr double PI_TWO_THIRDS = 0x1.129052c2c6753p+1;
constexpr double PI_FOUR_THIRDS = 0x1.2679428626b38p+2;
constexpr double PI_FIVE_THIRDS = 0x1.af48a3d921bdbp+2;

// --------------------------------------------- correctly rounded exp / pow
// The per-point masses come from exp(ln Mc), exp(ln Mt) and
// pow(Mc/Mt, 5/3) (QuickCBC.c:3705-3718); the Table-V fits in eta and chi
// amplify a one-ulp change of eta a hundredfold.  These calls are therefore
// evaluated in double-double and rounded once, which gives the correctly
// rounded result (as glibc's < 0.52-ulp exp and pow return it for all but
// vanishingly rare arguments).  Per point only; nothing per bin.
struct DD {
  double hi, lo;
};
__device__ __forceinline__ DD dd_two_sum(double a, double b) {
  double s = a + b, bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ DD dd_fast_two_sum(double a, double b) {
  double s = a + b;
  return {s, b - (s - a)};
}
__device__ __forceinline__ DD dd_add(DD a, DD b) {
  DD s = dd_two_sum(a.hi, b.hi);
  double t = s.lo + (a.lo + b.lo);
  return dd_fast_two_sum(s.hi, t);
}
__device__ __forceinline__ DD dd_mul(DD a, DD b) {
  double p = a.hi * b.hi;
  double e = fma(a.hi, b.hi, -p) + (a.hi * b.lo + a.lo * b.hi);
  return dd_fast_two_sum(p, e);
}
__device__ __forceinline__ DD dd_mul_d(DD a, double b) {
  double p = a.hi * b;
  double e = fma(a.hi, b, -p) + a.lo * b;
  return dd_fast_two_sum(p, e);
}
__device__ __forceinline__ DD dd_div_d(DD a, double b) {
  double q = a.hi / b;
  double p = q * b, pe = fma(q, b, -p);
  double r = ((a.hi - p) - pe + a.lo) / b;
  return dd_fast_two_sum(q, r);
}
// 2^(j/128), j = 0..127, in double-double (generated at 60 digits).
__device__ constexpr DD QCBC_EXP2_J[128] = {
    {0x1.0000000000000p+0, 0x0.0p+0},
    {0x1.0163da9fb3335p+0, 0x1.b61299ab8cdb7p-54},
    {0x1.02c9a3e778061p+0, -0x1.19083535b085dp-56},
    {0x1.04315e86e7f85p+0, -0x1.0a31c1977c96ep-54},
    {0x1.059b0d3158574p+0, 0x1.d73e2a475b465p-55},
    {0x1.0706b29ddf6dep+0, -0x1.c91dfe2b13c27p-55},
    {0x1.0874518759bc8p+0, 0x1.186be4bb284ffp-57},
    {0x1.09e3ecac6f383p+0, 0x1.1487818316136p-54},
    {0x1.0b5586cf9890fp+0, 0x1.8a62e4adc610bp-54},
    {0x1.0cc922b7247f7p+0, 0x1.01edc16e24f71p-54},
    {0x1.0e3ec32d3d1a2p+0, 0x1.03a1727c57b53p-59},
    {0x1.0fb66affed31bp+0, -0x1.b9bedc44ebd7bp-57},
    {0x1.11301d0125b51p+0, -0x1.6c51039449b3ap-54},
    {0x1.12abdc06c31ccp+0, -0x1.1b514b36ca5c7p-58},
    {0x1.1429aaea92de0p+0, -0x1.32fbf9af1369ep-54},
    {0x1.15a98c8a58e51p+0, 0x1.2406ab9eeab0ap-55},
    {0x1.172b83c7d517bp+0, -0x1.19041b9d78a76p-55},
    {0x1.18af9388c8deap+0, -0x1.11023d1970f6cp-54},
    {0x1.1a35beb6fcb75p+0, 0x1.e5b4c7b4968e4p-55},
    {0x1.1bbe084045cd4p+0, -0x1.95386352ef607p-54},
    {0x1.1d4873168b9aap+0, 0x1.e016e00a2643cp-54},
    {0x1.1ed5022fcd91dp+0, -0x1.1df98027bb78cp-54},
    {0x1.2063b88628cd6p+0, 0x1.dc775814a8495p-55},
    {0x1.21f49917ddc96p+0, 0x1.2a97e9494a5eep-55},
    {0x1.2387a6e756238p+0, 0x1.9b07eb6c70573p-54},
    {0x1.251ce4fb2a63fp+0, 0x1.ac155bef4f4a4p-55},
    {0x1.26b4565e27cddp+0, 0x1.2bd339940e9d9p-55},
    {0x1.284dfe1f56381p+0, -0x1.a4c3a8c3f0d7ep-54},
    {0x1.29e9df51fdee1p+0, 0x1.612e8afad1255p-55},
    {0x1.2b87fd0dad990p+0, -0x1.10adcd6381aa4p-59},
    {0x1.2d285a6e4030bp+0, 0x1.0024754db41d5p-54},
    {0x1.2ecafa93e2f56p+0, 0x1.1ca0f45d52383p-56},
    {0x1.306fe0a31b715p+0, 0x1.6f46ad23182e4p-55},
    {0x1.32170fc4cd831p+0, 0x1.a9ce78e18047cp-55},
    {0x1.33c08b26416ffp+0, 0x1.32721843659a6p-54},
    {0x1.356c55f929ff1p+0, -0x1.b5cee5c4e4628p-55},
    {0x1.371a7373aa9cbp+0, -0x1.63aeabf42eae2p-54},
    {0x1.38cae6d05d866p+0, -0x1.e958d3c9904bdp-54},
    {0x1.3a7db34e59ff7p+0, -0x1.5e436d661f5e3p-56},
    {0x1.3c32dc313a8e5p+0, -0x1.efff8375d29c3p-54},
    {0x1.3dea64c123422p+0, 0x1.ada0911f09ebcp-55},
    {0x1.3fa4504ac801cp+0, -0x1.7d023f956f9f3p-54},
    {0x1.4160a21f72e2ap+0, -0x1.ef3691c309278p-58},
    {0x1.431f5d950a897p+0, -0x1.1c7dde35f7999p-55},
    {0x1.44e086061892dp+0, 0x1.89b7a04ef80d0p-59},
    {0x1.46a41ed1d0057p+0, 0x1.c944bd1648a76p-54},
    {0x1.486a2b5c13cd0p+0, 0x1.3c1a3b69062f0p-56},
    {0x1.4a32af0d7d3dep+0, 0x1.9cb62f3d1be56p-54},
    {0x1.4bfdad5362a27p+0, 0x1.d4397afec42e2p-56},
    {0x1.4dcb299fddd0dp+0, 0x1.8ecdbbc6a7833p-54},
    {0x1.4f9b2769d2ca7p+0, -0x1.4b309d25957e3p-54},
    {0x1.516daa2cf6642p+0, -0x1.f768569bd93efp-55},
    {0x1.5342b569d4f82p+0, -0x1.07abe1db13cadp-55},
    {0x1.551a4ca5d920fp+0, -0x1.d689cefede59bp-55},
    {0x1.56f4736b527dap+0, 0x1.9bb2c011d93adp-54},
    {0x1.58d12d497c7fdp+0, 0x1.295e15b9a1de8p-55},
    {0x1.5ab07dd485429p+0, 0x1.6324c054647adp-54},
    {0x1.5c9268a5946b7p+0, 0x1.c4b1b816986a2p-60},
    {0x1.5e76f15ad2148p+0, 0x1.ba6f93080e65ep-54},
    {0x1.605e1b976dc09p+0, -0x1.3e2429b56de47p-54},
    {0x1.6247eb03a5585p+0, -0x1.383c17e40b497p-54},
    {0x1.6434634ccc320p+0, -0x1.c483c759d8933p-55},
    {0x1.6623882552225p+0, -0x1.bb60987591c34p-54},
    {0x1.68155d44ca973p+0, 0x1.038ae44f73e65p-57},
    {0x1.6a09e667f3bcdp+0, -0x1.bdd3413b26456p-54},
    {0x1.6c012750bdabfp+0, -0x1.2895667ff0b0dp-56},
    {0x1.6dfb23c651a2fp+0, -0x1.bbe3a683c88abp-57},
    {0x1.6ff7df9519484p+0, -0x1.83c0f25860ef6p-55},
    {0x1.71f75e8ec5f74p+0, -0x1.16e4786887a99p-55},
    {0x1.73f9a48a58174p+0, -0x1.0a8d96c65d53cp-54},
    {0x1.75feb564267c9p+0, -0x1.0245957316dd3p-54},
    {0x1.780694fde5d3fp+0, 0x1.866b80a02162dp-54},
    {0x1.7a11473eb0187p+0, -0x1.41577ee04992fp-55},
    {0x1.7c1ed0130c132p+0, 0x1.f124cd1164dd6p-54},
    {0x1.7e2f336cf4e62p+0, 0x1.05d02ba15797ep-56},
    {0x1.80427543e1a12p+0, -0x1.27c86626d972bp-54},
    {0x1.82589994cce13p+0, -0x1.d4c1dd41532d8p-54},
    {0x1.8471a4623c7adp+0, -0x1.8d684a341cdfbp-55},
    {0x1.868d99b4492edp+0, -0x1.fc6f89bd4f6bap-54},
    {0x1.88ac7d98a6699p+0, 0x1.994c2f37cb53ap-54},
    {0x1.8ace5422aa0dbp+0, 0x1.6e9f156864b27p-54},
    {0x1.8cf3216b5448cp+0, -0x1.0d55e32e9e3aap-56},
    {0x1.8f1ae99157736p+0, 0x1.5cc13a2e3976cp-55},
    {0x1.9145b0b91ffc6p+0, -0x1.dd6792e582524p-54},
    {0x1.93737b0cdc5e5p+0, -0x1.75fc781b57ebcp-57},
    {0x1.95a44cbc8520fp+0, -0x1.64b7c96a5f039p-56},
    {0x1.97d829fde4e50p+0, -0x1.d185b7c1b85d1p-54},
    {0x1.9a0f170ca07bap+0, -0x1.173bd91cee632p-54},
    {0x1.9c49182a3f090p+0, 0x1.c7c46b071f2bep-56},
    {0x1.9e86319e32323p+0, 0x1.824ca78e64c6ep-56},
    {0x1.a0c667b5de565p+0, -0x1.359495d1cd533p-54},
    {0x1.a309bec4a2d33p+0, 0x1.6305c7ddc36abp-54},
    {0x1.a5503b23e255dp+0, -0x1.d2f6edb8d41e1p-54},
    {0x1.a799e1330b358p+0, 0x1.bcb7ecac563c7p-54},
    {0x1.a9e6b5579fdbfp+0, 0x1.0fac90ef7fd31p-54},
    {0x1.ac36bbfd3f37ap+0, -0x1.f9234cae76cd0p-55},
    {0x1.ae89f995ad3adp+0, 0x1.7a1cd345dcc81p-54},
    {0x1.b0e07298db666p+0, -0x1.bdef54c80e425p-54},
    {0x1.b33a2b84f15fbp+0, -0x1.2805e3084d708p-57},
    {0x1.b59728de5593ap+0, -0x1.c71dfbbba6de3p-54},
    {0x1.b7f76f2fb5e47p+0, -0x1.5584f7e54ac3bp-56},
    {0x1.ba5b030a1064ap+0, -0x1.efcd30e54292ep-54},
    {0x1.bcc1e904bc1d2p+0, 0x1.23dd07a2d9e84p-55},
    {0x1.bf2c25bd71e09p+0, -0x1.efdca3f6b9c73p-54},
    {0x1.c199bdd85529cp+0, 0x1.11065895048ddp-55},
    {0x1.c40ab5fffd07ap+0, 0x1.b4537e083c60ap-54},
    {0x1.c67f12e57d14bp+0, 0x1.2884dff483cadp-54},
    {0x1.c8f6d9406e7b5p+0, 0x1.1acbc48805c44p-56},
    {0x1.cb720dcef9069p+0, 0x1.503cbd1e949dbp-56},
    {0x1.cdf0b555dc3fap+0, -0x1.dd83b53829d72p-55},
    {0x1.d072d4a07897cp+0, -0x1.cbc3743797a9cp-54},
    {0x1.d2f87080d89f2p+0, -0x1.d487b719d8578p-54},
    {0x1.d5818dcfba487p+0, 0x1.2ed02d75b3707p-55},
    {0x1.d80e316c98398p+0, -0x1.11ec18beddfe8p-54},
    {0x1.da9e603db3285p+0, 0x1.c2300696db532p-54},
    {0x1.dd321f301b460p+0, 0x1.2da5778f018c3p-54},
    {0x1.dfc97337b9b5fp+0, -0x1.1a5cd4f184b5cp-54},
    {0x1.e264614f5a129p+0, -0x1.7b627817a1496p-54},
    {0x1.e502ee78b3ff6p+0, 0x1.39e8980a9cc8fp-55},
    {0x1.e7a51fbc74c83p+0, 0x1.2d522ca0c8de2p-54},
    {0x1.ea4afa2a490dap+0, -0x1.e9c23179c2893p-54},
    {0x1.ecf482d8e67f1p+0, -0x1.c93f3b411ad8cp-54},
    {0x1.efa1bee615a27p+0, 0x1.dc7f486a4b6b0p-54},
    {0x1.f252b376bba97p+0, 0x1.3a1a5bf0d8e43p-54},
    {0x1.f50765b6e4540p+0, 0x1.9d3e12dd8a18bp-54},
    {0x1.f7bfdad9cbe14p+0, -0x1.dbb12d006350ap-54},
    {0x1.fa7c1819e90d8p+0, 0x1.74853f3a5931ep-55},
    {0x1.fd3c22b8f71f1p+0, 0x1.2eb74966579e7p-57},
};

// exp of a double-double (Tang's table method): x = (128 m + j) ln2/128 + r
// with |r| <= ln2/256, exp(x) = 2^m 2^(j/128) (1 + r + r^2/2 + r^3/6 + r^4 q(r)),
// the series through r^3 in double-double and r^4 q(r) (to r^8/8!, a term
// below 3e-12) in double: relative error below ~1e-27.
__device__ inline DD dd_exp(DD x) {
  const double L_HI = 0x1.62e4200000000p-8;  // ln2/128, 21 bits: k L_HI is exact
  const double L_LO = 0x1.fdf473de6af28p-29, L_LO2 = -0x1.c4c67fc0d0951p-83;
  const double kk = rint(x.hi * 0x1.71547652b82fep+7);
  DD r = dd_two_sum(x.hi - kk * L_HI, x.lo);
  r = dd_add(r, {-kk * L_LO, -fma(kk, L_LO, -kk * L_LO)});
  r = dd_add(r, {-kk * L_LO2, 0.0});
  const int k = (int)kk, j = k & 127, m = (k - j) / 128;
  const double h = r.hi;
  const double q = 1.0 / 24.0 + h * (1.0 / 120.0 + h * (1.0 / 720.0 + h * (1.0 / 5040.0 + h * (1.0 / 40320.0))));
  const DD r2 = dd_mul(r, r);
  const DD r3 = dd_mul(r2, r);
  DD em1 = dd_add(dd_mul(r3, {0x1.5555555555555p-3, 0x1.5555555555555p-57}), {(r2.hi * r2.hi) * q, 0.0});
  em1 = dd_add(em1, {0.5 * r2.hi, 0.5 * r2.lo});
  em1 = dd_add(em1, r);
  const DD T = QCBC_EXP2_J[j];
  const DD v = dd_add(T, dd_mul(T, em1));
  return {ldexp(v.hi, m), ldexp(v.lo, m)};
}
// log by one Newton step on dd_exp from the double log: y + (x e^-y - 1).
__device__ inline DD dd_log(double x) {
  const double y0 = log(x);
  DD t = dd_add(dd_mul_d(dd_exp({-y0, 0.0}), x), {-1.0, 0.0});
  return dd_add({y0, 0.0}, t);
}
__device__ inline double exp_cr(double x) { return dd_exp({x, 0.0}).hi; }
__device__ inline double pow_cr(double x, double y) { return dd_exp(dd_mul_d(dd_log(x), y)).hi; }

// Slopes of the natural cubic splines through the QNM tables (row 0 f_ring,
// row 1 f_damp), one copy per translation unit (device globals do not link
// across units without relocatable device code): each unit that evaluates the
// set-up fills its copy once per device with a kernel calling
// qnm_init_tables.
[[maybe_unused]] static __device__ double g_qnm_slope[2][QCBC_QNM_N];
[[maybe_unused]] static __device__ double g_qnm_sweep[2][QCBC_QNM_N];  // modified super-diagonal of the solve

// The tridiagonal system h_i m_{i-1} + 2(h_{i-1}+h_i) m_i + h_{i-1} m_{i+1}
// = 3(h_i s_{i-1} + h_{i-1} s_i), with natural end rows, by the Thomas
// algorithm; thread w = 0 builds the f_ring spline, w = 1 the f_damp spline.
__device__ inline void qnm_init_tables(int w) {
  if (w > 1) return;
  const double* x = qcbc_qnm_a;
  const double* y = w ? qcbc_qnm_fdamp : qcbc_qnm_fring;
  double* m = g_qnm_slope[w];
  const int n = QCBC_QNM_N;
  double* cprime = g_qnm_sweep[w];
  double dprime;
  double prev_c = 0.0, prev_d = 0.0;
  for (int i = 0; i < n; i++) {
    double a, b, cc, r;
    if (i == 0) {
      double h0 = x[1] - x[0];
      a = 0.0; b = 2.0; cc = 1.0; r = 3.0 * (y[1] - y[0]) / h0;
    } else if (i == n - 1) {
      double h = x[n - 1] - x[n - 2];
      a = 1.0; b = 2.0; cc = 0.0; r = 3.0 * (y[n - 1] - y[n - 2]) / h;
    } else {
      double hl = x[i] - x[i - 1], hr = x[i + 1] - x[i];
      double sl = (y[i] - y[i - 1]) / hl, sr = (y[i + 1] - y[i]) / hr;
      a = hr; b = 2.0 * (hl + hr); cc = hl; r = 3.0 * (hr * sl + hl * sr);
    }
    double den = b - a * prev_c;
    prev_c = cc / den;
    dprime = (r - a * prev_d) / den;
    prev_d = dprime;
    cprime[i] = prev_c;
    m[i] = dprime;
  }
  for (int i = n - 2; i >= 0; i--) m[i] -= cprime[i] * m[i + 1];
}

// ------------------------------------------------------------- helpers
__device__ __forceinline__ double fit11(int w, double eta, double chi) {
  const double* c = qcbc_fit[w];
  double xi = -1 + chi;
  double xi2 = xi * xi;
  double xi3 = xi2 * xi;
  double eta2 = eta * eta;
  return c[0] + c[1] * eta + (c[2] + c[3] * eta + c[4] * eta2) * xi +
         (c[5] + c[6] * eta + c[7] * eta2) * xi2 + (c[8] + c[9] * eta + c[10] * eta2) * xi3;
}

// Natural cubic spline through the QNM table in Hermite (slope) form.  The
// knots are spaced 0.002 apart except the two outermost intervals at each
// end (0.001), so the interval follows from the abscissa directly.
__device__ inline double qnm_spline(int which, double x) {
  const double* ya = which ? qcbc_qnm_fdamp : qcbc_qnm_fring;
  int lo = 2 + (int)floor((x + 0.996) * 500.0);
  lo = lo < 0 ? 0 : (lo > QCBC_QNM_N - 2 ? QCBC_QNM_N - 2 : lo);
  if (lo > 0 && qcbc_qnm_a[lo] > x) lo--;
  if (lo > 0 && qcbc_qnm_a[lo] > x) lo--;
  if (lo < QCBC_QNM_N - 2 && qcbc_qnm_a[lo + 1] <= x) lo++;
  if (lo < QCBC_QNM_N - 2 && qcbc_qnm_a[lo + 1] <= x) lo++;
  double h = qcbc_qnm_a[lo + 1] - qcbc_qnm_a[lo];
  double t = (x - qcbc_qnm_a[lo]) / h;
  double m0 = g_qnm_slope[which][lo] * h, m1 = g_qnm_slope[which][lo + 1] * h;
  double y0 = ya[lo], y1 = ya[lo + 1];
  // cubic Hermite basis in nested form
  double dy = y1 - y0;
  double a = m0;
  double b = 3.0 * dy - 2.0 * m0 - m1;
  double c = m0 + m1 - 2.0 * dy;
  return y0 + t * (a + t * (b + t * c));
}

// Final spin and radiated energy fits, arXiv:1508.07250 Eqs. 3.6-3.8
// (IMRPhenomD_internals.c:415-473).
__device__ inline double final_spin(double eta, double chi1, double chi2) {
  const double* c = qcbc_finalspin;
  double Seta = sqrt(1.0 - 4.0 * eta);
  double m1 = 0.5 * (1.0 + Seta), m2 = 0.5 * (1.0 - Seta);
  double s = m1 * m1 * chi1 + m2 * m2 * chi2;
  double e2 = eta * eta, e3 = e2 * eta, e4 = e3 * eta;
  double s2 = s * s, s3 = s2 * s, s4 = s3 * s;
  return c[0] * eta + c[1] * e2 + c[2] * e3 + c[3] * e4 + (c[4] + c[5] * eta + c[6] * e2) * s +
         (c[7] * eta + c[8] * e2) * s2 + (c[9] * eta + c[10] * e2) * s3 + (c[11] * eta + c[12] * e2) * s4;
}

__device__ inline double erad(double eta, double chi1, double chi2) {
  const double* c = qcbc_erad;
  double Seta = sqrt(1.0 - 4.0 * eta);
  double m1 = 0.5 * (1.0 + Seta), m2 = 0.5 * (1.0 - Seta);
  double m1s = m1 * m1, m2s = m2 * m2;
  double s = (m1s * chi1 + m2s * chi2) / (m1s + m2s);
  double e2 = eta * eta, e3 = e2 * eta, e4 = e3 * eta;
  return ((c[0] * eta + c[1] * e2 + c[2] * e3 + c[3] * e4) * (c[4] + (c[5] + c[6] * eta + c[7] * e2) * s)) /
         (c[8] + (c[9] + c[10] * eta + c[11] * e2) * s);
}

// ----------------------------------------------------- frequency powers
// Powers of Mf shared by amplitude and phase (UsefulPowers,
// IMRPhenomD_internals.c:515-533) plus the inverse powers the phase needs.
struct Pow {
  double Mf, s6, third, two_thirds, four_thirds, five_thirds, two, seven_thirds, eight_thirds;
  double i_s6, i_third, i_two_thirds, i_Mf, i_five_thirds;
  double logMf, logv;
};

// Direct form: one pow() per frequency, as the reference.
__device__ __forceinline__ Pow pow_direct(double Mf) {
  Pow p;
  double x = Mf < 0.0 ? 1.0e-10 : Mf;
  p.Mf = Mf;
  p.s6 = pow(x, 1 / 6.0);
  p.third = p.s6 * p.s6;
  p.two_thirds = x / p.third;
  p.four_thirds = x * p.third;
  p.five_thirds = p.four_thirds * p.third;
  p.two = x * x;
  p.seven_thirds = p.third * p.two;
  p.eight_thirds = p.two_thirds * p.two;
  p.i_s6 = 1.0 / p.s6;
  p.i_third = 1.0 / p.third;
  p.i_two_thirds = 1.0 / p.two_thirds;
  p.i_Mf = 1.0 / Mf;
  p.i_five_thirds = 1.0 / p.five_thirds;
  p.logMf = log(Mf);
  p.logv = log(p.third * PI_THIRD);  // v = (pi Mf)^(1/3)
  return p;
}

// ------------------------------------------------------ amplitude pieces
// AmpInsAnsatz (IMRPhenomD_internals.c:601-610, arXiv:1508.07253 Eqs. 29-30)
__device__ __forceinline__ double amp_ins(const Point& c, const Pow& p) {
  return 1.0 + p.two_thirds * c.ai_23 + p.Mf * c.ai_1 + p.four_thirds * c.ai_43 + p.five_thirds * c.ai_53 +
         p.two * c.ai_2 + p.seven_thirds * c.ai_73 + p.eight_thirds * c.ai_83 + p.Mf * p.two * c.ai_3;
}
// AmpMRDAnsatz (:757-767, Eq. 19)
__device__ __forceinline__ double amp_mrd(const Point& c, double f) {
  double df = f - c.fRD;
  return exp(-df * c.mrd_rate) * c.mrd_num / (df * df + c.fDMg3 * c.fDMg3);
}
// AmpIntAnsatz (:816-821, Eq. 21), evaluated in Newton form.
__device__ __forceinline__ double amp_int(const Point& c, double f) {
  double x1 = f - c.col_f1;
  return c.col_c0 + x1 * (c.col_c1 + x1 * (c.col_c2 + (f - c.col_f2) * (c.col_c3 + (f - c.col_f3) * c.col_c4)));
}
// IMRPhenDAmplitude (:1107-1135) without the amp0 scale.
__device__ __forceinline__ double amplitude(const Point& c, const Pow& p) {
  double pre = c.amp_pre * p.i_Mf * p.i_s6;
  double f = p.Mf;
  if (!(f >= AMP_JOIN)) return pre * amp_ins(c, p);
  if (f >= c.fmaxCalc) return pre * amp_mrd(c, f);
  return pre * amp_int(c, f);
}

// ----------------------------------------------------------- phase pieces
// PhiInsAnsatzInt (:1399-1425, Eqs. 27-28)
__device__ __forceinline__ double phi_ins(const Point& c, const Pow& p) {
  double ph = c.pi_init;
  ph += c.pi_23 * p.two_thirds;
  ph += c.pi_13 * p.third;
  ph += c.pi_13L * p.logv * p.third;
  ph += c.pi_L * p.logv;
  ph += c.pi_m13 * p.i_third;
  ph += c.pi_m23 * p.i_two_thirds;
  ph += c.pi_m1 * p.i_Mf;
  ph += c.pi_m53 * p.i_five_thirds;
  ph += (c.pi_1 * p.Mf + c.pi_43 * p.four_thirds + c.pi_53 * p.five_thirds + c.pi_2 * p.two) * c.inv_eta;
  return ph;
}
// PhiIntAnsatz (:1298-1303, Eq. 16)
__device__ __forceinline__ double phi_int(const Point& c, double f, double logf) {
  return c.beta1 * f - c.beta3 / (3. * (f * f * f)) + c.beta2 * logf;
}
// PhiMRDAnsatzInt (:1222-1233, Eq. 14)
__device__ __forceinline__ double phi_mrd(const Point& c, double f) {
  double sq = sqrt(f);
  double f075 = sqrt(f * sq);
  return -(c.alpha2 / f) + (4.0 / 3.0) * (c.alpha3 * f075) + c.alpha1 * f + c.alpha4 * atan((f - c.a5fRD) / c.fDM);
}
// IMRPhenDPhase (:1590-1612), then the time/phase shift of IMRPhenomD.c:550.
__device__ __forceinline__ double phase(const Point& c, const Pow& p) {
  double f = p.Mf, ph;
  if (!(f >= PHI_JOIN))
    ph = phi_ins(c, p);
  else if (f >= c.fPhiMRD)
    ph = c.inv_eta * phi_mrd(c, f) + c.C1MRD + c.C2MRD * f;
  else
    ph = c.inv_eta * phi_int(c, f, p.logMf) + c.C1Int + c.C2Int * f;
  return ph - (c.t0 * (f - c.MfRef) + c.phi_precalc);
}

// ------------------------------------------------ detector geometry
// F_ant (Utilities.c:995-1076): the detectors' expanded polynomials,
// table-driven.  x holds the nine trigonometric factors of one sky position
// and polarisation (shared by all detectors).
__device__ inline void antenna_factors(double psi, double alpha, double sindelta, double gmst, double* x) {
  sincos(psi, &x[1], &x[0]);
  x[2] = sqrt(1.0 - sindelta * sindelta);
  x[3] = sindelta;
  sincos(alpha - gmst, &x[5], &x[4]);
  x[6] = x[0] * x[0];
  x[7] = x[3] * x[3];
  x[8] = x[4] * x[4];
}

// Polynomial K evaluated term by term in source order.  The tables are
// compile-time constants and every loop is unrolled, so each factor index
// folds to a register operand (a run-time index into x[] would put x in
// scratch memory).
template <int K>
__device__ __forceinline__ double antenna_poly(const double* x) {
  double s = 0.0;
#pragma unroll
  for (int t = 0; t < QCBC_ANT_MAXTERMS; t++) {
    if (t < qcbc_ant_nterms[K]) {
      double term = qcbc_ant_coef[K][t];
#pragma unroll
      for (int j = 0; j < QCBC_ANT_MAXFAC; j++)
        if (qcbc_ant_fac[K][t][j] >= 0) term *= x[qcbc_ant_fac[K][t][j]];
      s = (t == 0) ? term : s + term;
    }
  }
  return s;
}

__device__ __forceinline__ void antenna_id(int id, const double* x, double* fp, double* fc) {
  switch (id) {
    case 0: *fp = antenna_poly<0>(x); *fc = antenna_poly<1>(x); break;
    case 1: *fp = antenna_poly<2>(x); *fc = antenna_poly<3>(x); break;
    default: *fp = antenna_poly<4>(x); *fc = antenna_poly<5>(x); break;
  }
}

__device__ inline void antenna(double psi, double alpha, double sindelta, double gmst, int id, double* fp,
                               double* fc) {
  double x[9];
  antenna_factors(psi, alpha, sindelta, gmst, x);
  antenna_id(id, x, fp, fc);
}

// Times (Utilities.c:1078-1109): geocentre-to-detector delay per label.
__device__ inline double delay_cs(double cosd, double sindelta, double ca, double sa, int id) {
  double v0, v1, v2;
  switch (id) {
    case 0: v0 = qcbc_vertex[0][0]; v1 = qcbc_vertex[0][1]; v2 = qcbc_vertex[0][2]; break;
    case 1: v0 = qcbc_vertex[1][0]; v1 = qcbc_vertex[1][1]; v2 = qcbc_vertex[1][2]; break;
    default: v0 = qcbc_vertex[2][0]; v1 = qcbc_vertex[2][1]; v2 = qcbc_vertex[2][2]; break;
  }
  return -(v0 * cosd * ca + v1 * cosd * sa + v2 * sindelta) / CLIGHT;
}
__device__ inline double delay(double alpha, double sindelta, double gmst, int id) {
  double sa, ca;
  sincos(gmst - alpha, &sa, &ca);
  return delay_cs(sqrt(1.0 - sindelta * sindelta), sindelta, ca, sa, id);
}

// fbegin (Utilities.c:587-629): 3PN frequency at the segment start.
// fbegin (Utilities.c:587-629) given exp(prm[0]) and exp(prm[1]).
__device__ inline double fbegin_m(const double* prm, double mc_SI, double mt_SI) {
  const double gamma_E = 0.5772156649;
  double Mc = mc_SI / MSUN_SI * TSUN;
  double M = mt_SI / MSUN_SI * TSUN;
  double eta = pow_cr(Mc / M, 5.0 / 3.0);
  if (eta > 0.25) eta = 0.25;
  double dm = sqrt(1.0 - 4.0 * eta);
  double m1 = M * (1.0 + dm) / 2.0, m2 = M * (1.0 - dm) / 2.0;
  double chi1 = prm[2], chi2 = prm[3];
  double chi = (m1 * chi1 + m2 * chi2) / M;
  double th = pow_cr(eta * prm[5] / (5.0 * M), -1.0 / 8.0);
  double th2 = th * th, th3 = th2 * th, th4 = th2 * th2, th5 = th2 * th3, th6 = th3 * th3;
  double PN1 = (11.0 / 32.0 * eta + 743.0 / 2688.0) * th2;
  double PN15 = -3.0 * PI / 10.0 * th3 + (1.0 / 160.0) * (113.0 * chi - 38.0 * eta * (chi1 + chi2)) * th3;
  double PN2 = (1855099.0 / 14450688.0 + 56975.0 / 258048.0 * eta + 371.0 / 2048.0 * eta * eta) * th4 +
               (1.0 / 14450688.0) * (-3386880.0 * chi * chi + 1512.0 * chi1 * chi2) * th4;
  double PN25 = -(7729.0 / 21504.0 - 13.0 / 256.0 * eta) * PI * th5;
  double PN3 = (-720817631400877.0 / 288412611379200.0 + 53.0 / 200.0 * PI * PI + 107.0 / 280.0 * gamma_E +
                (25302017977.0 / 4161798144.0 - 451.0 / 2048.0 * PI * PI) * eta - 30913.0 / 1835008.0 * eta * eta +
                235925.0 / 1769472.0 * eta * eta * eta + 107.0 / 280.0 * dd_log(2.0 * th).hi) *
               th6;
  return th3 / (8.0 * M * PI) * (1.0 + PN1 + PN15 + PN2 + PN25 + PN3);
}
__device__ inline double fbegin(const double* prm) { return fbegin_m(prm, exp_cr(prm[0]), exp_cr(prm[1])); }

// --------------------------------------------------------- PN phasing
// TaylorF2 coefficients (PNPhasing_F2, IMRPhenomD_internals.c:271-365;
// arXiv:0907.0700 Eq. 3.18, arXiv:1303.7412 Eqs. 3.15-16), already scaled by
// pfaN, with the 3PN spin-spin term subtracted again (IMRPhenomD.c:493).
struct PN {
  double v[8], vl5, vl6;
};

__device__ inline double pn_ss3(double eta, double m1M, double m2M, double chi1, double chi2) {
  double ss = (326.75 / 1.12 + 557.5 / 1.8 * eta) * eta * chi1 * chi2;
  ss += ((4703.5 / 8.4 + 2935. / 6. * m1M - 120. * m1M * m1M) + (-4108.25 / 6.72 - 108.5 / 1.2 * m1M + 125.5 / 3.6 * m1M * m1M)) *
        m1M * m1M * (chi1 * chi1);
  ss += ((4703.5 / 8.4 + 2935. / 6. * m2M - 120. * m2M * m2M) + (-4108.25 / 6.72 - 108.5 / 1.2 * m2M + 125.5 / 3.6 * m2M * m2M)) *
        m2M * m2M * (chi2 * chi2);
  return ss;
}

__device__ inline PN pn_phasing(double m1, double m2, double chi1, double chi2) {
  PN pn;
  const double mtot = m1 + m2;
  const double d = (m1 - m2) / mtot;
  const double eta = m1 * m2 / mtot / mtot;
  const double m1M = m1 / mtot, m2M = m2 / mtot;
  const double SL = m1M * m1M * chi1 + m2M * m2M * chi2;
  const double dSL = d * (m2M * chi2 - m1M * chi1);
  const double pfaN = 3.0 / (128.0 * eta);
  const double e2 = eta * eta;
  double v0 = 1.0, v2, v3, v4, v5, v6, v7, vl5, vl6;
  v2 = 5.0 * (743.0 / 84.0 + 11.0 * eta) / 9.0;
  v3 = -16.0 * PI;
  v4 = 5.0 * (3058.673 / 7.056 + 5429.0 / 7.0 * eta + 617.0 * e2) / 72.0;
  v5 = 5.0 / 9.0 * (7729.0 / 84.0 - 13.0 * eta) * PI;
  vl5 = 5.0 / 3.0 * (7729.0 / 84.0 - 13.0 * eta) * PI;
  v6 = (11583.231236531 / 4.694215680 - 640.0 / 3.0 * PI * PI - 6848.0 / 21.0 * EULER_GAMMA) +
       eta * (-15737.765635 / 3.048192 + 2255. / 12. * PI * PI) + e2 * 76055.0 / 1728.0 - e2 * eta * 127825.0 / 1296.0;
  v6 += (-6848.0 / 21.0) * 1.3862943611198906;  // log(4)
  vl6 = -6848.0 / 21.0;
  v7 = PI * (77096675.0 / 254016.0 + 378515.0 / 1512.0 * eta - 74045.0 / 756.0 * e2);
  double sigma = eta * (721.0 / 48.0 * chi1 * chi2 - 247.0 / 48.0 * (chi1 * chi2));
  sigma += 719.0 / 96.0 * m1M * m1M * chi1 * chi1 + 719.0 / 96.0 * m2M * m2M * chi2 * chi2;
  sigma -= 233.0 / 96.0 * m1M * m1M * (chi1 * chi1) + 233.0 / 96.0 * m2M * m2M * (chi2 * chi2);
  double ss3 = pn_ss3(eta, m1M, m2M, chi1, chi2);
  double gam = (554345.0 / 1134.0 + 110.0 * eta / 9.0) * SL + (13915.0 / 84.0 - 10.0 * eta / 3.0) * dSL;
  v7 += (-8980424995.0 / 762048.0 + 6586595.0 * eta / 756.0 - 305.0 * e2 / 36.0) * SL -
        (170978035.0 / 48384.0 - 2876425.0 * eta / 672.0 - 4735.0 * e2 / 144.0) * dSL;
  v6 += PI * (3760.0 * SL + 1490.0 * dSL) / 3.0 + ss3;
  v5 += -gam;
  vl5 += -3.0 * gam;
  v4 += -10.0 * sigma;
  v3 += 188.0 * SL / 3.0 + 25.0 * dSL;
  pn.v[0] = v0 * pfaN;
  pn.v[1] = 0.0;
  pn.v[2] = v2 * pfaN;
  pn.v[3] = v3 * pfaN;
  pn.v[4] = v4 * pfaN;
  pn.v[5] = v5 * pfaN;
  pn.v[6] = v6 * pfaN;
  pn.v[7] = v7 * pfaN;
  pn.vl5 = vl5 * pfaN;
  pn.vl6 = vl6 * pfaN;
  pn.v[6] -= pn_ss3(m1 * m2 / (mtot * mtot), m1M, m2M, chi1, chi2) * pn.v[0];
  return pn;
}

// DPhiInsAnsatzInt (:1461-1501): dPhi_ins/dMf.
__device__ inline double dphi_ins(const Point& c, const PN& pn, double Mf, double pi_third, double pi_two_thirds,
                                  double s1, double s2, double s3, double s4) {
  const double v = cbrt(PI * Mf);
  const double logv = log(v);
  const double v2 = v * v, v3 = v * v2, v4 = v * v3, v5 = v * v4, v6 = v * v5, v7 = v * v6, v8 = v * v7;
  double D = 2.0 * pn.v[7] * v7 + (pn.v[6] + pn.vl6 * (1.0 + logv)) * v6 + pn.vl5 * v5 - pn.v[4] * v4 -
             2.0 * pn.v[3] * v3 - 3.0 * pn.v[2] * v2 - 5.0 * pn.v[0];
  D /= v8 * 3.0 / PI;
  D += (s1 + s2 * v / pi_third + s3 * v2 / pi_two_thirds + (s4 / PI) * v3) * c.inv_eta;
  return D;
}

__device__ __forceinline__ double dphi_int(const Point& c, double f) {
  double f2 = f * f;
  return (c.beta1 + c.beta3 / (f2 * f2) + c.beta2 / f) * c.inv_eta;
}

__device__ __forceinline__ double dphi_mrd(const Point& c, double f) {
  double x = f - c.a5fRD;
  return (c.alpha1 + c.alpha2 / (f * f) + c.alpha3 / pow(f, 0.25) + c.alpha4 / (c.fDM * (1 + (x * x) / (c.fDM * c.fDM)))) *
         c.inv_eta;
}

// d/dMf of AmpInsAnsatz, from its prefactors (the quantity of DAmpInsAnsatz,
// IMRPhenomD_internals.c:671-702).
__device__ inline double damp_ins(const Point& c, double Mf) {
  double t = cbrt(Mf);
  return (2.0 / 3.0) * c.ai_23 / t + c.ai_1 + (4.0 / 3.0) * c.ai_43 * t + (5.0 / 3.0) * c.ai_53 * t * t +
         2.0 * c.ai_2 * Mf + (7.0 / 3.0) * c.ai_73 * Mf * t + (8.0 / 3.0) * c.ai_83 * Mf * t * t + 3.0 * c.ai_3 * Mf * Mf;
}

// DAmpMRDAnsatz (:772-787)
__device__ inline double damp_mrd(const Point& c, double f, double gamma1, double gamma2, double gamma3) {
  double df = f - c.fRD;
  double ex = exp((df * gamma2) / c.fDMg3);
  double s = df * df + c.fDMg3 * c.fDMg3;
  return (-2 * c.fDM * df * gamma3 * gamma1) / (ex * (s * s)) - (gamma2 * gamma1) / (ex * s);
}

// The prefix both halves of IMRPhenomD's set-up need: the ordered masses and
// spins, eta, chi_PN and delta = sqrt(1 - 4 eta).
struct PDPre {
  double m1, m2, chi1, chi2, eta, chi, delta_m;
};

// IMRPhenomD's set-up in four parts: the prefix (domain checks, masses,
// scales, final spin, ringdown and damping frequencies), the amplitude
// coefficients, the phase coefficients with their C1 connection, and the
// time shift and reference phase, which need both halves.  phenomd_setup runs
// them in order in one thread; k_setup runs the two middle parts in two waves.
__device__ inline int phenomd_prefix(double m1_SI, double m2_SI, double chi1_in, double chi2_in, double distance,
                                     double fref, double fmin0, Point& c, PDPre& o) {
  // IMRPhenomD.c:197-242
  const double m1_in = m1_SI / MSUN_SI, m2_in = m2_SI / MSUN_SI;
  if (!(m1_in > 0) || !(m2_in > 0) || !(distance > 0) || !(fref >= 0) || !(fmin0 > 0)) return ST_EDOM;
  if (!(chi1_in <= 1.0 && chi1_in >= -1.0 && chi2_in <= 1.0 && chi2_in >= -1.0)) return ST_EDOM;
  if (F_CUT / ((m1_in + m2_in) * MTSUN_SI) <= fmin0) return ST_EDOM;
  // IMRPhenomD.c:429-439: m1 > m2
  double m1, m2, chi1, chi2;
  if (m1_in > m2_in) { m1 = m1_in; m2 = m2_in; chi1 = chi1_in; chi2 = chi2_in; }
  else { m1 = m2_in; m2 = m1_in; chi1 = chi2_in; chi2 = chi1_in; }
  const double M = m1 + m2;
  const double eta = m1 * m2 / (M * M);
  if (eta > 0.25 || eta < 0.0) return ST_EDOM;
  c.eta = eta;
  c.inv_eta = 1.0 / eta;
  c.Msec = M * MTSUN_SI;
  c.amp0 = M * MRSUN_SI * M * MTSUN_SI / distance;
  c.fCut = F_CUT / c.Msec;
  c.s6_scale = pow(c.Msec, 1.0 / 6.0);
  c.is6_scale = 1.0 / c.s6_scale;
  c.v_scale = cbrt(c.Msec);
  c.logMsec = log(c.Msec);
  // final state (IMRPhenomD_internals.c:415-513)
  const double finspin = final_spin(eta, chi1, chi2);
  c.finspin = finspin;
  if (!(finspin <= 1.0 && finspin >= -1.0)) return ST_EDOM;
  const double er = 1.0 - erad(eta, chi1, chi2);
  c.fRD = qnm_spline(0, finspin) / er;
  c.fDM = qnm_spline(1, finspin) / er;
  const double delta_m = sqrt(1.0 - 4.0 * eta);
  const double chi = ((chi1 + chi2) / 2.0) * (1.0 - eta * 76.0 / 113.0) + delta_m * ((chi1 - chi2) / 2.0);  // chiPN
  o = PDPre{m1, m2, chi1, chi2, eta, chi, delta_m};
  return ST_OK;
}

// fmaxCalc (IMRPhenomD_internals.c:793-808, set at :1090): the peak of the
// merger-ringdown amplitude, from the ringdown and damping frequencies and the
// fits gamma2, gamma3
__device__ __forceinline__ double fmax_calc(double fRD, double fDM, double g2, double g3) {
  return (g2 <= 1) ? fabs(fRD + (fDM * (-1 + sqrt(1 - g2 * g2)) * g3) / g2) : fabs(fRD + (fDM * (-1) * g3) / g2);
}

__device__ inline void phenomd_amp(const PDPre& o, Point& c) {
  const double chi1 = o.chi1, chi2 = o.chi2, eta = o.eta, chi = o.chi, delta_m = o.delta_m;
  const double pi_s6 = PI_S6, pi_two_thirds = PI_TWO_THIRDS;
  const double pi_four_thirds = PI_FOUR_THIRDS, pi_five_thirds = PI_FIVE_THIRDS;
  const double pi2 = PI * PI;
  // amplitude coefficients (:1072-1100)
  const double g1 = fit11(QCBC_FIT_GAMMA1, eta, chi), g2 = fit11(QCBC_FIT_GAMMA2, eta, chi),
               g3 = fit11(QCBC_FIT_GAMMA3, eta, chi);
  c.fDMg3 = c.fDM * g3;
  c.mrd_num = c.fDMg3 * g1;
  c.mrd_rate = g2 / c.fDMg3;
  c.fmaxCalc = fmax_calc(c.fRD, c.fDM, g2, g3);
  const double rho1 = fit11(QCBC_FIT_RHO1, eta, chi), rho2 = fit11(QCBC_FIT_RHO2, eta, chi),
               rho3 = fit11(QCBC_FIT_RHO3, eta, chi);
  // init_amp_ins_prefactors (:612-665)
  {
    const double c12 = chi1 * chi1, c22 = chi2 * chi2, e2 = eta * eta, e3 = eta * e2, S = delta_m;
    c.amp_pre = (sqrt(2.0 / 3.0) * sqrt(eta)) / pi_s6;
    c.ai_23 = ((-969 + 1804 * eta) * pi_two_thirds) / 672.;
    c.ai_1 = ((chi1 * (81 * (1 + S) - 44 * eta) + chi2 * (81 - 81 * S - 44 * eta)) * PI) / 48.;
    c.ai_43 = ((-27312085.0 - 10287648 * c22 - 10287648 * c12 * (1 + S) + 10287648 * c22 * S +
                24 * (-1975055 + 857304 * c12 - 994896 * chi1 * chi2 + 857304 * c22) * eta + 35371056 * e2) *
               pi_four_thirds) / 8.128512e6;
    c.ai_53 = (pi_five_thirds * (chi2 * (-285197 * (-1 + S) + 4 * (-91902 + 1579 * S) * eta - 35632 * e2) +
                                 chi1 * (285197 * (1 + S) - 4 * (91902 + 1579 * S) * eta - 35632 * e2) +
                                 42840 * (-1.0 + 4 * eta) * PI)) / 32256.;
    c.ai_2 = -(pi2 * (-336 * (-3248849057.0 + 2943675504 * c12 - 3339284256 * chi1 * chi2 + 2943675504 * c22) * e2 -
                      324322727232 * e3 -
                      7 * (-177520268561 + 107414046432 * c22 + 107414046432 * c12 * (1 + S) - 107414046432 * c22 * S +
                           11087290368 * (chi1 + chi2 + chi1 * S - chi2 * S) * PI) +
                      12 * eta * (-545384828789 - 176491177632 * chi1 * chi2 + 202603761360 * c22 +
                                  77616 * c12 * (2610335 + 995766 * S) - 77287373856 * c22 * S +
                                  5841690624 * (chi1 + chi2) * PI + 21384760320 * pi2))) / 6.0085960704e10;
    c.ai_73 = rho1;
    c.ai_83 = rho2;
    c.ai_3 = rho3;
  }
  // collocation of the intermediate amplitude (:1005-1063): the quartic with
  // values v1, v2, v3 at f1, f2, f3 and slopes d1 at f1, d2 at f3, built by
  // Hermite divided differences on the nodes (f1, f1, f2, f3, f3).
  {
    const double f1 = AMP_JOIN, f3 = c.fmaxCalc, f2 = f1 + (f3 - f1) / 2.0;
    c.col_f1 = f1; c.col_f2 = f2; c.col_f3 = f3;
    Pow p1 = pow_direct(f1);
    const double v1 = amp_ins(c, p1);
    const double d1 = damp_ins(c, f1);
    const double v3 = amp_mrd(c, f3);
    const double d2 = damp_mrd(c, f3, g1, g2, g3);
    const double v2 = fit11(QCBC_FIT_V2, eta, chi);
    if (f3 > f1) {
      double a01 = d1, a12 = (v2 - v1) / (f2 - f1), a23 = (v3 - v2) / (f3 - f2), a34 = d2;
      double b02 = (a12 - a01) / (f2 - f1), b13 = (a23 - a12) / (f3 - f1), b24 = (a34 - a23) / (f3 - f2);
      double c03 = (b13 - b02) / (f3 - f1), c14 = (b24 - b13) / (f3 - f1);
      double d04 = (c14 - c03) / (f3 - f1);
      c.col_c0 = v1; c.col_c1 = a01; c.col_c2 = b02; c.col_c3 = c03; c.col_c4 = d04;
    } else {  // the intermediate region is empty
      c.col_c0 = c.col_c1 = c.col_c2 = c.col_c3 = c.col_c4 = 0.0;
    }
  }
}

__device__ inline void phenomd_phase(const PDPre& o, Point& c) {
  const double m1 = o.m1, m2 = o.m2, chi1 = o.chi1, chi2 = o.chi2, eta = o.eta, chi = o.chi;
  const double pi_third = PI_THIRD, pi_two_thirds = PI_TWO_THIRDS, pi_five_thirds = PI_FIVE_THIRDS;
  // phase coefficients (:1509-1540)
  const double sg1 = fit11(QCBC_FIT_SIGMA1, eta, chi), sg2 = fit11(QCBC_FIT_SIGMA2, eta, chi),
               sg3 = fit11(QCBC_FIT_SIGMA3, eta, chi), sg4 = fit11(QCBC_FIT_SIGMA4, eta, chi);
  c.beta1 = fit11(QCBC_FIT_BETA1, eta, chi);
  c.beta2 = fit11(QCBC_FIT_BETA2, eta, chi);
  c.beta3 = fit11(QCBC_FIT_BETA3, eta, chi);
  c.alpha1 = fit11(QCBC_FIT_ALPHA1, eta, chi);
  c.alpha2 = fit11(QCBC_FIT_ALPHA2, eta, chi);
  c.alpha3 = fit11(QCBC_FIT_ALPHA3, eta, chi);
  c.alpha4 = fit11(QCBC_FIT_ALPHA4, eta, chi);
  c.a5fRD = fit11(QCBC_FIT_ALPHA5, eta, chi) * c.fRD;
  const PN pn = pn_phasing(m1, m2, chi1, chi2);
  // init_phi_ins_prefactors (:1427-1456)
  c.pi_init = pn.v[5] - PI_4;
  c.pi_23 = pn.v[7] * pi_two_thirds;
  c.pi_13 = pn.v[6] * pi_third;
  c.pi_13L = pn.vl6 * pi_third;
  c.pi_L = pn.vl5;
  c.pi_m13 = pn.v[4] / pi_third;
  c.pi_m23 = pn.v[3] / pi_two_thirds;
  c.pi_m1 = pn.v[2] / PI;
  c.pi_m53 = pn.v[0] / pi_five_thirds;
  c.pi_1 = sg1;
  c.pi_43 = sg2 * 3.0 / 4.0;
  c.pi_53 = sg3 * 3.0 / 5.0;
  c.pi_2 = sg4 / 2.0;
  // C1 connection (:1547-1584)
  {
    const double fj = PHI_JOIN;
    c.C2Int = dphi_ins(c, pn, fj, pi_third, pi_two_thirds, sg1, sg2, sg3, sg4) - dphi_int(c, fj);
    Pow pj = pow_direct(fj);
    c.C1Int = phi_ins(c, pj) - c.inv_eta * phi_int(c, fj, log(fj)) - c.C2Int * fj;
    c.fPhiMRD = 0.5 * c.fRD;
    const double fm = c.fPhiMRD;
    const double phiT = c.inv_eta * phi_int(c, fm, log(fm)) + c.C1Int + c.C2Int * fm;
    const double dphiT = c.C2Int + dphi_int(c, fm);
    c.C2MRD = dphiT - dphi_mrd(c, fm);
    c.C1MRD = phiT - c.inv_eta * phi_mrd(c, fm) - c.C2MRD * fm;
  }
}

// time shift and reference phase (IMRPhenomD.c:504-521) with fmaxCalc given:
// reads only the prefix and phase fields of c
__device__ inline void phenomd_refphase_at(double phi0, double fref, double fmax, Point& c) {
  c.t0 = dphi_mrd(c, fmax);
  c.MfRef = c.Msec * fref;
  c.phi_precalc = 0.0;
  {
    Pow pr = pow_direct(c.MfRef);
    c.phi_precalc = 2. * phi0 + phase(c, pr);  // phase() subtracts t0*0 + 0 here
  }
}

__device__ inline void phenomd_refphase(double phi0, double fref, Point& c) {
  phenomd_refphase_at(phi0, fref, c.fmaxCalc, c);
}

__device__ inline int phenomd_setup(double m1_SI, double m2_SI, double chi1_in, double chi2_in, double distance,
                                    double phi0, double fref, double fmin0, Point& c) {
  PDPre o;
  const int st = phenomd_prefix(m1_SI, m2_SI, chi1_in, chi2_in, distance, fref, fmin0, c, o);
  if (st != ST_OK) return st;
  phenomd_amp(o, c);
  phenomd_phase(o, c);
  phenomd_refphase(phi0, fref, c);
  return ST_OK;
}

// Smallest k in [lo, hi) with Msec (k / Tobs) >= thr, else hi: the region
// tests of amplitude() and phase() on a bin's own Mf = Msec f_k, with f_k
// as the band tables hold it; Mf does not decrease with k.
__device__ inline int first_bin_at(double thr, double Msec, double Tobs, int lo, int hi) {
  const double e = ceil(thr / Msec * Tobs);
  int k = !(e == e) ? hi : (e < (double)lo ? lo : (e > (double)hi ? hi : (int)e));
  while (k > lo && Msec * ((double)(k - 1) / Tobs) >= thr) k--;
  while (k < hi && !(Msec * ((double)k / Tobs) >= thr)) k++;
  return k;
}

// Constants of the fused likelihood kernels: region boundaries as bin
// indices, and the amplitude and the projected argument with every
// per-point constant folded in (the projected detector-0 argument
// f tshift_0 + (lambda_0 - Psi)/2 pi in turns, Psi = phi(Mf) - t0 (Mf - MfRef)
// - phi_precalc as phase() evaluates it; the kernels add the f tshift_0 term,
// up to ~3e4 turns, with one rounding).
__device__ inline void fused_setup(const AnalysisConst& an, Point& c) {
  c.seg_a1 = first_bin_at(AMP_JOIN, c.Msec, an.Tobs, c.kstart, c.kend);
  c.seg_a2 = first_bin_at(c.fmaxCalc, c.Msec, an.Tobs, c.kstart, c.kend);
  c.seg_p1 = first_bin_at(PHI_JOIN, c.Msec, an.Tobs, c.kstart, c.kend);
  c.seg_p2 = first_bin_at(c.fPhiMRD, c.Msec, an.Tobs, c.kstart, c.kend);
  const double K = -INV_TPI;
  const double G = (LOG_PI + c.logMsec) / 3.0;  // log v = G + log(f) / 3
  const double base = (c.lambda[0] + c.phi_precalc - c.t0 * c.MfRef) * INV_TPI;
  // Powers of Mf in bin units: Mf^(1/3) = s u and Mf^(-1/3) = sg^2 E with
  // u = f^(1/3), E = f^(-1/3) = e^2, e = f^(-1/6) (the band planes), s =
  // Msec^(1/3), sg = Msec^(-1/6); each coefficient below carries its power
  // of s or sg, so the kernels evaluate the polynomials in u and E directly.
  const double sg = c.is6_scale, sg2 = sg * sg, sg3 = sg2 * sg, sg4 = sg2 * sg2, sg6 = sg3 * sg3;
  const double sg7 = sg6 * sg, sg10 = sg6 * sg4, sg18 = sg6 * sg6 * sg6;
  const double s1 = c.v_scale, s2 = s1 * s1, s4 = s2 * s2, s5 = s4 * s1, s6 = s4 * s2;
  // inspiral (PhiInsAnsatzInt)
  c.ins_q0 = base - (c.pi_init + c.pi_L * G) * INV_TPI;
  c.ins_q1 = (c.t0 - c.pi_1 * c.inv_eta) * INV_TPI;
  c.ins_q13 = K * (c.pi_13 + c.pi_13L * G) * s1;
  c.ins_q13L = K * c.pi_13L / 3.0 * s1;
  c.ins_qL = K * c.pi_L / 3.0;
  c.ins_q23 = K * c.pi_23 * s2;
  c.ins_q43 = K * c.inv_eta * c.pi_43 * s4;
  c.ins_q53 = K * c.inv_eta * c.pi_53 * s5;
  c.ins_q2 = K * c.inv_eta * c.pi_2 * s6;
  c.ins_qm13 = K * c.pi_m13 * sg2;
  c.ins_qm23 = K * c.pi_m23 * sg4;
  c.ins_qm1 = K * c.pi_m1 * sg6;
  c.ins_qm53 = K * c.pi_m53 * sg10;
  // intermediate: inv_eta (beta1 Mf - beta3 / (3 Mf^3) + beta2 log Mf) + C1Int + C2Int Mf
  c.int_qL = K * c.inv_eta * c.beta2;
  c.int_q0 = base - c.C1Int * INV_TPI + c.int_qL * c.logMsec;
  c.int_q1 = (c.t0 - c.C2Int - c.inv_eta * c.beta1) * INV_TPI;
  c.int_qm3 = c.inv_eta * c.beta3 * INV_TPI / 3.0 * sg18;
  // merger-ringdown: inv_eta (-alpha2 / Mf + 4/3 alpha3 Mf^(3/4) + alpha1 Mf
  //                  + alpha4 atan((Mf - a5fRD) / fDM)) + C1MRD + C2MRD Mf
  c.mrd_q0 = base - c.C1MRD * INV_TPI;
  c.mrd_q1 = (c.t0 - c.C2MRD - c.inv_eta * c.alpha1) * INV_TPI;
  c.mrd_qm1 = c.inv_eta * c.alpha2 * INV_TPI;
  c.mrd_q34 = K * c.inv_eta * (4.0 / 3.0) * c.alpha3;
  c.mrd_qat = K * c.inv_eta * c.alpha4;
  c.mrd_a5 = c.a5fRD;
  c.mrd_ifDM = 1.0 / c.fDM;
  // amplitude, amp_pre and the bin-unit powers folded in: inspiral
  // Mf^(-7/6) amp_pre (1 + ai_23 Mf^(2/3) + ... + ai_3 Mf^3) = e [E (ia0 E^2 +
  // ia23) + ia1 + ia43 u + ia53 u^2 + ia2 u^3 + ia73 u^4 + ia83 u^5 + ia3 u^6];
  // intermediate Mf^(-7/6) amp_pre N(Mf) = e E^3 N'(f) with the Newton form
  // N' on the nodes col_f* in Hz, rewritten as a quartic in x1 = f - nf1:
  //   N' = ib0 + x1 [ib1 + x1 [ib2 + (x1 + d2) [ib3 + ib4 (x1 + d3)]]],
  //   d2 = f1 - f2, d3 = f1 - f3 (nodes in Hz), so the x1^2 and x1^3 coefficients are
  //   ib2 + d2 (ib3 + ib4 d3) and ib3 + ib4 (d2 + d3)
  c.ia0 = c.amp_pre * sg7;
  c.ia23 = c.amp_pre * c.ai_23 * sg3;
  c.ia1 = c.amp_pre * c.ai_1 * sg;
  c.ia43 = c.amp_pre * c.ai_43 * sg * s1;
  c.ia53 = c.amp_pre * c.ai_53 * sg * s2;
  c.ia2 = c.amp_pre * c.ai_2 * sg * (s2 * s1);
  c.ia73 = c.amp_pre * c.ai_73 * sg * s4;
  c.ia83 = c.amp_pre * c.ai_83 * sg * s5;
  c.ia3 = c.amp_pre * c.ai_3 * sg * s6;
  const double M = c.Msec;
  c.ib0 = c.amp_pre * c.col_c0 * sg7;
  c.ib1 = c.amp_pre * c.col_c1 * sg7 * M;
  c.ib2 = c.amp_pre * c.col_c2 * sg7 * (M * M);
  c.ib3 = c.amp_pre * c.col_c3 * sg7 * (M * M * M);
  c.ib4 = c.amp_pre * c.col_c4 * sg7 * ((M * M) * (M * M));
  c.nf1 = c.col_f1 / M;
  {
    const double d2 = c.nf1 - c.col_f2 / M, d3 = c.nf1 - c.col_f3 / M;
    c.im2 = c.ib2 + d2 * (c.ib3 + c.ib4 * d3);
    c.im3 = c.ib3 + c.ib4 * (d2 + d3);
  }
  c.mnum = c.amp_pre * c.mrd_num;
  c.mg2 = c.fDMg3 * c.fDMg3;
  for (int d = 0; d < 3; d++) {
    const bool on = d < an.n_det;
    c.dt[d] = on ? c.tshift[d] - c.tshift[0] : 0.0;
    c.dl[d] = on ? (c.lambda[d] - c.lambda[0]) * INV_TPI : 0.0;
    sincospi(128.0 * an.inv_T * c.dt[d], &c.ws[d], &c.wc[d]);
  }
}

// fulltemplates' parameter handling (QuickCBC.c:3700-3718) and IMRPhenomD's
// set-up for one 11-entry parameter vector.
__device__ inline int point_setup_model(const double* prm, const AnalysisConst& an, Point& c) {
  c.status = ST_OK;
  c.kstart = c.kend = an.imin;
  const double mc = exp_cr(prm[0]), mt = exp_cr(prm[1]);
  const double eta_in = pow_cr(mc / mt, 5.0 / 3.0);
  const double dm = eta_in > 0.25 ? 0.0 : sqrt(1.0 - 4.0 * eta_in);
  const double m1_SI = mt * (1.0 + dm) / 2.0, m2_SI = mt * (1.0 - dm) / 2.0;
  int st = phenomd_setup(m1_SI, m2_SI, prm[2], prm[3], exp_cr(prm[6]), 0.5 * prm[4], an.fref, an.fmin0, c);
  if (st != ST_OK) c.status = st;
  return st;
}

// The same in halves for two threads: each converts the masses and runs the
// prefix (both write the prefix fields, with the same values), then half 0
// the amplitude coefficients and half 1 the phase coefficients.  After both,
// point_setup_refphase completes the model.
__device__ inline int point_setup_model_half(const double* prm, const AnalysisConst& an, Point& c, int half) {
  c.status = ST_OK;
  c.kstart = c.kend = an.imin;
  const double mc = exp_cr(prm[0]), mt = exp_cr(prm[1]);
  const double eta_in = pow_cr(mc / mt, 5.0 / 3.0);
  const double dm = eta_in > 0.25 ? 0.0 : sqrt(1.0 - 4.0 * eta_in);
  const double m1_SI = mt * (1.0 + dm) / 2.0, m2_SI = mt * (1.0 - dm) / 2.0;
  PDPre o;
  const int st = phenomd_prefix(m1_SI, m2_SI, prm[2], prm[3], exp_cr(prm[6]), an.fref, an.fmin0, c, o);
  if (st != ST_OK) {
    c.status = st;
    return st;
  }
  if (half == 0) phenomd_amp(o, c);
  else phenomd_phase(o, c);
  return ST_OK;
}

__device__ inline void point_setup_refphase(const double* prm, const AnalysisConst& an, Point& c) {
  phenomd_refphase(0.5 * prm[4], an.fref, c);
}

// Half 1 of the model (the prefix and the phase coefficients) followed by the
// time shift and reference phase, with fmaxCalc recomputed from the same fits
// as half 0 computes it (the same operations on the same values): after half
// 0, fbegin and the projection the record holds everything the heterodyned
// likelihood reads except the final detector scales (Ascale amp0 / sqrt(Tobs),
// which its evaluation applies).
__device__ inline int point_setup_phase_ref(const double* prm, const AnalysisConst& an, Point& c) {
  c.status = ST_OK;
  c.kstart = c.kend = an.imin;
  const double mc = exp_cr(prm[0]), mt = exp_cr(prm[1]);
  const double eta_in = pow_cr(mc / mt, 5.0 / 3.0);
  const double dm = eta_in > 0.25 ? 0.0 : sqrt(1.0 - 4.0 * eta_in);
  const double m1_SI = mt * (1.0 + dm) / 2.0, m2_SI = mt * (1.0 - dm) / 2.0;
  PDPre o;
  const int st = phenomd_prefix(m1_SI, m2_SI, prm[2], prm[3], exp_cr(prm[6]), an.fref, an.fmin0, c, o);
  if (st != ST_OK) {
    c.status = st;
    return st;
  }
  phenomd_phase(o, c);
  const double fmax =
      fmax_calc(c.fRD, c.fDM, fit11(QCBC_FIT_GAMMA2, o.eta, o.chi), fit11(QCBC_FIT_GAMMA3, o.eta, o.chi));
  phenomd_refphase_at(0.5 * prm[4], an.fref, fmax, c);
  return ST_OK;
}

// The projection constants (QuickCBC.c:3734-3750); Ascale without amp0,
// which point_setup_finish applies.
__device__ inline void point_setup_projection(const double* prm, const AnalysisConst& an, Point& c) {
  const double ts = an.Tobs - prm[5];
  const double alpha = prm[7], sindelta = prm[8], psi = prm[9], ciota = prm[10];
  const double Ap = (1.0 + ciota * ciota) / 2.0, Ac = -ciota;
  double x[9];
  antenna_factors(psi, alpha, sindelta, an.gmst, x);
  // Times uses cos/sin(gmst - alpha) (Utilities.c:1090-1091): the same
  // angle with the opposite sign as the antenna patterns' alpha - gmst
  const double sa = -x[5], ca = x[4];
  for (int d = 0; d < 3; d++) {
    c.Fs[d] = c.tshift[d] = c.lambda[d] = c.Ascale[d] = 0.0;
    if (d >= an.n_det) continue;
    int id = an.labels[d];
    double Fp, Fc;
    antenna_id(id, x, &Fp, &Fc);
    double Fs = sqrt(Ap * Ap * Fp * Fp + Ac * Ac * Fc * Fc);
    double lam = atan2(Ac * Fc, Ap * Fp);
    if (lam < 0.0) lam += TPI;
    c.Fs[d] = Fs;
    c.lambda[d] = lam;
    c.tshift[d] = ts - delay_cs(x[2], sindelta, ca, sa, id);
    c.Ascale[d] = Fs * H22FAC;
  }
}

// templates' projection (QuickCBC.c:3571-3607), the reference-frame form of
// MCMC_intrinsic: detector 0 unprojected (amplitude ratio 1, phase and time
// offsets 0), detector j > 0 with amplitude ratio, phase offset and time
// offset q[7 + 3 (j-1) + {2, 0, 1}] relative to it; A = Ar h22fac amp / sqrt(Tobs)
// and x = 2 pi f (ts - td) + pd - Psi, so Ar, pd and td take the places of
// Fs, lambda and the light-travel delay.
__device__ inline void point_setup_projection_offsets(const double* q, const AnalysisConst& an, Point& c) {
  const double ts = an.Tobs - q[5];
  for (int d = 0; d < 3; d++) {
    c.Fs[d] = c.tshift[d] = c.lambda[d] = c.Ascale[d] = 0.0;
    if (d >= an.n_det) continue;
    double Ar = 1.0, pd = 0.0, td = 0.0;
    if (d == 1) {
      pd = q[7];
      td = q[8];
      Ar = q[9];
    } else if (d == 2) {
      pd = q[10];
      td = q[11];
      Ar = q[12];
    }
    c.Fs[d] = Ar;
    c.lambda[d] = pd;
    c.tshift[d] = ts - td;
    c.Ascale[d] = Ar * H22FAC;
  }
}

// After the model, fbegin (c.fs) and the projection: the detector scales,
// the contributing bins and the fused kernels' constants.
__device__ inline void point_setup_finish(const AnalysisConst& an, Point& c) {
  for (int d = 0; d < an.n_det && d < 3; d++) c.Ascale[d] = c.Ascale[d] * c.amp0 / an.sqT;
  // bins that contribute: f_k > fs (QuickCBC.c:3758) and f_k <= fCut
  // (IMRPhenomD.c:530), inside the band; f_k = k / Tobs as QuickCBC.c:419.
  {
    int lo = an.imin, hi = an.imax;
    double kf = floor(c.fs * an.Tobs) - 1.0;
    if (kf != kf) kf = (double)hi;  // fs is NaN: no bin passes f > fs
    int k0 = kf > (double)hi ? hi : (kf > (double)lo ? (int)kf : lo);
    while (k0 < hi && !((double)k0 / an.Tobs > c.fs)) k0++;
    double kc = floor(c.fCut * an.Tobs) + 2.0;
    int k1 = kc < (double)hi ? (int)kc : hi;
    if (k1 < k0) k1 = k0;
    while (k1 > k0 && (double)(k1 - 1) / an.Tobs > c.fCut) k1--;
    c.kstart = k0;
    c.kend = k1;
  }
  fused_setup(an, c);
}

// fulltemplates' parameter handling and projection constants
// (QuickCBC.c:3700-3750) for one 11-entry parameter vector, in one thread.
// k_setup runs the model, fbegin and the projection in three waves instead.
__device__ inline int point_setup(const double* prm, const AnalysisConst& an, Point& c) {
  int st = point_setup_model(prm, an, c);
  c.fs = fbegin(prm);
  if (st != ST_OK) return st;
  point_setup_projection(prm, an, c);
  point_setup_finish(an, c);
  return ST_OK;
}

}  // namespace qcbc
