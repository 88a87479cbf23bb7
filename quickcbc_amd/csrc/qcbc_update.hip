// QuickCBC's own PT step on the GPU, compiled with -ffp-contract=off: the
// proposal mix of update() (QuickCBC.c:2238-2405) and the step of MCMC_all
// (QuickCBC.c:2029-2092) for an ensemble of independent chain columns, each
// column one of the reference's temperature ladders.
//
// Per column and step s, as MCMC_all's iteration: with probability 0.2 (more
// than one temperature) one swap of a random adjacent pair of temperatures,
// otherwise update() at every temperature: the proposal y made at the end of
// step s - 1, set up and its heterodyned logL evaluated at the start of step
// s, is accepted when
//   H = (logL_y - logL_x)/T + pMcM(y) + 3 lnDL_y - 3 lnDL_x - pMcM(x) > log u
// and y lies in the prior (bounds, eta <= 1/4, eta >= etamin), with
// pMcM = 2 ln Mt + ln eta - ln(1 - 4 eta)/2 the Jacobian that makes the prior
// flat in m1, m2 (:2345-2358).  On update steps with s % hist_every == 0 the
// column's states enter its history (:2071-2080).  Then the proposals of
// step s + 1, by update()'s thresholds on one uniform alpha:
//   alpha > a       Fisher jump: y = x + sqrt(T) ej[i] N(0,1) ev[i], i uniform in 0..10
//   alpha > b       differential evolution (de_jump, :2407-2425): y = x + g (h_i - h_j),
//                   i != j from the temperature's history, g = N(0, 0.5) with
//                   probability 0.9, else 1
//   alpha > c       jiggle: y = x + 0.01 10^-floor(3u) sqrt(T) N(0,1) per parameter
//   otherwise       the sky ring (:2293-2330): a new sky location on the ring of
//                   constant time delay between two detectors (Ring_all), the
//                   extrinsics that keep both detectors' signals (skymap_all), the
//                   log density of that map (skydensity_all) added to H, and phase
//                   and polarisation shifted by half a period with probability 1/2
// then the angles wrapped into range (:2336-2342).
//
// Random numbers are counter-based (SplitMix64 of (seed, step, point,
// draw)), the point number p = t n_chains + c whatever the storage order,
// so every number of column groups and every graph replay gives the same
// chains, and the oracle replays them from the same counters.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "qcbc_device.h"
#include "qcbc_het_device.h"
#include "qcbc_update_internal.h"
#include "qcbc_wrap.h"

namespace qcbc {

namespace {

constexpr double TPI_U = 6.283185307179586476925286766559;
constexpr double PI_U = 3.141592653589793238462643383279;

__device__ __forceinline__ uint64_t mix(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
// uniform in (0, 1) from the counters: 53 random bits, centred in their cell
__device__ __forceinline__ double uni(uint64_t seed, uint64_t salt, uint64_t a, uint64_t k) {
  const uint64_t r = mix((seed ^ salt) ^ mix(a * 0x2545F4914F6CDD1DULL + k));
  return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
// standard normal by Box-Muller from draws k and k + 1
__device__ __forceinline__ double gauss(uint64_t seed, uint64_t salt, uint64_t a, uint64_t k) {
  const double u1 = uni(seed, salt, a, k), u2 = uni(seed, salt, a, k + 1);
  return sqrt(-2.0 * log(u1)) * cos(TPI_U * u2);
}

// update()'s prior checks (:2345-2363) and the mass Jacobian; returns true
// when y is outside the prior
__device__ bool prior_flag(const double* y, const UpdateConst& u, double* pMcM) {
  bool flag = false;
  for (int i = 0; i < 11; i++)  // (a NaN fails too)
    if (!(y[i] <= u.max[i] && y[i] >= u.min[i])) flag = true;
  const double leta = (5.0 / 3.0) * (y[0] - y[1]);
  if (leta > log(0.25)) flag = true;
  *pMcM = 0.0;
  if (!flag) {
    const double eta = exp(leta);
    *pMcM = 2.0 * y[1] + leta - 0.5 * log(1.0 - 4.0 * eta);
    if (eta < u.etamin) flag = true;
  }
  return flag;
}

// the Jacobian term of a state already in the prior (eta <= 1/4)
__device__ double mass_jacobian(const double* x) {
  const double leta = (5.0 / 3.0) * (x[0] - x[1]);
  const double eta = exp(leta);
  return 2.0 * x[1] + leta - 0.5 * log(1.0 - 4.0 * eta);
}

// ---------------------------------------------------------------- sky ring
// The parameters a sky-ring move changes: phase [4], tc [5], ln DL [6],
// alpha [7], sin delta [8], psi [9], cos iota [10], passed by value so that
// nothing of the (rarely taken) sky path lives in memory.
struct Ext {
  double p4, p5, p6, p7, p8, p9, p10;
};

// Ring_all (Utilities.c:1445-1594): y's sky location rotated by 2 pi u_rot
// about the baseline of detectors d1, d2 (the vertex vectors of Times, in s).
// Called by all lanes of a point's segment (lane j): lane 0 takes the cosine
// and sine of alpha - gmst and lane 1 those of the rotation angle, in one
// evaluation, and the segment shares them.
__device__ Ext ring_all(Ext x, int d1, int d2, double gmst, double u_rot, int j) {
  double zv[3], uv[3], vv[3], kv[3], kvr[3];
  double zmag = 0.0;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    double a = 0.0, b = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {  // vertex rows picked without a run-time index
      if (d == d1) a = qcbc_vertex[d][i] / CLIGHT;
      if (d == d2) b = qcbc_vertex[d][i] / CLIGHT;
    }
    zv[i] = a - b;
  }
#pragma unroll
  for (int i = 0; i < 3; i++) zmag += zv[i] * zv[i];
  zmag = sqrt(zmag);
#pragma unroll
  for (int i = 0; i < 3; i++) zv[i] /= zmag;
  const double sind = x.p8, cosd = sqrt(1.0 - sind * sind);
  double rot = TPI_U * u_rot;
  const double arg = j == 1 ? rot : x.p7 - gmst;
  const double cv = cos(arg), sv = sin(arg);
  const double calpha = __shfl(cv, 0, 16), salpha = -__shfl(sv, 0, 16);
  const double cr = __shfl(cv, 1, 16), sr = __shfl(sv, 1, 16);
  kv[0] = cosd * calpha;
  kv[1] = cosd * salpha;
  kv[2] = sind;
  double cmu = 0.0;
#pragma unroll
  for (int i = 0; i < 3; i++) cmu += zv[i] * kv[i];
  if (cmu > 0.999999) cmu = 0.999999;
  if (cmu < -0.999999) cmu = -0.999999;
  const double smu = sqrt(1.0 - cmu * cmu);
#pragma unroll
  for (int i = 0; i < 3; i++) uv[i] = (kv[i] - cmu * zv[i]) / smu;
  vv[0] = (zv[1] * kv[2] - zv[2] * kv[1]) / smu;
  vv[1] = (zv[2] * kv[0] - zv[0] * kv[2]) / smu;
  vv[2] = (zv[0] * kv[1] - zv[1] * kv[0]) / smu;
#pragma unroll
  for (int i = 0; i < 3; i++) kvr[i] = cmu * zv[i] + smu * (cr * uv[i] + sr * vv[i]);
  Ext y = x;
  y.p8 = kvr[2];
  rot = atan2(kvr[1], kvr[0]);
  rot = gmst - rot;
  if (rot < 0.0) rot += TPI_U;
  y.p7 = rot;
  return y;
}

// What skymap_all needs besides x's extrinsics: the antenna patterns of
// detectors ifo1, ifo2 at x's and at y's sky location, detector ifo1's delay
// at both, and for x the cosine and sine of phi and of 2 psi and exp(ln DL).
struct SkyGeo {
  double fxp1, fxc1, fxp2, fxc2, fyp1, fyc1, fyp2, fyc2, dtx, dty;
  double cphi, sphi, c2p, s2p, eDL;
};

// skymap_all (QuickCBC.c:7085-7146) with uvwz_all (:7295-7323), uvwz_sol
// (:7526-7544) and exsolve_all (:7325-7407): the extrinsics at y's sky
// location that keep the signal in detectors ifo1 and ifo2, given the
// geometry g of x and y
__device__ Ext skymap_core(Ext x, Ext y, const SkyGeo& g) {
  const double fxp1 = g.fxp1, fxc1 = g.fxc1, fxp2 = g.fxp2, fxc2 = g.fxc2;
  const double fyp1 = g.fyp1, fyc1 = g.fyc1, fyp2 = g.fyp2, fyc2 = g.fyc2;
  const double dtx = g.dtx, dty = g.dty;
  // uvwz_all
  const double ciota = x.p10, Ap = 0.5 * (1.0 + ciota * ciota), Ac = -ciota;
  const double ecc0 = Ac / Ap, Amp = Ap / g.eDL;
  const double cphi = g.cphi, sphi = g.sphi, c2p = g.c2p, s2p = g.s2p;
  const double ux = Amp * (cphi * c2p + ecc0 * sphi * s2p);
  const double vx = Amp * (cphi * s2p - ecc0 * sphi * c2p);
  const double wx = Amp * (sphi * c2p - ecc0 * cphi * s2p);
  const double zx = Amp * (sphi * s2p + ecc0 * cphi * c2p);
  // uvwz_sol
  const double den = fyc1 * fyp2 - fyc2 * fyp1;
  const double uy = (vx * (fyc1 * fxc2 - fxc1 * fyc2) + ux * (fyc1 * fxp2 - fyc2 * fxp1)) / den;
  const double vy = (vx * (fxc1 * fyp2 - fxc2 * fyp1) + ux * (fyp2 * fxp1 - fyp1 * fxp2)) / den;
  const double wy = (zx * (fyc1 * fxc2 - fxc1 * fyc2) + wx * (fyc1 * fxp2 - fyc2 * fxp1)) / den;
  const double zy = (zx * (fxc1 * fyp2 - fxc2 * fyp1) + wx * (fyp2 * fxp1 - fyp1 * fxp2)) / den;
  // exsolve_all
  const double q = 2.0 * (uy * wy + vy * zy) / ((uy * uy + vy * vy) - (wy * wy + zy * zy));
  const double rad = sqrt(1.0 + q * q);
  double a = atan2(2.0 * (uy * wy + vy * zy), ((uy * uy + vy * vy) - (wy * wy + zy * zy)));
  if (a < 0.0) a += TPI_U;
  const double phiy = 0.5 * a;
  const bool flag1 = cos(a) > 0.0, flag2 = cos(a / 2.0) > 0.0, flag3 = sin(a / 2.0) > 0.0;
  const double p =
      ((rad + 1.0) * (uy * uy + vy * vy) + (rad - 1.0) * (wy * wy + zy * zy) + 2.0 * q * (uy * wy + vy * zy)) /
      (2.0 * rad);
  const double ecc = flag1 ? (uy * zy - vy * wy) / p : p / (uy * zy - vy * wy);
  const double ci = -(1.0 - sqrt(1.0 - ecc * ecc)) / ecc;
  // the four branches' arguments, then one atan2 (lanes of a segment take
  // different branches)
  double ya, xa;
  if (flag1 && flag2) {
    ya = (1.0 + rad) * vy + q * zy;
    xa = (1.0 + rad) * uy + q * wy;
  } else if (flag1 && !flag2) {
    ya = -((1.0 + rad) * vy + q * zy);
    xa = -((1.0 + rad) * uy + q * wy);
  } else if (!flag1 && flag3) {
    ya = ecc * ((1.0 + rad) * uy + q * wy);
    xa = -ecc * ((1.0 + rad) * vy + q * zy);
  } else {
    ya = -ecc * ((1.0 + rad) * uy + q * wy);
    xa = ecc * ((1.0 + rad) * vy + q * zy);
  }
  a = atan2(ya, xa);
  if (a < 0.0) a += TPI_U;
  const double Ay = flag1 ? sqrt(p) : sqrt(p) / fabs(ecc);
  y.p9 = 0.5 * a;
  y.p10 = ci;
  y.p6 = log((1.0 + ci * ci) / (2.0 * Ay));
  y.p4 = phiy;
  y.p5 = x.p5 + dtx - dty;
  return y;
}

// |det| of a 4 x 4 matrix as det (Utilities.c:953-988): LU with row pivoting
// on the transposed matrix, the product of the pivots (row exchanges as
// predicated swaps, so every index is a compile-time constant)
__device__ double det4_abs(const double (&A)[16]) {
  double m[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) m[i][j] = A[j * 4 + i];
#pragma unroll
  for (int j = 0; j < 3; j++) {
    int ip = j;
    double amax = fabs(m[j][j]);
#pragma unroll
    for (int i = j + 1; i < 4; i++)
      if (fabs(m[i][j]) > amax) {
        amax = fabs(m[i][j]);
        ip = i;
      }
#pragma unroll
    for (int i = j + 1; i < 4; i++)
      if (i == ip)
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const double t = m[j][k];
          m[j][k] = m[i][k];
          m[i][k] = t;
        }
    if (m[j][j] != 0.0)
#pragma unroll
      for (int i = j + 1; i < 4; i++) {
        const double l = m[i][j] / m[j][j];
        m[i][j] = l;
#pragma unroll
        for (int k = j + 1; k < 4; k++) m[i][k] -= l * m[j][k];
      }
  }
  double dx = 1.0;
#pragma unroll
  for (int i = 0; i < 4; i++) dx *= m[i][i];
  return fabs(dx);
}

// the component of the Jacobian's order (ln DL, psi, phi, cos iota)
__device__ __forceinline__ double jcomp(const Ext& e, int j) {
  return j == 0 ? e.p6 : (j == 1 ? e.p9 : (j == 2 ? e.p4 : e.p10));
}

// update()'s sky-ring proposal (QuickCBC.c:2293-2330) with skydensity_all
// (:7411-7490): the pair of detectors, Ring_all, skymap_all, the log of |det|
// of the one-sided (epsilon 1e-6) Jacobian of (ln DL, psi, phi, cos iota) at
// y with respect to x (each perturbed image mapped to y's solution branch),
// then the half-period shift of phase and polarisation with probability 1/2.
// Called by the 11 lanes of one point (lane j of its 16-lane segment): lane 0
// maps x onto the new sky location (skydensity_all's own second mapping of x
// onto y repeats the same operations on the same values, so it is made
// once) while lanes 1-4 each map one perturbed x; y and the four Jacobian
// rows are then exchanged by shuffles within the segment.
struct SkyOut {
  Ext y;
  double jac;
};
__device__ __noinline__ SkyOut sky_proposal(Ext x, double gmst, int n_det, int l0, int l1, int l2, uint64_t seed,
                                           uint64_t ctr, int j) {
  int i1 = 0, i2 = 1;
  if (n_det > 2) {
    i1 = (int)((double)n_det * uni(seed, SALT_PROPOSE, ctr, D_SKY));
    i2 = i1;
    for (int r = 0; r < DE_TRIES && i2 == i1; r++) i2 = (int)((double)n_det * uni(seed, SALT_PROPOSE, ctr, D_SKY + 1 + r));
    if (i2 == i1) i2 = (i1 + 1) % n_det;
  }
  const int id1 = i1 == 0 ? l0 : (i1 == 1 ? l1 : l2);
  const int id2 = i2 == 0 ? l0 : (i2 == 1 ? l1 : l2);
  const Ext yr = ring_all(x, id1, id2, gmst, uni(seed, SALT_PROPOSE, ctr, D_SKY_ROT), j);
  // skymap_all reads only the sky location of its second argument, so the
  // map of x (lane 0) and the perturbed maps (lanes 1-4) run side by side.
  // Their geometry is evaluated once per function across the lanes: lanes
  // 0-3 the antenna patterns at x's (0, 1) and y's (2, 3) sky location for
  // detectors id1 (0, 2) and id2 (1, 3), lanes 0 and 2 detector id1's delay
  // there; lanes 4-7 cos and sin of phi (4), phi + ep (5), 2 psi (6) and
  // 2 (psi + ep) (7); lanes 8, 9 exp of ln DL and ln DL + ep.
  constexpr double ep = 1.0e-6;
  SkyGeo g;
  {
    const bool ysky = (j & 2) != 0;
    const double al = ysky ? yr.p7 : x.p7, sd = ysky ? yr.p8 : x.p8;
    double fp, fc;
    antenna(0.0, al, sd, gmst, (j & 1) ? id2 : id1, &fp, &fc);
    const double dl = delay(al, sd, gmst, id1);
    const double ang = (j == 6 || j == 7) ? 2.0 * (j == 7 ? x.p9 + ep : x.p9) : (j == 5 ? x.p4 + ep : x.p4);
    const double cv = cos(ang), sv = sin(ang);
    const double ev = exp(j == 9 ? x.p6 + ep : x.p6);
    g.fxp1 = __shfl(fp, 0, 16);
    g.fxc1 = __shfl(fc, 0, 16);
    g.fxp2 = __shfl(fp, 1, 16);
    g.fxc2 = __shfl(fc, 1, 16);
    g.fyp1 = __shfl(fp, 2, 16);
    g.fyc1 = __shfl(fc, 2, 16);
    g.fyp2 = __shfl(fp, 3, 16);
    g.fyc2 = __shfl(fc, 3, 16);
    g.dtx = __shfl(dl, 0, 16);
    g.dty = __shfl(dl, 2, 16);
    const int lph = j == 3 ? 5 : 4, lps = j == 2 ? 7 : 6, lex = j == 1 ? 9 : 8;
    g.cphi = __shfl(cv, lph, 16);
    g.sphi = __shfl(sv, lph, 16);
    g.c2p = __shfl(cv, lps, 16);
    g.s2p = __shfl(sv, lps, 16);
    g.eDL = __shfl(ev, lex, 16);
  }
  Ext xp = x;
  if (j == 1) xp.p6 += ep;
  if (j == 2) xp.p9 += ep;
  if (j == 3) xp.p4 += ep;
  if (j == 4) xp.p10 += ep;
  const Ext ys = skymap_core(xp, yr, g);
  Ext y = yr;  // lane 0's map of x, to every lane
  y.p4 = __shfl(ys.p4, 0, 16);
  y.p5 = __shfl(ys.p5, 0, 16);
  y.p6 = __shfl(ys.p6, 0, 16);
  y.p9 = __shfl(ys.p9, 0, 16);
  y.p10 = __shfl(ys.p10, 0, 16);
  double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;  // this lane's Jacobian row (lanes 1-4)
  if (j >= 1 && j <= 4) {
    Ext yp = ys;
    if (fabs(yp.p9 + PI_U / 2.0 - y.p9) < fabs(yp.p9 - y.p9)) yp.p9 += PI_U / 2.0;
    if (fabs(yp.p9 - PI_U / 2.0 - y.p9) < fabs(yp.p9 - y.p9)) yp.p9 -= PI_U / 2.0;
    if (fabs(yp.p4 + PI_U - y.p4) < fabs(yp.p4 - y.p4)) yp.p4 += PI_U;
    if (fabs(yp.p4 - PI_U - y.p4) < fabs(yp.p4 - y.p4)) yp.p4 -= PI_U;
    r0 = (jcomp(yp, 0) - jcomp(y, 0)) / ep;
    r1 = (jcomp(yp, 1) - jcomp(y, 1)) / ep;
    r2 = (jcomp(yp, 2) - jcomp(y, 2)) / ep;
    r3 = (jcomp(yp, 3) - jcomp(y, 3)) / ep;
  }
  double J[16];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    J[i * 4 + 0] = __shfl(r0, 1 + i, 16);
    J[i * 4 + 1] = __shfl(r1, 1 + i, 16);
    J[i * 4 + 2] = __shfl(r2, 1 + i, 16);
    J[i * 4 + 3] = __shfl(r3, 1 + i, 16);
  }
  SkyOut o;
  o.jac = log(det4_abs(J));
  if (uni(seed, SALT_PROPOSE, ctr, D_SKY_FLIP) > 0.5) {
    y.p4 += PI_U;
    y.p9 += PI_U / 2.0;
  }
  o.y = y;
  return o;
}

// Component j of update()'s proposal for the state x at temperature T
// (sT = sqrt(T)) from the counters (seed, ctr); *typ gets the proposal type
// (1 Fisher, 2 DE, 3 jiggle).  One lane per component: the draws a proposal
// shares (its type, eigendirection, history rows, scale) are the same
// counters on every lane, so each component is computed exactly as a
// sequential loop over the components would compute it.  All of lanes 0-10
// of a point's 16-lane segment call it together (the sky ring shuffles within
// the segment).
__device__ double propose_component(const double* __restrict__ x, const double* __restrict__ ej,
                                    const double* __restrict__ ev, const double* __restrict__ hist, int n_hist,
                                    double sT, const UpdateConst& u, uint64_t seed, uint64_t ctr, int j, int* typ,
                                    double* jac) {
  const double alpha = uni(seed, SALT_PROPOSE, ctr, D_ALPHA);
  double y;
  *jac = 0.0;
  if (alpha > u.a) {  // Fisher jump along an eigendirection
    *typ = 1;
    const int i = (int)(uni(seed, SALT_PROPOSE, ctr, D_FISHER) * 11.0);
    const double beta = sT * ej[i] * gauss(seed, SALT_PROPOSE, ctr, D_FISHER + 1);
    y = x[j] + beta * ev[11 * i + j];
  } else if (alpha > u.b) {  // differential evolution from the temperature's history
    *typ = 2;
    const int i = (int)((double)n_hist * uni(seed, SALT_PROPOSE, ctr, D_DE));
    int k = i;
    for (int r = 0; r < DE_TRIES && k == i; r++) k = (int)((double)n_hist * uni(seed, SALT_PROPOSE, ctr, D_DE + 1 + r));
    if (k == i) k = (i + 1) % n_hist;
    double g = 1.0;
    if (uni(seed, SALT_PROPOSE, ctr, D_DE_SCALE) < 0.9) g = 0.5 * gauss(seed, SALT_PROPOSE, ctr, D_DE_SCALE + 1);
    y = x[j] + g * (hist[11 * i + j] - hist[11 * k + j]);
  } else if (alpha > u.c) {  // jiggle
    *typ = 3;
    const double beta = 0.01 * pow(10.0, -floor(3.0 * uni(seed, SALT_PROPOSE, ctr, D_JIGGLE))) * sT;
    y = x[j] + beta * gauss(seed, SALT_PROPOSE, ctr, D_JIGGLE + 1 + 2 * j);
  } else {  // sky ring: every lane maps the whole point and keeps its component
    *typ = 4;
    Ext e;
    e.p4 = x[4];
    e.p5 = x[5];
    e.p6 = x[6];
    e.p7 = x[7];
    e.p8 = x[8];
    e.p9 = x[9];
    e.p10 = x[10];
    const SkyOut o = sky_proposal(e, u.gmst, u.n_det, u.labels[0], u.labels[1], u.labels[2], seed, ctr, j);
    *jac = o.jac;
    y = j == 4 ? o.y.p4
               : j == 5 ? o.y.p5
               : j == 6 ? o.y.p6
               : j == 7 ? o.y.p7 : j == 8 ? o.y.p8 : j == 9 ? o.y.p9 : j == 10 ? o.y.p10 : x[j];
  }
  // angles back into range (:2336-2342), with the loops' own result (a run of
  // steps within one binade taken at once, so an angle of 1e4 rad costs ~40
  // turns instead of 1600).  Only an angle within +-ANGLE_WRAP_MAX rad is
  // wrapped; one beyond (or not finite) stays out of its prior range and the
  // proposal is rejected.
  if ((j == 4 || j == 7) && fabs(y) <= ANGLE_WRAP_MAX) {
    y = wrap_up(wrap_down(y, TPI_U), TPI_U);
  } else if (j == 9 && fabs(y) <= ANGLE_WRAP_MAX) {
    y = wrap_up(wrap_down(y, PI_U), PI_U);
  }
  return y;
}

// lanes per point of the proposal loops (11 used)
constexpr int PROP_LANES = 16;

// The proposals of step s for every temperature of the group's column b
// (storage lp = t ncols + b): thread l of the caller's nthreads is lane
// l % PROP_LANES of temperature l / PROP_LANES, then of every nthreads /
// PROP_LANES-th temperature after it.
__device__ void propose_column(const double* __restrict__ x, const double* __restrict__ ej,
                               const double* __restrict__ ev, const double* __restrict__ hist,
                               const double* __restrict__ tc, const UpdateConst& u, int n_temps, int n_chains,
                               int c0, int ncols, int b, uint64_t seed, uint64_t s, int l, int nthreads,
                               double* __restrict__ y, int* __restrict__ typ, double* __restrict__ jac) {
  const uint64_t P = (uint64_t)n_temps * n_chains;
  const int per = nthreads / PROP_LANES, sub = l / PROP_LANES, j = l % PROP_LANES;
  if (j >= 11) return;
  for (int t = sub; t < n_temps; t += per) {
    const int lp = t * ncols + b;
    const uint64_t gp = (uint64_t)t * n_chains + c0 + b;
    const double* h = hist + ((size_t)b * n_temps + t) * (size_t)u.n_hist * 11;
    int ty;
    double jc;
    const double yj = propose_component(x + 11 * lp, ej + 11 * lp, ev + 121 * lp, h, u.n_hist, tc[3 * t], u, seed,
                                        s * P + gp, j, &ty, &jc);
    y[11 * lp + j] = yj;
    if (j == 0) {
      typ[lp] = ty;
      jac[lp] = jc;
    }
  }
}

// Proposals of step *step for the group's points (used at creation and after
// a Fisher refresh; later steps get theirs from k_pt_steps).  One block per
// column.
__global__ __launch_bounds__(512) void k_update_propose(const double* __restrict__ x, const double* __restrict__ ej,
                                                        const double* __restrict__ ev, const double* __restrict__ hist,
                                                        const double* __restrict__ tc, UpdateConst u, int n_temps,
                                                        int n_chains, int c0, int ncols, uint64_t seed,
                                                        const unsigned long long* __restrict__ step,
                                                        double* __restrict__ y, int* __restrict__ typ,
                                                        double* __restrict__ jac) {
  propose_column(x, ej, ev, hist, tc, u, n_temps, n_chains, c0, ncols, blockIdx.x, seed, *step, threadIdx.x,
                 blockDim.x, y, typ, jac);
}

// On update steps with s % hist_every == 0 the column's states enter its
// history (:2071-2080).  All threads of the block; ends synchronised.
__device__ void column_history(const double* __restrict__ x, double* __restrict__ hist,
                               unsigned long long* __restrict__ hcount, const UpdateConst& u, int n_temps, int ncols,
                               int b, int c, uint64_t s) {
  if (s % (uint64_t)u.hist_every != 0) return;
  const uint64_t m = hcount[c];
  for (int t = threadIdx.x; t < n_temps; t += blockDim.x) {
    double* h = hist + (((size_t)b * n_temps + t) * u.n_hist + (size_t)(m % (uint64_t)u.n_hist)) * 11;
    for (int i = 0; i < 11; i++) h[i] = x[11 * (t * ncols + b) + i];
  }
  __syncthreads();
  if (threadIdx.x == 0) hcount[c] = m + 1;
}

// One swap of temperatures j, j + 1 of the column (:2038-2051), with the
// Fisher jumps that belong to the states.  All threads of the block; *sw is
// block-shared.
__device__ void column_swap(double* __restrict__ x, double* __restrict__ L, double* __restrict__ ej,
                            double* __restrict__ ev, const double* __restrict__ tc, int n_temps, int ncols, int b,
                            uint64_t seed, uint64_t cs, unsigned long long* __restrict__ cst, int* sw) {
  if (threadIdx.x == 0) {
    const int j = (int)((double)(n_temps - 1) * uni(seed, SALT_SWAP, cs, 1));
    const double dL = L[j * ncols + b] - L[(j + 1) * ncols + b];
    const double beta = exp(dL * tc[3 * (j + 1) + 1] - dL * tc[3 * j + 1]);
    *sw = beta > uni(seed, SALT_SWAP, cs, 2) ? j : -1;
    cst[1] += 1;
    if (*sw >= 0) cst[2] += 1;
  }
  __syncthreads();
  const int w = *sw;
  if (w >= 0) {
    const int la = w * ncols + b, lb = (w + 1) * ncols + b;
    for (int i = threadIdx.x; i < 11 + 11 + 121 + 1; i += blockDim.x) {
      double* pa;
      double* pb;
      if (i < 11) {
        pa = x + 11 * la + i;
        pb = x + 11 * lb + i;
      } else if (i < 22) {
        pa = ej + 11 * la + (i - 11);
        pb = ej + 11 * lb + (i - 11);
      } else if (i < 143) {
        pa = ev + 121 * la + (i - 22);
        pb = ev + 121 * lb + (i - 22);
      } else {
        pa = L + la;
        pb = L + lb;
      }
      const double tmp = *pa;
      *pa = *pb;
      *pb = tmp;
    }
  }
}

// The last block of a launch to finish advances the group's step counter.
__device__ __forceinline__ void finish_launch(unsigned long long* __restrict__ step, unsigned int* __restrict__ done,
                                              uint64_t next) {
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(done, 1u) == gridDim.x - 1) {
      *step = next;
      *done = 0u;
      __threadfence();
    }
  }
}

// ---------------------------------------------------------- fused steps
// temperatures per chunk of the fused step: their set-up records live in LDS
// (one 16-lane segment of the 512-thread block per temperature)
constexpr int PT_CHUNK = 32;

// The terms of update()'s acceptance (:2345-2372) that do not need the
// proposal's logL, for one temperature of the chunk: whether q lies outside
// the prior, and H's other terms, each formed as H = jc + (Lq - Lx) / T
// + pMcM(q) + 3 q[6] - 3 x[6] - pMcM(x) forms it; then log u.
struct AcceptPre {
  double pMcMy[PT_CHUNK], q6[PT_CHUNK], x6[PT_CHUNK], mjx[PT_CHUNK], logu[PT_CHUNK];
  int out[PT_CHUNK];
};

// One part of one point's set-up, written to its record c: role 0 the
// amplitude half of the model, 3 the phase half with the time shift and
// reference phase, 1 fbegin, 2 the projection.
__device__ __forceinline__ void setup_part(int role, const double* __restrict__ q, const AnalysisConst& an,
                                           Point& c) {
  double prm[11];
#pragma unroll
  for (int i = 0; i < 11; i++) prm[i] = q[i];
  if (role == 0)
    (void)point_setup_model_half(prm, an, c, 0);
  else if (role == 3)
    (void)point_setup_phase_ref(prm, an, c);
  else if (role == 1)
    c.fs = fbegin(prm);
  else
    point_setup_projection(prm, an, c);
}

// One chunk (temperatures t0 .. t0 + nt - 1) before its logL: waves 0-3 set
// up the proposals into the LDS records spt, one part per wave as k_setup
// spreads them, lane l the temperature t0 + l; waves 4-7 meanwhile form the
// acceptance terms that do not need the logL, one kind per wave.  The caller
// synchronises after it.
__device__ void chunk_setup(const double* __restrict__ x, const double* __restrict__ y, const AnalysisConst& an,
                            const UpdateConst& u, Point* spt, AcceptPre* pre, int t0, int nt, int n_chains, int ncols,
                            int b, int c, uint64_t seed, uint64_t s, uint64_t P) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = (int)(threadIdx.x & 63);
  if (l >= nt) return;
  const int t = t0 + l, lp = t * ncols + b;
  const double* q = y + 11 * lp;
  const double* xs = x + 11 * lp;
  if (wave < 4) {
    setup_part(wave, q, an, spt[l]);
  } else if (wave == 4) {
    double pMcMy;
    pre->out[l] = prior_flag(q, u, &pMcMy) ? 1 : 0;
    pre->pMcMy[l] = pMcMy;
  } else if (wave == 5) {
    pre->mjx[l] = mass_jacobian(xs);
  } else if (wave == 6) {
    pre->logu[l] = log(uni(seed, SALT_ACCEPT, s * P + (uint64_t)t * n_chains + c, 0));
  } else {
    pre->q6[l] = 3.0 * q[6];
    pre->x6[l] = 3.0 * xs[6];
  }
}

__device__ __forceinline__ double seg_sum(double v) {
#pragma unroll
  for (int o = PROP_LANES / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, PROP_LANES);
  return v;
}

// log_likelihood_het (QuickCBC.c:3871-3940) of one set-up point by the
// PROP_LANES lanes of its segment (lane j takes grid samples j, j +
// PROP_LANES, ...), as k_het_ll evaluates each sample: the h22 amplitude and
// phase from the sample planes, the taper, each detector's projection against
// the heterodyne's amplitude and phase, and the folded weights;
// logL = logLb - sum_d (HR_d + HH_d / 2), 0 for a point whose set-up failed.
// Every lane returns the sum.
template <int ND>
__device__ double het_ll_segment(const Point& c, const AnalysisConst& an, const HetArgs& h, int j) {
  double hr[ND], hh[ND];
#pragma unroll
  for (int d = 0; d < ND; d++) hr[d] = hh[d] = 0.0;
  const bool live = c.status == ST_OK;
  if (live) {
    const int M = h.M;
    const double i_Msec = 1.0 / c.Msec;
    double As[ND];  // the detector scales Ascale amp0 / sqrt(Tobs), as point_setup_finish forms them
#pragma unroll
    for (int d = 0; d < ND; d++) As[d] = c.Ascale[d] * c.amp0 / an.sqT;
    for (int m = j; m < M; m += PROP_LANES) {
      // the point's constants are read again per sample (LDS) rather than
      // held in registers across the loop: the segments of a wave hold
      // different points, so they would be vector registers, ~70 doubles
      asm volatile("" ::: "memory");
      const double f = h.freq[m];
      double a, p;
      h22_planes(c, i_Msec, f, h.pl, M, m, &a, &p);
      const double yt = taper(f, c.fs);
#pragma unroll
      for (int d = 0; d < ND; d++) {
        const double xd = TPI * f * c.tshift[d] + c.lambda[d] - p;
        const double ad = As[d] * a * yt;
        double sn, cs;
        sincos(h.hp[(size_t)d * M + m] - xd, &sn, &cs);
        const double cc = 2.0 * (h.ha[(size_t)d * M + m] - ad * cs);
        const double ss = 2.0 * ad * sn;
        hr[d] += h.wf[(size_t)(1 * ND + d) * M + m] * cc + h.wf[(size_t)(2 * ND + d) * M + m] * ss;
        hh[d] += h.wf[(size_t)(0 * ND + d) * M + m] * (cc * cc + ss * ss);
      }
    }
  }
  double Lv = h.logLb;
#pragma unroll
  for (int d = 0; d < ND; d++) {
    const double r = seg_sum(hr[d]), q = seg_sum(hh[d]);
    Lv -= (r + 0.5 * q);
  }
  return live ? Lv : 0.0;
}

// nsteps PT steps of a group of chain columns in one launch; block b runs
// column c0 + b through all of them (the columns are independent chains, so
// no block waits for another).  An update step takes the temperatures in
// chunks of PT_CHUNK: the chunk's set-up (and the acceptance terms without
// the logL), then each temperature's 16-lane segment evaluates its
// proposal's heterodyned logL and decides its acceptance.  A swap step makes
// one swap (its proposals are discarded unevaluated).  Then the history and
// the proposals of the next step.
template <int ND>
__global__ __launch_bounds__(512) void k_pt_steps(
    double* __restrict__ x, double* __restrict__ L, double* __restrict__ ej, double* __restrict__ ev,
    double* __restrict__ hist, unsigned long long* __restrict__ hcount, double* __restrict__ y,
    int* __restrict__ typ, double* __restrict__ jac, const double* __restrict__ tc, UpdateConst u, AnalysisConst an,
    HetArgs h, int n_temps, int n_chains, int c0, int ncols, uint64_t seed, int nsteps,
    unsigned long long* __restrict__ step, unsigned long long* __restrict__ stats, unsigned int* __restrict__ done) {
  __shared__ Point spt[PT_CHUNK];
  __shared__ AcceptPre pre;
  __shared__ int sw;
  const int b = blockIdx.x, c = c0 + b;
  const uint64_t P = (uint64_t)n_temps * n_chains, s0 = *step;
  unsigned long long* cst = stats + (size_t)c * UPD_STATS;
  const int j = (int)(threadIdx.x % PROP_LANES), seg = (int)(threadIdx.x / PROP_LANES);  // seg < PT_CHUNK
  for (int k = 0; k < nsteps; k++) {
    const uint64_t s = s0 + (uint64_t)k;
    const uint64_t cs = s * (uint64_t)n_chains + c;
    const bool swap_step = n_temps > 1 && uni(seed, SALT_SWAP, cs, 0) < 0.2;
    if (!swap_step) {
      for (int t0 = 0; t0 < n_temps; t0 += PT_CHUNK) {
        const int nt = n_temps - t0 < PT_CHUNK ? n_temps - t0 : PT_CHUNK;
        chunk_setup(x, y, an, u, spt, &pre, t0, nt, n_chains, ncols, b, c, seed, s, P);
        __syncthreads();
        if (seg < nt) {
          const int t = t0 + seg, lp = t * ncols + b;
          const double Lq = het_ll_segment<ND>(spt[seg], an, h, j);
          if (j == 0) {  // the acceptance, H summed in update()'s order
            const int kt = typ[lp];
            atomicAdd(&cst[3 + 2 * (kt - 1)], 1ULL);
            if (!(pre.out[seg] || spt[seg].status != ST_OK)) {
              const double H = jac[lp] + (Lq - L[lp]) * tc[3 * t + 1] + pre.pMcMy[seg] + pre.q6[seg] -
                               pre.x6[seg] - pre.mjx[seg];
              if (H > pre.logu[seg]) {
                for (int i = 0; i < 11; i++) x[11 * lp + i] = y[11 * lp + i];
                L[lp] = Lq;
                atomicAdd(&cst[4 + 2 * (kt - 1)], 1ULL);
              }
            }
          }
        }
        __syncthreads();
      }
      column_history(x, hist, hcount, u, n_temps, ncols, b, c, s);
      if (threadIdx.x == 0) cst[0] += 1;
    } else {
      column_swap(x, L, ej, ev, tc, n_temps, ncols, b, seed, cs, cst, &sw);
    }
    __syncthreads();
    propose_column(x, ej, ev, hist, tc, u, n_temps, n_chains, c0, ncols, b, seed, s + 1, threadIdx.x, blockDim.x,
                   y, typ, jac);
    __syncthreads();
  }
  finish_launch(step, done, s0 + (uint64_t)nsteps);
}

// this unit's copy of the QNM spline slopes (the fused step kernel's set-up)
__global__ void k_update_qnm_init() { qnm_init_tables(threadIdx.x); }

// The initial history of each column (MCMC_all, :1946-1954): row i of
// temperature t holds the initial state of a temperature q drawn uniformly
// from the column's.
__global__ void k_update_hist_init(const double* __restrict__ x, int n_temps, int n_chains, int c0, int ncols,
                                   int n_hist, uint64_t seed, double* __restrict__ hist) {
  const size_t n = (size_t)ncols * n_temps * n_hist;
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int i = (int)(r % n_hist);
  const int t = (int)((r / n_hist) % n_temps);
  const int b = (int)(r / ((size_t)n_hist * n_temps));
  const int c = c0 + b;
  const int q = (int)((double)n_temps * uni(seed, SALT_HIST, ((uint64_t)c * n_temps + t), (uint64_t)i));
  for (int k = 0; k < 11; k++) hist[r * 11 + k] = x[11 * (q * ncols + b) + k];
}

}  // namespace

// threads per column block: PROP_LANES per temperature, whole waves, at most 512
static int update_threads(int n_temps) {
  const int n = n_temps * PROP_LANES;
  return n >= 512 ? 512 : ((n + 63) / 64) * 64;
}

hipError_t launch_update_hist_init(const double* x, int n_temps, int n_chains, int c0, int ncols, int n_hist,
                                   uint64_t seed, double* hist, hipStream_t s) {
  const size_t n = (size_t)ncols * n_temps * n_hist;
  hipLaunchKernelGGL(k_update_hist_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, n_temps, n_chains, c0,
                     ncols, n_hist, seed, hist);
  return hipGetLastError();
}

hipError_t launch_update_propose(const double* x, const double* ej, const double* ev, const double* hist,
                                 const double* tc, const UpdateConst& u, int n_temps, int n_chains, int c0, int ncols,
                                 uint64_t seed, const unsigned long long* step, double* y, int* typ, double* jac,
                                 hipStream_t s) {
  hipLaunchKernelGGL(k_update_propose, dim3(ncols), dim3(update_threads(n_temps)), 0, s, x, ej, ev, hist, tc, u,
                     n_temps, n_chains, c0, ncols, seed, step, y, typ, jac);
  return hipGetLastError();
}

hipError_t launch_update_qnm_init(hipStream_t s) {
  hipLaunchKernelGGL(k_update_qnm_init, dim3(1), dim3(64), 0, s);
  return hipGetLastError();
}

hipError_t launch_pt_steps(double* x, double* L, double* ej, double* ev, double* hist, unsigned long long* hcount,
                           double* y, int* typ, double* jac, const double* tc, const UpdateConst& u,
                           const AnalysisConst& an, const HetArgs& h, int n_temps, int n_chains, int c0, int ncols,
                           uint64_t seed, int nsteps, unsigned long long* step, unsigned long long* stats,
                           unsigned int* done, hipStream_t s) {
  if (nsteps <= 0 || ncols <= 0) return hipSuccess;
  // 512 threads whatever n_temps: a chunk of PT_CHUNK 16-lane segments, and
  // the set-up's four parts and the acceptance terms on eight waves
  const dim3 g(ncols), blk(PT_CHUNK * PROP_LANES);
  if (h.nd == 1)
    hipLaunchKernelGGL(k_pt_steps<1>, g, blk, 0, s, x, L, ej, ev, hist, hcount, y, typ, jac, tc, u, an, h, n_temps,
                       n_chains, c0, ncols, seed, nsteps, step, stats, done);
  else if (h.nd == 2)
    hipLaunchKernelGGL(k_pt_steps<2>, g, blk, 0, s, x, L, ej, ev, hist, hcount, y, typ, jac, tc, u, an, h, n_temps,
                       n_chains, c0, ncols, seed, nsteps, step, stats, done);
  else
    hipLaunchKernelGGL(k_pt_steps<3>, g, blk, 0, s, x, L, ej, ev, hist, hcount, y, typ, jac, tc, u, an, h, n_temps,
                       n_chains, c0, ncols, seed, nsteps, step, stats, done);
  return hipGetLastError();
}

}  // namespace qcbc
