// HIP kernels of the waveform + likelihood path, written for gfx950.
//
// Data layout of one analysis in HBM ("band planes"): band bins
// j = k - imin, k in [imin, imax), padded to nbp; FP64 planes
//   0: f_k = k/Tobs   1: f_k^(1/3)   2: f_k^(-1/6)   3: log f_k
//   4+3d: D_re/S   5+3d: D_im/S   6+3d: 1/S          (detector slot d)
// and, for the mixed-precision kernel, FP32 planes scaled by w_ref (per
// detector, so 1/S fits in FP32):
//   3d: D_re/(S sqrt(w_ref))   3d+1: D_im/(S sqrt(w_ref))   3d+2: 1/(S w_ref)
#include "qcbc_device.h"
#include "qcbc_internal.h"

namespace qcbc {

// ------------------------------------------------------------ utilities
// One DPP step of a wave reduction on a double (two 32-bit DPP moves; lanes
// the control leaves without a source, or rows outside ROWMASK, add 0).  With
// every row written, the bound control alone supplies the zeros.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_f64(double v) {
  const unsigned long long u = __double_as_longlong(v);
  int lo, hi;
  if (ROWMASK == 0xf) {
    lo = __builtin_amdgcn_mov_dpp((int)(unsigned)u, CTRL, 0xf, 0xf, true);
    hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), CTRL, 0xf, 0xf, true);
  } else {
    lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, ROWMASK, 0xf, true);
    hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, ROWMASK, 0xf, true);
  }
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Sum over the 64 lanes of a wave, in lane 63 (the other lanes hold partial
// sums): shifts by 1, 2, 4, 8 within each row of 16 lanes leave each row's sum
// in its lane 15, row_bcast:15 adds row 0's into row 1 and row 2's into row 3,
// row_bcast:31 adds lane 31 into rows 2 and 3.  Register-to-register, with no
// LDS round trips.  Every lane must be active.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<0x111, 0xf>(v);  // row_shr:1
  v += dpp_f64<0x112, 0xf>(v);  // row_shr:2
  v += dpp_f64<0x114, 0xf>(v);  // row_shr:4
  v += dpp_f64<0x118, 0xf>(v);  // row_shr:8
  v += dpp_f64<0x142, 0xa>(v);  // row_bcast:15
  v += dpp_f64<0x143, 0xc>(v);  // row_bcast:31
  return v;
}
constexpr int SUM_LANE = 63;

// Lane exchanges of a double across the halves (permlane32) or across
// neighbouring rows of 16 lanes (permlane16) of two registers, gfx950's
// v_permlane*_swap: after swap_halves, x holds the lower halves of x and y
// and y their upper halves, so x + y has x's half-folded sums in lanes 0-31
// and y's in lanes 32-63; after swap_rows, x holds rows 0 and 2 of x and y
// and y their rows 1 and 3, so x + y holds x's row-pair sums in rows 0 and 2
// and y's in rows 1 and 3.
__device__ __forceinline__ void swap_halves(double& x, double& y) {
  const unsigned long long ux = __double_as_longlong(x), uy = __double_as_longlong(y);
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)ux, (unsigned)uy, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(ux >> 32), (unsigned)(uy >> 32), false, false);
  x = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
  y = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
}
__device__ __forceinline__ void swap_rows(double& x, double& y) {
  const unsigned long long ux = __double_as_longlong(x), uy = __double_as_longlong(y);
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)ux, (unsigned)uy, false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(ux >> 32), (unsigned)(uy >> 32), false, false);
  x = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
  y = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
}
// Sum over each row of 16 lanes, in the row's last lane.
__device__ __forceinline__ double row_sum(double v) {
  v += dpp_f64<0x111, 0xf>(v);  // row_shr:1
  v += dpp_f64<0x112, 0xf>(v);  // row_shr:2
  v += dpp_f64<0x114, 0xf>(v);  // row_shr:4
  v += dpp_f64<0x118, 0xf>(v);  // row_shr:8
  return v;
}

// The wave totals of the 2 nd per-lane values v[0..2nd) written to out[0..2nd)
// (when write): the values are folded pairwise, halves then rows, so that
// one row sum finishes four totals at once.  Value q's total lands in the
// last lane of row rho(q): for nd >= 2, v[0..3] in rows 0, 2, 1, 3; the last
// pair (v[4], v[5] at nd = 3, v[0], v[1] at nd = 1) in rows 0 and 2.  Every
// lane must be active.
template <int ND>
__device__ __forceinline__ void wave_sums_store(double (&v)[2 * ND], double* out, bool write) {
  const int lane = threadIdx.x & 63, row = lane >> 4;
  double t = 0.0, u = 0.0;
  if (ND >= 2) {
    double a0 = v[0], a1 = v[1], b0 = v[2], b1 = v[3];
    swap_halves(a0, a1);
    swap_halves(b0, b1);
    double a = a0 + a1, b = b0 + b1;  // (v0 | v1), (v2 | v3) by halves
    swap_rows(a, b);
    t = row_sum(a + b);               // rows: v0, v2, v1, v3
  }
  if (ND != 2) {
    double c0 = v[2 * ND - 2], c1 = v[2 * ND - 1];
    swap_halves(c0, c1);
    double c = c0 + c1, cc = c;
    swap_rows(c, cc);
    u = row_sum(c + cc);              // rows: last pair's first, same, second, same
  }
  if (write && (lane & 15) == 15) {
    if (ND >= 2) out[(row & 1) * 2 + (row >> 1)] = t;
    if (ND != 2 && (row & 1) == 0) out[2 * ND - 2 + (row >> 1)] = u;
  }
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Hides a constant from the optimiser in a scalar register, so that an FMA
// takes it as its (negated) third operand instead of copying it into a vector
// register first.  The asm body is empty.
__device__ __forceinline__ double sreg(double v) {
  asm("" : "+s"(v));
  return v;
}

// In-place FMAs for loop-carried sums: the result is written over acc's own
// registers, so a sum updated in a conditional block needs no copy back at
// its end.  y is wave-uniform (a scalar register).
__device__ __forceinline__ void fma_acc_x(double& acc, double y, double z) {  // acc = acc y + z
  asm("v_fma_f64 %0, %0, %1, %2" : "+v"(acc) : "s"(y), "v"(z));
}
__device__ __forceinline__ void fma_acc_z(double& acc, double x, double y) {  // acc = x y + acc
  asm("v_fma_f64 %0, %1, %2, %0" : "+v"(acc) : "v"(x), "s"(y));
}

// sin(2 pi r) and cos(2 pi r) for a phase r given in turns.  t = r - rint(r)
// (exact, |t| <= 1/2); degree-6 polynomials in t^2 give the quarter angle,
// sin(pi t / 2) = t P(t^2) and cos(pi t / 2) = Q(t^2), to below 2 ulp
// (Chebyshev fits in t^2 at 40 digits, generated by gen_sincos.py beside this
// file), and two complex squarings give the full angle, to a few ulp.  This
// replaces the general sincos, whose large-argument reduction the ~1e4-turn
// projected phase would otherwise pull in.
//
// The first Horner step of each polynomial adds two constants; a VALU
// operation reads at most one scalar register, so a bin loop keeps the second
// coefficients in vector registers (SinCosK, loaded once per block) instead
// of copying one of them every bin.
struct SinCosK {
  double p1, q1;
  double p2, p3, q2, q3, q4;  // sincos_amp_turns only: its next coefficients, taken from vector registers too
};

__device__ __forceinline__ double vreg_v(double v) {
  asm("" : "+v"(v));
  return v;
}

__device__ __forceinline__ SinCosK sincos_consts() {
  return SinCosK{vreg_v(-0x1.e30071afc3e59p-19), vreg_v(-0x1.a6c9c1be9eb49p-16)};
}

__device__ __forceinline__ void sincos_turns(double r, double* s, double* c, const SinCosK& k) {
  const double t = r - rint(r);
  const double z = t * t;
  double p = fma(z, sreg(0x1.e3f38399551bfp-25), k.p1);
  p = fma(p, z, -sreg(-0x1.50782fda12d96p-13));
  p = fma(p, z, -sreg(0x1.32d2cce2e5b19p-8));
  p = fma(p, z, -sreg(-0x1.466bc677587f8p-4));
  p = fma(p, z, -sreg(0x1.4abbce625be41p-1));
  p = fma(p, z, -sreg(-0x1.921fb54442d18p+0));
  double q = fma(z, sreg(0x1.f3dbcea61b1a4p-22), k.q1);
  q = fma(q, z, -sreg(-0x1.e1f4fb60281f6p-11));
  q = fma(q, z, -sreg(0x1.55d3c7dbfd139p-6));
  q = fma(q, z, -sreg(-0x1.03c1f081b0780p-2));
  q = fma(q, z, -sreg(0x1.3bd3cc9be458bp+0));
  q = fma(q, z, 1.0);
  double sn = t * p, cs = q;
#pragma unroll
  for (int k2 = 0; k2 < 2; k2++) {  // (cs + i sn)^2, twice
    const double sc = sn * cs;
    cs = fma(cs, cs, -(sn * sn));
    sn = sc + sc;
  }
  *s = sn;
  *c = cs;
}

__device__ __forceinline__ void sincos_turns(double r, double* s, double* c) {
  sincos_turns(r, s, c, SinCosK{-0x1.e30071afc3e59p-19, -0x1.a6c9c1be9eb49p-16});
}

// a sin and a cos of 2 pi r.  sincos_turns' quarter-angle polynomials with
// every coefficient times 2^(3/4) (gen_sincos.py: PS, QS, each rounded once),
// one double-angle step at that scale, (2^(3/4) s)(2^(3/4) c) = sqrt2 sin 2x
// and (2^(3/4) c)^2 - sqrt2 = sqrt2 cos 2x, and the last step folded with the
// amplitude: with P = a sqrt2 cos y, a sin 2y = P (sqrt2 sin y) and
// a cos 2y = P (sqrt2 cos y) - a, three operations for the last step and the
// amplitude's two products.  k: sincos_consts_amp().
__device__ __forceinline__ void sincos_amp_turns(double r, double a, double* as, double* ac, const SinCosK& k) {
  const double t = r - rint(r);
  const double z = t * t;
  double p = fma(z, sreg(0x1.96f3ee17e49d0p-24), k.p1);
  p = fma(p, z, k.p2);
  p = fma(p, z, k.p3);
  p = fma(p, z, -sreg(-0x1.127c7aa3ad157p-3));
  p = fma(p, z, -sreg(0x1.161cdace59eb0p+0));
  p = fma(p, z, -sreg(-0x1.5224febe4ed1bp+1));
  double q = fma(z, sreg(0x1.a4544e88d2cbfp-21), k.q1);
  q = fma(q, z, k.q2);
  q = fma(q, z, k.q3);
  q = fma(q, z, k.q4);
  q = fma(q, z, -sreg(0x1.0993fcf8017d5p+1));
  q = fma(q, z, -sreg(-0x1.ae89f995ad3adp+0));
  const double sn = t * p;  // 2^(3/4) sin and cos of the quarter angle: sn, q
  const double s1 = sn * q, c1 = fma(q, q, -sreg(0x1.6a09e667f3bcdp+0));  // sqrt2 sin, cos of the half angle
  const double pa = a * c1;
  *as = pa * s1;
  *ac = fma(pa, c1, -a);
}

__device__ __forceinline__ SinCosK sincos_consts_amp() {
  return SinCosK{vreg_v(-0x1.9627888bf363fp-18), vreg_v(-0x1.638560cc2646fp-15),
                 vreg_v(0x1.1aef9c5dc3c56p-12),  vreg_v(-0x1.0201b90b986b1p-7),
                 vreg_v(0x1.9546a01b8e092p-10),  vreg_v(-0x1.1f70faa628642p-5),
                 vreg_v(0x1.b4dbc2030e65dp-2)};
}


// ------------------------------------------------- fused likelihood blocks
// A block works on one group of cpb points of one analysis (one wave per
// point, so every per-point constant is wave-uniform) and one split of the
// band.  It walks the LDS tiles of the split that any of its points needs:
// each tile (frequency tables and whitened data of every detector) is read
// from HBM once and used by all cpb waves.  Lane l of a point's wave takes
// the band bins j = ks + l + 64 n of the point's range [ks, ke) whatever the
// tiles, so a point's sums do not depend on the tile grid or on the other
// points of its block.  Within a tile a wave walks the point's segments:
// runs of bins over which neither the amplitude nor the phase changes region
// (the bin indices come from the set-up, on the same Mf = Msec f_k tests as
// the piecewise waveform), each with a loop of its own that holds only that
// region pair's constants.  Per-lane sums, one wave reduction, one partial
// per (point, split).

// Bins of the split that the block's points need: [lo, hi), with lo rounded
// down to the tile grid.
template <int TILE>
__device__ __forceinline__ void block_range(const LLInst& in, int b, bool active, int j0, int j1, int* s_lohi,
                                            int* ks, int* ke, int* lo, int* hi) {
  const int lane = threadIdx.x & 63;
  *ks = 0;
  *ke = 0;
  if (active) {
    *ks = max(in.pts[b].kstart - in.imin, j0);
    *ke = min(in.pts[b].kend - in.imin, j1);
  }
  __syncthreads();  // s_lohi may still be read by the previous work item
  if (threadIdx.x == 0) { s_lohi[0] = j1; s_lohi[1] = j0; }
  __syncthreads();
  if (lane == 0 && active && *ke > *ks) {
    atomicMin(&s_lohi[0], *ks);
    atomicMax(&s_lohi[1], *ke);
  }
  __syncthreads();
  *lo = j0 + ((s_lohi[0] - j0) / TILE) * TILE;
  *hi = s_lohi[1];
}

// Makes v opaque to the optimiser in place (empty asm): later subtractions of
// scalar constants from it stay single subtractions instead of FMAs that need
// the constant copied into a vector register.
__device__ __forceinline__ void vreg(double& v) { asm("" : "+v"(v)); }

// One Horner step on negated partial sums: -(c + x p) from np = -p.  The
// negated third operand keeps the FMA in its three-address form (the
// two-address form would first copy c into the destination register).
__device__ __forceinline__ double hstep(double np, double x, double c) { return fma(np, x, -c); }

// e^x for the merger-ringdown amplitude's exponent (|x| below ~50 there, so
// no overflow or underflow handling): Cody and Waite's reduction x = k ln2 + r
// with k = rint(x / ln2) and ln2 split into a 32-bit head (k ln2_hi exact)
// and a tail, then e^r by its Taylor series through r^13 (|r| <= ln2/2, so
// the truncation is below 5e-18 relative) in Horner form, and 2^k by ldexp.
__device__ __forceinline__ double exp_mrd(double x) {
  const double k = rint(x * sreg(0x1.71547652b82fep+0));  // 1 / ln2
  double r = fma(k, -sreg(0x1.62e42ff000000p-1), x);      // ln2 head
  r = fma(k, sreg(0x1.718432a1b0e26p-35), r);             // minus the (negative) tail
  double p = r * sreg(0x1.6124613a86d09p-33) + sreg(0x1.1eed8eff8d898p-29);  // 1/13!, 1/12!
  p = fma(p, r, -sreg(-0x1.ae64567f544e4p-26));  // 1/11!
  p = fma(p, r, -sreg(-0x1.27e4fb7789f5cp-22));  // 1/10!
  p = fma(p, r, -sreg(-0x1.71de3a556c734p-19));  // 1/9!
  p = fma(p, r, -sreg(-0x1.a01a01a01a01ap-16));  // 1/8!
  p = fma(p, r, -sreg(-0x1.a01a01a01a01ap-13));  // 1/7!
  p = fma(p, r, -sreg(-0x1.6c16c16c16c17p-10));  // 1/6!
  p = fma(p, r, -sreg(-0x1.1111111111111p-7));   // 1/5!
  p = fma(p, r, -sreg(-0x1.5555555555555p-5));   // 1/4!
  p = fma(p, r, -sreg(-0x1.5555555555555p-3));   // 1/3!
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)k);
}

// 1/d for a positive normal d: the hardware reciprocal estimate and two
// Newton steps (each squares the relative error), to about an ulp.
__device__ __forceinline__ double rcp_pos(double d) {
  double y = __builtin_amdgcn_rcp(d);
  double e = fma(-d, y, 1.0);
  y = fma(y, e, y);
  e = fma(-d, y, 1.0);
  return fma(y, e, y);
}

// atan(x) for the merger-ringdown phase: |x| is reduced to [0, 1] by
// atan a = pi/2 - atan(1/a), then to [0, tan(pi/12)] by atan t = pi/6 +
// atan((sqrt3 t - 1)/(t + sqrt3)), where atan t = t R(t^2) with R of degree 9
// (Chebyshev fit in t^2 at 40 digits, generated by gen_atan.py beside this
// file; about 2 ulp overall).  Both reductions are selects, so the lanes of a
// wave take one path.
__device__ __forceinline__ double atan_mrd(double x) {
  const double a = fabs(x);
  const bool inv = a > 1.0;
  const double t1 = inv ? rcp_pos(a) : a;
  const double s3 = sreg(0x1.bb67ae8584caap+0);  // sqrt 3
  const bool big = t1 > sreg(0x1.126145e9ecd56p-2);  // tan(pi/12)
  const double t2 = big ? fma(t1, s3, -1.0) * rcp_pos(t1 + s3) : t1;
  const double z = t2 * t2;
  double p = z * sreg(-0x1.3a23485e6c654p-5) + sreg(0x1.d11bfd68cc957p-5);
  p = fma(p, z, -sreg(0x1.106138ecc7d04p-4));
  p = fma(p, z, -sreg(-0x1.3b0ae034ec65bp-4));
  p = fma(p, z, -sreg(0x1.745cd0c6674e0p-4));
  p = fma(p, z, -sreg(-0x1.c71c706e9f431p-4));
  p = fma(p, z, -sreg(0x1.24924922ada3cp-3));
  p = fma(p, z, -sreg(-0x1.9999999996e3ep-3));
  p = fma(p, z, -sreg(0x1.5555555555549p-2));
  p = fma(p, z, 1.0);
  double r = t2 * p;
  if (big) r += sreg(0x1.0c152382d7366p-1);        // pi/6
  if (inv) r = sreg(0x1.921fb54442d18p+0) - r;     // pi/2
  return copysign(r, x);
}

// The point records as the fused FP64 kernel reads them: through the
// constant address space (read-only while a kernel runs).  Their wave-uniform
// fields then stay scalar loads even after the compiler-opaque waits of the
// tile staging, which would otherwise send them down the vector memory path.
typedef const __attribute__((address_space(4))) Point CPoint;

// Point constants a segment's bin loop keeps in vector registers: the first
// Horner step of the inspiral amplitude, the intermediate amplitude and the
// inspiral phase, and the inspiral amplitude's w^2 term, each add two point
// constants, one of which would otherwise be copied into a vector register
// every bin.
template <int AR, int PR>
struct VConst {
  double a0 = 0.0, a1 = 0.0, p0 = 0.0, pL = 0.0;
  template <class PT>
  __device__ __forceinline__ explicit VConst(const PT& c) {
    if (AR == 0) {
      a0 = vreg_v(c.ia83);
      a1 = vreg_v(c.ia23);
    }
    if (AR == 1) a0 = vreg_v(c.im3);
    if (PR == 0) {
      p0 = vreg_v(c.ins_q53);
      pL = vreg_v(c.ins_qL);
    }
  }
};

// A wave-uniform double pinned in scalar registers (readfirstlane first, so
// that a value the compiler placed in vector registers can still go there).
__device__ __forceinline__ double spin(double v) {
  const unsigned long long u = __double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u), hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
  return sreg(__longlong_as_double((long long)(((unsigned long long)hi << 32) | lo)));
}

// The point constants of one region pair, read once per segment into scalar
// registers (the empty asm of sreg keeps each one there for the whole bin
// loop instead of re-loading it from the point record every bin: the records
// of the ~24 points resident on a CU would not stay in the scalar cache).
// Field names follow Point, so amp_turns reads either.
template <int AR, int PR>
struct SConst {
  double Msec, is6_scale, tshift[1];
  double ia0, ia23, ia1, ia43, ia53, ia2, ia73, ia3;
  double nf1, ib0, ib1, im2, ib4;
  double fRD, mrd_rate, mnum, mg2;
  double ins_q0, ins_q1, ins_q13, ins_q13L, ins_q23, ins_q43, ins_q2, ins_qm13, ins_qm23, ins_qm1, ins_qm53;
  double int_q0, int_q1, int_qL, int_qm3;
  double mrd_q0, mrd_q1, mrd_qm1, mrd_q34, mrd_qat, mrd_a5, mrd_ifDM;
  double wc[3], ws[3], dt[3], dl[3];
  double tsq;  // tshift_0 plus the phase's Mf-linear coefficient times Msec: the argument's f-linear term
  // ND detectors; the Horner rotation (wc, ws) for the FP64 kernel, the
  // detector phase offsets (dt, dl) for the mixed one
  template <int ND, bool ROT = true, class PT>
  __device__ __forceinline__ void load(const PT& c) {
    if (AR == 2 || PR == 2) Msec = spin(c.Msec);
    tshift[0] = spin(c.tshift[0]);
    tsq = spin(fma(PR == 0 ? c.ins_q1 : PR == 1 ? c.int_q1 : c.mrd_q1, c.Msec, c.tshift[0]));
    if (AR == 2 || PR == 2) is6_scale = spin(c.is6_scale);
    if (AR == 0) {
      ia0 = spin(c.ia0); ia1 = spin(c.ia1); ia43 = spin(c.ia43);
      ia53 = spin(c.ia53); ia2 = spin(c.ia2); ia73 = spin(c.ia73); ia3 = spin(c.ia3);
    } else if (AR == 1) {
      nf1 = spin(c.nf1);
      ib0 = spin(c.ib0); ib1 = spin(c.ib1); im2 = spin(c.im2); ib4 = spin(c.ib4);
    } else {
      fRD = spin(c.fRD); mrd_rate = spin(c.mrd_rate); mnum = spin(c.mnum); mg2 = spin(c.mg2);
    }
    if (PR == 0) {
      ins_q0 = spin(c.ins_q0); ins_q1 = spin(c.ins_q1); ins_q13 = spin(c.ins_q13); ins_q13L = spin(c.ins_q13L);
      ins_q23 = spin(c.ins_q23); ins_q43 = spin(c.ins_q43); ins_q2 = spin(c.ins_q2); ins_qm13 = spin(c.ins_qm13);
      ins_qm23 = spin(c.ins_qm23); ins_qm1 = spin(c.ins_qm1); ins_qm53 = spin(c.ins_qm53);
    } else if (PR == 1) {
      int_q0 = spin(c.int_q0); int_q1 = spin(c.int_q1); int_qL = spin(c.int_qL); int_qm3 = spin(c.int_qm3);
    } else {
      mrd_q0 = spin(c.mrd_q0); mrd_q1 = spin(c.mrd_q1); mrd_qm1 = spin(c.mrd_qm1); mrd_q34 = spin(c.mrd_q34);
      mrd_qat = spin(c.mrd_qat); mrd_a5 = spin(c.mrd_a5); mrd_ifDM = spin(c.mrd_ifDM);
    }
#pragma unroll
    for (int d = 1; d < ND; d++) {
      if (ROT) {
        wc[d] = spin(c.wc[d]);
        ws[d] = spin(c.ws[d]);
      } else {
        dt[d] = spin(c.dt[d]);
        dl[d] = spin(c.dl[d]);
      }
    }
  }
};

// amp_turns (below): amplitude without the detector scale, and the projected
// detector-0 argument in turns, of one bin in amplitude region AR and phase
// region PR (0 inspiral, 1 intermediate, 2 merger-ringdown), from the folded
// constants of the point and the bin's f, u = f^(1/3), e = f^(-1/6), log f.
// The inspiral and intermediate pieces are polynomials in u and E = e^2 =
// f^(-1/3) (Horner) and a quartic in f - f1 (Hz), with the point's powers of
// Msec folded into their coefficients, so they need no power of Mf; the
// merger-ringdown pieces work in Mf = Msec f.
template <int AR, int PR, class K>
__device__ __forceinline__ void amp_turns(const K& c, const VConst<AR, PR>& k, double f, double u,
                                          double e, double logf, double* a, double* r) {
  // powers of the bin's f: u = f^(1/3), e = f^(-1/6) from the planes, E = f^(-1/3)
  const double E = e * e;
  const double E2 = E * E;
  double E3 = 0.0;
  if (AR == 1 || PR == 1) E3 = E2 * E;
  // the merger-ringdown pieces work in Mf and its powers
  double Mf = 0.0, i_s6 = 0.0, w = 0.0, w3 = 0.0;
  if (AR == 2 || PR == 2) {
    Mf = c.Msec * f;
    vreg(Mf);  // Mf - node stays one subtraction with the node in a scalar register
    i_s6 = c.is6_scale * e;  // Mf^(-1/6)
    w = i_s6 * i_s6;         // Mf^(-1/3)
    w3 = (w * w) * w;        // 1 / Mf
  }
  if (AR == 0) {
    // e [E (ia0 E^2 + ia23) + ia1 + ia43 u + ... + ia3 u^6], the u part by
    // Horner on the negated partial sums
    double np = fma(u, -c.ia3, -k.a0);
    np = hstep(np, u, c.ia73);
    np = hstep(np, u, c.ia2);
    np = hstep(np, u, c.ia53);
    np = hstep(np, u, c.ia43);
    np = hstep(np, u, c.ia1);
    *a = e * fma(E, fma(E2, c.ia0, k.a1), -np);
  } else if (AR == 1) {
    // e E^3 times the quartic in x1 = f - f1 (the Newton form on the nodes
    // (f1, f1, f2, f3) expanded about f1), Horner on negated partials
    const double x1 = f - c.nf1;
    double np = fma(x1, -c.ib4, -k.a0);
    np = hstep(np, x1, c.im2);
    np = hstep(np, x1, c.ib1);
    *a = (e * E3) * -hstep(np, x1, c.ib0);
  } else {
    const double df = Mf - c.fRD;
    *a = (w3 * i_s6) * ((c.mnum * exp_mrd(-df * c.mrd_rate)) * rcp_pos(fma(df, df, c.mg2)));  // Mf^(-7/6)
  }
  // the projected argument: f tsq (tshift_0 plus the region's Mf-linear
  // coefficient times Msec, per point) in the last FMA, plus the rest of the
  // region's function of f
  double g;
  if (PR == 0) {
    // q0 + sum_k q_k u^k (k = 1, 2, 4, 5, 6) + sum_k q_-k E^k (k = 1, 2, 3, 5)
    // + (qL + q13L u) log f
    double np = fma(u, -c.ins_q2, -k.p0);
    np = hstep(np, u, c.ins_q43);
    np = np * u;  // no u^3 term: q1 Mf is added below
    np = hstep(np, u, c.ins_q23);
    np = hstep(np, u, c.ins_q13);  // -(q13 + q23 u + q43 u^3 + q53 u^4 + q2 u^5)
    double nn = -(E * c.ins_qm53);
    nn = hstep(nn, E, c.ins_qm1);
    nn = hstep(nn, E, c.ins_qm23);
    nn = hstep(nn, E, c.ins_qm13);  // -(qm13 + qm23 E + qm1 E^2 + qm53 E^4)
    // q0 rides in the log term's FMA, and the E chain's last step takes the rest as its addend
    const double lg = fma(fma(u, c.ins_q13L, k.pL), logf, c.ins_q0);
    g = fma(-nn, E, fma(-np, u, lg));  // q1 Mf is in tsq
  } else if (PR == 1) {
    g = fma(c.int_qm3, (E3 * E3) * E3, fma(c.int_qL, logf, c.int_q0));  // qm3 E^9 = qm3' / Mf^3
  } else {
    const double sq = Mf * (w * i_s6);  // Mf^(1/2)
    const double f34 = sqrt(Mf * sq);
    g = fma(c.mrd_qat, atan_mrd((Mf - c.mrd_a5) * c.mrd_ifDM), fma(c.mrd_q34, f34, c.mrd_qm1 * w3)) + c.mrd_q0;
  }
  *r = fma(f, c.tsq, g);
}

// Walks the segments of [pos, hi): SEG<AR, PR>::run(st, pos, end) for each
// run [pos, end) of one region pair.  Band-index region boundaries a1..p2.
template <template <int, int> class SEG, class ST>
__device__ __forceinline__ int run_segments(ST& st, int pos, int hi, int a1, int a2, int p1, int p2) {
  while (pos < hi) {
    const int ar = pos < a1 ? 0 : (pos >= a2 ? 2 : 1);
    const int pr = pos < p1 ? 0 : (pos >= p2 ? 2 : 1);
    int end = hi;
    if (a1 > pos && a1 < end) end = a1;
    if (a2 > pos && a2 < end) end = a2;
    if (p1 > pos && p1 < end) end = p1;
    if (p2 > pos && p2 < end) end = p2;
    switch (ar * 3 + pr) {
      case 0: SEG<0, 0>::run(st, pos, end); break;
      case 1: SEG<0, 1>::run(st, pos, end); break;
      case 2: SEG<0, 2>::run(st, pos, end); break;
      case 3: SEG<1, 0>::run(st, pos, end); break;
      case 4: SEG<1, 1>::run(st, pos, end); break;
      case 5: SEG<1, 2>::run(st, pos, end); break;
      case 6: SEG<2, 0>::run(st, pos, end); break;
      case 7: SEG<2, 1>::run(st, pos, end); break;
      default: SEG<2, 2>::run(st, pos, end); break;
    }
    pos = end;
  }
  return pos;
}

// ---- FP64
//
// Detector d > 0 sees the phase of detector 0 plus theta_d(f) = f dt_d + dl_d
// turns, linear in f.  A lane steps 64 bins at a time, so along its bins
// theta_d advances by the constant phi_d = 64 dt_d / Tobs, and the lane's sum
//   sum_m y_m e^{-i theta_d,m},  y_m = a_m D_m e^{-i x0_m} / S_m
// is a polynomial in e^{-i phi_d} evaluated by Horner's rule,
//   H <- H e^{i phi_d} + y_m,  sum = e^{-i theta_d,last} H,
// which takes one complex multiply-add per bin and detector and one sincos
// per lane and detector at the end.
//
// Bins and tiles.  The block's waves walk the split in 64-bin chunks on the
// split's own grid (chunk n holds bins j0 + 64 n + lane), so lane l of a
// point's wave takes every bin j = j0 + l (mod 64) of the point's range, in
// order, whatever the tiles or the other points of the block.  The planes
// come through LDS in tiles of HT = TILE / 2 bins, two buffers: while the
// waves read one tile, the next one arrives in the other buffer by
// buffer-to-LDS transfers (no registers, no LDS stores), and one barrier per
// tile hands it over.  A wave walks its point's segments (runs of bins over
// which neither the amplitude nor the phase changes region) each in a loop
// of its own that holds only that region pair's constants, and moves to the
// next tile from inside that loop, so a segment's constants are set up once
// per segment rather than once per tile.  Every wave takes every tile
// hand-over of the block in the same order (waves whose point has no bins
// left take the remaining hand-overs after their segments), so the barriers
// match.
// Buffer-to-LDS transfer of np planes of HT elements of E bytes (plane
// stride nbp elements, from element t on) into LDS at dst, plane after plane.
// Plane p goes to wave (p + rot) mod cpb as PER transfers of 64 lanes x LB
// bytes; the k-th carries the instruction offset k x 64 x LB, which the
// hardware adds to the global and the LDS address alike.  The whole byte
// offset but that is in the vector offset (the range check counts the scalar
// offset too), so elements past the last plane read as 0.  0 <= rot < cpb.
template <int E, int HT>
__device__ __forceinline__ void lds_fetch_planes(char* dst, const void* src, int np, int nbp, int t, int wave, int cpb,
                                                 int rot, unsigned lane16) {
  constexpr int LB = (HT * E) % 1024 == 0 ? 16 : 4;  // bytes per lane
  constexpr int PER = HT * E / (64 * LB);            // transfers per plane
  static_assert(PER == 1 || PER == 2, "one or two transfers a plane");
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(src), (short)0, np * nbp * E, 0x00020000);
  int p = wave - rot;
  if (p < 0) p += cpb;
  unsigned voff = (LB == 16 ? lane16 : lane16 / 4) + (unsigned)E * (unsigned)(p * nbp + t);
  const unsigned step = (unsigned)E * (unsigned)(cpb * nbp);
  for (; p < np; p += cpb, voff += step) {
    auto d = (__attribute__((address_space(3))) void*)(dst + p * HT * E);
    if constexpr (LB == 16) {
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, d, 16, voff, 0, 0, 0);
      if constexpr (PER == 2) __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, d, 16, voff, 0, 64 * 16, 0);
    } else {
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, d, 4, voff, 0, 0, 0);
      if constexpr (PER == 2) __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, d, 4, voff, 0, 64 * 4, 0);
    }
  }
}

template <int ND, int TILE>
struct F64State {
  static constexpr int NP = 4 + 3 * ND;
  static constexpr int HT = TILE / 2;  // bins per LDS buffer
  CPoint* c;
  double* lds;                // two buffers of NP x HT doubles
  const double* p64;          // the band planes from bin 0
  int nbp;                    // plane stride (doubles)
  int j0;                     // split start: origin of the 64-bin chunks
  int t0;                     // first band bin of the tile being read
  int cur;                    // its buffer
  int bend;                   // the block's last band bin + 1
  int wave, cpb;
  unsigned lane16;            // 16 x lane
  double hh[ND];     // sum a^2 / S_d
  double dh0;        // sum a Re(D_0 e^{-i x0}) / S_0
  double hr[ND], hi[ND];  // Horner sums of detectors d > 0
  SinCosK sk;
};

// Buffer-to-LDS transfer of the tile at band bin t into buffer q (the
// planes' layout in LDS: buffer, plane, bin).
template <int NP, int HT>
__device__ __forceinline__ void f64_fetch(double* lds, const double* p64, int nbp, int wave, int cpb, unsigned lane16,
                                          int t, int q) {
  lds_fetch_planes<8, HT>((char*)(lds + q * NP * HT), p64, NP, nbp, t, wave, cpb, 0, lane16);
}

template <int ND, int TILE>
__device__ __forceinline__ void f64_fetch(F64State<ND, TILE>& st, int t, int q) {
  f64_fetch<F64State<ND, TILE>::NP, F64State<ND, TILE>::HT>(st.lds, st.p64, st.nbp, st.wave, st.cpb, st.lane16, t, q);
}

// Hand-over to the next tile: this wave's transfers have landed and it is
// done with the current tile; after the barrier so are every wave's, the
// next tile is complete and the buffer just read can take the one after.
template <int ND, int TILE>
__device__ __forceinline__ void f64_advance(F64State<ND, TILE>& st) {
  constexpr int HT = F64State<ND, TILE>::HT;
  asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
  st.t0 += HT;
  st.cur ^= 1;
  if (st.t0 + HT < st.bend) f64_fetch(st, st.t0 + HT, st.cur ^ 1);
}

template <int ND, int TILE>
struct F64Seg {
  template <int AR, int PR>
  struct S {
    static __device__ __forceinline__ void bin(F64State<ND, TILE>& st, const double* L, const SConst<AR, PR>& c,
                                               const VConst<AR, PR>& k) {
      constexpr int HT = F64State<ND, TILE>::HT;
      double a, r;
      amp_turns<AR, PR>(c, k, L[0], L[HT], L[2 * HT], L[3 * HT], &a, &r);
      double us, uc;
      sincos_amp_turns(r, a, &us, &uc, st.sk);  // a sin, a cos
      // h_d = Ascale_d a e^{i x_d}: |h_d|^2 = Ascale_d^2 a^2, so the sums carry
      // a^2 / S and a D e^{-i x_d} / S, and k_finalize applies the per-point
      // detector scale once.
      const double a2 = a * a;
#pragma unroll
      for (int d = 0; d < ND; d++) {
        const double Dre = L[(4 + 3 * d) * HT], Dim = L[(5 + 3 * d) * HT];
        st.hh[d] = fma(a2, L[(6 + 3 * d) * HT], st.hh[d]);
        if (d == 0) {
          st.dh0 = fma(Dre, uc, fma(Dim, us, st.dh0));
        } else {
          const double yr = fma(Dre, uc, Dim * us), yi = fma(Dim, uc, -(Dre * us));
          const double tr = fma(-st.hi[d], c.ws[d], yr);
          fma_acc_x(st.hi[d], c.wc[d], yi);       // hi wc + yi
          fma_acc_z(st.hi[d], st.hr[d], c.ws[d]);  // + hr ws
          fma_acc_x(st.hr[d], c.wc[d], tr);       // hr wc - hi ws + yr
        }
      }
    }
    // The segment's bins [pos, end), chunk by chunk; lanes outside it idle.
    static __device__ __forceinline__ void run(F64State<ND, TILE>& st, int pos, int end) {
      constexpr int NP = F64State<ND, TILE>::NP, HT = F64State<ND, TILE>::HT;
      const VConst<AR, PR> k(*st.c);
      SConst<AR, PR> c;
      c.template load<ND>(*st.c);
      const int lane = threadIdx.x & 63;
      const unsigned span = (unsigned)(end - pos);
      const int ce = uniform(st.j0 + 64 * ((end - 1 - st.j0) >> 6));
      const double* Ll = st.lds + lane;  // the lane's bin of a chunk at the start of buffer 0
      for (int jb = uniform(st.j0 + 64 * ((pos - st.j0) >> 6)); jb <= ce; jb += 64) {
        while (jb >= st.t0 + HT) f64_advance(st);
        if ((unsigned)(jb + lane - pos) < span) bin(st, Ll + uniform(st.cur * NP * HT + (jb - st.t0)), c, k);
      }
    }
  };
};

template <int ND, int TILE>
__device__ void ll_f64_block(const LLInst& in, int group, int split, double* lds, int* s_lohi) {
  constexpr int NP = 4 + 3 * ND, HT = TILE / 2;
  const int cpb = blockDim.x >> 6;
  const int wave = uniform(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int b = group * cpb + wave;
  const int j0 = split * in.split_len, j1 = min(in.nb, j0 + in.split_len);
  const bool active = b < in.B && in.pts[b].status == ST_OK;
  int ks, ke, blo, bhi;
  block_range<HT>(in, b, active, j0, j1, s_lohi, &ks, &ke, &blo, &bhi);
  F64State<ND, TILE> st;
  st.c = (CPoint*)(in.pts + uniform(active ? b : 0));
  st.lds = lds;
  st.p64 = in.p64;
  st.nbp = in.nbp;
  st.j0 = j0;
  st.t0 = blo;
  st.cur = 0;
  st.bend = bhi;
  st.wave = wave;
  st.cpb = uniform(cpb);
  st.lane16 = 16u * (unsigned)lane;
  st.sk = sincos_consts_amp();
  st.dh0 = 0.0;
#pragma unroll
  for (int d = 0; d < ND; d++) { st.hh[d] = 0.0; st.hr[d] = 0.0; st.hi[d] = 0.0; }
  if (blo < bhi) {
    // first tile, then the second on its way while the waves read the first
    f64_fetch(st, blo, 0);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    if (blo + HT < bhi) f64_fetch(st, blo + HT, 1);
  }
  if (active && ks < ke) {
    CPoint& c = *st.c;
    run_segments<F64Seg<ND, TILE>::template S>(st, ks, ke, c.seg_a1 - in.imin, c.seg_a2 - in.imin,
                                               c.seg_p1 - in.imin, c.seg_p2 - in.imin);
  }
  while (st.t0 + HT < bhi) f64_advance(st);  // the hand-overs after this wave's last bin
  double acc[2 * ND];
#pragma unroll
  for (int d = 0; d < ND; d++) {
    double dh = st.dh0;
    if (d > 0) {
      // back from the lane's last bin: e^{-i theta_d,last} H (zero if the lane took no bin)
      dh = 0.0;
      int jl = j0 + 64 * ((ke - 1 - j0) >> 6) + lane;
      if (jl >= ke) jl -= 64;
      if (active && jl >= ks) {
        CPoint& c = *st.c;
        double s, co;
        sincos_turns(fma(in.p64[jl], c.dt[d], c.dl[d]), &s, &co);
        dh = fma(st.hr[d], co, st.hi[d] * s);
      }
    }
    acc[2 * d] = st.hh[d];
    acc[2 * d + 1] = dh;
  }
  wave_sums_store<ND>(acc, in.partial + ((size_t)b * in.splits + split) * (2 * ND), b < in.B);
}

// ---- mixed precision.  FP64: the waveform as amp_turns evaluates it (the
// powers of f, the amplitude and the projected detector-0 argument
// f (ts - td) + (lambda - Psi)/2 pi in turns), the argument's reduction to a
// fraction of a turn (v_fract_f64), which keeps 2 pi f t_c + Psi ~ 1e5 rad
// exact to ~1e-11, the detectors' Horner sums (their phases theta_d, linear
// in f, enter exactly, as in the FP64 kernel) and the per-lane sums.  FP32:
// sin/cos of the reduced detector-0 turn (hardware v_sin/v_cos take turns;
// one pair per bin, shared by every detector), the amplitude rounded once,
// and its products with the scaled FP32 data that feed the FP64 sums.  (An
// FP32 amplitude, its coefficients rounded to FP32, biases (h|h) by ~6e-8 of
// itself: 0.019 in logL at the loudest prior draws.)
//
// The block walks the split as the FP64 block does (lanes on the split's
// 64-bin grid, segment loops outermost, two LDS buffers of HT = TILE / 2 bins
// filled by buffer-to-LDS transfers, one barrier per tile).  A buffer holds
// the four FP64 frequency planes, then the 3 nd FP32 planes.
template <int ND, int TILE>
struct MixState {
  static constexpr int HT = TILE / 2;
  static constexpr int BYTES64 = 4 * HT * 8;             // the FP64 planes of a buffer
  static constexpr int BUFB = BYTES64 + 3 * ND * HT * 4;  // bytes per buffer
  const Point* c;
  char* lds;  // two buffers
  const double* p64;
  const float* p32;
  int nbp, j0, t0, cur, bend, wave, cpb;
  unsigned lane16;
  double hh[ND];          // sum a^2 / (S w_ref)
  double dh0;             // sum a Re(D_0 e^{-i x0}) / (S sqrt(w_ref))
  double hr[ND], hi[ND];  // Horner sums of detectors d > 0
};

template <int ND, int TILE>
__device__ __forceinline__ void mix_fetch(MixState<ND, TILE>& st, int t, int q) {
  constexpr int HT = MixState<ND, TILE>::HT;
  char* buf = st.lds + q * MixState<ND, TILE>::BUFB;
  // the four FP64 planes to waves 0-3, the FP32 planes from wave 4 on
  lds_fetch_planes<8, HT>(buf, st.p64, 4, st.nbp, t, st.wave, st.cpb, 0, st.lane16);
  lds_fetch_planes<4, HT>(buf + MixState<ND, TILE>::BYTES64, st.p32, 3 * ND, st.nbp, t, st.wave, st.cpb, 4 % st.cpb,
                          st.lane16);
}

template <int ND, int TILE>
__device__ __forceinline__ void mix_advance(MixState<ND, TILE>& st) {
  constexpr int HT = MixState<ND, TILE>::HT;
  asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
  st.t0 += HT;
  st.cur ^= 1;
  if (st.t0 + HT < st.bend) mix_fetch(st, st.t0 + HT, st.cur ^ 1);
}

template <int ND, int TILE>
struct MixSeg {
  template <int AR, int PR>
  struct S {
    static __device__ __forceinline__ void run(MixState<ND, TILE>& st, int pos, int end) {
      constexpr int HT = MixState<ND, TILE>::HT;
      const VConst<AR, PR> k(*st.c);
      SConst<AR, PR> c;
      c.template load<ND>(*st.c);
      const int lane = threadIdx.x & 63;
      const unsigned span = (unsigned)(end - pos);
      const int ce = uniform(st.j0 + 64 * ((end - 1 - st.j0) >> 6));
      for (int jb = uniform(st.j0 + 64 * ((pos - st.j0) >> 6)); jb <= ce; jb += 64) {
        while (jb >= st.t0 + HT) mix_advance(st);
        if ((unsigned)(jb + lane - pos) < span) {
          const char* buf = st.lds + uniform(st.cur * MixState<ND, TILE>::BUFB);
          const double* L = reinterpret_cast<const double*>(buf) + (jb - st.t0) + lane;
          const float* M = reinterpret_cast<const float*>(buf + MixState<ND, TILE>::BYTES64) + (jb - st.t0) + lane;
          const double f = L[0];
          double a64, r;
          amp_turns<AR, PR>(c, k, f, L[HT], L[2 * HT], L[3 * HT], &a64, &r);
          const float a = (float)a64;
          const float t = (float)__builtin_amdgcn_fract(r);  // the turn, reduced in FP64
          const float uc = a * __builtin_amdgcn_cosf(t), us = a * __builtin_amdgcn_sinf(t), a2 = a * a;
#pragma unroll
          for (int d = 0; d < ND; d++) {
            const float Dre = M[(3 * d) * HT], Dim = M[(3 * d + 1) * HT];
            st.hh[d] += (double)(a2 * M[(3 * d + 2) * HT]);
            if (d == 0) {
              st.dh0 += (double)fmaf(Dre, uc, Dim * us);
            } else {
              const double yr = (double)fmaf(Dre, uc, Dim * us), yi = (double)fmaf(Dim, uc, -(Dre * us));
              const double tr = fma(-st.hi[d], c.ws[d], yr);
              fma_acc_x(st.hi[d], c.wc[d], yi);       // hi wc + yi
              fma_acc_z(st.hi[d], st.hr[d], c.ws[d]);  // + hr ws
              fma_acc_x(st.hr[d], c.wc[d], tr);       // hr wc - hi ws + yr
            }
          }
        }
      }
    }
  };
};

template <int ND, int TILE>
__device__ void ll_mixed_block(const LLInst& in, int group, int split, double* lds, int* s_lohi) {
  constexpr int HT = MixState<ND, TILE>::HT;
  const int cpb = blockDim.x >> 6;
  const int wave = uniform(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int b = group * cpb + wave;
  const int j0 = split * in.split_len, j1 = min(in.nb, j0 + in.split_len);
  const bool active = b < in.B && in.pts[b].status == ST_OK;
  int ks, ke, blo, bhi;
  block_range<HT>(in, b, active, j0, j1, s_lohi, &ks, &ke, &blo, &bhi);
  MixState<ND, TILE> st;
  st.c = in.pts + uniform(active ? b : 0);
  st.lds = reinterpret_cast<char*>(lds);
  st.p64 = in.p64;
  st.p32 = in.p32;
  st.nbp = in.nbp;
  st.j0 = j0;
  st.t0 = blo;
  st.cur = 0;
  st.bend = bhi;
  st.wave = wave;
  st.cpb = uniform(cpb);
  st.lane16 = 16u * (unsigned)lane;
  st.dh0 = 0.0;
#pragma unroll
  for (int d = 0; d < ND; d++) { st.hh[d] = 0.0; st.hr[d] = 0.0; st.hi[d] = 0.0; }
  if (blo < bhi) {
    mix_fetch(st, blo, 0);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    if (blo + HT < bhi) mix_fetch(st, blo + HT, 1);
  }
  if (active && ks < ke) {
    const Point& c = *st.c;
    run_segments<MixSeg<ND, TILE>::template S>(st, ks, ke, c.seg_a1 - in.imin, c.seg_a2 - in.imin,
                                               c.seg_p1 - in.imin, c.seg_p2 - in.imin);
  }
  while (st.t0 + HT < bhi) mix_advance(st);  // the hand-overs after this wave's last bin
  double acc[2 * ND];
#pragma unroll
  for (int d = 0; d < ND; d++) {
    double dh = st.dh0;
    if (d > 0) {
      // back from the lane's last bin: e^{-i theta_d,last} H (zero if the lane took no bin)
      dh = 0.0;
      int jl = j0 + 64 * ((ke - 1 - j0) >> 6) + lane;
      if (jl >= ke) jl -= 64;
      if (active && jl >= ks) {
        const Point& c = *st.c;
        double sn, co;
        sincos_turns(fma(in.p64[jl], c.dt[d], c.dl[d]), &sn, &co);
        dh = fma(st.hr[d], co, st.hi[d] * sn);
      }
    }
    acc[2 * d] = st.hh[d];
    acc[2 * d + 1] = dh;
  }
  wave_sums_store<ND>(acc, in.partial + ((size_t)b * in.splits + split) * (2 * ND), b < in.B);
}

template <int PREC, int TILE>
__device__ __forceinline__ void ll_block(const LLInst& in, int group, int split, double* lds, int* s_lohi) {
  if (PREC == 0) {
    if (in.nd == 1) ll_f64_block<1, TILE>(in, group, split, lds, s_lohi);
    else if (in.nd == 2) ll_f64_block<2, TILE>(in, group, split, lds, s_lohi);
    else ll_f64_block<3, TILE>(in, group, split, lds, s_lohi);
  } else {
    if (in.nd == 1) ll_mixed_block<1, TILE>(in, group, split, lds, s_lohi);
    else if (in.nd == 2) ll_mixed_block<2, TILE>(in, group, split, lds, s_lohi);
    else ll_mixed_block<3, TILE>(in, group, split, lds, s_lohi);
  }
}

// One analysis: grid x = point groups, y = splits; up to 16 points (waves)
// per block.
template <int ND, int PREC, int TILE>
__global__ __launch_bounds__(1024) void k_ll(LLInst in) {
  extern __shared__ double lds[];
  __shared__ int s_lohi[2];
  if (PREC == 0)
    ll_f64_block<ND, TILE>(in, blockIdx.x, blockIdx.y, lds, s_lohi);
  else
    ll_mixed_block<ND, TILE>(in, blockIdx.x, blockIdx.y, lds, s_lohi);
}

// Many analyses in one persistent launch: every block takes (analysis, point
// group, split) items from a queue until it is empty.  The exit test reads a
// block-uniform item index, so every wave of every block reaches it.
// The queue kernel holds its arguments and loop state in scalar registers
// beside the segment constants, and the compiler places the vector register
// that takes the scalar spills above the others: the occupancy hint keeps it
// within the 80 VGPRs of 6 waves per SIMD, the count the LDS tile allows.
template <int PREC, int TILE>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(6))) void k_ll_queue(const LLInst* __restrict__ insts, const LLItem* __restrict__ items,
                                                  int n_items, int* __restrict__ counter) {
  extern __shared__ double lds[];
  __shared__ int s_lohi[2];
  __shared__ int s_item;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    const int it = uniform(s_item);
    if (it >= n_items) break;
    const LLItem w = items[it];
    const LLInst in = insts[w.inst];
    ll_block<PREC, TILE>(in, w.group, w.split, lds, s_lohi);
  }
}

// ------------------------------------------------------------ finalize
// logL = sum_d [HD_d - HH_d/2], rho_d = HD_d / sqrt(HH_d) with HH = 4 sum |h|^2/S,
// HD = 4 sum Re(d h*)/S (QuickCBC.c:4259-4266, Utilities.c:716-736).  Splits are summed in index order, so a point's
// result does not depend on what else shares the launch.  mode FIN_RAW: the
// partials are the sums themselves; FIN_FP64: they lack the factors Ascale_d^2
// and Ascale_d; FIN_MIXED: also the plane scales w_ref and sqrt(w_ref).
__device__ __forceinline__ void finalize_point(const double* __restrict__ partial, int splits, int nd,
                                               const Point& c, int b, int mode, const double* __restrict__ wref,
                                               double* __restrict__ logL, double* __restrict__ rho) {
  if (c.status != ST_OK) {
    logL[b] = 0.0;
    for (int d = 0; d < nd; d++) rho[(size_t)b * nd + d] = 0.0;
    return;
  }
  double L = 0.0;
  for (int d = 0; d < nd; d++) {
    double HH = 0.0, HD = 0.0;
    for (int s = 0; s < splits; s++) {
      HH += partial[((size_t)b * splits + s) * (2 * nd) + 2 * d];
      HD += partial[((size_t)b * splits + s) * (2 * nd) + 2 * d + 1];
    }
    if (mode == FIN_FP64) {
      const double A = c.Ascale[d];
      HH *= A * A;
      HD *= A;
    } else if (mode == FIN_MIXED) {
      const double A = c.Ascale[d];
      HH *= (A * A) * wref[d];
      HD *= A * sqrt(wref[d]);
    }
    HH *= 4.0;  // fourier_nwip returns 4 sum (Utilities.c:734)
    HD *= 4.0;
    rho[(size_t)b * nd + d] = HD / sqrt(HH);
    L += HD - 0.5 * HH;
  }
  logL[b] = L;
}

__global__ void k_finalize(const double* __restrict__ partial, int splits, int nd, const Point* __restrict__ pts,
                           int B, int mode, const double* __restrict__ wref, double* __restrict__ logL,
                           double* __restrict__ rho) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  finalize_point(partial, splits, nd, pts[b], b, mode, wref, logL, rho);
}

// log_likelihood_intrinsic's sums (QuickCBC.c:4113-4136) from the FP64
// partials: per detector x = HD - HH/2 and snrsq = 2 x, logL the sum of the x;
// with more than one detector logL = 0 unless at least two have snrsq > 25.
__global__ void k_finalize_intrinsic(const double* __restrict__ partial, int splits, int nd,
                                     const Point* __restrict__ pts, int B, double* __restrict__ logL,
                                     double* __restrict__ snrsq) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const Point& c = pts[b];
  if (c.status != ST_OK) {
    logL[b] = 0.0;
    for (int d = 0; d < nd; d++) snrsq[(size_t)b * nd + d] = 0.0;
    return;
  }
  double L = 0.0;
  int loud = 0;
  for (int d = 0; d < nd; d++) {
    double HH = 0.0, HD = 0.0;
    for (int s = 0; s < splits; s++) {
      HH += partial[((size_t)b * splits + s) * (2 * nd) + 2 * d];
      HD += partial[((size_t)b * splits + s) * (2 * nd) + 2 * d + 1];
    }
    const double A = c.Ascale[d];
    HH *= A * A;
    HD *= A;
    HH *= 4.0;  // fourier_nwip returns 4 sum (Utilities.c:734)
    HD *= 4.0;
    const double x = HD - 0.5 * HH;
    snrsq[(size_t)b * nd + d] = 2.0 * x;
    if (2.0 * x > 25.0) loud++;
    L += x;
  }
  if (nd > 1 && loud < 2) L = 0.0;  // a signal in at least two detectors
  logL[b] = L;
}

hipError_t launch_finalize_intrinsic(const double* partial, int splits, int nd, const Point* pts, int B,
                                     double* logL, double* snrsq, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_finalize_intrinsic, dim3((B + 127) / 128), dim3(128), 0, s, partial, splits, nd, pts, B, logL,
                     snrsq);
  return hipGetLastError();
}

// ------------------------------------------------------------ launchers
size_t ll_lds_bytes(int precision, int nd, int tile) {
  if (precision == 0) return (size_t)(4 + 3 * nd) * tile * sizeof(double);
  return (size_t)4 * tile * sizeof(double) + (size_t)3 * nd * tile * sizeof(float);
}

bool ll_tile_ok(int tile) { return tile == 128 || tile == 256 || tile == 512; }

template <int TILE>
static void launch_ll_t(int precision, const LLInst& in, dim3 grid, dim3 block, size_t lds, hipStream_t s) {
  if (precision == 0) {
    switch (in.nd) {
      case 1: hipLaunchKernelGGL((k_ll<1, 0, TILE>), grid, block, lds, s, in); break;
      case 2: hipLaunchKernelGGL((k_ll<2, 0, TILE>), grid, block, lds, s, in); break;
      default: hipLaunchKernelGGL((k_ll<3, 0, TILE>), grid, block, lds, s, in); break;
    }
  } else {
    switch (in.nd) {
      case 1: hipLaunchKernelGGL((k_ll<1, 1, TILE>), grid, block, lds, s, in); break;
      case 2: hipLaunchKernelGGL((k_ll<2, 1, TILE>), grid, block, lds, s, in); break;
      default: hipLaunchKernelGGL((k_ll<3, 1, TILE>), grid, block, lds, s, in); break;
    }
  }
}

hipError_t launch_ll(int precision, const LLInst& in, int cpb, int tile, hipStream_t s) {
  if (in.B <= 0) return hipSuccess;
  if (!ll_tile_ok(tile) || in.nd < 1 || in.nd > 3) return hipErrorInvalidValue;
  dim3 grid((in.B + cpb - 1) / cpb, in.splits), block(64 * cpb);
  size_t lds = ll_lds_bytes(precision, in.nd, tile);
  if (tile == 128) launch_ll_t<128>(precision, in, grid, block, lds, s);
  else if (tile == 256) launch_ll_t<256>(precision, in, grid, block, lds, s);
  else launch_ll_t<512>(precision, in, grid, block, lds, s);
  return hipGetLastError();
}

template <int TILE>
static void launch_queue_t(int precision, const LLInst* d_insts, const LLItem* d_items, int n_items, int* d_counter,
                           dim3 grid, dim3 block, size_t lds, hipStream_t s) {
  if (precision == 0)
    hipLaunchKernelGGL((k_ll_queue<0, TILE>), grid, block, lds, s, d_insts, d_items, n_items, d_counter);
  else
    hipLaunchKernelGGL((k_ll_queue<1, TILE>), grid, block, lds, s, d_insts, d_items, n_items, d_counter);
}

hipError_t launch_ll_queue(int precision, const LLInst* d_insts, const LLItem* d_items, int n_items, int* d_counter,
                           int max_nd, int cpb, int tile, int blocks, hipStream_t s) {
  if (!ll_tile_ok(tile)) return hipErrorInvalidValue;
  size_t lds = ll_lds_bytes(precision, max_nd, tile);
  dim3 grid(blocks), block(64 * cpb);
  if (tile == 128) launch_queue_t<128>(precision, d_insts, d_items, n_items, d_counter, grid, block, lds, s);
  else if (tile == 256) launch_queue_t<256>(precision, d_insts, d_items, n_items, d_counter, grid, block, lds, s);
  else launch_queue_t<512>(precision, d_insts, d_items, n_items, d_counter, grid, block, lds, s);
  return hipGetLastError();
}

hipError_t launch_finalize(const double* partial, int splits, int nd, const Point* pts, int B, int mode,
                           const double* wref, double* logL, double* rho, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_finalize, dim3((B + 127) / 128), dim3(128), 0, s, partial, splits, nd, pts, B, mode, wref, logL, rho);
  return hipGetLastError();
}

__global__ void k_fill_i32(int* p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_fill_i32(int* p, int n, int v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_i32, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

// ------------------------------------------------------ PT ensemble step
// Counter-based random numbers: SplitMix64 finaliser of (seed, step, point,
// draw), so a step's draws depend only on its counters (graph replays and
// reruns give the same chain).  Points are numbered p = t n_chains + c
// (temperature t, chain column c) for the random counters whatever the
// storage order: a group of chain columns [c0, c0 + ncols) stores its points
// as t ncols + (c - c0), and every group sees the draws the whole ensemble
// would.

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t seed, uint64_t a, uint64_t b) {
  uint64_t r = mix64(seed ^ mix64(a * 0x2545F4914F6CDD1DULL + b));
  return ((r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// Per-temperature constants, tc[3 t .. 3 t + 2] = sqrt(T_t), 1 / T_t,
// ladder^-t with T_t = ladder^t.
__global__ void k_pt_temps(double* __restrict__ tc, int n_temps, double ladder) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_temps) return;
  const double T = pow(ladder, (double)t);
  tc[3 * t] = sqrt(T);
  tc[3 * t + 1] = 1.0 / T;
  tc[3 * t + 2] = pow(ladder, -(double)t);
}

// Gaussian jiggle proposal (QuickCBC.c:2290-2291 shape): x' = x + sqrt(T) s z,
// z by Box-Muller from the counters (seed, step, point); parameters i and
// i + 1 (i even) share one pair of uniforms.  lp: storage index, gp: point
// number of the random counters.
__device__ __forceinline__ void propose_pair(const double* __restrict__ x, double* __restrict__ y,
                                             const double* __restrict__ sigma, double sT, int lp, uint64_t gp, int i,
                                             uint64_t P, uint64_t seed, uint64_t st) {
  double u1 = u01(seed, st * P + gp, 2 * i), u2 = u01(seed, st * P + gp, 2 * i + 1);
  double r = sqrt(-2.0 * log(u1));
  double s, c;
  sincospi(2.0 * u2, &s, &c);
  y[11 * lp + i] = x[11 * lp + i] + sT * sigma[i] * r * c;
  if (i + 1 < 11) y[11 * lp + i + 1] = x[11 * lp + i + 1] + sT * sigma[i + 1] * r * s;
}

// Proposals of one group at step *step (used once, at creation; afterwards
// k_pt_post makes them).
__global__ void k_pt_propose(const double* __restrict__ x, double* __restrict__ y, const double* __restrict__ sigma,
                             const double* __restrict__ tc, int n_temps, int n_chains, int c0, int ncols,
                             uint64_t seed, const unsigned long long* step) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 6 * n_temps * ncols) return;
  const int lp = q / 6, t = lp / ncols, c = c0 + lp % ncols;
  propose_pair(x, y, sigma, tc[3 * t], lp, (uint64_t)t * n_chains + c, 2 * (q % 6), (uint64_t)n_temps * n_chains,
               seed, *step);
}

__device__ __forceinline__ bool in_prior(const double* q) {
  return q[2] > -1.0 && q[2] < 1.0 && q[3] > -1.0 && q[3] < 1.0 && q[8] > -1.0 && q[8] < 1.0 && q[10] > -1.0 &&
         q[10] < 1.0 && q[0] < q[1] + (-0.8317766166719343);  // eta <= 1/4: ln Mc <= ln Mt + 0.6 ln(1/4)
}

// Everything of a PT step after the likelihood partial sums, in one launch,
// for the group of chain columns [c0, c0 + ncols): block b owns column
// c0 + b (its n_temps points).  logL of the proposals (as k_finalize),
// Metropolis acceptance at inverse temperature 1/T with flat priors inside
// the bounds, the adjacent-temperature swaps of the step (even pairs on even
// steps, odd pairs on odd steps), then the proposals of the next step.  The
// last block to finish advances the group's step counter and re-arms the
// completion count, so the kernel replays in a graph with no fill or memset
// node.
__global__ __launch_bounds__(1024) void k_pt_post(
    const double* __restrict__ partial, int splits, int nd, const Point* __restrict__ pts, int mode,
    const double* __restrict__ wref, double* __restrict__ x, double* __restrict__ L, double* __restrict__ y,
    double* __restrict__ Ly, double* __restrict__ rho, const int* __restrict__ st_y, const double* __restrict__ sigma,
    const double* __restrict__ tc, int n_temps, int n_chains, int c0, int ncols, uint64_t seed,
    unsigned long long* __restrict__ step, unsigned long long* __restrict__ acc,
    unsigned long long* __restrict__ swaps, unsigned int* __restrict__ done) {
  const int b = blockIdx.x, c = c0 + b;
  const uint64_t P = (uint64_t)n_temps * n_chains;
  const uint64_t s = *step;
  for (int t = threadIdx.x; t < n_temps; t += blockDim.x) {
    const int lp = t * ncols + b;
    const uint64_t gp = (uint64_t)t * n_chains + c;
    if (mode != FIN_DONE) finalize_point(partial, splits, nd, pts[lp], lp, mode, wref, Ly, rho);
    const double u = u01(seed ^ 0x5bd1e995ULL, s * P + gp, 99);
    const double* q = y + 11 * lp;
    if (st_y[lp] == 0 && in_prior(q) && log(u) < tc[3 * t + 1] * (Ly[lp] - L[lp])) {
      for (int i = 0; i < 11; i++) x[11 * lp + i] = q[i];
      L[lp] = Ly[lp];
      acc[lp] += 1;
    }
  }
  __syncthreads();
  for (int t = (int)(s & 1) + 2 * (int)threadIdx.x; t + 1 < n_temps; t += 2 * (int)blockDim.x) {
    const int la = t * ncols + b, lb = (t + 1) * ncols + b;
    const double u = u01(seed ^ 0x27d4eb2fULL, s * P + (uint64_t)t * n_chains + c, 7);
    if (log(u) < (tc[3 * t + 2] - tc[3 * t + 5]) * (L[lb] - L[la])) {
      for (int i = 0; i < 11; i++) {
        double tmp = x[11 * la + i];
        x[11 * la + i] = x[11 * lb + i];
        x[11 * lb + i] = tmp;
      }
      double tl = L[la];
      L[la] = L[lb];
      L[lb] = tl;
      atomicAdd(&swaps[c], 1ULL);
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < 6 * n_temps; q += blockDim.x) {
    const int t = q / 6;
    propose_pair(x, y, sigma, tc[3 * t], t * ncols + b, (uint64_t)t * n_chains + c, 2 * (q % 6), P, seed, s + 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(done, 1u) == gridDim.x - 1) {
      *step = s + 1;
      *done = 0u;
      __threadfence();
    }
  }
}

hipError_t launch_pt_temps(double* tc, int n_temps, double ladder, hipStream_t s) {
  hipLaunchKernelGGL(k_pt_temps, dim3((n_temps + 63) / 64), dim3(64), 0, s, tc, n_temps, ladder);
  return hipGetLastError();
}

hipError_t launch_pt_propose(const double* x, double* y, const double* sigma, const double* tc, int n_temps,
                             int n_chains, int c0, int ncols, uint64_t seed, const unsigned long long* step,
                             hipStream_t s) {
  const int n = 6 * n_temps * ncols;
  hipLaunchKernelGGL(k_pt_propose, dim3((n + 127) / 128), dim3(128), 0, s, x, y, sigma, tc, n_temps, n_chains, c0,
                     ncols, seed, step);
  return hipGetLastError();
}

hipError_t launch_pt_post(const LLInst& in, int mode, const double* wref, double* x, double* L, double* y, double* Ly,
                          double* rho, const int* st_y, const double* sigma, const double* tc, int n_temps,
                          int n_chains, int c0, int ncols, uint64_t seed, unsigned long long* step,
                          unsigned long long* acc, unsigned long long* swaps, unsigned int* done, hipStream_t s) {
  const int threads = 6 * n_temps >= 1024 ? 1024 : ((6 * n_temps + 63) / 64) * 64;
  hipLaunchKernelGGL(k_pt_post, dim3(ncols), dim3(threads), 0, s, in.partial, in.splits, in.nd, in.pts, mode, wref,
                     x, L, y, Ly, rho, st_y, sigma, tc, n_temps, n_chains, c0, ncols, seed, step, acc, swaps, done);
  return hipGetLastError();
}

__global__ void k_fill_u64(unsigned long long* p, int n, unsigned long long v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_fill_u64(unsigned long long* p, int n, unsigned long long v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u64, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

}  // namespace qcbc

namespace qcbc {
// Plain device copy, 16 bytes per lane: the calibration reference for the HBM
// byte counters (a copy of n bytes reads n and writes n).
__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
  size_t n16 = bytes / 16;
  hipLaunchKernelGGL(k_copy, dim3(2048), dim3(256), 0, s, (const uint4*)src, (uint4*)dst, n16);
  return hipGetLastError();
}
}  // namespace qcbc
