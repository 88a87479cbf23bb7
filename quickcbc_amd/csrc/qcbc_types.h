// Host/device-neutral records of the waveform + likelihood path.
#pragma once

namespace qcbc {

enum { ST_OK = 0, ST_EDOM = 33 };

// Analysis-wide constants the per-point set-up needs.
struct AnalysisConst {
  double Tobs, inv_T, sqT, gmst, fref, fmin0;  // fmin0 = freq[0] = 1/Tobs
  int n_det, labels[3];
  int imin, imax, N;
};

// Frequency-independent state of one parameter point (~130 doubles).  The
// fields the fused likelihood kernels read in their bin loops come first,
// grouped by region so that a segment's constants sit on few 64-byte lines
// (the scalar cache holds the records of every point resident on a CU).
struct alignas(16) Point {
  int status, kstart, kend, region_flags;
  // ---- fused likelihood kernels: the same waveform with its constants
  // folded.  First bins (absolute index k) at which the amplitude leaves the
  // inspiral (Mf >= 0.014) and enters the merger-ringdown (Mf >= fmaxCalc),
  // and the phase leaves the inspiral (Mf >= 0.018) and enters the
  // merger-ringdown (Mf >= fPhiMRD); clipped to [kstart, kend].
  int seg_a1, seg_a2, seg_p1, seg_p2;
  double Msec, s6_scale, is6_scale, v_scale;  // Msec^(1/6), Msec^(-1/6), Msec^(1/3)
  // projection: detector time shifts (s), and for detector d the rotation of
  // its phase relative to detector 0 over 64 bins, exp(i 2 pi 64 dt_d / Tobs)
  // = wc_d + i ws_d
  double tshift[3];
  double wc[3], ws[3];
  // Projected detector-0 argument in turns, f tshift_0 + (lambda_0 - Psi)/2 pi,
  // is f tshift_0 plus a function of f per phase region, in the bin's
  // u = f^(1/3) and E = f^(-1/3) (Mf^(1/3) = Msec^(1/3) u, Mf^(-1/3) =
  // Msec^(-1/3) E: the coefficients carry those powers):
  //   inspiral:      q0 + q1 Mf + u (q13 + q13L log f) + qL log f + q23 u^2
  //                  + q43 u^4 + q53 u^5 + q2 u^6 + qm13 E + qm23 E^2 + qm1 E^3 + qm53 E^5
  //   intermediate:  q0 + q1 Mf + qL log f + qm3 E^9
  //   merger:        q0 + q1 Mf + qm1 / Mf + q34 Mf^(3/4) + qat atan((Mf - a5) ifDM)
  double ins_q0, ins_q1, ins_q13, ins_q13L, ins_qL, ins_q23, ins_q43, ins_q53, ins_q2;
  double ins_qm13, ins_qm23, ins_qm1, ins_qm53;
  // amplitude without Ascale, with e = f^(-1/6):
  //   inspiral:      e [E (ia0 E^2 + ia23) + ia1 + ia43 u + ia53 u^2 + ia2 u^3 + ia73 u^4 + ia83 u^5 + ia3 u^6]
  //                  (Mf^(-7/6) amp_pre (1 + ai_23 Mf^(2/3) + ... + ai_3 Mf^3))
  //   intermediate:  e E^3 x the quartic in x1 = f - nf1 (Hz): ib0 + ib1 x1 + im2 x1^2 + im3 x1^3 + ib4 x1^4
  //   merger:        Mf^(-7/6) mnum exp(-(Mf - fRD) mrd_rate) / ((Mf - fRD)^2 + fDMg3^2)
  double ia0, ia23, ia1, ia43, ia53, ia2, ia73, ia83, ia3;
  double int_q0, int_q1, int_qL, int_qm3;
  // intermediate amplitude in Newton form on nodes (f1, f1, f2, f3) in Mf
  double col_f1, col_f2, col_f3;
  double ib0, ib1, ib2, ib3, ib4;
  double nf1;       // the first node in Hz (col_f1 / Msec)
  double im2, im3;  // the Newton form's x1^2 and x1^3 coefficients as a polynomial in x1 = f - nf1
  double fRD, mrd_rate, mnum, mg2;  // mrd_rate = g2/(fDM g3)
  double mrd_q0, mrd_q1, mrd_qm1, mrd_q34, mrd_qat, mrd_a5, mrd_ifDM;
  // detector d relative to detector 0: f dt_d + dl_d turns
  double dt[3], dl[3];
  // ---- the reference's set-up (IMRPhenomD coefficients, projection)
  double fs, fCut, amp0;
  double logMsec, inv_eta;
  // amplitude: amp_pre = sqrt(2/3) sqrt(eta) / pi^(1/6) (amp0Func)
  double amp_pre, ai_23, ai_1, ai_43, ai_53, ai_2, ai_73, ai_83, ai_3;
  double fmaxCalc;
  double col_c0, col_c1, col_c2, col_c3, col_c4;
  double fDM, fDMg3, mrd_num;  // mrd_num = fDM g3 g1
  // phase
  double pi_init, pi_23, pi_13, pi_13L, pi_L, pi_m13, pi_m23, pi_m1, pi_m53;
  double pi_1, pi_43, pi_53, pi_2;
  double fPhiMRD;
  double beta1, beta2, beta3, C1Int, C2Int;
  double alpha1, alpha2, alpha3, alpha4, a5fRD, C1MRD, C2MRD;
  double t0, MfRef, phi_precalc;
  // projection per detector slot
  double Fs[3], lambda[3], Ascale[3];
  double eta, finspin;
};

// One analysis as the fused likelihood kernels see it.  Bin splits are
// split_len band bins long (a multiple of 64; tiles start at each split's
// first bin) so that tile and split boundaries depend only on the analysis,
// never on the batch.
struct LLInst {
  const double* p64;
  const float* p32;
  const Point* pts;
  double* partial;   // (B x splits x 2 nd) partial (h|h), (d|h)
  double inv_T;
  int nbp, imin, nb, nd;
  int B, split_len, splits, pad;
};

// One work item of the persistent kernel: (analysis, group of points, split).
struct LLItem {
  int inst, group, split, pad;
};

}  // namespace qcbc
