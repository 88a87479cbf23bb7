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

// Frequency-independent state of one parameter point (~90 doubles).
struct alignas(16) Point {
  int status, kstart, kend, region_flags;
  double Msec, fs, fCut, amp0;
  double s6_scale, is6_scale, logMsec, inv_eta;
  // amplitude: amp_pre = sqrt(2/3) sqrt(eta) / pi^(1/6) (amp0Func)
  double amp_pre, ai_23, ai_1, ai_43, ai_53, ai_2, ai_73, ai_83, ai_3;
  double fmaxCalc;
  // intermediate amplitude in Newton form on nodes (f1, f1, f2, f3)
  double col_f1, col_f2, col_f3, col_c0, col_c1, col_c2, col_c3, col_c4;
  double fRD, fDM, fDMg3, mrd_num, mrd_rate;  // mrd_num = fDM g3 g1, mrd_rate = g2/(fDM g3)
  // phase
  double pi_init, pi_23, pi_13, pi_13L, pi_L, pi_m13, pi_m23, pi_m1, pi_m53;
  double pi_1, pi_43, pi_53, pi_2;
  double fPhiMRD;
  double beta1, beta2, beta3, C1Int, C2Int;
  double alpha1, alpha2, alpha3, alpha4, a5fRD, C1MRD, C2MRD;
  double t0, MfRef, phi_precalc;
  // projection per detector slot
  double Fs[3], tshift[3], lambda[3], Ascale[3];
  double eta, finspin;
};

// One analysis as the fused likelihood kernels see it.  Bin splits are
// split_len band bins long (a multiple of the LDS tile) so that tile and
// split boundaries depend only on the analysis, never on the batch.
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
