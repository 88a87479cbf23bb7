/*
 * CPU oracle for the QuickCBC waveform + likelihood path.  TEST
 * INFRASTRUCTURE ONLY: nothing in the product links or calls this; it is the
 * checker that the HIP path is compared against (and, timed, the CPU
 * baseline).  Plain C99, FP64, GSL-free; the QuickCBC function each routine
 * restates is cited beside it.
 *
 * Parity status of each part (what pins it) is listed in qcbc_oracle.c.
 */
#ifndef QCBC_ORACLE_H
#define QCBC_ORACLE_H

#include <stddef.h>

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-point IMRPhenomD state: every frequency-independent quantity that the
 * reference computes before its per-bin loop (IMRPhenomD.c:413-523). */
typedef struct {
  double m1, m2, chi1, chi2;       /* after the m1 > m2 swap, solar masses */
  double M, eta, Msec, amp0, fCut, finspin;
  double chi;                      /* chiPN */
  double fRD, fDM;
  /* amplitude (IMRPhenomD_internals.h:739-776) */
  double gamma1, gamma2, gamma3, fmaxCalc, rho1, rho2, rho3;
  double f1, f2, f3, v1, v2, v3, d1, d2;
  double delta[5];
  /* inspiral amplitude prefactors (IMRPhenomD_internals.h:891-906) */
  double ai_amp0, ai_23, ai_1, ai_43, ai_53, ai_2, ai_73, ai_83, ai_3;
  /* phase (IMRPhenomD_internals.h:781-818) */
  double sigma1, sigma2, sigma3, sigma4, beta1, beta2, beta3;
  double alpha1, alpha2, alpha3, alpha4, alpha5;
  double C1Int, C2Int, C1MRD, C2MRD, fPhiIns, fPhiMRD;
  double pn_v[13], pn_vlogv[13];
  /* inspiral phase prefactors (IMRPhenomD_internals.h:914-929) */
  double pi_init, pi_23, pi_13, pi_13L, pi_L, pi_m13, pi_m23, pi_m1, pi_m53;
  double pi_1, pi_43, pi_53, pi_2;
  double t0, MfRef, phifRef, phi_precalc;
} qo_phenomd;

/* status codes: 0 ok, 33 domain (IMRPhenomD_internals.h:675-682) */
int qo_phenomd_setup(qo_phenomd *p, double m1_msun, double m2_msun, double chi1,
                     double chi2, double distance, double phi0, double fref);
void qo_phenomd_eval(const qo_phenomd *p, double f, double *amp, double *phase);
double qo_amp_ins_ansatz(const qo_phenomd *p, double Mf);
double qo_amp_int_ansatz(const qo_phenomd *p, double Mf);
double qo_amp_mrd_ansatz(const qo_phenomd *p, double Mf);
double qo_damp_ins_ansatz(const qo_phenomd *p, double Mf);
double qo_damp_mrd_ansatz(const qo_phenomd *p, double Mf);
double qo_phase_region(const qo_phenomd *p, double Mf, int region); /* 0 ins, 1 int, 2 mrd */
double qo_final_spin(double eta, double chi1, double chi2);
double qo_erad(double eta, double chi1, double chi2);
int qo_qnm(double finspin, double *fring_raw, double *fdamp_raw);
double qo_fit(int which, double eta, double chi);

/* IMRPhenomDGenerateh22FDAmpPhase restated (IMRPhenomD.c:186-248, :413-563).
 * m1_SI, m2_SI in kg; distance in m.  Returns 0 or an error code. */
int qo_h22_ampphase(const double *freq, size_t n, double phi0, double fref,
                    double m1_SI, double m2_SI, double chi1, double chi2,
                    double distance, double *amp, double *phase);

double qo_fbegin(const double *params);                          /* Utilities.c:587-629 */
void qo_times(double alpha, double sindelta, double gmst, double *dtimes); /* Utilities.c:1078-1109 */
void qo_f_ant(double psi, double alpha, double sindelta, double gmst,
              double *fplus, double *fcross, int id);             /* Utilities.c:995-1076 */
double qo_fourier_nwip(const double *a, const double *b, const double *Sn,
                       int imin, int imax, int N);                 /* Utilities.c:716-736 */

/* fulltemplates (QuickCBC.c:3683-3777): hwave is nifo rows of N doubles,
 * halfcomplex.  freq has N/2 entries with freq[0]=1/Tobs. */
int qo_fulltemplates(int nifo, const int *labels, double gmst, double fref,
                     double **hwave, const double *freq, const double *params, int N);

/* log_likelihood_full (QuickCBC.c:4236-4273) with run-time band edges. */
double qo_log_likelihood_full(int nifo, const int *labels, double gmst, double fref,
                              double fmin, double fmax, double **D, const double *params,
                              const double *freq, double **SN, double *rho, int N, double Tobs,
                              int *status);

/* OpenMP batch over points, as QuickCBC.c:2063 parallelises over chains.
 * D and SN are nifo rows; params is B x 11; rho is B x nifo. */
void qo_loglike_batch(int nifo, const int *labels, double gmst, double fref, double fmin,
                      double fmax, double **D, double **SN, const double *freq, int N,
                      double Tobs, const double *params, int B, double *logL, double *rho,
                      int *status, int nthreads);

/* Flat-array conveniences for the ctypes wrapper. */
double qo_loglike_flat(int nifo, const int *labels, double gmst, double fref, double fmin,
                       double fmax, const double *D, const double *SN, int N, double Tobs,
                       const double *params, double *rho, int *status);
void qo_loglike_batch_flat(int nifo, const int *labels, double gmst, double fref, double fmin,
                           double fmax, const double *D, const double *SN, int N, double Tobs,
                           const double *params, int B, double *logL, double *rho, int *status,
                           int nthreads);
int qo_fulltemplates_flat(int nifo, const int *labels, double gmst, double fref,
                          double *hwave, int N, double Tobs, const double *params);
int qo_h22_flat(const double *freq, size_t n, double phi0, double fref, double m1_SI,
                double m2_SI, double chi1, double chi2, double distance, double *amp,
                double *phase);
int qo_setup_flat(double m1_msun, double m2_msun, double chi1, double chi2, double distance,
                  double phi0, double fref, double *out, int nout);
int qo_spline_eval_flat(int which, const double *x, int n, double *y);
int qo_waveform_params(const double *params, int N, double Tobs, double fref, double *amp, double *phase);
double qo_probe(int which, double m1, double m2, double chi1, double chi2, double Mf);
double qo_exp_cr(double x);
double qo_pow_cr(double x, double y);
double qo_log_cr(double x);
int qo_num_threads(void);

/* Heterodyned likelihood (QuickCBC.c:3780-3869, 5055-5444; Utilities.c:1862).
 * Grids and set-up arrays are flat and row-major, layouts in qcbc_oracle.c. */
void qo_legendre(int J, int U, double *P);
void qo_inverse(const double *M, double *IM, int d);
int qo_fullphaseamp_flat(int nifo, const int *labels, double gmst, double fref, double Tobs,
                         const double *freq, int n, const double *params, double *amp, double *phase);
int qo_heterodyne_flat(int nifo, const int *labels, double gmst, double fref, const double *D,
                       const double *SN, const double *params, int N, double Tobs, const int *fgrid,
                       int M, int J, int MN, int MM, double *amp, double *phase, double *IP,
                       double *SL, double *lc, double *ls, double *rc, double *rs, double *aa,
                       double *logLb);
void qo_loglike_het_batch_flat(int nifo, const int *labels, double gmst, double fref, double Tobs,
                               int M, int J, const double *freq, const double *amp,
                               const double *phase, const double *IP, const double *SL,
                               const double *lc, const double *ls, const double *rc,
                               const double *rs, const double *aa, double logLb,
                               const double *params, int B, double *logL, int *status,
                               int nthreads);

/* Fisher_Het (QuickCBC.c:5446-5617) and FisherEvec (QuickCBC.c:2449-2546):
 * fish row-major 11 x 11 per point; ej d, ev d x d (row i = eigenvector i). */
void qo_fisher_het_batch_flat(int nifo, const int *labels, double gmst, double fref, double Tobs,
                              int M, int J, const double *freq, const double *amp,
                              const double *phase, const double *IP, const double *SL,
                              const double *aa, const double *params, int B, double *fish,
                              int *status, int nthreads);
int qo_fisher_evec(const double *fish, double *ej, double *ev, int d);
void qo_fisher_evec_batch(const double *fish, int B, int d, double *ej, double *ev, int *sweeps);

/* QuickCBC's PT step (update(), QuickCBC.c:2238-2405; MCMC_all, :2029-2092)
 * with the counter-based random numbers of the GPU step. */
double qo_uni(uint64_t seed, uint64_t salt, uint64_t a, uint64_t k);
int qo_update_propose(const double *x, const double *ej, const double *ev, const double *hist, int n_hist,
                      double sT, const double *abc, uint64_t seed, uint64_t ctr, double *y, double *jac);
void qo_ring_all(const double *px, double *py, int d1, int d2, double gmst, double u_rot);
void qo_skymap_all(const double *px, double *py, double gmst, int ifo1, int ifo2);
double qo_skydensity_all(const double *px, double *py, double gmst, int ifo1, int ifo2);
void qo_update_hist_init(const double *x, int n_temps, int n_chains, int n_hist, uint64_t seed, double *hist);
int qo_update_prior(const double *y, const double *uconst, double *pMcM);
void qo_pt_update_run(int nifo, const int *labels, double gmst, double fref, double Tobs, int M, int J,
                      const double *freq, const double *amp, const double *phase, const double *IP,
                      const double *SL, const double *lc, const double *ls, const double *rc, const double *rs,
                      const double *aa, double logLb, const double *uconst, int n_hist, int hist_every, int n_temps,
                      int n_chains, double ladder, uint64_t seed, int nsteps, long long *step, double *x, double *L,
                      double *ej, double *ev, double *hist, long long *hcount, double *y, int *typ, double *jac,
                      long long *stats);

/* templates (QuickCBC.c:3527-3611) and log_likelihood_intrinsic (:4093-4145):
 * params 7 + 3 (nifo - 1) per point (reference-frame form). */
int qo_templates_flat(int nifo, double fref, double *hwave, int N, double Tobs, const double *params);
void qo_loglike_intrinsic_batch_flat(int nifo, double fref, double fmin, double fmax, const double *D,
                                     const double *SN, int N, double Tobs, const double *params, int B,
                                     double *logL, double *snrsq, int *status);

/* makespec (Utilities.c:1908-1959) */
int qo_makespec(int Nspline, int Nlines, const double *ffit, const double *Sspline, const double *linef,
                const double *lineh, const double *lineQ, const double *linew, const double *deltafmax, double *SN,
                double Tobs, int N);

int qo_amp_phase(double *Af, double *Pf, const double *freq, const double *params, double fref, int N);
int qo_fisher_full_flat(int nifo, double fref, double fmin, double fmax, const double *SN, const double *params, int d,
                        int N, double Tobs, double *fish);
int qo_collocation_probe(double m1, double m2, double chi1, double chi2, double *out);
int qo_read_framed(const char *path, int cap, double *time, double *x);
int qo_read_spec(const char *path, int n2, double *SN, double *SM);

#ifdef __cplusplus
}
#endif
#endif
