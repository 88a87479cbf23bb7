/*
 * libqcbc: MI355X (gfx950) implementation of QuickCBC's waveform + likelihood
 * path.  Plain C ABI: plain pointers and sizes, no C++ or device types.
 *
 * Part 1 re-exports QuickCBC's own entry points with identical names and
 * signatures, so a QuickCBC build links against this library instead of the
 * reference objects.  Part 2 is the batch API the drop-ins are built on: one
 * analysis (detector set, data, PSD) is made resident on the GPU once and
 * then evaluated for batches of parameter points.  Parts 3-5 add the PT
 * ensemble step, the multi-analysis persistent batch and the serving loop;
 * Part 6 the heterodyned likelihood.
 *
 * Every computation runs on the GPU; if no GPU (or no gfx950 code object) is
 * usable every entry point reports QCBC_ENODEV (the drop-ins print the error
 * and exit, the reference's error convention).
 */
#ifndef QCBC_H
#define QCBC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ *
 * ABI types, layout-identical to QuickCBC's                           *
 * ------------------------------------------------------------------ */
#ifndef QCBC_NO_REFERENCE_TYPES
/* QuickCBC.h:13-25 */
struct Net {
  int Nifo;
  double Tobs;
  double tmax;
  double tmin;
  double GMST;
  int *labels;
  double *tds;
  double **delays;
  int *lstart;
  int *lstop;
};
/* IMRPhenomD_internals.h:662-665 */
typedef struct tagRealVector {
  double *data;
  size_t length;
} RealVector;
/* IMRPhenomD_internals.h:649-654 */
typedef struct tagAmpPhaseFDWaveform {
  double *freq;
  double *amp;
  double *phase;
  size_t length;
} AmpPhaseFDWaveform;
/* QuickCBC.h:27-50: a heterodyne.  Arrays are the reference's
 * double_matrix / double_tensor shapes (Utilities.c:2981-3025: n+1 rows of
 * m+1, each malloc'd), so its freehet releases them. */
struct Het {
  int M;
  int MM;
  int MN;
  int ell;
  int J;
  int NR;
  double hh;
  double hr;
  double logLb;
  double ***IP;
  double ***SL;
  double **amp;
  double **phase;
  double **rc;
  double **rs;
  double **aa;
  double ***lc;
  double ***ls;
  double *pref;
  int *fgrid;
  RealVector *freq;
};
#endif

/* ------------------------------------------------------------------ *
 * Part 1: drop-in replacements of QuickCBC entry points               *
 * ------------------------------------------------------------------ */

/* Replaces log_likelihood_full, QuickCBC.c:4236-4273 (prototype QuickCBC.h:110).
 * logL = sum_d [(d|h) - (h|h)/2], rho[d] = (d|h)/sqrt((h|h)), band
 * [(int)(Tobs*fmin), min((int)(Tobs*fmax), N/2)).  fmin/fmax default to the
 * reference's compiled 16/1024 Hz (ConstCBC.h:33-34); see qcbc_set_band.
 * D and SN are cached on the GPU keyed by their row pointers; call
 * qcbc_invalidate() after changing their contents in place. */
double log_likelihood_full(struct Net *net, double **D, double *params, RealVector *freq,
                           double **SN, double *rho, int N, double Tobs);

/* Replaces fulltemplates, QuickCBC.c:3683-3777 (prototype QuickCBC.h:97).
 * hwave: net->Nifo rows of N doubles, half-complex. */
void fulltemplates(struct Net *net, double **hwave, RealVector *freq, double *params, int N);

/* Replaces IMRPhenomDGenerateh22FDAmpPhase, IMRPhenomD.c:186-248 (prototype
 * IMRPhenomD.h:20-30).  *h22 must be NULL; it is allocated here and freed by
 * the caller with DestroyAmpPhaseFDWaveform.  Returns 0 (PD_SUCCESS) or, in
 * non-exiting mode, the reference's error code. */
int IMRPhenomDGenerateh22FDAmpPhase(AmpPhaseFDWaveform **h22, RealVector *freq, double phi0,
                                    double fRef_in, double m1_SI, double m2_SI, double chi1,
                                    double chi2, double distance);

/* Replaces fourier_nwip, Utilities.c:716-736 (prototypes Utilities.h:38, QuickCBC.h:84). */
double fourier_nwip(double *a, double *b, double *Sn, int imin, int imax, int N);

/* Replace F_ant (Utilities.c:995-1000, prototype Utilities.h:46), Times
 * (Utilities.c:1078-1109, Utilities.h:50) and fbegin (Utilities.c:587-629,
 * Utilities.h:34). */
void F_ant(double psi, double alpha, double sindelta, double GMST, double *Fplus, double *Fcross,
           int id);
void Times(double alpha, double sindelta, double GMST, double *dtimes);
double fbegin(double *param);

/* Replaces templates, QuickCBC.c:3527-3611 (prototype QuickCBC.h:95): the
 * reference-frame templates of MCMC_intrinsic.  params: the seven
 * quasi-intrinsic parameters (ln Mc, ln Mt, chi1, chi2, phi0 = params[4]
 * itself, tc, ln DL) then, for detector j = 1, 2, its phase offset, time
 * offset and amplitude ratio relative to detector 0 at params[7 + 3 (j-1)].
 * hwave as fulltemplates. */
void templates(struct Net *net, double **hwave, RealVector *freq, double *params, int N);

/* Replaces log_likelihood_intrinsic, QuickCBC.c:4093-4145 (prototype
 * QuickCBC.h:109): sum over detectors of (d|h) - (h|h)/2 with templates();
 * with more than one detector 0 unless at least two have 2 ((d|h) - (h|h)/2)
 * above 25. */
double log_likelihood_intrinsic(struct Net *net, double **D, double *params, RealVector *freq, double **SN,
                                int N, double Tobs);

/* Replaces fullphaseamp, QuickCBC.c:3780-3869 (prototype QuickCBC.h:98): per
 * detector the projected amplitude (tapered by exp(10 (f - fs)/fs) below
 * fbegin) and phase 2 pi f (ts - td) + lambda - Psi on freq->data[0..N/2).
 * amp, phase: net->Nifo rows of at least N/2. */
void fullphaseamp(struct Net *net, double **amp, double **phase, RealVector *freq, double *params, int N);

/* Replaces heterodyne, QuickCBC.c:5112-5299 (prototype QuickCBC.h:106): the
 * set-up of the heterodyned likelihood at params on the grid het_space chose
 * (het->M, MN, MM, J, NR, fgrid, freq).  Allocates and fills het's IP, SL,
 * amp, phase, rc, rs, aa, lc, ls and pref in the reference's shapes, sets
 * logLb and prints it as the reference does.  The GPU copy stays cached for
 * log_likelihood_het. */
void heterodyne(struct Net *net, struct Het *het, double **D, double *params, RealVector *freq, double **SN,
                int N, double Tobs);

/* Replaces log_likelihood_het, QuickCBC.c:5303-5444 (prototype QuickCBC.h:107).
 * het may come from heterodyne() above or from the reference's own: it is
 * copied to the GPU on first use and cached by its array pointers and pref;
 * call qcbc_invalidate() after changing its arrays in place. */
double log_likelihood_het(struct Net *net, struct Het *het, double *params, double Tobs);

/* Replaces Fisher_Het, QuickCBC.c:5446-5617 (prototype QuickCBC.h:151): the
 * Fisher matrix of the heterodyned likelihood at params, central differences
 * of fullphaseamp with epsilon 1e-5 on het's grid.  fish: 11 rows of 11.
 * With |chi1| or |chi2| >= 0.9999, 1e4 times the identity.  het is found or
 * imported as log_likelihood_het finds it. */
void Fisher_Het(struct Net *net, struct Het *het, double **fish, double *params, double Tobs);

/* Replaces FisherEvec, QuickCBC.c:2449-2546 (prototype QuickCBC.h:164): the
 * eigen-decomposition of the symmetric d x d matrix fish (d <= 11) ordered by
 * |eigenvalue| ascending, eval[i] = 1/sqrt(max(lambda_i, 10)) (the 1-sigma
 * jump along eigenvector i) and evec[i][j] its component j.  GSL's
 * eigen_symmv is replaced by the cyclic Jacobi method: the same eigenpairs
 * to rounding, each eigenvector signed so that its largest-magnitude
 * component is positive.  A diagonal entry below 1e-16 gives 1e4 and the
 * unit vectors, as in the reference. */
void FisherEvec(double **fish, double *eval, double **evec, int d);

/* The heterodyne's grid choice.  SN/SM: net->Nifo rows of N/2 (the PSD);
 * freq the analysis grid (as log_likelihood_full's).  fmin/fmax are the
 * drop-in band (qcbc_set_band).  The reference's prints and its diagnostic
 * files (ratio.dat, fspace.dat, df.dat, df_flat.dat) are not written. */
/* Replaces limits, QuickCBC.c:4275-4385 (prototype QuickCBC.h:101): stst[0],
 * stst[1] the band over which the network SNR^2 of params passes from 0.001
 * to within 0.001 of its total; ratio[j] = 1 / (SNR^2 density at bin j over
 * the band average) for j = 1 .. M-1 (M twice the ringdown frequency, at
 * least 64 Hz, in bins), 1 beyond. */
void limits(struct Net *net, double *params, RealVector *freq, double **SN, int *stst, double *ratio, int N,
            double Tobs);
/* Replaces Fisher_All, QuickCBC.c:2695-2817 (prototype QuickCBC.h:99): the
 * Fisher matrix of the full likelihood over the band, central differences of
 * fullphaseamp with epsilon 1e-5.  fish: 11 rows of 11. */
void Fisher_All(struct Net *net, double **fish, double *params, RealVector *freq, double **SN, int N, double Tobs);
/* Replaces Fisher_Full, QuickCBC.c:2551-2688 (prototype QuickCBC.h:150):
 * MCMC_intrinsic's Fisher matrix of the first d (<= 7) reference-frame
 * parameters over the band, central differences of AmpPhase (the bare h22,
 * :3395-3444) with epsilon 1e-5, detector l weighted by 4 (l = 0) or 4 r_l^2
 * (r_l = params[7 + 3 (l-1) + 2]); diagonal entries at least 1; 1e4 I when a
 * spin is at 0.9999 or more.  fish: d rows of d. */
void Fisher_Full(struct Net *net, double **fish, int d, double *params, RealVector *freq, double **SN, int N,
                 double Tobs);
/* Replaces diff, QuickCBC.c:4913-4964 (prototype QuickCBC.h:103): sum over
 * detectors and bins [max(fmin, fbegin) Tobs, 2 f_ringdown Tobs) of
 * 4 |h(params) - h(dparams)|^2 / S. */
double diff(struct Net *net, double *params, double *dparams, RealVector *freq, double **SM, int N, double Tobs);
/* Replaces het_diff, QuickCBC.c:4966-5053 (prototype QuickCBC.h:104): diff's
 * sum on the heterodyne het (found or imported as log_likelihood_het finds it). */
double het_diff(struct Net *net, struct Het *het, double *params, double *dparams);
/* Replaces efix, QuickCBC.c:4387-4567 (prototype QuickCBC.h:100): each jump
 * eval[i] along evec[i] (11 rows of 11) rescaled so that the displaced
 * point's distance (diff when hflag is 0, het_diff on het otherwise) comes
 * to zs; a direction that never comes within zs gets 0.  The prior box
 * min/max and etamin 0.05 (ConstCBC.h:57) bound the displaced points. */
void efix(struct Net *net, struct Het *het, double *params, double *min, double *max, RealVector *freq, double **SM,
          int N, double Tobs, double *eval, double **evec, double zs, int hflag);
/* Replaces het_space, QuickCBC.c:4570-4910 (prototype QuickCBC.h:105): the
 * heterodyne grid for params (limits, Fisher_All, FisherEvec, efix to 30,
 * the linear-span scan with tolerance 0.003, J = 2): sets het->MN, MM, J,
 * NR, M and allocates het->fgrid (M ints) and het->freq (CreateRealVector(M)),
 * as the reference does; heterodyne() then sets it up. */
void het_space(struct Net *net, struct Het *het, double *params, double *min, double *max, RealVector *freq,
               double **SN, int N, double Tobs);

/* The search likelihoods of CBC_start.  params: the reference-frame
 * parameters of templates() (7 + 3 (Nifo - 1)), moved in place onto the
 * maximum as the reference moves them.  The band is the drop-in band
 * (qcbc_set_band).  pflag = 1's diagnostic prints are not made. */
/* Replaces log_likelihood_APmax, QuickCBC.c:5672-5762 (prototype QuickCBC.h:112). */
double log_likelihood_APmax(struct Net *net, double **D, double *params, RealVector *freq, double **SN, int N,
                            double Tobs);
/* Replaces log_likelihood_max, QuickCBC.c:5765-6231 (prototype QuickCBC.h:113):
 * tc in [tmin, tmax]; uses net->tmax (PDwave's mask) and net->delays. */
double log_likelihood_max(struct Net *net, double **D, double *params, RealVector *freq, double **SN, int N,
                          double Tobs, double tmin, double tmax, int pflag);
/* Replaces log_likelihood_penalized, QuickCBC.c:6233-6408 (prototype QuickCBC.h:114):
 * one detector's D, twave (half-complex N) and SN (N/2). */
double log_likelihood_penalized(double *params, double *D, double *twave, double *SN, double *Ax, double *Px, int N,
                                double Tobs, int pflag);
/* Replaces PDwave, QuickCBC.c:3448-3525 (prototype QuickCBC.h:160). */
void PDwave(double *wavef, RealVector *freq, double tcmin, double *params, int N);
/* Replaces pbt_shift, Utilities.c:687-713 (prototype Utilities.h:37): N a power of 2. */
void pbt_shift(double *corr, double *corrf, double *data1, double *data2, double *Sn, int imin, int imax, int N);
/* Replaces tukey, Utilities.c:635-654 (prototype Utilities.h:35): the window
 * applied in place to N samples. */
void tukey(double *data, double alpha, int N);
/* Replaces tukey_scale, Utilities.c:656-683 (prototype Utilities.h:36): the
 * window's mean (s1) and root mean square (s2). */
void tukey_scale(double *s1, double *s2, double alpha, int N);
/* Replaces whiten, Utilities.c:294-309 (prototype Utilities.h:24): bins 1 ..
 * N/2 - 1 of a half-complex row divided by sqrt(Sn), the DC and Nyquist
 * bins set to 0. */
void whiten(double *data, double *Sn, int N);

/* Allocators of the ABI types, IMRPhenomD_internals.c:141-181. */
AmpPhaseFDWaveform *CreateAmpPhaseFDWaveform(size_t length);
void DestroyAmpPhaseFDWaveform(AmpPhaseFDWaveform *wf);
RealVector *CreateRealVector(size_t length);
void DestroyRealVector(RealVector *v);

/* Drop-in configuration (no counterpart in the reference, whose values are
 * compile-time constants): band edges used by log_likelihood_full, the
 * reference frequency (ConstCBC.h:31), and whether domain errors exit(-1) as
 * the reference's ERROR() does (default 1) or return silently with logL = 0. */
void qcbc_set_band(double fmin, double fmax);
void qcbc_set_fref(double fref);
void qcbc_set_exit_on_error(int on);
/* Forget cached GPU copies of D/SN (their contents changed in place). */
void qcbc_invalidate(void);

/* ------------------------------------------------------------------ *
 * Part 2: resident-analysis batch API                                 *
 * ------------------------------------------------------------------ */
enum {
  QCBC_OK = 0,
  QCBC_EINVAL = 1,
  QCBC_ENOMEM = 12,   /* as PD_ENOMEM */
  QCBC_EFAULT = 14,   /* as PD_EFAULT */
  QCBC_EDOM = 33,     /* as PD_EDOM: a parameter point outside the model */
  QCBC_EHIP = 100,    /* a HIP runtime call failed */
  QCBC_ENODEV = 101,  /* no usable gfx950 device */
  QCBC_ENOTBOUND = 102,
  QCBC_EBUSY = 103,   /* admission refused: memory budget or slots */
  QCBC_ECANCELED = 104,
  QCBC_ETIMEDOUT = 105,
  QCBC_ERANGE = 34,   /* as ERANGE: a batch or file larger than a capacity */
  QCBC_EIO = 5        /* as EIO: a file that cannot be read or written */
};

enum { QCBC_FP64 = 0, QCBC_MIXED = 1 };

typedef struct {
  int n_det;        /* 1..3 */
  int labels[3];    /* 0 = H, 1 = L, 2 = V (F_ant / Times labels) */
  double gmst;
  int N;            /* samples per segment; N/2 frequency bins */
  double Tobs;      /* segment length (s) */
  double fmin, fmax;/* likelihood band (Hz) */
  double fref;      /* IMRPhenomD reference frequency (Hz) */
} qcbc_desc;

typedef struct qcbc_ctx qcbc_ctx;

int qcbc_ctx_create(const qcbc_desc *desc, qcbc_ctx **out);
int qcbc_ctx_destroy(qcbc_ctx *ctx);
/* Device bytes a context holds for this shape (data, PSD, tables), excluding
 * per-batch scratch. */
size_t qcbc_ctx_footprint(const qcbc_desc *desc);

/* Bind data and PSD.  D: n_det rows of N doubles (half-complex, QuickCBC's
 * D layout); SN: n_det rows of N/2 doubles (one-sided PSD).  Row-major flat
 * arrays; the _rows variant takes QuickCBC's double** row pointers. */
int qcbc_bind_data(qcbc_ctx *ctx, const double *D, const double *SN);
int qcbc_bind_rows(qcbc_ctx *ctx, double *const *D, double *const *SN);

/* logL and rho for B points (params: B x 11, host memory).  status[b] is 0
 * or QCBC_EDOM for a point outside the model (its logL is then 0).
 * precision: QCBC_FP64 (parity path) or QCBC_MIXED. */
int qcbc_loglike(qcbc_ctx *ctx, const double *params, int B, int precision, double *logL,
                 double *rho, int *status);
/* Same with device pointers, asynchronous on the context's stream. */
int qcbc_loglike_device(qcbc_ctx *ctx, const double *d_params, int B, int precision,
                        double *d_logL, double *d_rho, int *d_status);
/* No counterpart in the reference: qcbc_loglike_device for callers that
 * cannot allocate device memory (code under HIP stream capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The same kernels and
 * results, but the point records and partial sums live in a fixed device
 * array of the library (one per GPU, about 17 MB, held until the process
 * ends) instead of the context's scratch, so the call allocates nothing.
 * At most QCBC_NOALLOC_MAX_POINTS points per call (QCBC_ERANGE beyond).
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (from different contexts or threads, before the previous call's work
 * on its stream has finished) race and must be serialised by the caller. */
#define QCBC_NOALLOC_MAX_POINTS 4096
int qcbc_loglike_device_noalloc(qcbc_ctx *ctx, const double *d_params, int B, int precision,
                                double *d_logL, double *d_rho, int *d_status);
/* The plain FP64 GPU path: per point, the reference's three loops as three
 * launches (waveform -> amp/phase in HBM, projection -> templates in HBM,
 * inner products), one point after another.  Host pointers. */
int qcbc_loglike_plain(qcbc_ctx *ctx, const double *params, int B, double *logL, double *rho,
                       int *status);

/* IMRPhenomD h22 amplitude and phase on the context's N/2 grid (freq[0] =
 * 1/Tobs, freq[i] = i/Tobs) for B points: amp, phase are B x N/2. */
int qcbc_waveform(qcbc_ctx *ctx, const double *params, int B, double *amp, double *phase,
                  int *status);
/* Projected half-complex templates (fulltemplates) for B points: hwave is
 * B x n_det x N. */
int qcbc_templates(qcbc_ctx *ctx, const double *params, int B, double *hwave, int *status);
/* The per-point set-up records (IMRPhenomD coefficients, projection
 * constants, contributing bin range) the kernels use, copied to host memory:
 * B records of qcbc_point_record_size() bytes each, for inspection and
 * set-up parity tests. */
size_t qcbc_point_record_size(void);
int qcbc_point_records(qcbc_ctx *ctx, const double *params, int B, void *records);

/* Kernel launch shape of the fused likelihood for this context:
 * chains per block (1..16), bins per LDS tile (128, 256 or 512), bin splits (0 = choose). */
int qcbc_ctx_set_shape(qcbc_ctx *ctx, int precision, int chains_per_block, int tile, int splits);
int qcbc_ctx_get_shape(qcbc_ctx *ctx, int precision, int *chains_per_block, int *tile,
                       int *splits);
/* Time candidate shapes on synthetic points and keep the fastest. */
int qcbc_autotune(qcbc_ctx *ctx, int precision, int B);
/* The fused likelihood kernel alone, for measurement (no counterpart in the
 * reference): set-up of B points (device memory) once, then reps launches
 * back to back on the context's stream between two HIP events;
 * *ms_per_launch = their average duration, launch gaps included. */
int qcbc_time_loglike_kernel(qcbc_ctx *ctx, const double *d_params, int B, int precision, int reps,
                             double *ms_per_launch);
/* Cumulative device time (ms) and launch count of the fused likelihood
 * kernel, measured with HIP events on the context's stream. */
int qcbc_ctx_timing(qcbc_ctx *ctx, int enable, double *kernel_ms, long long *launches);
int qcbc_ctx_sync(qcbc_ctx *ctx);

/* log_likelihood_intrinsic (QuickCBC.c:4093-4145) for B points of
 * templates' reference-frame parameters (B x (7 + 3 (n_det - 1)), host
 * memory), fused FP64 path on the bound data: logL[b], and snrsq (B x n_det,
 * may be NULL) the per-detector 2 ((d|h) - (h|h)/2). */
int qcbc_loglike_intrinsic(qcbc_ctx *ctx, const double *params, int B, double *logL, double *snrsq,
                           int *status);
/* templates() (QuickCBC.c:3527-3611) for B points: hwave B x n_det x N. */
int qcbc_templates_intrinsic(qcbc_ctx *ctx, const double *params, int B, double *hwave, int *status);

/* Data conditioning (QuickCBC.c:268-318) on the GPU, then binding: each
 * detector's N time samples x (n_det x N, host memory) windowed by tukey()
 * (Utilities.c:635-654, alpha = 2 t_rise / Tobs; the reference uses t_rise
 * = 0.4 s), Fourier transformed (N a power of 2, as its radix-2 FFT), in the
 * half-complex layout and scaled by sqrt(Tobs)/N, bound with the PSD SN
 * (n_det x N/2) as qcbc_bind_data.  D (n_det x N, may be NULL) receives the
 * conditioned data.  Allocates its device buffers and FFT plan per call (a
 * once-per-analysis step). */
int qcbc_bind_time_series(qcbc_ctx *ctx, const double *x, const double *SN, double t_rise, double *D);

/* The conditioning of QuickCBC's main (QuickCBC.c:250-318) and of CBC_start
 * after glitch cleaning (:807-823) on the GPU, then binding.  x (n_det x N
 * time samples, host) windowed by tukey (alpha = 2 t_rise / Tobs) and
 * transformed; D (n_det x N, may be NULL) receives the half-complex data
 * times sqrt(Tobs)/N, which is bound with the PSD SN (n_det x N/2).
 * fmin_floor > 0 binds the PSD as 1 below bin (int)(fmin_floor Tobs), as
 * CBC_start does (SN itself is not written).  main's outputs from the same
 * transform (each may be NULL, n_det rows): DW (N) the whitened series of
 * dataw.dat (whiten, the half-complex inverse, / (sqrt(N/2)/sqrt(Tobs))),
 * pspec (N/2) the periodogram of pspec_*.dat, (2 Tobs/N^2)(Re^2 + Im^2),
 * and wcheck (N/2) = pspec / SN of wcheck_*.dat.  The bin-0 entry of pspec
 * takes the DC term only (the reference adds the square of its row's spare
 * element).  qcbc_bind_time_series(ctx, x, SN, t_rise, D) is
 * qcbc_condition(ctx, x, SN, t_rise, 0, D, NULL, NULL, NULL). */
int qcbc_condition(qcbc_ctx *ctx, const double *x, const double *SN, double t_rise, double fmin_floor, double *D,
                   double *DW, double *pspec, double *wcheck);

/* makespec (Utilities.c:1908-1959) on the GPU for the analysis's N and Tobs:
 * SN (N/2) = exp(natural cubic spline of ln S through (ffit, Sspline)) times
 * (1 + the Lorentzian lines at linef with heights lineh, quality lineQ,
 * half-widths linew and taper scales deltafmax), SN[0] and SN[N/2-1] the end
 * knots' values.  The knots must rise and cover bins 1 .. N/2-2 (the
 * reference's GSL spline aborts outside them); QCBC_EINVAL otherwise.
 * Rebind with qcbc_bind_data to use the draw (noise marginalisation,
 * QuickCBC.c:1994-2010).  Allocates its device buffers per call. */
int qcbc_makespec(qcbc_ctx *ctx, int Nspline, int Nlines, const double *ffit, const double *Sspline,
                  const double *linef, const double *lineh, const double *lineQ, const double *linew,
                  const double *deltafmax, double *SN);

const char *qcbc_strerror(int code);
int qcbc_device_info(char *name, size_t len, int *cu_count, double *mem_gb);

/* ------------------------------------------------------------------ *
 * Part 3: parallel-tempered ensemble step (resident state)            *
 * ------------------------------------------------------------------ */
typedef struct qcbc_pt qcbc_pt;
typedef struct qcbc_het qcbc_het; /* a heterodyne, Part 6 */
/* n_temps x n_chains points; temperatures ladder^t (t = 0..n_temps-1).
 * init: (n_temps*n_chains) x 11.  sigma: 11 per-parameter jump scales at
 * temperature 1 (scaled by sqrt(T)). */
int qcbc_pt_create(qcbc_ctx *ctx, int n_temps, int n_chains, double ladder, const double *init,
                   const double *sigma, uint64_t seed, int precision, qcbc_pt **out);
/* As qcbc_pt_create, with the chain columns split into `groups` (1..4, at
 * most n_chains; 0 = 3 when n_chains >= 3 and n_temps*n_chains >= 1536, else
 * 1; on the heterodyned likelihood, qcbc_pt_create_het and
 * qcbc_pt_create_update, 0 = 1) groups that step on streams of their own, so that one group's
 * latency-bound kernels and likelihood tail overlap another group's
 * likelihood kernel.  The chains do not depend on the number of groups. */
int qcbc_pt_create_groups(qcbc_ctx *ctx, int n_temps, int n_chains, double ladder, const double *init,
                          const double *sigma, uint64_t seed, int precision, int groups,
                          qcbc_pt **out);
int qcbc_pt_groups(qcbc_pt *pt, int *groups);
/* The same ensemble on the heterodyned likelihood of het (Part 6), which must
 * belong to ctx and outlive the ensemble: the production sampler's
 * likelihood (QuickCBC.c:2372) in the resident step.  FP64. */
int qcbc_pt_create_het(qcbc_ctx *ctx, qcbc_het *het, int n_temps, int n_chains, double ladder,
                       const double *init, const double *sigma, uint64_t seed, int groups,
                       qcbc_pt **out);
/* QuickCBC's own PT step on the heterodyned likelihood: update()'s proposal
 * mix (QuickCBC.c:2238-2405) and MCMC_all's step (:2029-2092), resident on
 * the GPU.  Each chain column is one of the reference's temperature ladders:
 * per step and column, with probability 0.2 one swap of a random adjacent
 * pair of temperatures, else update() at every temperature (Fisher
 * eigen-jumps above alpha > a, differential evolution from the
 * temperature's history above b, jiggle above c, below c the sky ring:
 * Ring_all, skymap_all and skydensity_all's log density), accepted by
 * H = dlogL/T + the mass Jacobian of a prior flat in m1, m2 + 3 dlnDL (+ the
 * sky-ring density), inside the prior box with eta in [etamin, 1/4].  The
 * sky ring needs two detectors (c = 0 turns it off: the reference's
 * alternative setting, :2256-2260).  The Fisher jumps (Fisher_Het +
 * FisherEvec, then efix when desc->efix_zs > 0) are computed at creation and
 * by qcbc_pt_refresh_fisher (the reference refreshes them every 50000
 * iterations, :2019-2027). */
typedef struct {
  double min[11], max[11]; /* prior box (update()'s min, max) */
  double a, b, c;          /* proposal thresholds: 0.5, 0.3, 0.2 */
  double etamin;           /* lower limit of eta (0.05, ConstCBC.h:57) */
  int n_hist;              /* history rows per temperature (NH = 1000, ConstCBC.h:72) */
  int hist_every;          /* update steps between history entries (20, QuickCBC.c:2071) */
  double efix_zs;          /* > 0: after each Fisher refresh, efix with the heterodyne
                              distance to this target (the reference's 1.0,
                              QuickCBC.c:1927, :2025); 0: the jumps as FisherEvec gives them */
} qcbc_update_desc;
int qcbc_pt_create_update(qcbc_ctx *ctx, qcbc_het *het, int n_temps, int n_chains, double ladder,
                          const double *init, const qcbc_update_desc *desc, uint64_t seed, int groups,
                          qcbc_pt **out);
/* Fisher jumps at every chain's current state, and the pending proposals
 * remade with them. */
int qcbc_pt_refresh_fisher(qcbc_pt *pt);
/* State of a QuickCBC-step ensemble, in point order (t*n_chains + c): jump
 * sizes (P x 11), directions (P x 11 x 11), the pending proposals (P x 11),
 * their types (1 Fisher, 2 DE, 3 jiggle, 4 sky ring) and sky-ring log
 * densities (P, 0 for the others); stats n_chains x 11 per column: update
 * steps, swap steps, swaps accepted, then proposed and accepted for each
 * type.  Any pointer may be NULL. */
int qcbc_pt_get_update(qcbc_pt *pt, double *ej, double *ev, double *proposals, int *types, double *jac,
                       long long *stats);
/* The histories (n_chains x n_temps x n_hist x 11) and their fill counts per
 * column (row count % n_hist is written next). */
int qcbc_pt_get_history(qcbc_pt *pt, double *hist, long long *counts);
/* nsteps PT steps: propose, evaluate, Metropolis accept, adjacent-temperature
 * swaps, all on the GPU.  use_graph: replay one captured step as a hipGraph. */
int qcbc_pt_run(qcbc_pt *pt, int nsteps, int use_graph);
int qcbc_pt_get(qcbc_pt *pt, double *params, double *logL, long long *accepted,
                long long *swaps);
/* Per-step wall latencies (ms) of the last qcbc_pt_run with timing on. */
int qcbc_pt_latencies(qcbc_pt *pt, double *ms, int max, int *n);
int qcbc_pt_set_step_timing(qcbc_pt *pt, int on);
int qcbc_pt_destroy(qcbc_pt *pt);

/* ------------------------------------------------------------------ *
 * Part 4: one persistent launch over many analyses                    *
 * ------------------------------------------------------------------ */
/* Evaluate n batches (ctxs[i], d_params[i], B[i]) in one persistent kernel
 * draining an (analysis, chain group, bin range) work queue.  Device
 * pointers; synchronous.  One-shot form of the plan below. */
int qcbc_multi_loglike(qcbc_ctx *const *ctxs, const double *const *d_params, const int *B,
                       int n, int precision, double *const *d_logL, double *const *d_rho,
                       int *const *d_status);

/* A reusable plan for a fixed set of analyses and batch sizes: the work
 * queue is built once; each run launches set-up per analysis, the persistent
 * queue kernel and finalize per analysis on the plan's stream, synchronously.
 * use_graph: capture that sequence once as a hipGraph and replay it (captured
 * again if the buffer pointers change).  The contexts must not be used by
 * other calls while a plan built on them runs. */
typedef struct qcbc_multi qcbc_multi;
int qcbc_multi_create(qcbc_ctx *const *ctxs, const int *B, int n, int precision,
                      qcbc_multi **out);
/* points per block (1..8), bins per LDS tile (128, 256 or 512), bins per split */
int qcbc_multi_set_shape(qcbc_multi *m, int chains_per_block, int tile, int split_bins);
int qcbc_multi_get_shape(qcbc_multi *m, int *chains_per_block, int *tile, int *split_bins,
                         int *n_items);
int qcbc_multi_run(qcbc_multi *m, const double *const *d_params, double *const *d_logL,
                   double *const *d_rho, int *const *d_status, int use_graph);
/* Time candidate shapes (reps runs each) and keep the fastest. */
int qcbc_multi_autotune(qcbc_multi *m, const double *const *d_params, double *const *d_logL,
                        double *const *d_rho, int reps, int *best_cpb, int *best_tile,
                        int *best_split, double *best_ms);
/* Host wall time (ms) of the last run, launch to completion. */
int qcbc_multi_last_ms(qcbc_multi *m, double *ms);
int qcbc_multi_destroy(qcbc_multi *m);

/* Device memory helpers for callers without their own allocator. */
int qcbc_dev_alloc(size_t bytes, void **ptr);
int qcbc_dev_free(void *ptr);
int qcbc_memcpy_h2d(void *dst, const void *src, size_t bytes);
int qcbc_memcpy_d2h(void *dst, const void *src, size_t bytes);
/* Device-to-device copy by a plain kernel (bytes a multiple of 16): the
 * reference of known size for hardware byte counters. */
int qcbc_dev_copy(void *dst, const void *src, size_t bytes);

/* ------------------------------------------------------------------ *
 * Part 5: serving loop                                                *
 * ------------------------------------------------------------------ */
typedef struct qcbc_server qcbc_server;
/* budget_bytes: device memory the resident analyses may hold. */
int qcbc_server_create(size_t budget_bytes, int precision, qcbc_server **out);
/* Admit an analysis (copies D/SN to the GPU).  QCBC_EBUSY if it does not fit
 * the remaining budget. */
int qcbc_server_admit(qcbc_server *s, const qcbc_desc *desc, const double *D, const double *SN,
                      int *analysis_id);
int qcbc_server_evict(qcbc_server *s, int analysis_id);
/* Submit B points (host memory, copied) for an analysis; returns a request id. */
int qcbc_server_submit(qcbc_server *s, int analysis_id, const double *params, int B,
                       int64_t *request_id);
/* Wait for a request; logL, rho (B x n_det) and status are written to host
 * memory; latency_ms is submit-to-completion.  timeout_ms < 0 waits forever;
 * QCBC_ETIMEDOUT leaves the request pending (wait again), QCBC_ECANCELED
 * reports a request whose analysis was evicted. */
int qcbc_server_wait(qcbc_server *s, int64_t request_id, double *logL, double *rho, int *status,
                     double *latency_ms, int timeout_ms);
int qcbc_server_stats(qcbc_server *s, size_t *bytes_in_use, int *n_analyses, long long *evals,
                      long long *batches);
/* Admit an analysis with a resident parallel-tempered ensemble: n_temps x
 * n_chains points (init (n_temps*n_chains) x 11, point t*n_chains + c,
 * temperatures ladder^t, jiggle scales sigma (11) at temperature 1, random
 * numbers from seed and the ensemble's step counter), charged to the budget
 * with the analysis.  Its states, logL, pending proposals and counters stay
 * on the GPU.  At most 65536 points. */
int qcbc_server_admit_pt(qcbc_server *s, const qcbc_desc *desc, const double *D, const double *SN,
                         int n_temps, int n_chains, double ladder, const double *init,
                         const double *sigma, uint64_t seed, int *analysis_id);
/* Advance an admitted ensemble by nsteps PT steps (propose, evaluate,
 * Metropolis accept, adjacent-temperature swaps; the steps of one ensemble
 * run in order, one per batch, beside other analyses' steps and likelihood
 * requests).  Wait with qcbc_server_wait (logL, rho, status NULL; latency is
 * submit to the last step's completion).  An ensemble's chain does not
 * depend on what else the server runs. */
int qcbc_server_submit_steps(qcbc_server *s, int analysis_id, int nsteps, int64_t *request_id);
/* An ensemble's state between steps: params ((n_temps*n_chains) x 11), logL,
 * accepted moves per point, swaps per chain column, steps run (each may be
 * NULL). */
int qcbc_server_get_pt(qcbc_server *s, int analysis_id, double *params, double *logL, long long *accepted,
                       long long *swaps, long long *steps);
int qcbc_server_destroy(qcbc_server *s);

/* ------------------------------------------------------------------ *
 * Part 6: heterodyned likelihood (QuickCBC.c:5112-5444)               *
 * ------------------------------------------------------------------ */
/* A heterodyne lives on the GPU and belongs to the analysis it was made
 * for (destroy it before the analysis).  Its grid is M = NR J + 1 bin
 * indices fgrid[] (strictly increasing, inside (0, N/2)), frequencies
 * fgrid / Tobs, in NR regions of J steps; QuickCBC's het_space chooses it.
 * Set-up arrays, row-major as flat copies of struct Het:
 *   amp, phase   n_det x M
 *   IP           NR x (J+1) x (J+1)
 *   SL, lc, ls   NR x n_det x (J+1)
 *   rc, rs, aa   n_det x NR */
/* heterodyne() at params (11) on a bound analysis; logLb over bins [MN, MM).
 * J <= 8.  QCBC_EDOM if the reference point is outside the model. */
int qcbc_het_create(qcbc_ctx *ctx, const double *params, const int *fgrid, int M, int J, int MN,
                    int MM, qcbc_het **out);
/* A heterodyne from set-up arrays made elsewhere (the reference's struct
 * Het, flattened); freq has M entries.  The analysis need not be bound. */
int qcbc_het_import(qcbc_ctx *ctx, int M, int J, const double *freq, const double *amp,
                    const double *phase, const double *IP, const double *SL, const double *lc,
                    const double *ls, const double *rc, const double *rs, const double *aa,
                    double logLb, qcbc_het **out);
int qcbc_het_info(qcbc_het *het, int *M, int *J, int *NR, int *n_det, double *logLb);
/* Copies of the set-up arrays (any pointer may be NULL). */
int qcbc_het_export(qcbc_het *het, double *freq, double *amp, double *phase, double *IP, double *SL,
                    double *lc, double *ls, double *rc, double *rs, double *aa);
/* The evaluation's per-sample weights (3 x n_det x M: wH, wC, wS): the
 * linear form log_likelihood_het applies to the samples' cc, hc, hs. */
int qcbc_het_weights(qcbc_het *het, double *w);
/* log_likelihood_het for B points (B x 11, host memory).  status[b] is 0 or
 * QCBC_EDOM (then logL[b] = 0). */
int qcbc_het_loglike(qcbc_het *het, const double *params, int B, double *logL, int *status);
/* The same on device memory, asynchronous on the heterodyne's stream; wait
 * with qcbc_het_sync.  d_status may be NULL. */
int qcbc_het_loglike_device(qcbc_het *het, const double *d_params, int B, double *d_logL,
                            int *d_status);
int qcbc_het_sync(qcbc_het *het);
/* HIP-event timing of the evaluation kernel, as qcbc_ctx_timing. */
int qcbc_het_timing(qcbc_het *het, int enable, double *kernel_ms, long long *launches);
/* The evaluation kernel alone: B points (device memory) set up once, then
 * reps launches back to back between two HIP events; the mean per launch. */
int qcbc_het_time_kernel(qcbc_het *het, const double *d_params, int B, int reps, double *ms_per_launch);
/* Fisher_Het (QuickCBC.c:5446-5617) at B points (B x 11, host memory) on
 * the heterodyne: fish B x 11 x 11 row-major.  When ej and ev are both given,
 * also FisherEvec (QuickCBC.c:2449-2546) of each matrix on the GPU: ej B x 11
 * jump sizes, ev B x 11 x 11 (row i = eigenvector i, as qcbc_fisher_evec).
 * status[b] (may be NULL) is 0 or QCBC_EDOM (a perturbed point outside the
 * model; the matrix is then 1e4 I). */
int qcbc_het_fisher(qcbc_het *het, const double *params, int B, double *fish, double *ej, double *ev,
                    int *status);
/* The same on device memory, asynchronous on the heterodyne's stream. */
int qcbc_het_fisher_device(qcbc_het *het, const double *d_params, int B, double *d_fish, double *d_ej,
                           double *d_ev, int *d_status);
/* Kernel timing of the Fisher matrix for B points (device memory): mean ms
 * per launch over reps back-to-back launches of each stage, ms[0] the set-up
 * of the 23 B perturbed points, ms[1] the Fisher kernel, ms[2] the eigen
 * kernel. */
int qcbc_het_fisher_time(qcbc_het *het, const double *d_params, int B, int reps, double *ms);
/* FisherEvec (QuickCBC.c:2449-2546) of B symmetric d x d matrices (d <= 11,
 * host memory, row-major) on the analysis's GPU: ej B x d, ev B x d x d
 * (row i = eigenvector i), ordered by |eigenvalue| ascending; cyclic Jacobi
 * in place of GSL's eigen_symmv, each eigenvector signed so that its
 * largest-magnitude component is positive.  Allocates its device buffers per
 * call (qcbc_het_fisher reuses the heterodyne's). */
int qcbc_fisher_evec(qcbc_ctx *ctx, const double *fish, int B, int d, double *ej, double *ev);
int qcbc_het_destroy(qcbc_het *het);

/* ------------------------------------------------------------------ *
 * Part 7: the heterodyne's grid choice (QuickCBC.c:4275-5053, 2695-2817)
 * ------------------------------------------------------------------ */
/* Waveforms, Fisher sums and distances run on the GPU; limits' running sum,
 * efix's search over the jump scale (every direction of every point in
 * lockstep, one batched evaluation per step) and het_space's span scan run
 * on the host.  These allocate their device buffers per call (set-up steps,
 * run once per heterodyne).  The analysis must be bound (its PSD is used).
 * pmin, pmax: the prior box (11 each); etamin: 0.05 in the reference. */
/* limits at params: band[2] = {MN, MM}; ratio (N/2 entries, may be NULL). */
int qcbc_limits(qcbc_ctx *ctx, const double *params, int *band, double *ratio);
/* Fisher_All at B points (B x 11): fish B x 11 x 11, status B (may be NULL;
 * QCBC_EDOM for a perturbed point outside the model, the matrix then 1e4 I). */
int qcbc_fisher_all(qcbc_ctx *ctx, const double *params, int B, double *fish, int *status);
/* Fisher_Full (QuickCBC.c:2551-2688) of B reference-frame points (B x (7 + 3
 * (n_det - 1)), host memory) on the bound PSD over the analysis's band:
 * fish B x d x d (d <= 7), status B (may be NULL; QCBC_EDOM for a point
 * outside the model, whose matrix is 1e4 I). */
int qcbc_fisher_full(qcbc_ctx *ctx, const double *params, int B, int d, double *fish, int *status);

/* diff between start point params[of[k]] (of B) and px[k], k < K: z (K). */
int qcbc_diff(qcbc_ctx *ctx, const double *params, int B, const double *px, const int *of, int K, double *z);
/* efix on the full data for B points: ej (B x 11) rescaled in place along
 * ev (B x 11 x 11, row i = direction i) to distance zs; trials (B x 11,
 * may be NULL) the distances evaluated per direction. */
int qcbc_efix(qcbc_ctx *ctx, const double *params, int B, double *ej, const double *ev, const double *pmin,
              const double *pmax, double etamin, double zs, int *trials);
/* het_space at params: MN, MM, M = NR J + 1 (J = 2) grid bins into fgrid (room
 * for cap; QCBC_ERANGE if more), efix's jump sizes and directions (eval 11,
 * evec 11 x 11) and the eleven displaced points the grid was fitted to
 * (dparams 11 x 11); the last three may be NULL.  qcbc_het_create then
 * sets the heterodyne up. */
int qcbc_het_space(qcbc_ctx *ctx, const double *params, const double *pmin, const double *pmax, double etamin,
                   int *MN, int *MM, int *M, int *NR, int *fgrid, int cap, double *eval, double *evec,
                   double *dparams);
/* ------------------------------------------------------------------ *
 * Part 8: search likelihoods (QuickCBC.c:5672-6408)
 * ------------------------------------------------------------------ */
/* log_likelihood_APmax for B points (params B x (7 + 3 (n_det - 1)),
 * reference-frame, host memory, moved in place): templates, the penalized
 * likelihood per detector (8 Hz bands, those with excess power cut) with
 * amplitude and phase maximised, and a second pass for points that moved.
 * status[b] (may be NULL) 0 or QCBC_EDOM. */
int qcbc_loglike_apmax(qcbc_ctx *ctx, double *params, int B, double *logL, int *status);
/* log_likelihood_max for B points (as above): tc set to Tobs/2, the template
 * PDwave (masked at fbegin for tc = net_tmax), its matched filter over every
 * lag in [tmin, tmax] (one backward FFT per point and detector), the peaks,
 * each refined by the penalized likelihood, and the network maximum over
 * the detectors' peaks within delays (n_det x n_det light travel times). */
int qcbc_loglike_max(qcbc_ctx *ctx, double *params, int B, double tmin, double tmax, double net_tmax,
                     const double *delays, double *logL, int *status);

/* ------------------------------------------------------------------ *
 * Part 9: the input and output text formats (QuickCBC.c:195-248,
 * :267-306), host code
 * ------------------------------------------------------------------ */
/* The file name QuickCBC builds, "<kind>_<(int)Tobs>_<(int)ttrig>_<label>.dat"
 * (kind "framed" or "spec", QuickCBC.c:200, :225, :244).  QCBC_ERANGE when it
 * does not fit len bytes. */
int qcbc_data_file_name(char *buf, size_t len, const char *kind, double Tobs, double ttrig, int label);
/* A framed_* file (QuickCBC.c:200-230): whitespace-separated pairs (time,
 * strain).  *N is the reference's count: the pairs read before the end of
 * the file is seen, so a file ending right after its last number counts one
 * pair fewer.  time and x (each may be NULL) receive the first *N pairs;
 * QCBC_ERANGE when *N > cap (with *N set), QCBC_EINVAL on a token that is not
 * a number (the reference's loop would not end), QCBC_EIO when the file
 * cannot be read. */
int qcbc_read_framed(const char *path, int cap, double *time, double *x, int *N);
/* A spec_* file (QuickCBC.c:242-248): n2 rows of four numbers, the second
 * and third the full and smooth PSDs, into SN and SM (each may be NULL).
 * QCBC_EINVAL when the file holds fewer numbers. */
int qcbc_read_spec(const char *path, int n2, double *SN, double *SM);
/* A table as main writes dataw.dat, pspec_*.dat and wcheck_*.dat
 * (QuickCBC.c:280-306): nrows rows of ncols columns (cols[c][r]), each value
 * printed "%.*e " with digits digits (6 for dataw.dat's "%e", 15 for the
 * spectra's "%.15e"), then a newline.  QCBC_EIO on a write error. */
int qcbc_write_table(const char *path, int nrows, int ncols, const double *const *cols, int digits);

/* het_diff and efix (het_diff distance) on a heterodyne, as above. */
int qcbc_het_diff(qcbc_het *het, const double *params, int B, const double *px, const int *of, int K, double *z);
int qcbc_het_efix(qcbc_het *het, const double *params, int B, double *ej, const double *ev, const double *pmin,
                  const double *pmax, double etamin, double zs, int *trials);

#ifdef __cplusplus
}
#endif
#endif
