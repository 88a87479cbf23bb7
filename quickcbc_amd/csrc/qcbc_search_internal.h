// Launch wrappers of the search-likelihood kernels (qcbc_search.hip).
// Host-safe: no device code in here.
#pragma once

#include <hip/hip_runtime_api.h>

#include "qcbc_types.h"

namespace qcbc {

// One penalized likelihood: point b, detector d; mode 0 the template row of
// the templates buffer, mode 1 the point's PDwave row scaled by nm and
// shifted by psh - 2 pi f tsh (a peak of log_likelihood_max)
struct PenJob {
  int b, d, mode, pad;
  double nm, psh, tsh;
};

// fbegin of each point (ips parameters per point) with tc = tc_at (NaN: its own)
hipError_t launch_fbegin_at(const double* params, int ips, int B, double tc_at, double* fs, hipStream_t s);
// PDwave of B set-up points into h (B x N); fs the mask frequency, ts Tobs - tc per point
hipError_t launch_pdwave(const Point* pts, const double* fs, int B, int N, double Tobs, const double* ts, double* h,
                         hipStream_t s);
int xspec_blocks(int N);
// cross spectra z (B x nd rows of N complex doubles) and (h|h) per row;
// hh_part: B nd xspec_blocks(N) scratch
hipError_t launch_xspec(const double* h, const double* D, const double* SN, int B, int nd, int N, int imin, int imax,
                        void* z, double* hh_part, double* hh, hipStream_t s);
// peaks per row of the backward transforms: mask rows x (mx - mn) scratch,
// npk rows, pk rows x maxp x 4 (logL, lag time, phase, scale)
hipError_t launch_peaks(const void* z, const double* hh, int rows, int N, double dt, int mn, int mx, double lpmin,
                        double dtp, int maxp, unsigned char* mask, int* npk, double* pk, hipStream_t s);
int penalized_max_bands();
// nq penalized likelihoods: out nq x 3 (logL, A, P); imin per point
hipError_t launch_penalized(const PenJob* jobs, int nq, const double* hw, const double* h, const double* D,
                            const double* SN, const int* imin, int nd, int N, double Tobs, int imax, int bw,
                            double* out, hipStream_t s);

}  // namespace qcbc
